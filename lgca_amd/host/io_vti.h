// io_vti.h -- VTK XML image files (.vti) of a lattice's output arrays,
// written without VTK.
//
// lgca_hip::IoVti<LatticeT> keeps the interface of the reference's
// IoVti<Model> (src/lgca_io_vti.h, src/lgca_io_vti.cpp:36-144): it wraps the
// lattice's host output arrays and write(step, dir) produces
//   dir + "cell_res_<step>.vti": (dim_x + 1) x (dim_y + 1) x 1 points, CellData
//       "Cell density" (1 component) and "Cell momentum" (2, x/y interleaved);
//   dir + "mean_res_<step>.vti": coarse_dim_x x coarse_dim_y x 1 points,
//       PointData "Mean density" (1) and "Mean momentum" (2);
// the same image sizes, array names, components and order as the reference's
// vtkImageData objects (src/lgca_io_vti.cpp:50-87), compressed with LZ4 as
// the reference's writer is set up (SetCompressorTypeToLZ4, :131-139): raw
// appended binary, each array a VTK compressed-data header (UInt64 block
// count, block size, size of a partial last block or 0, compressed size of
// every block) followed by its 32 KiB blocks in the LZ4 block format, which
// this file encodes itself.  IoVti(lattice, scalars, false) writes the arrays
// uncompressed (UInt64 byte count, then the bytes).  The active scalars
// attribute is written for the section that holds an array of that name, as
// SetActiveScalars does.
//
// Host-only C++11; LatticeT needs the reference getters dim_x(), dim_y(),
// coarse_dim_x(), coarse_dim_y(), num_cells(), num_coarse_cells(),
// cell_density(), cell_momentum(), mean_density(), mean_momentum().
#ifndef LGCA_HIP_IO_VTI_H_
#define LGCA_HIP_IO_VTI_H_

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

namespace lgca_hip {

// LZ4 block format (lz4.github.io/lz4, "LZ4 Block Format Description"):
// sequences of a token (literal count << 4 | match length - 4, each 15 meaning
// "more bytes follow", 255 each), the literals, a 2-byte little-endian match
// offset and the match length's extra bytes; the last sequence is literals
// only.  Greedy matching on a hash of 4-byte prefixes; matches start at least
// 12 bytes and end at least 5 bytes before the end of the block, as the
// format requires.  dst needs n + n / 255 + 16 bytes.
inline size_t lz4_compress_block(const uint8_t* src, size_t n, uint8_t* dst)
{
    uint8_t* op = dst;
    auto put_len = [&](size_t len) {
        for (; len >= 255; len -= 255) *op++ = 255;
        *op++ = (uint8_t)len;
    };
    auto emit = [&](size_t anchor, size_t lit, size_t off, size_t mlen) {
        const size_t ml = mlen ? mlen - 4 : 0;
        *op++ = (uint8_t)((lit < 15 ? lit : 15) << 4 | (mlen ? (ml < 15 ? ml : 15) : 0));
        if (lit >= 15) put_len(lit - 15);
        std::memcpy(op, src + anchor, lit);
        op += lit;
        if (!mlen) return;
        *op++ = (uint8_t)(off & 0xFF);
        *op++ = (uint8_t)(off >> 8);
        if (ml >= 15) put_len(ml - 15);
    };
    size_t anchor = 0;
    if (n >= 13) {
        std::vector<uint32_t> table(1u << 12, 0);  // position + 1 of the last prefix with this hash
        const size_t mflimit = n - 12, matchlimit = n - 5;
        size_t ip = 0;
        while (ip < mflimit) {
            uint32_t seq;
            std::memcpy(&seq, src + ip, 4);
            const uint32_t h = (seq * 2654435761u) >> 20;
            const size_t ref = table[h];
            table[h] = (uint32_t)(ip + 1);
            uint32_t rseq = 0;
            if (ref) std::memcpy(&rseq, src + ref - 1, 4);
            if (ref && ip - (ref - 1) <= 65535 && rseq == seq) {
                const size_t r = ref - 1;
                size_t len = 4;
                while (ip + len < matchlimit && src[r + len] == src[ip + len]) ++len;
                emit(anchor, ip - anchor, ip - r, len);
                ip += len;
                anchor = ip;
            } else {
                ++ip;
            }
        }
    }
    emit(anchor, n - anchor, 0, 0);
    return (size_t)(op - dst);
}

struct VtiArray {
    std::string name;
    const float* data;
    uint64_t tuples;
    int components;
};

// The appended bytes of one array: LZ4 (VTK's compressed-data header, then
// the blocks) or raw (byte count, then the bytes).
inline std::string vti_payload(const VtiArray& a, bool lz4)
{
    const uint64_t n = a.tuples * (uint64_t)a.components * 4;
    const uint8_t* src = (const uint8_t*)a.data;
    std::string out;
    if (!lz4) {
        out.append((const char*)&n, 8);
        if (n) out.append((const char*)src, (size_t)n);
        return out;
    }
    const uint64_t block = 32768;
    const uint64_t nblocks = (n + block - 1) / block;
    const uint64_t last = n % block;
    std::vector<uint64_t> head(3 + nblocks);
    head[0] = nblocks;
    head[1] = block;
    head[2] = last;
    std::string body;
    std::vector<uint8_t> tmp(block + block / 255 + 16);
    for (uint64_t i = 0; i < nblocks; ++i) {
        const uint64_t len = std::min<uint64_t>(block, n - i * block);
        const size_t c = lz4_compress_block(src + i * block, (size_t)len, tmp.data());
        head[3 + i] = c;
        body.append((const char*)tmp.data(), c);
    }
    out.append((const char*)head.data(), head.size() * 8);
    out += body;
    return out;
}

// One ImageData file of points_x x points_y x 1 points whose arrays belong to
// the cells (cell_data) or the points.  Returns false on an I/O error.
inline bool write_vti(const std::string& path, unsigned points_x, unsigned points_y, bool cell_data,
                      const std::vector<VtiArray>& arrays, const std::string& scalars,
                      bool lz4 = false)
{
    const uint16_t probe = 1;
    if (*(const uint8_t*)&probe != 1) return false;  // the byte_order below says LittleEndian
    std::ostringstream x;
    const std::string extent = "0 " + std::to_string(points_x - 1) + " 0 " + std::to_string(points_y - 1) + " 0 0";
    const char* section = cell_data ? "CellData" : "PointData";
    bool has_scalars = false;
    for (const VtiArray& a : arrays) has_scalars |= a.name == scalars;
    x << "<?xml version=\"1.0\"?>\n"
      << "<VTKFile type=\"ImageData\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\""
      << (lz4 ? " compressor=\"vtkLZ4DataCompressor\"" : "") << ">\n"
      << "  <ImageData WholeExtent=\"" << extent << "\" Origin=\"0 0 0\" Spacing=\"1 1 1\">\n"
      << "    <Piece Extent=\"" << extent << "\">\n"
      << "      <" << section;
    if (has_scalars) x << " Scalars=\"" << scalars << "\"";
    x << ">\n";
    std::vector<std::string> blobs;
    uint64_t offset = 0;
    for (const VtiArray& a : arrays) {
        x << "        <DataArray type=\"Float32\" Name=\"" << a.name << "\" NumberOfComponents=\""
          << a.components << "\" format=\"appended\" offset=\"" << offset << "\"/>\n";
        blobs.push_back(vti_payload(a, lz4));
        offset += blobs.back().size();
    }
    x << "      </" << section << ">\n"
      << "    </Piece>\n"
      << "  </ImageData>\n"
      << "  <AppendedData encoding=\"raw\">\n   _";
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const std::string head = x.str();
    bool ok = std::fwrite(head.data(), 1, head.size(), f) == head.size();
    for (const std::string& b : blobs) ok = ok && std::fwrite(b.data(), 1, b.size(), f) == b.size();
    const char tail[] = "\n  </AppendedData>\n</VTKFile>\n";
    ok = ok && std::fwrite(tail, 1, sizeof tail - 1, f) == sizeof tail - 1;
    return std::fclose(f) == 0 && ok;
}

template <class LatticeT>
class IoVti {
public:
    IoVti(LatticeT* lattice, const std::string scalars, bool lz4 = true)
        : m_lattice(lattice), m_scalars(scalars), m_lz4(lz4)
    {
    }

    void set_scalars(const std::string scalars) { m_scalars = scalars; }

    // The arrays are read at write time, so there is nothing to refresh.
    void update() {}

    bool write(const size_t step, const std::string dir)
    {
        LatticeT& L = *m_lattice;
        const std::string s = std::to_string(step);
        std::vector<VtiArray> cell = {{"Cell density", L.cell_density(), (uint64_t)L.num_cells(), 1},
                                      {"Cell momentum", L.cell_momentum(), (uint64_t)L.num_cells(), 2}};
        std::vector<VtiArray> mean = {{"Mean density", L.mean_density(), (uint64_t)L.num_coarse_cells(), 1},
                                      {"Mean momentum", L.mean_momentum(), (uint64_t)L.num_coarse_cells(), 2}};
        const bool a = write_vti(dir + "cell_res_" + s + ".vti", L.dim_x() + 1, L.dim_y() + 1, true, cell,
                                 m_scalars, m_lz4);
        const bool b = write_vti(dir + "mean_res_" + s + ".vti", L.coarse_dim_x(), L.coarse_dim_y(), false,
                                 mean, m_scalars, m_lz4);
        return a && b;
    }

private:
    LatticeT* m_lattice;
    std::string m_scalars;
    bool m_lz4;
};

}  // namespace lgca_hip

#endif  // LGCA_HIP_IO_VTI_H_
