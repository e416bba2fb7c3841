// Writes the .vti pair of a small fake lattice with lgca_hip::IoVti, for the
// CPU tests of the writer:
//   vti_writer <dir> raw     uncompressed (compared byte for byte with Python)
//   vti_writer <dir> lz4     LZ4-compressed, as the reference's writer
//   vti_writer <dir> blocks  LZ4 blocks of assorted buffers, for the decoder
#include "io_vti.h"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

struct FakeLattice {
    unsigned dx = 6, dy = 4, cdx = 3, cdy = 2;
    std::vector<float> cd, cm, md, mm;
    FakeLattice() : cd(24), cm(48), md(6), mm(12)
    {
        for (size_t i = 0; i < cd.size(); ++i) cd[i] = 0.5f * (float)i;
        for (size_t i = 0; i < cm.size(); ++i) cm[i] = 0.25f * (float)i - 3.0f;
        for (size_t i = 0; i < md.size(); ++i) md[i] = 1.0f / (float)(i + 1);
        for (size_t i = 0; i < mm.size(); ++i) mm[i] = -0.125f * (float)i;
    }
    unsigned dim_x() const { return dx; }
    unsigned dim_y() const { return dy; }
    unsigned coarse_dim_x() const { return cdx; }
    unsigned coarse_dim_y() const { return cdy; }
    size_t num_cells() const { return 24; }
    size_t num_coarse_cells() const { return 6; }
    float* cell_density() { return cd.data(); }
    float* cell_momentum() { return cm.data(); }
    float* mean_density() { return md.data(); }
    float* mean_momentum() { return mm.data(); }
};

static int blocks(const std::string& dir)
{
    // assorted inputs: every size 0..40, random bytes, runs, periodic data,
    // float fields like the lattice's, and full 32 KiB blocks
    std::vector<std::vector<uint8_t>> in;
    uint32_t r = 12345;
    auto rnd = [&]() { r = r * 1664525u + 1013904223u; return (uint8_t)(r >> 24); };
    for (size_t n = 0; n <= 40; ++n) {
        std::vector<uint8_t> v(n);
        for (auto& b : v) b = (uint8_t)(rnd() & 3);
        in.push_back(v);
    }
    std::vector<uint8_t> v(32768);
    for (auto& b : v) b = rnd();
    in.push_back(v);
    std::fill(v.begin(), v.end(), 0);
    in.push_back(v);
    for (size_t i = 0; i < v.size(); ++i) v[i] = (uint8_t)(i % 7);
    in.push_back(v);
    for (size_t i = 0; i < v.size(); ++i) v[i] = (uint8_t)((i / 300) % 2 ? rnd() : 9);
    in.push_back(v);
    std::vector<float> f(8192);
    for (size_t i = 0; i < f.size(); ++i) f[i] = (float)(rnd() % 8);
    std::memcpy(v.data(), f.data(), 32768);
    in.push_back(v);
    FILE* fp = std::fopen((dir + "/blocks.bin").c_str(), "wb");
    if (!fp) return 1;
    for (const auto& b : in) {
        std::vector<uint8_t> out(b.size() + b.size() / 255 + 16);
        const uint64_t n = b.size(), c = lgca_hip::lz4_compress_block(b.data(), b.size(), out.data());
        std::fwrite(&n, 8, 1, fp);
        std::fwrite(&c, 8, 1, fp);
        if (n) std::fwrite(b.data(), 1, n, fp);
        if (c) std::fwrite(out.data(), 1, c, fp);
    }
    return std::fclose(fp) == 0 ? 0 : 1;
}

int main(int argc, char** argv)
{
    if (argc < 3) return 2;
    const std::string dir = argv[1], mode = argv[2];
    if (mode == "blocks") return blocks(dir);
    FakeLattice L;
    lgca_hip::IoVti<FakeLattice> io(&L, "Mean density", mode == "lz4");
    return io.write(7, dir + "/") ? 0 : 1;
}
