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
// vtkImageData objects (src/lgca_io_vti.cpp:50-87).  The reference writes
// them with VTK's LZ4 compressor; here the arrays are raw appended binary
// (UInt64 byte-count headers, little-endian), which VTK and ParaView read as
// they read the reference's files.  The active scalars attribute is written
// for the section that holds an array of that name, as SetActiveScalars does.
//
// Host-only C++11; LatticeT needs the reference getters dim_x(), dim_y(),
// coarse_dim_x(), coarse_dim_y(), num_cells(), num_coarse_cells(),
// cell_density(), cell_momentum(), mean_density(), mean_momentum().
#ifndef LGCA_HIP_IO_VTI_H_
#define LGCA_HIP_IO_VTI_H_

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

namespace lgca_hip {

struct VtiArray {
    std::string name;
    const float* data;
    uint64_t tuples;
    int components;
};

// One ImageData file of points_x x points_y x 1 points whose arrays belong to
// the cells (cell_data) or the points.  Returns false on an I/O error.
inline bool write_vti(const std::string& path, unsigned points_x, unsigned points_y, bool cell_data,
                      const std::vector<VtiArray>& arrays, const std::string& scalars)
{
    const uint16_t probe = 1;
    if (*(const uint8_t*)&probe != 1) return false;  // the byte_order below says LittleEndian
    std::ostringstream x;
    const std::string extent = "0 " + std::to_string(points_x - 1) + " 0 " + std::to_string(points_y - 1) + " 0 0";
    const char* section = cell_data ? "CellData" : "PointData";
    bool has_scalars = false;
    for (const VtiArray& a : arrays) has_scalars |= a.name == scalars;
    x << "<?xml version=\"1.0\"?>\n"
      << "<VTKFile type=\"ImageData\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n"
      << "  <ImageData WholeExtent=\"" << extent << "\" Origin=\"0 0 0\" Spacing=\"1 1 1\">\n"
      << "    <Piece Extent=\"" << extent << "\">\n"
      << "      <" << section;
    if (has_scalars) x << " Scalars=\"" << scalars << "\"";
    x << ">\n";
    uint64_t offset = 0;
    for (const VtiArray& a : arrays) {
        x << "        <DataArray type=\"Float32\" Name=\"" << a.name << "\" NumberOfComponents=\""
          << a.components << "\" format=\"appended\" offset=\"" << offset << "\"/>\n";
        offset += 8 + a.tuples * (uint64_t)a.components * 4;
    }
    x << "      </" << section << ">\n"
      << "    </Piece>\n"
      << "  </ImageData>\n"
      << "  <AppendedData encoding=\"raw\">\n   _";
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const std::string head = x.str();
    bool ok = std::fwrite(head.data(), 1, head.size(), f) == head.size();
    for (const VtiArray& a : arrays) {
        const uint64_t n = a.tuples * (uint64_t)a.components * 4;
        ok = ok && std::fwrite(&n, 8, 1, f) == 1;
        ok = ok && (n == 0 || std::fwrite(a.data, 1, (size_t)n, f) == (size_t)n);
    }
    const char tail[] = "\n  </AppendedData>\n</VTKFile>\n";
    ok = ok && std::fwrite(tail, 1, sizeof tail - 1, f) == sizeof tail - 1;
    return std::fclose(f) == 0 && ok;
}

template <class LatticeT>
class IoVti {
public:
    IoVti(LatticeT* lattice, const std::string scalars) : m_lattice(lattice), m_scalars(scalars) {}

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
                                 m_scalars);
        const bool b = write_vti(dir + "mean_res_" + s + ".vti", L.coarse_dim_x(), L.coarse_dim_y(), false,
                                 mean, m_scalars);
        return a && b;
    }

private:
    LatticeT* m_lattice;
    std::string m_scalars;
};

}  // namespace lgca_hip

#endif  // LGCA_HIP_IO_VTI_H_
