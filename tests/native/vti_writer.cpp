// Writes the .vti pair of a small fake lattice with lgca_hip::IoVti, for the
// CPU test that compares it byte for byte with the Python writer.
#include "io_vti.h"

#include <cstdio>
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

int main(int argc, char** argv)
{
    if (argc < 2) return 2;
    FakeLattice L;
    lgca_hip::IoVti<FakeLattice> io(&L, "Mean density");
    return io.write(7, std::string(argv[1])) ? 0 : 1;
}
