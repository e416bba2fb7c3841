// oracle/ref_models_dump.cpp -- dumps the reference's model rules and offset
// tables as JSON, compiled against src/lgca_models.h where it lies under the
// reference checkout (include path set by oracle/Makefile; nothing is copied).
//
// TEST INFRASTRUCTURE ONLY: the output becomes the fixtures under tests/golden/
// that pin the CPU oracle (oracle/cpu_ref.c).  It is built only where the
// reference checkout exists and never ships with the engine.
//
//   ref_models_dump tables                 rules of all four models
//   ref_models_dump offsets <m> <dx> <dy>  the ten offset tables of model m
#include "lgca_models.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>

using namespace lgca;

namespace {

template <typename T, size_t N>
void print_int_array(const char* name, const T (&a)[N], bool last = false)
{
    std::printf("\"%s\": [", name);
    for (size_t i = 0; i < N; ++i) std::printf("%s%d", i ? ", " : "", (int)a[i]);
    std::printf("]%s", last ? "" : ", ");
}

template <typename T, size_t N>
void print_float_bits(const char* name, const T (&a)[N])
{
    std::printf("\"%s\": [", name);
    for (size_t i = 0; i < N; ++i) {
        float v = a[i];
        unsigned bits;
        std::memcpy(&bits, &v, sizeof bits);
        std::printf("%s%u", i ? ", " : "", bits);
    }
    std::printf("], ");
}

// Runs one of the reference's per-cell rules on every input byte; the output
// buffer starts as a copy of the input, as in src/omp_lattice.cpp:185-190.
template <typename MD, typename F>
void print_table(const char* name, F rule, bool last = false)
{
    const unsigned q = MD::NUM_DIR;
    std::printf("\"%s\": [", name);
    for (unsigned s = 0; s < (1u << q); ++s) {
        unsigned char in[8], out[8];
        for (unsigned d = 0; d < q; ++d) in[d] = out[d] = (s >> d) & 1u;
        rule(in, out);
        unsigned r = 0;
        for (unsigned d = 0; d < q; ++d) r |= (unsigned)(out[d] & 1u) << d;
        std::printf("%s%u", s ? ", " : "", r);
    }
    std::printf("]%s", last ? "" : ", ");
}

template <Model M>
void dump_model(const char* name, bool last)
{
    using MD = ModelDescriptor<M>;
    std::printf("\"%s\": {\"num_dir\": %u, ", name, MD::NUM_DIR);
    print_int_array("inv_dir", MD::INV_DIR);
    print_int_array("mir_dir_x", MD::MIR_DIR_X);
    print_int_array("mir_dir_y", MD::MIR_DIR_Y);
    print_float_bits("lattice_vec_x", MD::LATTICE_VEC_X);
    print_float_bits("lattice_vec_y", MD::LATTICE_VEC_Y);
    print_table<MD>("collide_p0", [](unsigned char* i, unsigned char* o) { MD::collide(i, o, false); });
    print_table<MD>("collide_p1", [](unsigned char* i, unsigned char* o) { MD::collide(i, o, true); });
    print_table<MD>("bounce_back", [](unsigned char* i, unsigned char* o) { MD::bounce_back(i, o); });
    print_table<MD>("bounce_forward_x", [](unsigned char* i, unsigned char* o) { MD::bounce_forward_x(i, o); });
    print_table<MD>("bounce_forward_y", [](unsigned char* i, unsigned char* o) { MD::bounce_forward_y(i, o); }, true);
    std::printf("}%s\n", last ? "" : ",");
}

template <Model M>
void dump_offsets(unsigned dx, unsigned dy)
{
    using MD = ModelDescriptor<M>;
    MD md(dx, dy);
    std::printf("{\"dim_x\": %u, \"dim_y\": %u, \"num_dir\": %u, ", dx, dy, MD::NUM_DIR);
    print_int_array("inv_dir", MD::INV_DIR);
    print_int_array("neighbor_even", md.offset_to_neighbor_even);
    print_int_array("neighbor_odd", md.offset_to_neighbor_odd);
    print_int_array("eastern_even", md.offset_to_eastern_boundary_even);
    print_int_array("eastern_odd", md.offset_to_eastern_boundary_odd);
    print_int_array("northern_even", md.offset_to_northern_boundary_even);
    print_int_array("northern_odd", md.offset_to_northern_boundary_odd);
    print_int_array("western_even", md.offset_to_western_boundary_even);
    print_int_array("western_odd", md.offset_to_western_boundary_odd);
    print_int_array("southern_even", md.offset_to_southern_boundary_even);
    print_int_array("southern_odd", md.offset_to_southern_boundary_odd, true);
    std::printf("}\n");
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc >= 2 && std::strcmp(argv[1], "tables") == 0) {
        std::printf("{\n");
        dump_model<Model::HPP>("HPP", false);
        dump_model<Model::FHP_I>("FHP_I", false);
        dump_model<Model::FHP_II>("FHP_II", false);
        dump_model<Model::FHP_III>("FHP_III", true);
        std::printf("}\n");
        return 0;
    }
    if (argc == 5 && std::strcmp(argv[1], "offsets") == 0) {
        const int m = std::atoi(argv[2]);
        const unsigned dx = (unsigned)std::strtoul(argv[3], nullptr, 10);
        const unsigned dy = (unsigned)std::strtoul(argv[4], nullptr, 10);
        switch (m) {
        case 0: dump_offsets<Model::HPP>(dx, dy); return 0;
        case 1: dump_offsets<Model::FHP_I>(dx, dy); return 0;
        case 2: dump_offsets<Model::FHP_II>(dx, dy); return 0;
        case 3: dump_offsets<Model::FHP_III>(dx, dy); return 0;
        default: break;
        }
    }
    std::fprintf(stderr, "usage: %s tables | offsets <model 0-3> <dim_x> <dim_y>\n", argv[0]);
    return 2;
}
