// viewer_loop.cpp -- replays the pipe viewer's use of the lattice
// (apps/pipe/pipe_viewer.cpp:35-80 set-up, :100-184 ticks, :186-201 check)
// through the C++ drop-in HIP_Lattice, and writes what happened for the
// parity test (tests/test_native_dropin.py) to compare with the oracle: the
// tick log, the final state and coarse fields, the IoVti pair of .vti files
// of the last tick (as the viewer's record button writes them, :162-167) and a
// checkpoint.
//
//   viewer_loop <out_dir> <seed> <ticks>
#include "hip_lattice.h"
#include "io_vti.h"

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

using namespace lgca_hip;

int main(int argc, char** argv)
{
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s <out_dir> <seed> <ticks>\n", argv[0]);
        return 2;
    }
    const std::string out = argv[1];
    const uint64_t seed = std::strtoull(argv[2], nullptr, 10);
    const int ticks = std::atoi(argv[3]);
    const unsigned PP_INTERVAL = 5;

    HIP_Lattice<Model::FHP_III> lat("pipe", 80.0f, 0.3f, 10);
    lat.set_seed(seed);
    lat.apply_bc_pipe();
    lat.init_random();
    const unsigned long n0 = lat.get_n_particles();
    lat.copy_data_to_output_buffer();
    lat.post_process();
    int forcing = (int)lat.get_initial_forcing();
    lat.setup_parallel();

    FILE* log = std::fopen((out + "/ticks.txt").c_str(), "w");
    std::fprintf(log, "%u %u %lu %.9g %d\n", lat.dim_x(), lat.dim_y(), n0, (double)lat.u(), forcing);
    for (int t = 0; t < ticks; ++t) {
        const std::vector<Real> v = lat.get_mean_velocity();
        int applied = -1;
        if (v[0] < lat.u()) {
            if (v[0] > 0.9 * lat.u()) forcing = (int)lat.get_equilibrium_forcing();
            lat.apply_body_force(forcing);
            applied = forcing;
        }
        for (unsigned s = 0; s < PP_INTERVAL; ++s) lat.collide_and_propagate();
        lat.copy_data_to_output_buffer();
        lat.post_process();
        std::fprintf(log, "%d %.9g %.9g %d\n", t, (double)v[0], (double)v[1], applied);
    }
    const unsigned long n1 = lat.get_n_particles();
    std::fclose(log);
    if (n1 == n0) std::fprintf(stderr, "Error check PASSED: There is no difference in the number of particles.\n");
    else std::fprintf(stderr, "Error check FAILED: difference of %ld particles.\n", (long)n1 - (long)n0);

    std::vector<uint8_t> state(lat.num_cells());
    lat.get_state_u8(state.data());
    FILE* f = std::fopen((out + "/state.bin").c_str(), "wb");
    std::fwrite(state.data(), 1, state.size(), f);
    std::fclose(f);
    f = std::fopen((out + "/mean.bin").c_str(), "wb");
    std::fwrite(lat.mean_density(), sizeof(Real), lat.num_coarse_cells(), f);
    std::fwrite(lat.mean_momentum(), sizeof(Real), 2 * lat.num_coarse_cells(), f);
    std::fclose(f);
    IoVti<HIP_Lattice<Model::FHP_III>> vti(&lat, "Mean density");
    vti.update();
    if (!vti.write((size_t)ticks * PP_INTERVAL, out + "/")) {
        std::fprintf(stderr, "could not write the .vti files\n");
        return 1;
    }
    lat.save_checkpoint(out + "/checkpoint.bin");
    return n1 == n0 ? 0 : 1;
}
