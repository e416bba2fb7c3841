// hip_lattice.h -- C++ drop-in for OMP_Lattice<Model> over the C ABI of lgca.h.
//
// Header-only; compile with any C++11 compiler and link liblgca_hip.so.  The
// class keeps the reference's interface (src/lattice.h:32-238,
// src/omp_lattice.h:27-78): the same constructor arguments, the same set-up,
// compute, post-processing and getter methods with the same meaning, the same
// host arrays at stable addresses (cell_density(), cell_momentum(),
// mean_density(), mean_momentum() stay valid for the object's lifetime, as
// IoVti expects, src/lgca_io_vti.cpp:57-81), and the reference's error
// behaviour (print and abort).
//
// Differences, all deliberate:
//   * random choices come from the counter hash of (seed, step, cell) instead
//     of rand(); collide_and_propagate(p) ignores p, as OMP_Lattice does;
//   * cell types are kept on the host, as m_cell_type_cpu, and uploaded before
//     the next device operation that needs them;
//   * extras: explicit-size constructor, seed / step control, the fused
//     per-cell body force, uint8 state access, the device handle.
#ifndef LGCA_HIP_LATTICE_H_
#define LGCA_HIP_LATTICE_H_

#include "lgca.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

namespace lgca_hip {

typedef float Real;  // src/lgca_common.h:46

enum class Model { HPP, FHP_I, FHP_II, FHP_III };                        // :49-54
enum class CellType { FLUID = 0, SOLID_NO_SLIP = 1, SOLID_SLIP = 2 };   // :56-60

template <Model M> struct ModelInfo;
template <> struct ModelInfo<Model::HPP>     { static constexpr unsigned NUM_DIR = 4; static constexpr int ID = LGCA_HPP;     static constexpr int INV0 = 2; };
template <> struct ModelInfo<Model::FHP_I>   { static constexpr unsigned NUM_DIR = 6; static constexpr int ID = LGCA_FHP_I;   static constexpr int INV0 = 3; };
template <> struct ModelInfo<Model::FHP_II>  { static constexpr unsigned NUM_DIR = 7; static constexpr int ID = LGCA_FHP_II;  static constexpr int INV0 = 3; };
template <> struct ModelInfo<Model::FHP_III> { static constexpr unsigned NUM_DIR = 7; static constexpr int ID = LGCA_FHP_III; static constexpr int INV0 = 3; };

template <Model model_>
class HIP_Lattice {
public:
    static constexpr unsigned SPATIAL_DIM = 2;
    static constexpr unsigned NUM_DIR = ModelInfo<model_>::NUM_DIR;

    // Lattice(test_case, Re, Ma, r): sizes and physics as src/lattice.cpp:29-157.
    HIP_Lattice(const std::string test_case, const Real Re, const Real Ma_s,
                const int coarse_graining_radius, int device = 0)
        : m_test_case(test_case), m_Re(Re), m_Ma_s(Ma_s)
    {
        if (!(Re > 1.0e-06) || !(Ma_s > 1.0e-06)) die("Lattice: Re and Ma must be positive");
        derive_physics();
        unsigned dim_y = 0;
        if (test_case == "pipe") {
            dim_y = (unsigned)(int)((Re * m_nu_s) / m_u + 0.5);
        } else if (test_case == "karman") {
            const Real diameter = (Re * m_nu_s) / m_u;
            dim_y = (unsigned)(int)(3.0 * diameter + 0.5);
        } else if (test_case == "collision") {
            dim_y = 8;
        } else if (test_case == "diffusion" || test_case == "periodic" || test_case == "box") {
            dim_y = (unsigned)(int)Re;
        } else {
            std::printf("ERROR in Lattice::Lattice(): Invalid test case %s.\n", test_case.c_str());
            std::abort();
        }
        const unsigned r2 = 2u * (unsigned)coarse_graining_radius;
        dim_y += r2 - (dim_y % r2);  // src/lattice.cpp:100, applied even to multiples
        unsigned dim_x = (test_case == "pipe" || test_case == "karman" || test_case == "collision")
                             ? 2 * dim_y : dim_y;
        if (test_case == "collision") ++dim_x;
        m_bf_dir = (test_case == "pipe" || test_case == "karman") ? 'x' : 0;
        create(dim_x, dim_y, coarse_graining_radius, device);
    }

    // Explicit size; physics parameters default to zero (no Re / Ma).
    HIP_Lattice(unsigned dim_x, unsigned dim_y, int coarse_graining_radius, int device = 0,
                char bf_dir = 0)
        : m_test_case("explicit"), m_Re(0), m_Ma_s(0)
    {
        derive_physics();
        m_bf_dir = bf_dir;
        create(dim_x, dim_y, coarse_graining_radius, device);
    }

    virtual ~HIP_Lattice()
    {
        lgca_destroy(m_h);
        std::free(m_cell_type_cpu);
        std::free(m_cell_density_cpu);
        std::free(m_mean_density_cpu);
        std::free(m_cell_momentum_cpu);
        std::free(m_mean_momentum_cpu);
    }

    HIP_Lattice(const HIP_Lattice&) = delete;
    HIP_Lattice& operator=(const HIP_Lattice&) = delete;

    // -------------------------------------------------- boundary conditions
    void apply_cell_type_all(const CellType t)
    {
        for (size_t c = 0; c < m_num_cells; ++c) m_cell_type_cpu[c] = (uint8_t)t;
        m_types_dirty = true;
    }
    void apply_boundary_cell_type_east(const CellType t)
    {
        for (size_t c = m_dim_x - 1; c < m_num_cells; c += m_dim_x) m_cell_type_cpu[c] = (uint8_t)t;
        m_types_dirty = true;
    }
    void apply_boundary_cell_type_north(const CellType t)
    {
        for (size_t c = m_num_cells - m_dim_x; c < m_num_cells; ++c) m_cell_type_cpu[c] = (uint8_t)t;
        m_types_dirty = true;
    }
    void apply_boundary_cell_type_west(const CellType t)
    {
        for (size_t c = 0; c < m_num_cells; c += m_dim_x) m_cell_type_cpu[c] = (uint8_t)t;
        m_types_dirty = true;
    }
    void apply_boundary_cell_type_south(const CellType t)
    {
        for (size_t c = 0; c < m_dim_x; ++c) m_cell_type_cpu[c] = (uint8_t)t;
        m_types_dirty = true;
    }
    void apply_bc_periodic() { apply_cell_type_all(CellType::FLUID); }
    void apply_bc_pipe()
    {
        apply_cell_type_all(CellType::FLUID);
        apply_boundary_cell_type_north(CellType::SOLID_NO_SLIP);
        apply_boundary_cell_type_south(CellType::SOLID_NO_SLIP);
    }
    void apply_bc_reflecting(const std::string bounce_type)
    {
        apply_cell_type_all(CellType::FLUID);
        CellType t;
        if (bounce_type == "back") t = CellType::SOLID_NO_SLIP;
        else if (bounce_type == "forward") t = CellType::SOLID_SLIP;
        else {
            std::printf("ERROR in apply_bc_reflecting(): Invalid bounce type %s.\n", bounce_type.c_str());
            std::abort();
        }
        apply_boundary_cell_type_east(t);
        apply_boundary_cell_type_north(t);
        apply_boundary_cell_type_west(t);
        apply_boundary_cell_type_south(t);
    }
    // Pipe walls plus a NO_SLIP disc; the reference's karman set-up is the
    // disc of diameter dim_y/3 at (dim_x/6, dim_y/2) (src/lattice.cpp:252-280).
    void apply_bc_cylinder(int center_x, int center_y, Real diameter)
    {
        apply_bc_pipe();
        for (size_t c = 0; c < m_num_cells; ++c)
            if (in_disc(c, center_x, center_y, diameter)) m_cell_type_cpu[c] = (uint8_t)CellType::SOLID_NO_SLIP;
        m_types_dirty = true;
    }
    void apply_bc_karman_vortex_street()
    {
        apply_bc_cylinder((int)(m_dim_x / 6), (int)(m_dim_y / 2 + 1 / 10 * m_dim_y), (Real)(m_dim_y / 3));
    }

    // ------------------------------------------------------ initialisation
    void init_zero() {}
    void init_random()
    {
        sync_types();
        check(lgca_init_random(dev(), 1.0 / NUM_DIR));
    }
    void init_single(const std::vector<size_t> occupied_nodes)
    {
        std::vector<uint8_t> s(m_num_cells);
        check(lgca_get_state_u8(dev(), s.data()));
        for (size_t n : occupied_nodes) s[n / 8] |= (uint8_t)(1u << (n % 8));
        check(lgca_set_state_u8(dev(), s.data()));
    }
    void init_single_collision()
    {
        std::vector<size_t> nodes;
        nodes.push_back((m_dim_x * m_dim_y / 2 + 1) * 8);
        nodes.push_back((m_dim_x * m_dim_y / 2 + 5) * 8 + ModelInfo<model_>::INV0);
        init_single(nodes);
    }
    void init_diffusion()
    {
        sync_types();
        std::vector<uint8_t> region(m_num_cells);
        for (size_t c = 0; c < m_num_cells; ++c)
            region[c] = in_disc(c, (int)(m_dim_x / 2), (int)(m_dim_y / 2), (Real)(m_dim_y / 4)) ? 1 : 0;
        check(lgca_init_random_masked(dev(), 1.0 / NUM_DIR, region.data()));
    }

    unsigned long get_n_particles()
    {
        uint64_t n = 0;
        check(lgca_count_particles(dev(), &n));
        m_num_particles = n;
        return (unsigned long)n;
    }

    void print() {}
    void print_info()
    {
        std::printf("Parameter for test case \"%s\":\n\n", m_test_case.c_str());
        std::printf("Number of cells in x direction: %u\nNumber of cells in y direction: %u\n\n", m_dim_x, m_dim_y);
        std::printf("Number of coarse cells in x direction: %u\nNumber of coarse cells in y direction: %u\n\n",
                    m_coarse_dim_x, m_coarse_dim_y);
    }

    // src/lattice.cpp:445-461
    size_t get_initial_forcing() { return (size_t)(0.01 * m_num_cells); }
    size_t get_equilibrium_forcing()
    {
        const Real forcing = (8.0 * m_nu_s * m_Ma_s * m_c_s) / std::pow((Real)m_dim_y, 2.0);
        return (size_t)std::ceil(0.5 * m_num_cells * forcing);
    }

    // --------------------------------------------------------- compute path
    void setup_parallel() { std::printf("HIP configuration: one stream on GPU %d.\n\n", m_device); }

    // One step, as OMP_Lattice::collide_and_propagate.  Steps are counted and
    // enqueued together, as one lgca_step of up to kStepBatch steps, when
    // anything else touches the lattice: the tiled kernel then advances several
    // steps per pass over HBM instead of one (the viewers call this once per
    // step).  Results are those of single steps.
    void collide_and_propagate(const bool p = false)
    {
        (void)p;
        sync_types();
        if (++m_pending == kStepBatch) flush();
    }

    // Count-mode body force in direction m_bf_dir (src/omp_lattice.cpp:254-346).
    void apply_body_force(const int forcing)
    {
        if (m_bf_dir != 'x' && m_bf_dir != 'y') return;
        sync_types();
        int64_t reverted = 0;
        check(lgca_apply_body_force_count(dev(), m_bf_dir == 'x' ? LGCA_FORCE_X : LGCA_FORCE_Y,
                                          forcing, &reverted));
    }

    std::vector<Real> get_mean_velocity()
    {
        float v[2] = {0, 0};
        check(lgca_mean_velocity(dev(), v));
        return std::vector<Real>{v[0], v[1]};
    }

    void post_process()
    {
        check(lgca_post_process(dev(), m_cell_density_cpu, m_cell_momentum_cpu, m_mean_density_cpu,
                                m_mean_momentum_cpu));
    }

    void copy_data_to_device() { sync_types(); }
    void copy_data_from_device() { check(lgca_synchronize(dev())); }
    void copy_data_to_output_buffer() { check(lgca_snapshot(dev())); }

    // ---------------------------------------------------------------- extras
    void set_seed(uint64_t seed) { check(lgca_set_seed(dev(), seed)); }
    void step(uint64_t nsteps)
    {
        sync_types();
        check(lgca_step(dev(), nsteps));
    }
    void set_forcing(char dir, double p, uint32_t x0, uint32_t x1, uint32_t y0, uint32_t y1)
    {
        check(lgca_set_forcing(dev(), dir == 'x' ? LGCA_FORCE_X : (dir == 'y' ? LGCA_FORCE_Y : LGCA_FORCE_NONE),
                               p, x0, x1, y0, y1));
    }
    // Times the candidate launch shapes of lgca_step now, without changing the
    // state or the step, and keeps the fastest (lgca_tune).
    void tune()
    {
        sync_types();
        check(lgca_tune(dev()));
    }
    // Several passes per launch of lgca_step (lgca_set_pipeline).
    void set_pipeline(bool on) { check(lgca_set_pipeline(dev(), on ? 1 : 0)); }
    // The viewer's forcing controller on the GPU: a tick decides on the
    // snapshot's mean velocity, applies the count-mode force, runs `steps`
    // steps and snapshots, without the host waiting (lgca_controller_*).
    void set_forcing_controller(int64_t forcing, int64_t equilibrium)
    {
        check(lgca_controller_set(dev(), m_bf_dir == 'y' ? LGCA_FORCE_Y : LGCA_FORCE_X, m_u, forcing,
                                  equilibrium));
    }
    void controlled_tick(uint64_t steps)
    {
        sync_types();
        check(lgca_controller_tick(dev(), steps));
    }
    // Checkpoint of state, types, seed, step and forcing; resume with
    // lgca_load_checkpoint (include/lgca.h).
    void save_checkpoint(const std::string& path)
    {
        sync_types();
        check(lgca_save_checkpoint(dev(), path.c_str()));
    }
    void get_state_u8(uint8_t* out) { check(lgca_get_state_u8(dev(), out)); }
    void set_state_u8(const uint8_t* in) { check(lgca_set_state_u8(dev(), in)); }
    // the handle, with every step asked for so far enqueued
    lgca_lattice* handle() { return dev(); }

    // --------------------------------------------------------------- getters
    Real nu_s() const { return m_nu_s; }
    Real c_s() const { return m_c_s; }
    Real u() const { return m_u; }
    unsigned int dim_x() const { return m_dim_x; }
    unsigned int dim_y() const { return m_dim_y; }
    size_t num_cells() const { return m_num_cells; }
    unsigned int coarse_dim_x() const { return m_coarse_dim_x; }
    unsigned int coarse_dim_y() const { return m_coarse_dim_y; }
    size_t num_coarse_cells() const { return m_num_coarse_cells; }
    Real* cell_density() { return m_cell_density_cpu; }
    Real* mean_density() { return m_mean_density_cpu; }
    Real* cell_momentum() { return m_cell_momentum_cpu; }
    Real* mean_momentum() { return m_mean_momentum_cpu; }
    Real cell_density(const int x, const int y) { return m_cell_density_cpu[y * m_dim_x + x]; }
    Real mean_density(const int x, const int y) { return m_mean_density_cpu[y * m_dim_x + x]; }
    uint8_t* cell_types() { return m_cell_type_cpu; }

private:
    void die(const char* what)
    {
        std::printf("ERROR: %s\n", what);
        std::abort();
    }
    void check(lgca_status st)
    {
        if (st != LGCA_OK) {
            std::printf("ERROR in HIP_Lattice: %s (status %d)\n", lgca_last_error(m_h), st);
            std::abort();
        }
    }
    void derive_physics()
    {
        // src/lattice.cpp:52-67, with its float / double promotions
        m_d = m_rho / NUM_DIR;
        m_nu = 1.0 / 12.0 * 1.0 / (m_d * std::pow((1.0 - m_d), 3.0)) - 1.0 / 8.0;
        m_g = SPATIAL_DIM / (SPATIAL_DIM + 2.0) * (1.0 - 2.0 * m_d) / (1.0 - m_d);
        m_nu_s = m_nu / m_g;
        m_c_s = m_c / std::sqrt(SPATIAL_DIM);
        m_u = m_Ma_s * m_c_s;
    }
    void create(unsigned dim_x, unsigned dim_y, int r, int device)
    {
        m_device = device;
        if (lgca_create(&m_h, ModelInfo<model_>::ID, dim_x, dim_y, (uint32_t)r, device) != LGCA_OK) {
            std::printf("ERROR in HIP_Lattice: %s\n", lgca_last_error(nullptr));
            std::abort();
        }
        lgca_info info;
        check(lgca_get_info(m_h, &info));
        m_dim_x = info.dim_x;
        m_dim_y = info.dim_y;
        m_num_cells = info.num_cells;
        m_coarse_dim_x = info.coarse_dim_x;
        m_coarse_dim_y = info.coarse_dim_y;
        m_num_coarse_cells = info.num_coarse_cells;
        m_cell_type_cpu = (uint8_t*)std::calloc(m_num_cells, 1);
        m_cell_density_cpu = (Real*)std::calloc(m_num_cells, sizeof(Real));
        m_mean_density_cpu = (Real*)std::calloc(m_num_coarse_cells, sizeof(Real));
        m_cell_momentum_cpu = (Real*)std::calloc(SPATIAL_DIM * m_num_cells, sizeof(Real));
        m_mean_momentum_cpu = (Real*)std::calloc(SPATIAL_DIM * m_num_coarse_cells, sizeof(Real));
        if (!m_cell_type_cpu || !m_cell_density_cpu || !m_mean_density_cpu || !m_cell_momentum_cpu ||
            !m_mean_momentum_cpu)
            die("out of host memory");
    }
    void flush()
    {
        if (m_pending) {
            const uint64_t n = m_pending;
            m_pending = 0;
            check(lgca_step(m_h, n));
        }
    }
    // The handle for a device call: steps still counted are enqueued first.
    lgca_lattice* dev()
    {
        flush();
        return m_h;
    }
    void sync_types()
    {
        if (m_types_dirty) {
            check(lgca_set_cell_types(dev(), m_cell_type_cpu));
            m_types_dirty = false;
        }
    }
    // (float)sqrt(dx^2 + dy^2) < diameter / 2, as src/lattice.cpp:272-274
    bool in_disc(size_t cell, int cx, int cy, Real diameter) const
    {
        const int px = (int)(cell % m_dim_x), py = (int)(cell / m_dim_x);
        const Real dist = std::sqrt(std::pow((double)(px - cx), 2.0) + std::pow((double)(py - cy), 2.0));
        return dist < (diameter / 2.0);
    }

    lgca_lattice* m_h = nullptr;
    static constexpr uint64_t kStepBatch = 64;
    uint64_t m_pending = 0;  // collide_and_propagate steps not yet enqueued
    int m_device = 0;
    std::string m_test_case;
    const Real m_rho = 1.0, m_c = 1.0;
    Real m_Re, m_Ma_s, m_d = 0, m_nu = 0, m_g = 0, m_nu_s = 0, m_c_s = 0, m_u = 0;
    char m_bf_dir = 0;
    unsigned m_dim_x = 0, m_dim_y = 0, m_coarse_dim_x = 0, m_coarse_dim_y = 0;
    size_t m_num_cells = 0, m_num_coarse_cells = 0, m_num_particles = 0;
    uint8_t* m_cell_type_cpu = nullptr;
    bool m_types_dirty = false;
    Real* m_cell_density_cpu = nullptr;
    Real* m_mean_density_cpu = nullptr;
    Real* m_cell_momentum_cpu = nullptr;
    Real* m_mean_momentum_cpu = nullptr;
};

}  // namespace lgca_hip

#endif  // LGCA_HIP_LATTICE_H_
