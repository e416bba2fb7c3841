/*
 * oracle/cpu_ref.h -- CPU restatement of the keva92/lgca OpenMP lattice-gas path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine:
 * only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py load
 * it.  The product library (lgca_amd/liblgca_hip.so) never links or calls it.
 *
 * State layout (the reference's Bitset, src/lgca_bitset.h:229-231, indexed
 * dir + cell*8 at src/omp_lattice.cpp:179,237): one byte per cell, bit d set
 * when a particle moves in lattice direction d; cell = y*dim_x + x
 * (src/omp_lattice.cpp:123).  Cell types use the values of CellType
 * (src/lgca_common.h:56-60): 0 FLUID, 1 SOLID_NO_SLIP, 2 SOLID_SLIP.
 *
 * Random choices (collision chirality, per-cell forcing, initial fill) come
 * from Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers:
 * as easy as 1, 2, 3", SC'11) keyed by the 64-bit lattice seed, with key =
 * {seed & 0xffffffff, seed >> 32} and step taken mod 2^32:
 *   chirality word (y, w)   = philox(ctr = {w, y, step>>2, 0})[step & 3]
 *                             bit (x & 31) of word x>>5 is cell x's chirality
 *                             (one call serves a word for four steps)
 *   initial fill (x, y, d)  = philox(ctr = {x>>2, y, d, 2})[x & 3]; fires when
 *                             (uint64)draw < thr
 *   forcing word (y, w)     = 32 independent Bernoulli(thr / 2^32) bits drawn
 *                             by first-fire sampling (oracle_force_word):
 *                             uniform 0 is philox({w, y, step>>2, 1})[step & 3]
 *                             (one call per word and four steps); later
 *                             uniforms are the extras of the four-step group
 *                             G = step>>2, E_e = philox({w, y, G, 5 + 4(e>>2)})
 *                             [e & 3], e = 0, 1, ..., taken in order by the
 *                             group's steps in step order (a step starts after
 *                             the extras its group's earlier steps took);
 *                             with q = 2^32 - thr, Q_1 = q and
 *                             Q_{k+1} = floor(Q_k q / 2^32), the first of the
 *                             next n cells to fire is the smallest m < n with
 *                             u < 2^32 - Q_{m+1} (none when u >= 2^32 - Q_n);
 *                             bit (x & 31) forces cell x
 * where thr = threshold(p) = floor(p * 2^32), clamped to [0, 2^32].  This
 * replaces the reference's rand() calls (src/lgca_bitset.h:284-288,
 * src/utils.h:119-122, src/omp_lattice.cpp:269).
 *
 * Model ids: 0 HPP, 1 FHP-I, 2 FHP-II, 3 FHP-III (src/lgca_common.h:49-54).
 */
#ifndef LGCA_ORACLE_CPU_REF_H_
#define LGCA_ORACLE_CPU_REF_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORACLE_FORCE_NONE = 0, ORACLE_FORCE_X = 1, ORACLE_FORCE_Y = 2 };

void     oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
uint64_t oracle_threshold(double p);
uint32_t oracle_chirality_word(uint64_t seed, uint64_t step, uint32_t y, uint32_t w);
uint32_t oracle_force_word(uint64_t seed, uint64_t step, uint32_t y, uint32_t w, uint64_t thr);
/* The same word from a prepared table C[m] = 2^32 - Q_{m+1} (oracle_force_table)
 * and key {seed_lo, seed_hi}: shared by cpu_ref.c and cpu_fast.c. */
void     oracle_force_table(uint64_t thr, uint64_t C[32]);
uint32_t oracle_force_word_with(const uint32_t key[2], const uint64_t C[32], uint64_t step,
                                uint32_t y, uint32_t w);
uint32_t oracle_fill_draw(uint64_t seed, uint32_t x, uint32_t y, uint32_t d);

int      oracle_num_dir(int model);
uint8_t  oracle_collide(int model, uint8_t s, int p);
uint8_t  oracle_bounce_back(int model, uint8_t s);
uint8_t  oracle_bounce_forward_x(int model, uint8_t s);
uint8_t  oracle_bounce_forward_y(int model, uint8_t s);
void     oracle_pull_source(int model, uint32_t dim_x, uint32_t dim_y, uint32_t x, uint32_t y,
                            int d, uint32_t* sx, uint32_t* sy);
uint32_t oracle_lattice_vec_bits(int model, int d, int component);

void     oracle_init_random(int model, uint32_t dim_x, uint32_t dim_y, const uint8_t* types,
                            uint8_t* state, uint64_t seed, double p_fill);
void     oracle_step(int model, uint32_t dim_x, uint32_t dim_y, const uint8_t* types,
                     const uint8_t* in, uint8_t* out, uint64_t seed, uint64_t step,
                     int force_dir, uint64_t force_thr, uint32_t fx0, uint32_t fx1,
                     uint32_t fy0, uint32_t fy1);
void     oracle_run(int model, uint32_t dim_x, uint32_t dim_y, const uint8_t* types,
                    uint8_t* state, uint64_t seed, uint64_t step0, uint64_t nsteps,
                    int force_dir, uint64_t force_thr, uint32_t fx0, uint32_t fx1,
                    uint32_t fy0, uint32_t fy1, int nthreads);
/* oracle_run computed bit-sliced, 64 cells per word (cpu_fast.c): same
 * arguments, same result, many times faster.  Returns 0, or -1 (state
 * untouched) when its word-level collision disagrees with oracle_collide on
 * some input, which it checks on every call. */
int      oracle_fast_run(int model, uint32_t dim_x, uint32_t dim_y, const uint8_t* types,
                         uint8_t* state, uint64_t seed, uint64_t step0, uint64_t nsteps,
                         int force_dir, uint64_t force_thr, uint32_t fx0, uint32_t fx1,
                         uint32_t fy0, uint32_t fy1, int nthreads);
int64_t  oracle_body_force_count(int model, uint32_t dim_x, uint32_t dim_y, const uint8_t* types,
                                 uint8_t* state, uint64_t seed, uint64_t call_index,
                                 int force_dir, int32_t forcing);

void     oracle_cell_post_process(int model, uint32_t dim_x, uint32_t dim_y, const uint8_t* state,
                                  float* cell_density, float* cell_momentum);
void     oracle_mean_post_process(uint32_t dim_x, uint32_t dim_y, uint32_t r,
                                  const float* cell_density, const float* cell_momentum,
                                  float* mean_density, float* mean_momentum);
void     oracle_coarse_velocity(size_t n_coarse, const float* mean_density,
                                const float* mean_momentum, float* velocity);
uint64_t oracle_count_particles(size_t n_cells, const uint8_t* state);
void     oracle_mean_velocity(uint32_t dim_x, uint32_t dim_y, const uint8_t* types,
                              const float* cell_density, const float* cell_momentum,
                              double out[2]);

#ifdef __cplusplus
}
#endif

#endif /* LGCA_ORACLE_CPU_REF_H_ */
