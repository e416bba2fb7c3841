/*
 * oracle/cpu_ref.c -- CPU restatement of the keva92/lgca OpenMP lattice-gas path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine, never part of it.
 * Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may
 * load this library.  Conventions (layout, hash contract, model ids) are stated
 * in cpu_ref.h.
 *
 * What pins each function (tests/test_oracle.py):
 *   oracle_philox4x32_10      Random123 known-answer vectors (Salmon et al. SC'11)
 *   oracle_collide            exhaustive table of the reference's own
 *                             ModelDescriptor<M>::collide, compiled from
 *                             src/lgca_models.h by oracle/Makefile, plus mass and
 *                             momentum conservation over every input
 *   oracle_bounce_*           exhaustive tables of the compiled reference
 *   oracle_pull_source        the reference's offset tables (compiled
 *                             src/lgca_models.h constructors) composed as in
 *                             src/omp_lattice.cpp:141-179, for every cell of
 *                             several small lattices; single-particle trajectories
 *   oracle_lattice_vec_bits   compiled reference LATTICE_VEC_X/Y bit patterns
 *   oracle_step / _run        composition: conservation laws, straight-line
 *                             transport with torus wrap, hand-worked wall and
 *                             slip cases
 *   oracle_mean_post_process  indicator-field probes of the averaging window
 *                             (each cell's membership in each coarse cell) and
 *                             closed-form counts
 *   oracle_body_force_count   hand-worked cases of the stopping rule, and the
 *                             stated Philox outputs of the candidate stream; the
 *                             candidate sequence itself is this build's contract
 *                             (the reference draws rand(), which nothing pins)
 *   oracle_mean_velocity,     closed forms on uniform states
 *   oracle_coarse_velocity
 * oracle_fast_run (cpu_fast.c) is pinned by equality with oracle_run.
 */
#include "cpu_ref.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- Philox4x32-10 */

void oracle_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    /* Round function and Weyl key schedule as published in SC'11, Table 2. */
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; ++round) {
        if (round > 0) { k0 += W0; k1 += W1; }
        uint64_t p0 = (uint64_t)M0 * c0;
        uint64_t p1 = (uint64_t)M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline void seed_key(uint64_t seed, uint32_t key[2])
{
    key[0] = (uint32_t)seed;
    key[1] = (uint32_t)(seed >> 32);
}

uint64_t oracle_threshold(double p)
{
    if (!(p > 0.0)) return 0;
    if (p >= 1.0) return (uint64_t)1 << 32;
    return (uint64_t)floor(p * 4294967296.0);
}

uint32_t oracle_chirality_word(uint64_t seed, uint64_t step, uint32_t y, uint32_t w)
{
    uint32_t key[2], out[4];
    const uint32_t ctr[4] = { w, y, (uint32_t)step >> 2, 0u };
    seed_key(seed, key);
    oracle_philox4x32_10(ctr, key, out);
    return out[step & 3];
}

/* Per-cell force decisions of one 32-cell word, by first-fire sampling: the
 * gap to the next forced cell is drawn from its (truncated geometric) CDF,
 * C[m] = P(first fire within m + 1 cells) * 2^32 = 2^32 - Q_{m+1}, where
 * Q_k = floor(2^32 (1 - p)^k) is built by exact integer steps so every
 * platform gets the same table. */
void oracle_force_table(uint64_t thr, uint64_t C[32])
{
    const uint64_t q = ((uint64_t)1 << 32) - thr;  /* thr >= 1, so q < 2^32 */
    uint64_t Qk = q;                               /* Q_1 */
    for (int m = 0; m < 32; ++m) {
        C[m] = ((uint64_t)1 << 32) - Qk;           /* 2^32 - Q_{m+1} */
        Qk = (Qk * q) >> 32;
    }
}

/* The decision word of one step, walking from uniform u; *used counts the
 * group's extras taken so far and ext[] holds the current block of four. */
static uint32_t force_walk(const uint32_t key[2], const uint64_t C[32], uint32_t group, uint32_t y,
                           uint32_t w, uint32_t u, unsigned* used, uint32_t ext[4])
{
    uint32_t word = 0;
    unsigned pos = 0;
    for (;;) {
        const unsigned n = 32 - pos;
        if ((uint64_t)u >= C[n - 1]) break;       /* no fire among the next n cells */
        unsigned m = 0;
        while ((uint64_t)u >= C[m]) ++m;          /* first fire, m < n */
        word |= 1u << (pos + m);
        pos += m + 1;
        if (pos == 32) break;
        if ((*used & 3) == 0) {
            const uint32_t c1[4] = { w, y, group, 5u + 4u * (*used >> 2) };
            oracle_philox4x32_10(c1, key, ext);
        }
        u = ext[*used & 3];
        ++*used;
    }
    return word;
}

/* Walks the steps of the word's four-step group up to `step`: they share the
 * group's extras in step order. */
uint32_t oracle_force_word_with(const uint32_t key[2], const uint64_t C[32], uint64_t step,
                                uint32_t y, uint32_t w)
{
    uint32_t prim[4], ext[4] = { 0, 0, 0, 0 };
    const uint32_t group = (uint32_t)step >> 2;
    const uint32_t c0[4] = { w, y, group, 1u };
    oracle_philox4x32_10(c0, key, prim);
    unsigned used = 0;
    uint32_t word = 0;
    for (unsigned i = 0; i <= (unsigned)(step & 3); ++i)
        word = force_walk(key, C, group, y, w, prim[i], &used, ext);
    return word;
}

uint32_t oracle_force_word(uint64_t seed, uint64_t step, uint32_t y, uint32_t w, uint64_t thr)
{
    if (thr == 0) return 0;
    uint32_t key[2];
    uint64_t C[32];
    seed_key(seed, key);
    oracle_force_table(thr, C);
    return oracle_force_word_with(key, C, step, y, w);
}

uint32_t oracle_fill_draw(uint64_t seed, uint32_t x, uint32_t y, uint32_t d)
{
    uint32_t key[2], out[4];
    const uint32_t ctr[4] = { x >> 2, y, d, 2u };
    seed_key(seed, key);
    oracle_philox4x32_10(ctr, key, out);
    return out[x & 3];
}

/* ------------------------------------------------------------- model rules */

int oracle_num_dir(int model)
{
    /* NUM_DIR: src/lgca_models.h:41,234,441,659 */
    switch (model) {
    case 0: return 4;
    case 1: return 6;
    case 2: case 3: return 7;
    default: return 0;
    }
}

static inline int popc8(uint8_t v) { return __builtin_popcount((unsigned)v); }

/* Fluid-cell collision.  The reference writes each model's rule as boolean
 * products over the input bits (src/lgca_models.h:134-152 HPP, 367-395 FHP-I,
 * 568-613 FHP-II, 786-856 FHP-III); here the same rules are stated by particle
 * configuration:
 *   HPP:   an isolated head-on pair along x turns into one along y and back.
 *   FHP:   an isolated head-on pair {k, k+3} rotates by +60 degrees (chirality
 *          p = 0) or -60 degrees (p = 1); the symmetric triples {0,2,4} and
 *          {1,3,5} swap; the rest bit (dir 6) does not take part in either.
 *   FHP-II adds the rest-particle rules: rest + a lone particle in direction i
 *          becomes two particles in i-1 and i+1, and the inverse.
 *   FHP-III: every FHP-III-only product in the reference vanishes identically
 *          (each needs a direction both set and clear), so FHP-III collides
 *          exactly like FHP-II. */
uint8_t oracle_collide(int model, uint8_t s, int p)
{
    if (model == 0) {
        const uint8_t mv = s & 0x0F;
        if (mv == 0x05) return (uint8_t)((s & ~0x0F) | 0x0A);
        if (mv == 0x0A) return (uint8_t)((s & ~0x0F) | 0x05);
        return s;
    }
    const uint8_t mv = s & 0x3F;
    const uint8_t rest = s & 0x40;
    const uint8_t hi = s & 0x80;   /* bit 7 is never produced; pass it through */
    const int n = popc8(mv);
    if (n == 2) {
        for (int k = 0; k < 3; ++k) {
            if (mv == (uint8_t)((1u << k) | (1u << (k + 3)))) {
                const int j = p ? (k + 2) % 3 : (k + 1) % 3;
                return (uint8_t)(hi | rest | (1u << j) | (1u << (j + 3)));
            }
        }
    }
    if (mv == 0x15) return (uint8_t)(hi | rest | 0x2A);
    if (mv == 0x2A) return (uint8_t)(hi | rest | 0x15);
    if (model >= 2) {
        if (rest && n == 1) {
            const int i = __builtin_ctz((unsigned)mv);
            return (uint8_t)(hi | (1u << ((i + 5) % 6)) | (1u << ((i + 1) % 6)));
        }
        if (!rest && n == 2) {
            for (int i = 0; i < 6; ++i) {
                if (mv == (uint8_t)((1u << ((i + 5) % 6)) | (1u << ((i + 1) % 6))))
                    return (uint8_t)(hi | 0x40 | (1u << i));
            }
        }
    }
    return s;
}

/* Direction permutations (src/lgca_models.h:44-48, 239-243, 446-450, 664-668):
 *   inverse      HPP d -> d+2 (mod 4),   FHP d -> d+3 (mod 6), rest fixed
 *   mirror in x  HPP d -> -d  (mod 4),   FHP d -> -d  (mod 6)
 *   mirror in y  HPP d -> 2-d (mod 4),   FHP d -> 3-d (mod 6) */
static int perm_dir(int model, int kind, int d)
{
    if (model == 0) {
        switch (kind) {
        case 0: return (d + 2) & 3;
        case 1: return (4 - d) & 3;
        default: return (6 - d) & 3;
        }
    }
    if (d == 6) return 6;
    switch (kind) {
    case 0: return (d + 3) % 6;
    case 1: return (6 - d) % 6;
    default: return (9 - d) % 6;
    }
}

static uint8_t permute(int model, int kind, uint8_t s)
{
    /* out[d] = in[perm(d)] as in bounce_back / bounce_forward_x/y
     * (src/lgca_models.h:184-220, 397-427, 615-645, 858-888). */
    const int q = oracle_num_dir(model);
    uint8_t out = (uint8_t)(s & ~((1u << q) - 1u));
    for (int d = 0; d < q; ++d)
        out |= (uint8_t)(((s >> perm_dir(model, kind, d)) & 1u) << d);
    return out;
}

uint8_t oracle_bounce_back(int model, uint8_t s)      { return permute(model, 0, s); }
uint8_t oracle_bounce_forward_x(int model, uint8_t s) { return permute(model, 1, s); }
uint8_t oracle_bounce_forward_y(int model, uint8_t s) { return permute(model, 2, s); }

/* Source cell of the particle that arrives at (x, y) moving in direction d.
 * The reference pulls from cell + offset_to_neighbor_{even,odd}[INV_DIR[d]]
 * plus per-edge wrap corrections (src/omp_lattice.cpp:141-179 with the tables
 * of src/lgca_models.h:79-132, 292-365, 483-566, 701-784); those sums reduce to
 * the torus neighbour below.  Rows alternate: even rows are not shifted, odd
 * rows sit half a cell to the east. */
void oracle_pull_source(int model, uint32_t dim_x, uint32_t dim_y, uint32_t x, uint32_t y,
                        int d, uint32_t* sx, uint32_t* sy)
{
    int dx = 0, dy = 0;
    if (model == 0) {
        static const int hdx[4] = { -1, 0, 1, 0 };
        static const int hdy[4] = { 0, -1, 0, 1 };
        dx = hdx[d]; dy = hdy[d];
    } else {
        static const int edx[7] = { -1, -1, 0, 1, 0, -1, 0 };
        static const int odx[7] = { -1, 0, 1, 1, 1, 0, 0 };
        static const int fdy[7] = { 0, -1, -1, 0, 1, 1, 0 };
        dx = (y & 1u) ? odx[d] : edx[d];
        dy = fdy[d];
    }
    *sx = (uint32_t)(((int64_t)x + dx + dim_x) % dim_x);
    *sy = (uint32_t)(((int64_t)y + dy + dim_y) % dim_y);
}

uint32_t oracle_lattice_vec_bits(int model, int d, int component)
{
    /* LATTICE_VEC_X/Y (src/lgca_models.h:51-52, 246-247, 453-454, 671-672);
     * SIN = (float)sin(pi/3), folded by the compiler from the double value. */
    static const float hx[4] = { 1.0f, 0.0f, -1.0f, 0.0f };
    static const float hy[4] = { 0.0f, 1.0f, 0.0f, -1.0f };
    static const float fx[7] = { 1.0f, 0.5f, -0.5f, -1.0f, -0.5f, 0.5f, 0.0f };
    const float S = (float)sin(M_PI / 3.0);
    const float fy[7] = { 0.0f, S, S, 0.0f, -S, -S, 0.0f };
    float v;
    if (model == 0) v = component ? hy[d] : hx[d];
    else v = component ? fy[d] : fx[d];
    uint32_t bits;
    memcpy(&bits, &v, sizeof bits);
    return bits;
}

static float vec_f(int model, int d, int component)
{
    uint32_t b = oracle_lattice_vec_bits(model, d, component);
    float v;
    memcpy(&v, &b, sizeof v);
    return v;
}

/* -------------------------------------------------------------- initialisers */

void oracle_init_random(int model, uint32_t dim_x, uint32_t dim_y, const uint8_t* types,
                        uint8_t* state, uint64_t seed, double p_fill)
{
    /* init_random (src/lattice.cpp:198-217): every FLUID cell, every direction
     * occupied with probability p; solid cells are left as they are. */
    const int q = oracle_num_dir(model);
    const uint64_t thr = oracle_threshold(p_fill);
    uint32_t key[2];
    seed_key(seed, key);
#pragma omp parallel for schedule(static)
    for (long long yl = 0; yl < (long long)dim_y; ++yl) {
        const uint32_t y = (uint32_t)yl;
        uint32_t draws[7][4];
        for (uint32_t x = 0; x < dim_x; ++x) {
            if ((x & 3u) == 0 || x == 0) {
                /* one Philox call yields the draws of 4 consecutive cells */
                for (int d = 0; d < q; ++d) {
                    const uint32_t ctr[4] = { x >> 2, y, (uint32_t)d, 2u };
                    oracle_philox4x32_10(ctr, key, draws[d]);
                }
            }
            const size_t cell = (size_t)y * dim_x + x;
            if (types && types[cell] != 0) continue;
            uint8_t s = state[cell];
            for (int d = 0; d < q; ++d) {
                if ((uint64_t)draws[d][x & 3u] < thr) s |= (uint8_t)(1u << d);
                else s &= (uint8_t)~(1u << d);
            }
            state[cell] = s;
        }
    }
}

/* --------------------------------------------------------------------- step */

typedef struct {
    uint8_t lut[2][256];
} collide_lut_t;

static void build_lut(int model, collide_lut_t* t)
{
    for (int p = 0; p < 2; ++p)
        for (int s = 0; s < 256; ++s) t->lut[p][s] = oracle_collide(model, (uint8_t)s, p);
}

static inline uint8_t apply_force(int model, int dir, uint8_t s)
{
    /* Flip rules of apply_body_force (src/omp_lattice.cpp:293-337). */
    if (model == 0) {
        if (dir == ORACLE_FORCE_X) {
            if (!(s & 1u) && (s & 4u)) s = (uint8_t)((s | 1u) & ~4u);
        } else {
            if ((s & 2u) && !(s & 8u)) s = (uint8_t)((s | 8u) & ~2u);
        }
    } else {
        if (dir == ORACLE_FORCE_X) {
            if (!(s & 1u) && (s & 8u)) s = (uint8_t)((s | 1u) & ~8u);
        } else {
            if ((s & 2u) && !(s & 32u)) s = (uint8_t)((s | 32u) & ~2u);
            if ((s & 4u) && !(s & 16u)) s = (uint8_t)((s | 16u) & ~4u);
        }
    }
    return s;
}

static void step_rows(int model, uint32_t dim_x, uint32_t dim_y, const uint8_t* types,
                      const uint8_t* in, uint8_t* out, uint64_t seed, uint64_t step,
                      int force_dir, uint64_t force_thr, uint32_t fx0, uint32_t fx1,
                      uint32_t fy0, uint32_t fy1, const collide_lut_t* lut,
                      uint32_t y_begin, uint32_t y_end)
{
    const int q = oracle_num_dir(model);
    uint32_t key[2];
    uint64_t fcdf[32];
    seed_key(seed, key);
    if (force_thr != 0) oracle_force_table(force_thr, fcdf);
    for (uint32_t y = y_begin; y < y_end; ++y) {
        uint32_t chir[4] = { 0, 0, 0, 0 };
        uint32_t chir_word = UINT32_MAX;
        uint32_t fword = 0;
        uint32_t fword_w = UINT32_MAX;
        const int row_forced = force_dir != ORACLE_FORCE_NONE && force_thr != 0 && y >= fy0 && y < fy1;
        const int on_ns = (y == 0) || (y == dim_y - 1);
        /* Source row and x offset of every direction for this row, taken from
         * oracle_pull_source at x = 1 (the offset is the same for every x). */
        const uint8_t* src_row[7];
        int src_dx[7];
        for (int d = 0; d < q; ++d) {
            uint32_t sx, sy;
            oracle_pull_source(model, dim_x > 2 ? dim_x : 3, dim_y, 1, y, d, &sx, &sy);
            src_row[d] = in + (size_t)sy * dim_x;
            src_dx[d] = (int)sx - 1;
        }
        for (uint32_t x = 0; x < dim_x; ++x) {
            const size_t cell = (size_t)y * dim_x + x;
            /* Pull stream (src/omp_lattice.cpp:141-180). */
            uint8_t s = 0;
            for (int d = 0; d < q; ++d) {
                long sx = (long)x + src_dx[d];
                if (sx < 0) sx += dim_x;
                else if (sx >= (long)dim_x) sx -= dim_x;
                s |= (uint8_t)(src_row[d][sx] & (1u << d));
            }
            /* Destination-cell dispatch (src/omp_lattice.cpp:193-231). */
            const uint8_t type = types ? types[cell] : 0;
            uint8_t o;
            if (type == 0) {
                int p = 0;
                if (model != 0) {
                    const uint32_t w = x >> 5;
                    if (w != chir_word) {
                        const uint32_t ctr[4] = { w, y, (uint32_t)step >> 2, 0u };
                        oracle_philox4x32_10(ctr, key, chir);
                        chir_word = w;
                    }
                    p = (int)((chir[step & 3] >> (x & 31)) & 1u);
                }
                o = lut->lut[p][s];
                if (row_forced && x >= fx0 && x < fx1) {
                    if ((x >> 5) != fword_w) {
                        fword = oracle_force_word_with(key, fcdf, step, y, x >> 5);
                        fword_w = x >> 5;
                    }
                    if ((fword >> (x & 31)) & 1u) o = apply_force(model, force_dir, o);
                }
            } else if (type == 1) {
                o = oracle_bounce_back(model, s);
            } else if (type == 2) {
                o = s;
                if (on_ns) o = oracle_bounce_forward_x(model, s);
                if (x == 0 || x == dim_x - 1) o = oracle_bounce_forward_y(model, s);
            } else {
                o = s;
            }
            out[cell] = o;
        }
    }
}

void oracle_step(int model, uint32_t dim_x, uint32_t dim_y, const uint8_t* types,
                 const uint8_t* in, uint8_t* out, uint64_t seed, uint64_t step,
                 int force_dir, uint64_t force_thr, uint32_t fx0, uint32_t fx1,
                 uint32_t fy0, uint32_t fy1)
{
    collide_lut_t lut;
    build_lut(model, &lut);
    step_rows(model, dim_x, dim_y, types, in, out, seed, step, force_dir, force_thr,
              fx0, fx1, fy0, fy1, &lut, 0, dim_y);
}

void oracle_run(int model, uint32_t dim_x, uint32_t dim_y, const uint8_t* types,
                uint8_t* state, uint64_t seed, uint64_t step0, uint64_t nsteps,
                int force_dir, uint64_t force_thr, uint32_t fx0, uint32_t fx1,
                uint32_t fy0, uint32_t fy1, int nthreads)
{
    const size_t n = (size_t)dim_x * dim_y;
    uint8_t* tmp = (uint8_t*)malloc(n);
    uint8_t* a = state;
    uint8_t* b = tmp;
    collide_lut_t lut;
    build_lut(model, &lut);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    for (uint64_t t = 0; t < nsteps; ++t) {
        /* Rows are independent within a step, as the reference's cells are
         * (tbb::parallel_for over cell blocks, src/omp_lattice.cpp:114-118). */
#pragma omp parallel for schedule(static)
        for (long long y = 0; y < (long long)dim_y; ++y)
            step_rows(model, dim_x, dim_y, types, a, b, seed, step0 + t, force_dir, force_thr,
                      fx0, fx1, fy0, fy1, &lut, (uint32_t)y, (uint32_t)y + 1);
        uint8_t* c = a; a = b; b = c;
    }
    if (a != state) memcpy(state, a, n);
    free(tmp);
}

/* Count-mode body force (src/omp_lattice.cpp:254-346): draw candidate cells
 * until `forcing` particles are reverted or 2N candidates are spent; the loop
 * runs at least once (do-while).  As in the reference, the count of reverted
 * particles is an unsigned int compared with the int amount (:263, :345), so
 * the amount is taken modulo 2^32: a negative one never stops the loop before
 * the 2N candidates.  Candidate i of call c is
 *   cell = (philox({i_lo, i_hi, c, 3})[1] << 32 | [0]) mod N
 * in place of rand() % N. */
int64_t oracle_body_force_count(int model, uint32_t dim_x, uint32_t dim_y, const uint8_t* types,
                                uint8_t* state, uint64_t seed, uint64_t call_index,
                                int force_dir, int32_t forcing)
{
    const uint64_t n = (uint64_t)dim_x * dim_y;
    const uint64_t it_max = 2 * n;
    uint32_t key[2];
    seed_key(seed, key);
    uint64_t it = 0;
    const uint32_t target = (uint32_t)forcing;
    uint32_t reverted = 0;
    do {
        uint32_t out[4];
        const uint32_t ctr[4] = { (uint32_t)it, (uint32_t)(it >> 32), (uint32_t)call_index, 3u };
        oracle_philox4x32_10(ctr, key, out);
        const uint64_t cell = (((uint64_t)out[1] << 32) | out[0]) % n;
        ++it;
        if (!types || types[cell] == 0) {
            const uint8_t s = state[cell];
            const uint8_t o = apply_force(model, force_dir, s);
            reverted += popc8((uint8_t)(s ^ o)) / 2;
            state[cell] = o;
        }
    } while (reverted < target && it < it_max);
    return (int64_t)reverted;
}

/* ------------------------------------------------------------ post-process */

void oracle_cell_post_process(int model, uint32_t dim_x, uint32_t dim_y, const uint8_t* state,
                              float* cell_density, float* cell_momentum)
{
    /* cell_post_process (src/omp_lattice.cpp:360-394): density = occupied
     * directions; momentum = running FP32 sum over directions 0..Q-1 of
     * bit * lattice vector. */
    const int q = oracle_num_dir(model);
    float vx[7], vy[7];
    for (int d = 0; d < q; ++d) { vx[d] = vec_f(model, d, 0); vy[d] = vec_f(model, d, 1); }
    const size_t n = (size_t)dim_x * dim_y;
    for (size_t c = 0; c < n; ++c) {
        int rho = 0;
        float mx = 0.0f, my = 0.0f;
        for (int d = 0; d < q; ++d) {
            const int b = (state[c] >> d) & 1;
            rho += b;
            mx += (float)b * vx[d];
            my += (float)b * vy[d];
        }
        cell_density[c] = (float)rho;
        cell_momentum[2 * c] = mx;
        cell_momentum[2 * c + 1] = my;
    }
}

void oracle_mean_post_process(uint32_t dim_x, uint32_t dim_y, uint32_t r,
                              const float* cell_density, const float* cell_momentum,
                              float* mean_density, float* mean_momentum)
{
    /* mean_post_process (src/omp_lattice.cpp:397-454).  Coarse cell c has its
     * corner at (2r * (c % cdx), 2r * (c / cdx)); candidates are the (2r+1)^2
     * cells at linear offsets y*dim_x + x, x, y in [0, 2r], visited y-outer,
     * x-inner.  A candidate counts when its linear index is inside the lattice
     * and its column is within r of the corner column (so the window is r+1
     * columns by up to 2r+1 rows).  Sums are sequential FP32, then divided
     * once by the count. */
    const uint32_t cdx = dim_x / (2 * r);
    const uint32_t cdy = dim_y / (2 * r);
    const size_t n = (size_t)dim_x * dim_y;
    for (size_t c = 0; c < (size_t)cdx * cdy; ++c) {
        const size_t corner = (c % cdx) * (2 * r) + (c / cdx) * (2 * r) * (size_t)dim_x;
        const long cx = (long)(corner % dim_x);
        float sd = 0.0f, smx = 0.0f, smy = 0.0f;
        int cnt = 0;
        for (uint32_t yy = 0; yy <= 2 * r; ++yy) {
            for (uint32_t xx = 0; xx <= 2 * r; ++xx) {
                const size_t idx = corner + (size_t)yy * dim_x + xx;
                if (idx >= n) continue;
                const long nx = (long)(idx % dim_x);
                if (labs(nx - cx) > (long)r) continue;
                ++cnt;
                sd += cell_density[idx];
                smx += cell_momentum[2 * idx];
                smy += cell_momentum[2 * idx + 1];
            }
        }
        mean_density[c] = sd / (float)cnt;
        mean_momentum[2 * c] = smx / (float)cnt;
        mean_momentum[2 * c + 1] = smy / (float)cnt;
    }
}

void oracle_coarse_velocity(size_t n_coarse, const float* mean_density,
                            const float* mean_momentum, float* velocity)
{
    /* Coarse velocity = mean momentum / mean density (0 where empty); the
     * reference exposes only the two factors (src/lattice.h:220-234). */
    for (size_t c = 0; c < n_coarse; ++c) {
        const float rho = mean_density[c];
        velocity[2 * c] = rho > 0.0f ? mean_momentum[2 * c] / rho : 0.0f;
        velocity[2 * c + 1] = rho > 0.0f ? mean_momentum[2 * c + 1] / rho : 0.0f;
    }
}

uint64_t oracle_count_particles(size_t n_cells, const uint8_t* state)
{
    /* get_n_particles (src/lattice.cpp:180-195) counts all 8 bits of a cell. */
    uint64_t total = 0;
    for (size_t c = 0; c < n_cells; ++c) total += (uint64_t)popc8(state[c]);
    return total;
}

void oracle_mean_velocity(uint32_t dim_x, uint32_t dim_y, const uint8_t* types,
                          const float* cell_density, const float* cell_momentum,
                          double out[2])
{
    /* get_mean_velocity (src/omp_lattice.cpp:508-557): mean over FLUID cells of
     * momentum / density, cells with density <= 1e-6 contributing zero; the
     * reference's FP32 reduction order is unspecified, so this sums in double. */
    const size_t n = (size_t)dim_x * dim_y;
    double sx = 0.0, sy = 0.0;
    size_t fluid = 0;
    for (size_t c = 0; c < n; ++c) {
        if (types && types[c] != 0) continue;
        ++fluid;
        const float rho = cell_density[c];
        if (rho > 1.0e-6f) {
            sx += (double)cell_momentum[2 * c] / rho;
            sy += (double)cell_momentum[2 * c + 1] / rho;
        }
    }
    out[0] = fluid ? sx / (double)fluid : 0.0;
    out[1] = fluid ? sy / (double)fluid : 0.0;
}
