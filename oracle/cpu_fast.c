/*
 * oracle/cpu_fast.c -- bit-sliced CPU restatement of the same path, for parity
 * runs too long for cpu_ref.c (thousands of steps on the 32768 x 8192
 * lattice).
 *
 * TEST INFRASTRUCTURE ONLY, like cpu_ref.c: only tests/ load it, as a checker.
 *
 * It computes exactly what oracle_run (cpu_ref.c) computes, by another route:
 *   * 64 cells per machine word, one word per direction per 64 cells of a row
 *     (planes[(y * Q + d) * W + w], bit b = cell 64 w + b);
 *   * the pull of src/omp_lattice.cpp:141-180 as shifts of whole words, with
 *     the neighbour word's end bit carried in and the torus closed at dim_x and
 *     dim_y (the same pull sources oracle_pull_source states);
 *   * the collision by particle pattern (each isolated head-on pair, triple
 *     and rest-rule configuration matched word-wide, its directions flipped),
 *     in straight-line word logic the compiler vectorises (the occupancy-count
 *     logic of the GPU kernels is not used);
 *   * bounce-back and slip of src/omp_lattice.cpp:193-231 on the destination
 *     cell, from oracle_bounce_back / _forward_x / _forward_y permutations;
 *   * chirality and per-cell force words of the hash contract (cpu_ref.h), the
 *     force words from cpu_ref.c's own first-fire walk.
 * Pinned by equality with oracle_run on every model, geometry and force
 * setting (tests/test_oracle.py); it is a faster checker, not a new contract.
 */
#include "cpu_ref.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef uint64_t u64;

typedef struct {
    int model, q;
    uint32_t dim_x, dim_y, W;       /* W: 64-cell words per row */
    u64 last_mask;                  /* valid bits of word W-1 */
    uint32_t nb;                    /* valid bits of word W-1, 1..64 */
    int perm[3][7];                 /* bounce_back, forward_x, forward_y: out d <- in perm[d] */
    int pull_dx[2][7], pull_dy[2][7];  /* by row parity: source offset of direction d */
} fast_t;

static void fast_setup(fast_t* f, int model, uint32_t dim_x, uint32_t dim_y)
{
    memset(f, 0, sizeof *f);
    f->model = model;
    f->q = oracle_num_dir(model);
    f->dim_x = dim_x;
    f->dim_y = dim_y;
    f->W = (dim_x + 63) / 64;
    f->nb = dim_x - 64 * (f->W - 1);
    f->last_mask = f->nb == 64 ? ~(u64)0 : (((u64)1 << f->nb) - 1);
    const int q = f->q;
    /* permutations, read back from the byte-level rules one direction at a time */
    for (int d = 0; d < q; ++d) {
        for (int e = 0; e < q; ++e) {
            const uint8_t in = (uint8_t)(1u << e);
            if (oracle_bounce_back(model, in) & (1u << d)) f->perm[0][d] = e;
            if (oracle_bounce_forward_x(model, in) & (1u << d)) f->perm[1][d] = e;
            if (oracle_bounce_forward_y(model, in) & (1u << d)) f->perm[2][d] = e;
        }
    }
    /* pull offsets by row parity, read from oracle_pull_source away from the
     * torus seams (they are the same for every cell of a parity) */
    const uint32_t px = 3, pw = 8, ph = 8;
    for (int par = 0; par < 2; ++par) {
        for (int d = 0; d < q; ++d) {
            uint32_t sx, sy;
            oracle_pull_source(model, pw, ph, px, 2 + par, d, &sx, &sy);
            f->pull_dx[par][d] = (int)sx - (int)px;
            f->pull_dy[par][d] = (int)sy - (int)(2 + par);
        }
    }
}

/* Decision bits of the 64 cells of word w for one step: two 32-cell force
 * words, restricted to the force columns. */
static inline u64 force_bits(const fast_t* f, const uint32_t key[2], const uint64_t C[32], uint64_t step,
                             uint32_t y, uint32_t w, uint32_t fx0, uint32_t fx1)
{
    u64 r = 0;
    for (uint32_t h = 0; h < 2; ++h) {
        const uint32_t w32 = 2 * w + h, x0 = 32 * w32;
        if (x0 >= f->dim_x || x0 >= fx1 || x0 + 32 <= fx0) continue;
        uint32_t word = oracle_force_word_with(key, C, step, y, w32);
        for (uint32_t b = 0; b < 32; ++b) {
            const uint32_t x = x0 + b;
            if (x < fx0 || x >= fx1 || x >= f->dim_x) word &= ~(1u << b);
        }
        r |= (u64)word << (32 * h);
    }
    return r;
}

/* Collision of n words of pulled planes s[d][0..n) into o[d][0..n), by
 * particle pattern, each pattern with every other moving direction empty:
 *   HPP: the head-on pair along x or along y turns into the other.
 *   FHP: head-on pair {k, k+3} -> {k+1, k+4} with chirality 0, {k+2, k+5}
 *   with chirality 1; the triples {0,2,4} and {1,3,5} swap; FHP-II also turns
 *   rest + lone particle i into particles i-1 and i+1, and the inverse.
 * Straight-line word logic, so the compiler vectorises the loops.  It must
 * equal oracle_collide, which the tests check through oracle_run. */
static void collide_hpp_words(u64* const* s, u64* const* o, uint32_t n)
{
    for (uint32_t w = 0; w < n; ++w) {
        const u64 a0 = s[0][w], a1 = s[1][w], a2 = s[2][w], a3 = s[3][w];
        const u64 x = a0 & a2 & ~a1 & ~a3, y = a1 & a3 & ~a0 & ~a2, c = x | y;
        o[0][w] = a0 ^ c;
        o[1][w] = a1 ^ c;
        o[2][w] = a2 ^ c;
        o[3][w] = a3 ^ c;
    }
}

static void collide_fhp_words(int rest, u64* const* s, u64* const* o, const u64* p, uint32_t n)
{
    for (uint32_t w = 0; w < n; ++w) {
        u64 a[6];
        for (int d = 0; d < 6; ++d) a[d] = s[d][w];
        const u64 r = rest ? s[6][w] : 0;
        const u64 pw = p[w];
        /* ex[i]: some moving direction other than i is occupied */
        u64 pre[7], suf[7], ex[6];
        pre[0] = 0;
        for (int d = 0; d < 6; ++d) pre[d + 1] = pre[d] | a[d];
        suf[6] = 0;
        for (int d = 5; d >= 0; --d) suf[d] = suf[d + 1] | a[d];
        for (int i = 0; i < 6; ++i) ex[i] = pre[i] | suf[i + 1];
        /* isolated head-on pairs: both set, the four others empty */
        u64 h[3];
        for (int k = 0; k < 3; ++k)
            h[k] = a[k] & a[k + 3] & ~(a[(k + 1) % 6] | a[(k + 2) % 6] | a[(k + 4) % 6] | a[(k + 5) % 6]);
        const u64 t = (a[0] & a[2] & a[4] & ~(a[1] | a[3] | a[5])) |
                      (a[1] & a[3] & a[5] & ~(a[0] | a[2] | a[4]));
        u64 fl[7];
        for (int k = 0; k < 3; ++k) {
            /* class k flips when its pair leaves or a pair rotates into it */
            const u64 in_pair = (h[(k + 2) % 3] & ~pw) | (h[(k + 1) % 3] & pw);
            fl[k] = fl[k + 3] = h[k] | in_pair | t;
        }
        fl[6] = 0;
        if (rest) {
            for (int i = 0; i < 6; ++i) {
                const int m1 = (i + 5) % 6, p1 = (i + 1) % 6;
                /* rest + lone i, or the pair i-1, i+1 alone without rest */
                const u64 lone = r & a[i] & ~ex[i];
                const u64 two = ~r & a[m1] & a[p1] & ~(a[i] | a[(i + 2) % 6] | a[(i + 3) % 6] | a[(i + 4) % 6]);
                const u64 ev = lone | two;
                fl[m1] |= ev;
                fl[i] |= ev;
                fl[p1] |= ev;
                fl[6] |= ev;
            }
        }
        for (int d = 0; d < 6; ++d) o[d][w] = a[d] ^ fl[d];
        if (rest) o[6][w] = r ^ fl[6];
    }
}

/* The word logic against oracle_collide on every state and chirality (256
 * cells, one per input): 0 when they agree. */
static int collide_self_check(const fast_t* f)
{
    const int q = f->q;
    u64 sv[7][4], ov[7][4], pv[4];
    memset(sv, 0, sizeof sv);
    memset(pv, 0, sizeof pv);
    for (int c = 0; c < 256; ++c) {
        const int st = c & ((1 << q) - 1), p = c >> 7;
        for (int d = 0; d < q; ++d)
            if ((st >> d) & 1) sv[d][c >> 6] |= (u64)1 << (c & 63);
        if (p) pv[c >> 6] |= (u64)1 << (c & 63);
    }
    u64* sp[7];
    u64* op[7];
    for (int d = 0; d < 7; ++d) {
        sp[d] = sv[d];
        op[d] = ov[d];
    }
    if (f->model == 0) collide_hpp_words(sp, op, 4);
    else collide_fhp_words(q == 7, sp, op, pv, 4);
    for (int c = 0; c < 256; ++c) {
        const int st = c & ((1 << q) - 1), p = c >> 7;
        uint8_t got = 0;
        for (int d = 0; d < q; ++d) got |= (uint8_t)(((ov[d][c >> 6] >> (c & 63)) & 1u) << d);
        if (got != oracle_collide(f->model, (uint8_t)st, f->model == 0 ? 0 : p)) return -1;
    }
    return 0;
}

/* One row: pull into row buffers, collide, then the solid cells and the
 * force of the words that have them. */
static void fast_row(const fast_t* f, const u64* in, u64* out, const u64* ns_pl, const u64* sl_pl,
                     const u64* pt_pl, const uint32_t (*chir)[4], uint64_t step, const uint32_t key[2],
                     const uint64_t* C, int force_dir, uint32_t fx0, uint32_t fx1, uint32_t fy0, uint32_t fy1,
                     uint32_t y, u64* scratch)
{
    const int q = f->q;
    const uint32_t W = f->W, dim_y = f->dim_y, W32 = (f->dim_x + 31) / 32;
    const int par = f->model == 0 ? 0 : (int)(y & 1u);
    u64* sb[7];
    u64* ob[7];
    for (int d = 0; d < q; ++d) {
        sb[d] = scratch + (size_t)d * W;
        ob[d] = out + ((size_t)y * q + d) * W;
    }
    u64* pw = scratch + (size_t)7 * W;
    /* pull stream: source rows and x shifts of this row's parity */
    for (int d = 0; d < q; ++d) {
        const uint32_t sy = (uint32_t)(((int64_t)y + f->pull_dy[par][d] + dim_y) % dim_y);
        const u64* row = in + ((size_t)sy * q + d) * W;
        const int dx = f->pull_dx[par][d];
        u64* dst = sb[d];
        if (dx == 0) {
            memcpy(dst, row, (size_t)W * sizeof(u64));
        } else if (dx < 0) {   /* out x = in x - 1 */
            for (uint32_t w = 1; w < W; ++w) dst[w] = (row[w] << 1) | (row[w - 1] >> 63);
            dst[0] = (row[0] << 1) | ((row[W - 1] >> (f->nb - 1)) & 1u);
            dst[W - 1] &= f->last_mask;
        } else {               /* out x = in x + 1 */
            for (uint32_t w = 0; w + 1 < W; ++w) dst[w] = (row[w] >> 1) | (row[w + 1] << 63);
            dst[W - 1] = (row[W - 1] >> 1) | ((row[0] & 1u) << (f->nb - 1));
        }
    }
    if (f->model == 0) {
        collide_hpp_words(sb, ob, W);
    } else {
        for (uint32_t w = 0; w < W; ++w) {
            const uint32_t lo = chir[2 * w][step & 3];
            const uint32_t hi = 2 * w + 1 < W32 ? chir[2 * w + 1][step & 3] : 0u;
            pw[w] = (u64)lo | ((u64)hi << 32);
        }
        collide_fhp_words(q == 7, sb, ob, pw, W);
    }
    /* solid cells: bounce-back, slip, pass-through (src/omp_lattice.cpp:193-231) */
    const int ns_row = y == 0 || y == dim_y - 1;
    if (ns_pl) {
        for (uint32_t w = 0; w < W; ++w) {
            const u64 nsw = ns_pl[(size_t)y * W + w], slw = sl_pl[(size_t)y * W + w];
            const u64 ptw = pt_pl[(size_t)y * W + w];
            if (!(nsw | slw | ptw)) continue;
            const u64 fluid = ~(nsw | slw | ptw);
            u64 edge = 0;
            if (w == 0) edge |= 1u;
            if (w == W - 1) edge |= (u64)1 << (f->nb - 1);
            for (int d = 0; d < q; ++d) {
                u64 sp = ns_row ? sb[f->perm[1][d]][w] : sb[d][w];
                sp = (edge & sb[f->perm[2][d]][w]) | (~edge & sp);
                ob[d][w] = (fluid & ob[d][w]) | (nsw & sb[f->perm[0][d]][w]) | (slw & sp) | (ptw & sb[d][w]);
            }
        }
    }
    /* per-cell force on the fluid cells of the region (src/omp_lattice.cpp:293-337) */
    if (force_dir != ORACLE_FORCE_NONE && C && y >= fy0 && y < fy1 && fx0 < fx1) {
        for (uint32_t w = fx0 / 64; w < W && 64 * w < fx1; ++w) {
            u64 fb = force_bits(f, key, C, step, y, w, fx0, fx1);
            if (ns_pl) fb &= ~(ns_pl[(size_t)y * W + w] | sl_pl[(size_t)y * W + w] | pt_pl[(size_t)y * W + w]);
            if (!fb) continue;
            if (f->model == 0) {
                if (force_dir == ORACLE_FORCE_X) {
                    const u64 m = fb & ~ob[0][w] & ob[2][w]; ob[0][w] ^= m; ob[2][w] ^= m;
                } else {
                    const u64 m = fb & ob[1][w] & ~ob[3][w]; ob[1][w] ^= m; ob[3][w] ^= m;
                }
            } else if (force_dir == ORACLE_FORCE_X) {
                const u64 m = fb & ~ob[0][w] & ob[3][w]; ob[0][w] ^= m; ob[3][w] ^= m;
            } else {
                const u64 m1 = fb & ob[1][w] & ~ob[5][w], m2 = fb & ob[2][w] & ~ob[4][w];
                ob[1][w] ^= m1; ob[5][w] ^= m1; ob[2][w] ^= m2; ob[4][w] ^= m2;
            }
        }
    }
}

int oracle_fast_run(int model, uint32_t dim_x, uint32_t dim_y, const uint8_t* types, uint8_t* state,
                    uint64_t seed, uint64_t step0, uint64_t nsteps, int force_dir, uint64_t force_thr,
                    uint32_t fx0, uint32_t fx1, uint32_t fy0, uint32_t fy1, int nthreads)
{
    fast_t f;
    fast_setup(&f, model, dim_x, dim_y);
    if (f.q == 0 || dim_x == 0 || dim_y == 0 || collide_self_check(&f) != 0) return -1;
    const int q = f.q;
    const uint32_t W = f.W, W32 = (dim_x + 31) / 32;
    const size_t plane_words = (size_t)dim_y * q * W;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    u64* a = (u64*)calloc(plane_words, sizeof(u64));
    u64* b = (u64*)calloc(plane_words, sizeof(u64));
    u64* ns = NULL;
    u64* sl = NULL;
    u64* pt = NULL;  /* cells of other types pass their pulled state through */
    int any_solid = 0;
    if (types)
        for (size_t c = 0; c < (size_t)dim_x * dim_y && !any_solid; ++c) any_solid = types[c] != 0;
    if (any_solid) {
        ns = (u64*)calloc((size_t)dim_y * W, sizeof(u64));
        sl = (u64*)calloc((size_t)dim_y * W, sizeof(u64));
        pt = (u64*)calloc((size_t)dim_y * W, sizeof(u64));
    }
    uint32_t (*chir)[4] = (uint32_t (*)[4])calloc((size_t)dim_y * W32 + 1, sizeof *chir);
#ifdef _OPENMP
    const int nthr = omp_get_max_threads();
#else
    const int nthr = 1;
#endif
    /* per thread: seven pulled plane rows and the chirality words of a row */
    u64* scratch_all = (u64*)malloc((size_t)nthr * 8 * W * sizeof(u64));
    /* bytes -> planes (and the type masks) */
#pragma omp parallel for schedule(static)
    for (long long y = 0; y < (long long)dim_y; ++y) {
        for (uint32_t x = 0; x < dim_x; ++x) {
            const size_t cell = (size_t)y * dim_x + x;
            const uint32_t w = x >> 6;
            const u64 bit = (u64)1 << (x & 63);
            for (int d = 0; d < q; ++d)
                if ((state[cell] >> d) & 1u) a[((size_t)y * q + d) * W + w] |= bit;
            if (any_solid) {
                if (types[cell] == 1) ns[(size_t)y * W + w] |= bit;
                else if (types[cell] == 2) sl[(size_t)y * W + w] |= bit;
                else if (types[cell] != 0) pt[(size_t)y * W + w] |= bit;
            }
        }
    }
    uint32_t key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) };
    uint64_t C[32];
    const uint64_t* Cp = NULL;
    if (force_dir != ORACLE_FORCE_NONE && force_thr != 0) {
        oracle_force_table(force_thr, C);
        Cp = C;
    }
    if (fx1 > dim_x) fx1 = dim_x;
    if (fy1 > dim_y) fy1 = dim_y;
    for (uint64_t t = 0; t < nsteps; ++t) {
        const uint64_t step = step0 + t;
#pragma omp parallel for schedule(static)
        for (long long yy = 0; yy < (long long)dim_y; ++yy) {
            const uint32_t y = (uint32_t)yy;
#ifdef _OPENMP
            u64* scratch = scratch_all + (size_t)omp_get_thread_num() * 8 * W;
#else
            u64* scratch = scratch_all;
#endif
            uint32_t (*cr)[4] = chir + (size_t)y * W32;
            /* one chirality call per 32-cell word serves its four-step group */
            if (model != 0 && (t == 0 || (step & 3) == 0)) {
                for (uint32_t w32 = 0; w32 < W32; ++w32) {
                    const uint32_t ctr[4] = { w32, y, (uint32_t)step >> 2, 0u };
                    oracle_philox4x32_10(ctr, key, cr[w32]);
                }
            }
            fast_row(&f, a, b, ns, sl, pt, (const uint32_t (*)[4])cr, step, key, Cp, force_dir, fx0, fx1, fy0,
                     fy1, y, scratch);
        }
        u64* c = a; a = b; b = c;
    }
    /* planes -> bytes */
#pragma omp parallel for schedule(static)
    for (long long y = 0; y < (long long)dim_y; ++y) {
        for (uint32_t x = 0; x < dim_x; ++x) {
            const uint32_t w = x >> 6, sh = x & 63;
            uint8_t v = 0;
            for (int d = 0; d < q; ++d) v |= (uint8_t)(((a[((size_t)y * q + d) * W + w] >> sh) & 1u) << d);
            state[(size_t)y * dim_x + x] = v;
        }
    }
    free(a);
    free(b);
    free(ns);
    free(sl);
    free(pt);
    free(chir);
    free(scratch_all);
    return 0;
}
