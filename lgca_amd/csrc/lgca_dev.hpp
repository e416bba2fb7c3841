// lgca_dev.hpp -- device building blocks of the lattice-gas engine (gfx950).
//
// Cells are multi-spin coded: one bit per cell in 32-bit words, one word per
// direction ("plane") per 32 consecutive cells of a row.  Bit b of word w of a
// row is cell x = 32*w + b.  Every rule below is pure bitwise logic on such
// words, so one VALU op advances 32 cells; gfx950 fuses most 3-input
// expressions into a single v_bitop3_b32.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lgca {

// Models (src/lgca_common.h:49-54).  FHP-III collides exactly like FHP-II in the
// reference (every FHP-III-only term of src/lgca_models.h:816-839 is
// identically zero), so kernels are instantiated for kinds 0..2 only.
enum : int { kHPP = 0, kFHP1 = 1, kFHP2 = 2 };

// Four consecutive words of one plane-row (128 cells).  Bitwise operators act
// element-wise, so every rule below is written once for T = uint32_t (one
// word) and T = u32x4 (a quad); the compiler emits one VALU op per element.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ constexpr int num_dir_of_kind(int kind) { return kind == kHPP ? 4 : (kind == kFHP1 ? 6 : 7); }

// ------------------------------------------------------- three-input logic
// Any bitwise function of three words is one v_bitop3_b32 on gfx950.  Its
// truth table LUT holds f(x, y, z) at bit 4x + 2y + z, so writing f on the
// operand masks kX, kY, kZ gives the table: bit3<kX & ~kY & kZ>(x, y, z) is
// x & ~y & z.  Left to itself the compiler splits many such expressions into
// two-input ops; the rules below are written directly as three-input steps.
constexpr uint32_t kX = 0xF0u, kY = 0xCCu, kZ = 0xAAu;

template <uint32_t LUT>
__host__ __device__ __forceinline__ uint32_t bit3(uint32_t x, uint32_t y, uint32_t z)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(x, y, z, (unsigned char)(LUT & 0xFFu));
#else
    uint32_t r = 0;
    for (int i = 0; i < 8; ++i)
        if ((LUT >> i) & 1u) r |= ((i & 4) ? x : ~x) & ((i & 2) ? y : ~y) & ((i & 1) ? z : ~z);
    return r;
#endif
}

template <uint32_t LUT>
__host__ __device__ __forceinline__ u32x4 bit3(u32x4 x, u32x4 y, u32x4 z)
{
    u32x4 r;
    r.x = bit3<LUT>(x.x, y.x, z.x);
    r.y = bit3<LUT>(x.y, y.y, z.y);
    r.z = bit3<LUT>(x.z, y.z, z.z);
    r.w = bit3<LUT>(x.w, y.w, z.w);
    return r;
}

constexpr uint32_t kXor3 = kX ^ kY ^ kZ;
constexpr uint32_t kMaj = (kX & kY) | (kZ & (kX | kY));
constexpr uint32_t kAnd3 = kX & kY & kZ;
constexpr uint32_t kOr3 = kX | kY | kZ;
constexpr uint32_t kMux = (kX & kY) | (~kX & kZ);  // x ? y : z

// --------------------------------------------------------------- Philox4x32-10
// Salmon, Moraes, Dror, Shaw, SC'11.  The key schedule is wave-uniform, so the
// per-round Weyl increments stay in scalar registers.
struct U4 { uint32_t x, y, z, w; };

__host__ __device__ __forceinline__ U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                            uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        // the first two rounds often work on wave-uniform counter words,
        // which the compiler keeps in scalar registers; later rounds use
        // one three-input xor each
        const uint32_t n0 = r < 2 ? ((uint32_t)(p1 >> 32) ^ c1 ^ k0) : bit3<kXor3>((uint32_t)(p1 >> 32), c1, k0);
        const uint32_t n2 = r < 2 ? ((uint32_t)(p0 >> 32) ^ c3 ^ k1) : bit3<kXor3>((uint32_t)(p0 >> 32), c3, k1);
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
    }
    return U4{c0, c1, c2, c3};
}

// ------------------------------------------------------------------- collision
// HPP (src/lgca_models.h:134-152): an isolated head-on pair along one axis
// turns into a pair along the other, which happens exactly when s0 == s2,
// s1 == s3 and s0 != s1.
template <typename T>
__host__ __device__ __forceinline__ void collide_hpp(T s[4])
{
    const T k = bit3<~(kX ^ kY) & (kX ^ kZ)>(s[0], s[2], s[1]);  // s0 == s2 != s1
    const T h = bit3<kX & ~(kY ^ kZ)>(k, s[1], s[3]);             // and s1 == s3
    s[0] ^= h; s[1] ^= h; s[2] ^= h; s[3] ^= h;
}

// FHP-I / FHP-II (src/lgca_models.h:367-395, 568-613), restated by occupancy.
// s[0..5] moving directions, s[6] rest (FHP-II), p chirality (1 = clockwise).
//   * exactly two moving particles, head-on {k, k+3}: rotate by +60 degrees
//     (p = 0) or -60 degrees (p = 1), rest untouched;
//   * exactly three, {0,2,4} or {1,3,5}: swap, rest untouched;
//   * FHP-II: rest + exactly one moving particle i -> particles i-1 and i+1;
//     exactly two moving at i-1 and i+1 without rest -> rest + particle i.
// Every step is one three-input operation (43 for FHP-II, 25 for FHP-I).
template <bool WITH_REST, typename T>
__host__ __device__ __forceinline__ void collide_fhp(T s[7], T p)
{
    const T a0 = s[0], a1 = s[1], a2 = s[2], a3 = s[3], a4 = s[4], a5 = s[5];
    // occupancy n = x1 + x2 + 2 (c1 + c2), from the sums of the two triples
    const T x1 = bit3<kXor3>(a0, a1, a2), c1 = bit3<kMaj>(a0, a1, a2);
    const T x2 = bit3<kXor3>(a3, a4, a5), c2 = bit3<kMaj>(a3, a4, a5);
    const T u = x1 ^ x2;  // n odd
    const T v = c1 ^ c2;  // exactly one triple holds two or more
    // n == 2: x1 == x2 and either both are set with no carries, or neither
    // is and exactly one carry is
    const T w = bit3<(kX & ~(kY | kZ)) | (~kX & (kY ^ kZ))>(x1, c1, c2);
    const T o2 = bit3<~kX & kY>(u, w, w);
    // n == 3 with a0 == a2 == a4: the symmetric triples
    const T eq = bit3<~(kX ^ kY) & ~(kY ^ kZ)>(a0, a2, a4);
    const T t = bit3<kAnd3>(u, v, eq);
    // head-on pairs with nothing else moving
    const T h0 = bit3<kAnd3>(o2, a0, a3), h1 = bit3<kAnd3>(o2, a1, a4), h2 = bit3<kAnd3>(o2, a2, a5);
    // class m = d mod 3 flips when its own pair leaves or a pair rotates in:
    // p = 0 moves pair k to class k+1, p = 1 to class k+2.
    const T A0 = bit3<kOr3>(t, h0, bit3<kMux>(p, h1, h2));
    const T A1 = bit3<kOr3>(t, h1, bit3<kMux>(p, h2, h0));
    const T A2 = bit3<kOr3>(t, h2, bit3<kMux>(p, h0, h1));
    if (WITH_REST) {
        const T r = s[6];
        const T o1 = bit3<kX & ~(kY | kZ)>(u, c1, c2);  // n == 1
        // n == 2 with both particles in one of the triples {0,2,4}, {1,3,5}:
        // the pair i-1, i+1 of the rest rule
        const T y = bit3<kX & (kY | kZ)>(o2, bit3<kMaj>(a0, a2, a4), bit3<kMaj>(a1, a3, a5));
        // a rest-rule event: with rest, a lone particle i moves to i-1 and
        // i+1, so direction d takes a[d-1] | a[d+1]; without, the pair i-1,
        // i+1 merges into i, so d takes a[d-1] & a[d+1].  Both are
        // maj(r, a[d-1], a[d+1]), and the rest bit flips.
        const T Z = bit3<kMux>(r, o1, y);
        s[0] = bit3<kMux>(Z, bit3<kMaj>(r, a5, a1), a0) ^ A0;
        s[1] = bit3<kMux>(Z, bit3<kMaj>(r, a0, a2), a1) ^ A1;
        s[2] = bit3<kMux>(Z, bit3<kMaj>(r, a1, a3), a2) ^ A2;
        s[3] = bit3<kMux>(Z, bit3<kMaj>(r, a2, a4), a3) ^ A0;
        s[4] = bit3<kMux>(Z, bit3<kMaj>(r, a3, a5), a4) ^ A1;
        s[5] = bit3<kMux>(Z, bit3<kMaj>(r, a4, a0), a5) ^ A2;
        s[6] = r ^ Z;
    } else {
        s[0] = a0 ^ A0; s[1] = a1 ^ A1; s[2] = a2 ^ A2;
        s[3] = a3 ^ A0; s[4] = a4 ^ A1; s[5] = a5 ^ A2;
    }
}

template <int KIND, typename T>
__host__ __device__ __forceinline__ void collide(T s[7], T p)
{
    if (KIND == kHPP) collide_hpp<T>(s);
    else collide_fhp<KIND == kFHP2, T>(s, p);
}

// ---------------------------------------------------------- solid-cell rules
// Direction permutations (src/lgca_models.h:44-48, 239-243, 446-450):
//   inverse: HPP d+2, FHP d+3; mirror in x: -d; mirror in y: HPP 2-d, FHP 3-d
//   (mod 4 or 6; the rest direction maps to itself).
template <int KIND> __host__ __device__ constexpr int inv_dir(int d)
{ return KIND == kHPP ? ((d + 2) & 3) : (d == 6 ? 6 : (d + 3) % 6); }
template <int KIND> __host__ __device__ constexpr int mirx_dir(int d)
{ return KIND == kHPP ? ((4 - d) & 3) : (d == 6 ? 6 : (6 - d) % 6); }
template <int KIND> __host__ __device__ constexpr int miry_dir(int d)
{ return KIND == kHPP ? ((6 - d) & 3) : (d == 6 ? 6 : (9 - d) % 6); }

// Applies the destination-cell dispatch of src/omp_lattice.cpp:193-231 to one
// word: fluid bits keep the collided value c, NO_SLIP bits take the bounce-back
// of the pulled value in, SLIP bits take the x-mirror on the north/south rows,
// the y-mirror on the east/west columns (y wins at corners), and pass through
// elsewhere.  ns, sl: NO_SLIP and SLIP masks; edge: cells on x = 0 or
// x = dim_x - 1; ns_row: the word lies on row 0 or dim_y - 1.
template <int KIND, typename T>
__host__ __device__ __forceinline__ void apply_solids(T c[7], const T in[7], T ns, T sl, T edge,
                                                      bool ns_row)
{
    constexpr int Q = num_dir_of_kind(KIND);
    const T fl = ~(ns | sl);
    T o[7];
#pragma unroll
    for (int d = 0; d < Q; ++d) {
        T sp = ns_row ? in[mirx_dir<KIND>(d)] : in[d];
        sp = (edge & in[miry_dir<KIND>(d)]) | (~edge & sp);
        o[d] = (fl & c[d]) | (ns & in[inv_dir<KIND>(d)]) | (sl & sp);
    }
#pragma unroll
    for (int d = 0; d < Q; ++d) c[d] = o[d];
}

// apply_solids where no cell of the word is SLIP (sl == 0): a NO_SLIP bit
// takes the bounce-back of the pulled value, a fluid bit the collided value,
// one three-input select per plane.
template <int KIND, typename T>
__host__ __device__ __forceinline__ void apply_no_slip(T c[7], const T in[7], T ns)
{
    constexpr int Q = num_dir_of_kind(KIND);
    T o[7];
#pragma unroll
    for (int d = 0; d < Q; ++d) o[d] = bit3<kMux>(ns, in[inv_dir<KIND>(d)], c[d]);
#pragma unroll
    for (int d = 0; d < Q; ++d) c[d] = o[d];
}

// Body-force flip rules of apply_body_force (src/omp_lattice.cpp:293-337) on
// the cells selected by f (already restricted to FLUID cells).
template <int KIND, typename T>
__host__ __device__ __forceinline__ void apply_force(T s[7], T f, int dir)
{
    if (KIND == kHPP) {
        if (dir == 1) { const T m = f & ~s[0] & s[2]; s[0] ^= m; s[2] ^= m; }
        else          { const T m = f & s[1] & ~s[3]; s[1] ^= m; s[3] ^= m; }
    } else {
        if (dir == 1) { const T m = f & ~s[0] & s[3]; s[0] ^= m; s[3] ^= m; }
        else {
            const T m1 = f & s[1] & ~s[5];
            const T m2 = f & s[2] & ~s[4];
            s[1] ^= m1; s[5] ^= m1;
            s[2] ^= m2; s[4] ^= m2;
        }
    }
}

// Initial fill of plane d of word w of row y (hash stream 2): bit b is set
// when philox({x >> 2, y, d, 2})[x & 3] < thr for x = 32w + b < dim_x.
__device__ __forceinline__ uint32_t fill_word(uint32_t w, uint32_t y, uint32_t d, uint32_t dim_x,
                                              uint32_t key0, uint32_t key1, uint64_t thr)
{
    uint32_t word = 0;
    for (uint32_t gi = 0; gi < 8; ++gi) {
        const uint32_t x = 32 * w + 4 * gi;
        if (x >= dim_x) break;
        const U4 r = philox4x32_10(x >> 2, y, d, 2u, key0, key1);
        word |= (uint32_t)((uint64_t)r.x < thr) << (4 * gi + 0);
        word |= (uint32_t)((uint64_t)r.y < thr) << (4 * gi + 1);
        word |= (uint32_t)((uint64_t)r.z < thr) << (4 * gi + 2);
        word |= (uint32_t)((uint64_t)r.w < thr) << (4 * gi + 3);
    }
    return word;
}

// Per-cell force decisions of word w of row y at one step of the four-step
// group G = step >> 2: 32 independent Bernoulli(thr / 2^32) bits by
// first-fire sampling.  u0 is the word's first uniform for this step,
// philox({w, y, G, 1})[step & 3] (one call serves the group).  Further
// uniforms, needed only after a fire, are the group's extras
// E_e = philox({w, y, G, 5 + 4 (e >> 2)})[e & 3], e = 0, 1, ..., shared by the
// group's steps in step order: `used` counts the extras the group's earlier
// steps took and `ext` holds block used >> 2 (the caller loads block 0 when
// any step of the group fires).  Qt[k] = Q_{k+1}, where Q_1 = q = 2^32 - thr
// and Q_{k+1} = floor(Q_k q / 2^32) (force_table); the first of the next n
// cells to fire is the smallest m < n with u + Q_{m+1} < 2^32, and there is
// none when u + Q_n >= 2^32.
__device__ __forceinline__ bool force_fires(uint32_t u0, const uint32_t* Qt)
{
    return (uint64_t)u0 + Qt[31] < ((uint64_t)1 << 32);
}

// Offset of the first fire among 32 cells for uniform u, 32 when none fires:
// the number of k in [0, 32) with u + Qt[k] >= 2^32, a prefix because Qt
// decreases.  Five fixed lifting steps find it up to 31, and Qt[31] adds the
// last; no branches, so the lanes of a wave stay in step.  For the next n
// cells of a word the first fire is this offset when it is < n, else there is
// none (the prefix covers the first n cells).
__device__ __forceinline__ uint32_t force_first(uint32_t u, const uint32_t* Qt)
{
    constexpr uint64_t kTwo32 = (uint64_t)1 << 32;
    uint32_t m = 0;
#pragma unroll
    for (uint32_t step = 16; step; step >>= 1)
        m += ((uint64_t)u + Qt[m + step - 1] >= kTwo32) ? step : 0u;
    return m + (((uint64_t)u + Qt[31] >= kTwo32) ? 1u : 0u);
}

__device__ __forceinline__ uint32_t force_word(uint32_t u0, uint32_t w, uint32_t y, uint32_t group,
                                               uint32_t key0, uint32_t key1, const uint32_t* Qt,
                                               uint32_t& used, U4& ext)
{
    if (!force_fires(u0, Qt)) return 0;  // no fire among all 32 cells
    uint32_t word = 0, pos = 0, u = u0;
    for (;;) {
        const uint32_t m = force_first(u, Qt);
        if (m >= 32 - pos) break;  // no fire among the word's remaining cells
        word |= 1u << (pos + m);
        pos += m + 1;
        if (pos == 32) break;
        if ((used & 3u) == 0 && used != 0) ext = philox4x32_10(w, y, group, 5u + used, key0, key1);
        const uint32_t i = used & 3u;
        u = i == 0 ? ext.x : (i == 1 ? ext.y : (i == 2 ? ext.z : ext.w));
        ++used;
    }
    return word;
}

// Fills Qt[0..31] = Q_1..Q_32 for q = 2^32 - thr (exact integer steps, the
// same on every platform) and makes it visible to the whole workgroup.
__device__ __forceinline__ void force_table(uint32_t* Qt, uint32_t q)
{
    if (threadIdx.x == 0) {
        uint32_t Qk = q;
        for (int k = 0; k < 32; ++k) {
            Qt[k] = Qk;
            Qk = __umulhi(Qk, q);
        }
    }
    __syncthreads();
}

// Bits of word w inside the forced columns [fx0, fx1) and the lattice.
__device__ __forceinline__ uint32_t force_columns(uint32_t w, uint32_t fx0, uint32_t fx1, uint32_t dim_x)
{
    const uint32_t x0 = 32 * w;
    const uint32_t hi = min(min(fx1, dim_x), x0 + 32);
    const uint32_t lo = max(fx0, x0);
    if (lo >= hi) return 0;
    const uint32_t nbits = hi - lo, sh = lo - x0;
    return (nbits == 32 ? 0xFFFFFFFFu : ((1u << nbits) - 1u)) << sh;
}

// One thread's share of a force draw: the decision words of word w of row y
// for the steps of group `group` (steps sharing (step & ~3), counted from
// step0 & ~3) that fall in [step0, step0 + nsteps), stored at dec[k * plane]
// for window index k; cols masks the forced columns of the word.  Steps of
// the group before step0 are walked too: they take extras before the later
// steps do.
__device__ __forceinline__ void draw_group(uint32_t* dec, uint64_t plane, uint32_t nsteps,
                                           uint32_t step0, uint32_t group, uint32_t w, uint32_t y,
                                           uint32_t key0, uint32_t key1, const uint32_t* Qt,
                                           uint32_t cols)
{
    const uint32_t first = step0 & 3u;
    const uint32_t G = ((step0 & ~3u) + 4u * group) >> 2;  // of the group's steps mod 2^32
    uint32_t u[4] = {0, 0, 0, 0};
    U4 ext{0, 0, 0, 0};
    if (cols) {
        const U4 r = philox4x32_10(w, y, G, 1u, key0, key1);
        u[0] = r.x; u[1] = r.y; u[2] = r.z; u[3] = r.w;
        // one call for the first four extras of the whole group, made up front
        // so that a wave whose lanes fire at different steps makes it once
        if (force_fires(u[0], Qt) || force_fires(u[1], Qt) || force_fires(u[2], Qt) ||
            force_fires(u[3], Qt))
            ext = philox4x32_10(w, y, G, 5u, key0, key1);
    }
    uint32_t used = 0;
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
        const uint32_t k = 4u * group + i - first;  // window index of step 4G + i
        if (4u * group + i >= first + nsteps) break;
        const uint32_t word = cols ? (force_word(u[i], w, y, G, key0, key1, Qt, used, ext) & cols) : 0u;
        if (4u * group + i >= first) dec[(size_t)k * plane] = word;
    }
}

}  // namespace lgca
