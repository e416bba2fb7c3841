// lgca_kernels.hip -- hand-written gfx950 kernels of the lattice-gas engine.
//
// The hot kernel is step_kernel: one fused pass per time step that pulls the
// hexagonal (or square) neighbourhood out of the bit-planes, collides, applies
// the solid-cell rules and the body force, and stores the next state.  Every
// (row, plane) word is read exactly once per step: a thread owns one 16-byte
// quad (128 cells) of one output row, loads the planes it pulls from rows
// y-1, y, y+1, and gets the one word of x-halo each shifted plane needs from
// the neighbouring lane (ds_bpermute), falling back to a global load only at
// workgroup-row edges.  The rest are small helper kernels (codec, masks, hash
// fill, force draws, post-processing, reductions).
#include "lgca_dev.hpp"
#include "lgca_kernels.hpp"

namespace lgca {

static inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

__device__ __forceinline__ uint32_t valid_mask(const Geometry& g, uint32_t w)
{
    if (w + 1 < g.W) return 0xFFFFFFFFu;
    if (w + 1 == g.W) return g.nb == 32 ? 0xFFFFFFFFu : ((1u << g.nb) - 1u);
    return 0u;
}

__device__ __forceinline__ u32x4 ldq(const uint32_t* p) { return *(const u32x4*)p; }

__device__ __forceinline__ u32x4 valid_quad(const Geometry& g, uint32_t q)
{
    return u32x4{valid_mask(g, 4 * q), valid_mask(g, 4 * q + 1), valid_mask(g, 4 * q + 2),
                 valid_mask(g, 4 * q + 3)};
}

// ------------------------------------------------------------------ step

// Shifts one plane-row quad by one cell along x, cyclically over dim_x cells.
//   m1: out cell x = in cell x-1 (particles arriving from the west side)
//   p1: out cell x = in cell x+1
// `row` points at word 0 of the plane-row (for the halo fallback loads).
struct ShiftCtx {
    uint32_t lane, q, W, nb;
    bool has_left, has_right, has_last;  // has_last: quad holds word W-1
};

__device__ __forceinline__ u32x4 shift_quad(const u32x4 v, bool m1, const uint32_t* row,
                                            const ShiftCtx& c)
{
    const uint32_t send = m1 ? v.w : v.x;
    const int src = m1 ? (int)c.lane - 1 : (int)c.lane + 1;
    const uint32_t nbw = (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)send);
    u32x4 o;
    if (m1) {
        uint32_t L;
        if (c.has_left) L = nbw;
        else if (c.q == 0) L = row[c.W - 1] << (32u - c.nb);  // cell dim_x-1 wraps to x = 0
        else L = row[4 * c.q - 1];
        o.x = __builtin_amdgcn_alignbit(v.x, L, 31);
        o.y = __builtin_amdgcn_alignbit(v.y, v.x, 31);
        o.z = __builtin_amdgcn_alignbit(v.z, v.y, 31);
        o.w = __builtin_amdgcn_alignbit(v.w, v.z, 31);
    } else {
        uint32_t R = 0;
        if (c.has_right) R = nbw;
        else if (4 * c.q + 4 < c.W) R = row[4 * c.q + 4];
        o.x = __builtin_amdgcn_alignbit(v.y, v.x, 1);
        o.y = __builtin_amdgcn_alignbit(v.z, v.y, 1);
        o.z = __builtin_amdgcn_alignbit(v.w, v.z, 1);
        o.w = __builtin_amdgcn_alignbit(R, v.w, 1);
        if (c.has_last) {
            // cell 0 wraps to x = dim_x - 1, bit nb-1 of word W-1
            const uint32_t k = c.W - 1 - 4 * c.q;
            const uint32_t carry = (row[0] & 1u) << (c.nb - 1u);
            if (k == 0) o.x = (v.x >> 1) | carry;
            else if (k == 1) o.y = (v.y >> 1) | carry;
            else if (k == 2) o.z = (v.z >> 1) | carry;
            else o.w = (v.w >> 1) | carry;
        }
    }
    return o;
}

template <int KIND>
__global__ __launch_bounds__(256) void step_kernel(StepArgs a)
{
    constexpr int Q = num_dir_of_kind(KIND);
    const Geometry& g = a.g;
    const uint32_t qb = 1u << g.qb_log;
    const uint32_t lane_row = threadIdx.x & (qb - 1u);
    const uint32_t q = (blockIdx.x << g.qb_log) | lane_row;
    const uint32_t y = blockIdx.y * (256u >> g.qb_log) + (threadIdx.x >> g.qb_log);
    if (q >= g.Wq || y >= g.dim_y) return;

    ShiftCtx c;
    c.lane = threadIdx.x & 63u;
    c.q = q;
    c.W = g.W;
    c.nb = g.nb;
    c.has_left = lane_row != 0;
    c.has_right = lane_row != qb - 1u && q + 1 < g.Wq;
    c.has_last = 4 * q <= g.W - 1 && g.W - 1 < 4 * q + 4;

    const uint32_t ym = y ? y - 1 : g.dim_y - 1;
    const uint32_t yp = (y + 1 == g.dim_y) ? 0 : y + 1;
    const bool odd = (y & 1u) != 0;
    const uint32_t* src = a.src;
    auto rowp = [&](uint32_t row, int d) { return src + ((size_t)row * Q + d) * g.P; };

    // Pull stream (src/omp_lattice.cpp:141-180; sources per direction as
    // the torus neighbours of the reference's offset tables).
    u32x4 in[7];
    if (KIND == kHPP) {
        in[0] = shift_quad(ldq(rowp(y, 0) + 4 * q), true, rowp(y, 0), c);   // from (x-1, y)
        in[1] = ldq(rowp(ym, 1) + 4 * q);                                    // from (x, y-1)
        in[2] = shift_quad(ldq(rowp(y, 2) + 4 * q), false, rowp(y, 2), c);  // from (x+1, y)
        in[3] = ldq(rowp(yp, 3) + 4 * q);                                    // from (x, y+1)
    } else {
        // dir 0 from (x-1, y), dir 3 from (x+1, y), rest in place.
        // even rows: 1 from (x-1, y-1), 2 from (x, y-1), 4 from (x, y+1), 5 from (x-1, y+1)
        // odd rows:  1 from (x, y-1),   2 from (x+1, y-1), 4 from (x+1, y+1), 5 from (x, y+1)
        const u32x4 a0 = ldq(rowp(y, 0) + 4 * q), a3 = ldq(rowp(y, 3) + 4 * q);
        const u32x4 b1 = ldq(rowp(ym, 1) + 4 * q), b2 = ldq(rowp(ym, 2) + 4 * q);
        const u32x4 c4 = ldq(rowp(yp, 4) + 4 * q), c5 = ldq(rowp(yp, 5) + 4 * q);
        if (Q == 7) in[6] = ldq(rowp(y, 6) + 4 * q);
        in[0] = shift_quad(a0, true, rowp(y, 0), c);
        in[3] = shift_quad(a3, false, rowp(y, 3), c);
        const u32x4 sb = shift_quad(odd ? b2 : b1, !odd, rowp(ym, odd ? 2 : 1), c);
        const u32x4 sc = shift_quad(odd ? c4 : c5, !odd, rowp(yp, odd ? 4 : 5), c);
        in[1] = odd ? b1 : sb;
        in[2] = odd ? sb : b2;
        in[4] = odd ? sc : c4;
        in[5] = odd ? c5 : sc;
    }

    u32x4 s[7];
#pragma unroll
    for (int d = 0; d < Q; ++d) s[d] = in[d];

    // Chirality word of word w: philox({w, y, step >> 2, 0})[step & 3].
    u32x4 p = {0, 0, 0, 0};
    if (KIND != kHPP) {
        const uint32_t sel = a.step & 3u;
        uint32_t pw[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const U4 ch = philox4x32_10(4 * q + k, y, a.step >> 2, 0u, a.key0, a.key1);
            pw[k] = sel == 0 ? ch.x : (sel == 1 ? ch.y : (sel == 2 ? ch.z : ch.w));
        }
        p = u32x4{pw[0], pw[1], pw[2], pw[3]};
    }
    collide<KIND, u32x4>(s, p);

    // Solid cells (rows without non-fluid cells skip the mask load).
    u32x4 fluid = {~0u, ~0u, ~0u, ~0u};
    if (a.mask) {
        const int lo = a.row_range[y], hi = a.row_range[g.dim_y + y];
        if ((int)(4 * q) < hi && (int)(4 * q + 4) > lo) {
            const u32x4 ns = ldq(a.mask + ((size_t)y * 2 + 0) * g.P + 4 * q);
            const u32x4 sl = ldq(a.mask + ((size_t)y * 2 + 1) * g.P + 4 * q);
            u32x4 edge = {q == 0 ? 1u : 0u, 0u, 0u, 0u};
            if (c.has_last) {
                const uint32_t k = g.W - 1 - 4 * q, b = 1u << (g.nb - 1u);
                edge.x |= k == 0 ? b : 0u;
                edge.y |= k == 1 ? b : 0u;
                edge.z |= k == 2 ? b : 0u;
                edge.w |= k == 3 ? b : 0u;
            }
            apply_solids<KIND, u32x4>(s, in, ns, sl, edge, y == 0 || y + 1 == g.dim_y);
            fluid = ~(ns | sl);
        }
    }

    // Body force on this quad's FLUID cells (decisions drawn ahead of time).
    if (a.fdec && y >= a.fy0 && y < a.fy1 && 4 * q < a.fw0 + a.fwn && 4 * q + 4 > a.fw0) {
        const uint32_t* drow = a.fdec + (size_t)(y - a.fy0) * a.fwn;
        u32x4 fd;
        fd.x = (4 * q + 0 >= a.fw0 && 4 * q + 0 < a.fw0 + a.fwn) ? drow[4 * q + 0 - a.fw0] : 0u;
        fd.y = (4 * q + 1 >= a.fw0 && 4 * q + 1 < a.fw0 + a.fwn) ? drow[4 * q + 1 - a.fw0] : 0u;
        fd.z = (4 * q + 2 >= a.fw0 && 4 * q + 2 < a.fw0 + a.fwn) ? drow[4 * q + 2 - a.fw0] : 0u;
        fd.w = (4 * q + 3 >= a.fw0 && 4 * q + 3 < a.fw0 + a.fwn) ? drow[4 * q + 3 - a.fw0] : 0u;
        apply_force<KIND, u32x4>(s, fd & fluid, a.force_dir);
    }

    const u32x4 vm = valid_quad(g, q);
    uint32_t* dst = a.dst;
#pragma unroll
    for (int d = 0; d < Q; ++d) *(u32x4*)(dst + ((size_t)y * Q + d) * g.P + 4 * q) = s[d] & vm;
}

hipError_t launch_step(int kind, const StepArgs& a, hipStream_t s)
{
    const uint32_t qb = 1u << a.g.qb_log;
    const dim3 grid(div_up(a.g.Wq, qb), div_up(a.g.dim_y, 256u / qb));
    switch (kind) {
    case kHPP: hipLaunchKernelGGL(step_kernel<kHPP>, grid, dim3(256), 0, s, a); break;
    case kFHP1: hipLaunchKernelGGL(step_kernel<kFHP1>, grid, dim3(256), 0, s, a); break;
    default: hipLaunchKernelGGL(step_kernel<kFHP2>, grid, dim3(256), 0, s, a); break;
    }
    return hipGetLastError();
}

// ------------------------------------------------------------ force draws
// Decision words of steps [step0, step0 + nsteps) of the forced rows and
// words: bit b of word (step, y, w) forces cell 32w + b (see force_word).
// A thread owns one word of one row for one group of four steps that share
// the word's first Philox call; groups follow step >> 2 (steps mod 2^32).

__global__ __launch_bounds__(256) void force_draw_kernel(ForceDrawArgs a)
{
    // grid: x over the (row, word) pairs of the force region, y over groups
    __shared__ uint32_t Qt[32];
    force_table(Qt, a.q);
    const uint64_t plane = (uint64_t)(a.fy1 - a.fy0) * a.fwn;
    const uint64_t idx = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= plane) return;
    uint32_t yy, ww;
    if (plane < ((uint64_t)1 << 32)) {  // 32-bit division, the usual case
        yy = (uint32_t)idx / a.fwn;
        ww = (uint32_t)idx - yy * a.fwn;
    } else {
        yy = (uint32_t)(idx / a.fwn);
        ww = (uint32_t)(idx - (uint64_t)yy * a.fwn);
    }
    const uint32_t w = a.fw0 + ww;
    draw_group(a.dec + idx, plane, a.nsteps, a.step0, blockIdx.y, w, a.fy0 + yy, a.key0, a.key1, Qt,
               force_columns(w, a.fx0, a.fx1, a.dim_x));
}

hipError_t launch_force_draw(const ForceDrawArgs& a, hipStream_t s)
{
    const uint32_t groups = ((a.step0 & 3u) + a.nsteps + 3u) / 4u;
    const uint64_t n = (uint64_t)(a.fy1 - a.fy0) * a.fwn;
    if ((n + 255) / 256 >= ((uint64_t)1 << 31) || groups > 65535) return hipErrorInvalidValue;
    hipLaunchKernelGGL(force_draw_kernel, dim3(div_up(n, 256), groups), dim3(256), 0, s, a);
    return hipGetLastError();
}

// ------------------------------------------------------------------ codec

__global__ __launch_bounds__(256) void pack_kernel(const uint8_t* __restrict__ st,
                                                   uint32_t* __restrict__ planes, Geometry g,
                                                   uint32_t* err)
{
    const uint32_t w = blockIdx.x * 256 + threadIdx.x;
    const uint32_t y = blockIdx.y;
    if (w >= g.P) return;
    uint32_t acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t bad = 0;
    const uint8_t* row = st + (size_t)y * g.dim_x;
    for (uint32_t b = 0; b < 32; ++b) {
        const uint32_t x = 32 * w + b;
        if (x >= g.dim_x) break;
        const uint32_t v = row[x];
        bad |= v >> g.Q;
#pragma unroll
        for (int d = 0; d < 8; ++d) acc[d] |= ((v >> d) & 1u) << b;
    }
    for (uint32_t d = 0; d < g.Q; ++d) planes[((size_t)y * g.Q + d) * g.P + w] = acc[d];
    if (bad) atomicOr(err, 1u);
}

hipError_t launch_pack(const uint8_t* state, uint32_t* planes, const Geometry& g, uint32_t* err,
                       hipStream_t s)
{
    hipLaunchKernelGGL(pack_kernel, dim3(div_up(g.P, 256), g.dim_y), dim3(256), 0, s, state,
                       planes, g, err);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void unpack_kernel(const uint32_t* __restrict__ planes,
                                                     uint8_t* __restrict__ st, Geometry g)
{
    const uint32_t w = blockIdx.x * 256 + threadIdx.x;
    const uint32_t y = blockIdx.y;
    if (w >= g.W) return;
    uint32_t pl[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t d = 0; d < g.Q; ++d) pl[d] = planes[((size_t)y * g.Q + d) * g.P + w];
    uint8_t* row = st + (size_t)y * g.dim_x;
    for (uint32_t b = 0; b < 32; ++b) {
        const uint32_t x = 32 * w + b;
        if (x >= g.dim_x) break;
        uint32_t v = 0;
#pragma unroll
        for (int d = 0; d < 8; ++d) v |= ((pl[d] >> b) & 1u) << d;
        row[x] = (uint8_t)v;
    }
}

hipError_t launch_unpack(const uint32_t* planes, uint8_t* state, const Geometry& g, hipStream_t s)
{
    hipLaunchKernelGGL(unpack_kernel, dim3(div_up(g.W, 256), g.dim_y), dim3(256), 0, s, planes,
                       state, g);
    return hipGetLastError();
}

// ------------------------------------------------------------- cell types

__global__ __launch_bounds__(256) void types_kernel(const uint8_t* __restrict__ types,
                                                    uint32_t* __restrict__ mask,
                                                    int32_t* __restrict__ row_range, Geometry g,
                                                    uint32_t* err)
{
    const uint32_t w = blockIdx.x * 256 + threadIdx.x;
    const uint32_t y = blockIdx.y;
    if (w >= g.P) return;
    uint32_t ns = 0, sl = 0, bad = 0;
    const uint8_t* row = types + (size_t)y * g.dim_x;
    for (uint32_t b = 0; b < 32; ++b) {
        const uint32_t x = 32 * w + b;
        if (x >= g.dim_x) break;
        const uint32_t t = row[x];
        ns |= (uint32_t)(t == 1) << b;
        sl |= (uint32_t)(t == 2) << b;
        bad |= t > 2;
    }
    mask[((size_t)y * 2 + 0) * g.P + w] = ns;
    mask[((size_t)y * 2 + 1) * g.P + w] = sl;
    if (ns | sl) {
        atomicMin(row_range + y, (int32_t)w);
        atomicMax(row_range + g.dim_y + y, (int32_t)w + 1);
    }
    if (bad) atomicOr(err, 1u);
}

hipError_t launch_types(const uint8_t* types, uint32_t* mask, int32_t* row_range,
                        const Geometry& g, uint32_t* err, hipStream_t s)
{
    hipLaunchKernelGGL(types_kernel, dim3(div_up(g.P, 256), g.dim_y), dim3(256), 0, s, types,
                       mask, row_range, g, err);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void types_unpack_kernel(const uint32_t* __restrict__ mask,
                                                           uint8_t* __restrict__ types, Geometry g)
{
    const uint32_t w = blockIdx.x * 256 + threadIdx.x;
    const uint32_t y = blockIdx.y;
    if (w >= g.W) return;
    const uint32_t ns = mask ? mask[((size_t)y * 2 + 0) * g.P + w] : 0u;
    const uint32_t sl = mask ? mask[((size_t)y * 2 + 1) * g.P + w] : 0u;
    uint8_t* row = types + (size_t)y * g.dim_x;
    for (uint32_t b = 0; b < 32; ++b) {
        const uint32_t x = 32 * w + b;
        if (x >= g.dim_x) break;
        row[x] = (uint8_t)(((ns >> b) & 1u) | (((sl >> b) & 1u) << 1));
    }
}

hipError_t launch_types_unpack(const uint32_t* mask, uint8_t* types, const Geometry& g,
                               hipStream_t s)
{
    hipLaunchKernelGGL(types_unpack_kernel, dim3(div_up(g.W, 256), g.dim_y), dim3(256), 0, s,
                       mask, types, g);
    return hipGetLastError();
}

// ------------------------------------------------------------ hash fill
// init_random (src/lattice.cpp:198-217) with the draw of cell (x, y),
// direction d = philox({x>>2, y, d, 2})[x & 3] < thr, on FLUID cells only.

__global__ __launch_bounds__(256) void init_fill_kernel(uint32_t* __restrict__ planes,
                                                        const uint32_t* __restrict__ mask,
                                                        const uint8_t* __restrict__ region,
                                                        Geometry g, uint32_t key0, uint32_t key1,
                                                        uint64_t thr)
{
    const uint32_t w = blockIdx.x * 256 + threadIdx.x;
    const uint32_t y = blockIdx.y;
    if (w >= g.W) return;
    uint32_t fluid = valid_mask(g, w);
    if (mask)
        fluid &= ~(mask[((size_t)y * 2 + 0) * g.P + w] | mask[((size_t)y * 2 + 1) * g.P + w]);
    if (region) {
        uint32_t sel = 0;
        for (uint32_t b = 0; b < 32; ++b) {
            const uint32_t x = 32 * w + b;
            if (x >= g.dim_x) break;
            sel |= (uint32_t)(region[(size_t)y * g.dim_x + x] != 0) << b;
        }
        fluid &= sel;
    }
    for (uint32_t d = 0; d < g.Q; ++d) {
        const uint32_t word = fill_word(w, y, d, g.dim_x, key0, key1, thr);
        uint32_t* p = planes + ((size_t)y * g.Q + d) * g.P + w;
        *p = (word & fluid) | (*p & ~fluid);
    }
}

hipError_t launch_init_fill(uint32_t* planes, const uint32_t* mask, const uint8_t* region,
                            const Geometry& g, uint32_t key0, uint32_t key1, uint64_t thr,
                            hipStream_t s)
{
    hipLaunchKernelGGL(init_fill_kernel, dim3(div_up(g.W, 256), g.dim_y), dim3(256), 0, s, planes,
                       mask, region, g, key0, key1, thr);
    return hipGetLastError();
}

// ---------------------------------------------------------- post-process
// Lattice vectors (src/lgca_models.h:51-52, 246-247, 453-454): per cell the
// momentum is an exact multiple of (1/2, sin(pi/3)) for FHP and of 1 for HPP,
// so the reference's running FP32 sums over directions are reproduced by
// integer weights times one FP32 constant.
__device__ constexpr float kSin60 = 0.866025388240814208984375f;  // (float)sin(pi/3)

template <int KIND>
__device__ __forceinline__ void cell_values(const uint32_t* pl, uint32_t b, int& rho, float& mx,
                                            float& my)
{
    uint32_t s[7];
#pragma unroll
    for (int d = 0; d < num_dir_of_kind(KIND); ++d) s[d] = (pl[d] >> b) & 1u;
    if (KIND == kHPP) {
        rho = (int)(s[0] + s[1] + s[2] + s[3]);
        mx = (float)((int)s[0] - (int)s[2]);
        my = (float)((int)s[1] - (int)s[3]);
    } else {
        rho = (int)(s[0] + s[1] + s[2] + s[3] + s[4] + s[5]);
        if (KIND == kFHP2) rho += (int)s[6];
        const int kx = 2 * (int)s[0] + (int)s[1] - (int)s[2] - 2 * (int)s[3] - (int)s[4] + (int)s[5];
        const int ky = (int)s[1] + (int)s[2] - (int)s[4] - (int)s[5];
        mx = 0.5f * (float)kx;
        my = (float)ky * kSin60;
    }
}

template <int KIND>
__global__ __launch_bounds__(256) void cell_fields_kernel(const uint32_t* __restrict__ planes,
                                                          Geometry g, float* __restrict__ dens,
                                                          float* __restrict__ mom)
{
    constexpr int Q = num_dir_of_kind(KIND);
    const uint32_t w = blockIdx.x * 256 + threadIdx.x;
    const uint32_t y = blockIdx.y;
    if (w >= g.W) return;
    uint32_t pl[7];
#pragma unroll
    for (int d = 0; d < Q; ++d) pl[d] = planes[((size_t)y * Q + d) * g.P + w];
    for (uint32_t b = 0; b < 32; ++b) {
        const uint32_t x = 32 * w + b;
        if (x >= g.dim_x) break;
        int rho;
        float mx, my;
        cell_values<KIND>(pl, b, rho, mx, my);
        const size_t cell = (size_t)y * g.dim_x + x;
        if (dens) dens[cell] = (float)rho;
        if (mom) { mom[2 * cell] = mx; mom[2 * cell + 1] = my; }
    }
}

hipError_t launch_cell_fields(int kind, const uint32_t* planes, const Geometry& g, float* density,
                              float* momentum, hipStream_t s)
{
    const dim3 grid(div_up(g.W, 256), g.dim_y);
    switch (kind) {
    case kHPP: hipLaunchKernelGGL(cell_fields_kernel<kHPP>, grid, dim3(256), 0, s, planes, g, density, momentum); break;
    case kFHP1: hipLaunchKernelGGL(cell_fields_kernel<kFHP1>, grid, dim3(256), 0, s, planes, g, density, momentum); break;
    default: hipLaunchKernelGGL(cell_fields_kernel<kFHP2>, grid, dim3(256), 0, s, planes, g, density, momentum); break;
    }
    return hipGetLastError();
}

// Coarse graining (src/omp_lattice.cpp:397-454).  Coarse cell c has its corner
// at (2r * (c % cdx), 2r * (c / cdx)); it sums, y-outer and x-inner, the cells
// of columns x0..x0+r on rows y0..y0+2r that exist, plus -- only when
// dim_x == 2r -- the candidate x = 2r of each row, which wraps to column 0 of
// the next row.  Sums are sequential FP32 in that order; one divide by the
// count.  One thread per coarse cell keeps the order.
template <int KIND>
__global__ __launch_bounds__(64) void coarse_kernel(const uint32_t* __restrict__ planes,
                                                    Geometry g, uint32_t r,
                                                    float* __restrict__ md, float* __restrict__ mm,
                                                    float* __restrict__ vel)
{
    constexpr int Q = num_dir_of_kind(KIND);
    const uint32_t cdx = g.dim_x / (2 * r), cdy = g.dim_y / (2 * r);
    const uint32_t c = blockIdx.x * 64 + threadIdx.x;
    if (c >= cdx * cdy) return;
    const uint32_t x0 = (c % cdx) * 2 * r, y0 = (c / cdx) * 2 * r;
    const bool wrap_col = g.dim_x == 2 * r;
    float sd = 0.0f, sx = 0.0f, sy = 0.0f;
    int cnt = 0;
    uint32_t pl[7];
    auto visit = [&](uint32_t x, uint32_t y) {
        int rho;
        float mx, my;
        cell_values<KIND>(pl, x & 31u, rho, mx, my);
        sd += (float)rho;
        sx += mx;
        sy += my;
        ++cnt;
    };
    for (uint32_t yy = 0; yy <= 2 * r; ++yy) {
        const uint32_t y = y0 + yy;
        if (y >= g.dim_y) break;
        uint32_t cur_w = 0xFFFFFFFFu;
        for (uint32_t x = x0; x <= x0 + r; ++x) {
            if ((x >> 5) != cur_w) {
                cur_w = x >> 5;
#pragma unroll
                for (int d = 0; d < Q; ++d) pl[d] = planes[((size_t)y * Q + d) * g.P + cur_w];
            }
            visit(x, y);
        }
        if (wrap_col && y + 1 < g.dim_y) {
#pragma unroll
            for (int d = 0; d < Q; ++d) pl[d] = planes[((size_t)(y + 1) * Q + d) * g.P];
            visit(0, y + 1);
        }
    }
    const float fc = (float)cnt;
    const float rho = sd / fc, ux = sx / fc, uy = sy / fc;
    md[c] = rho;
    mm[2 * c] = ux;
    mm[2 * c + 1] = uy;
    if (vel) {
        vel[2 * c] = rho > 0.0f ? ux / rho : 0.0f;
        vel[2 * c + 1] = rho > 0.0f ? uy / rho : 0.0f;
    }
}

hipError_t launch_coarse(int kind, const uint32_t* planes, const Geometry& g, uint32_t r,
                         float* mean_density, float* mean_momentum, float* velocity, hipStream_t s)
{
    const uint32_t n = (g.dim_x / (2 * r)) * (g.dim_y / (2 * r));
    const dim3 grid(div_up(n, 64));
    switch (kind) {
    case kHPP: hipLaunchKernelGGL(coarse_kernel<kHPP>, grid, dim3(64), 0, s, planes, g, r, mean_density, mean_momentum, velocity); break;
    case kFHP1: hipLaunchKernelGGL(coarse_kernel<kFHP1>, grid, dim3(64), 0, s, planes, g, r, mean_density, mean_momentum, velocity); break;
    default: hipLaunchKernelGGL(coarse_kernel<kFHP2>, grid, dim3(64), 0, s, planes, g, r, mean_density, mean_momentum, velocity); break;
    }
    return hipGetLastError();
}

// ------------------------------------------------------------ reductions

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__global__ __launch_bounds__(256) void count_kernel(const uint32_t* __restrict__ words,
                                                    uint64_t n, unsigned long long* out)
{
    unsigned long long acc = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * 256)
        acc += (unsigned long long)__popc(words[i]);
    acc = wave_sum_u64(acc);
    if ((threadIdx.x & 63u) == 0) atomicAdd(out, acc);
}

hipError_t launch_count(const uint32_t* words, uint64_t nwords, unsigned long long* out,
                        hipStream_t s)
{
    uint32_t blocks = div_up(nwords, 256 * 8);
    if (blocks > 4096) blocks = 4096;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(count_kernel, dim3(blocks), dim3(256), 0, s, words, nwords, out);
    return hipGetLastError();
}

// get_mean_velocity (src/omp_lattice.cpp:508-557), summed exactly: with
// density rho in 1..7, 840 * (momentum / rho) is an integer multiple of the
// unit (1 for HPP x and y, 1/2 for FHP x, sin(pi/3) for FHP y), so three
// 64-bit integer sums (x, y, FLUID count) give an order-independent result.
template <int KIND>
__global__ __launch_bounds__(256) void mean_velocity_kernel(const uint32_t* __restrict__ planes,
                                                            const uint32_t* __restrict__ mask,
                                                            Geometry g, long long* acc)
{
    constexpr int Q = num_dir_of_kind(KIND);
    const uint32_t w = blockIdx.x * 256 + threadIdx.x;
    const uint32_t y = blockIdx.y;
    long long sx = 0, sy = 0, nf = 0;
    if (w < g.W) {
        uint32_t fluid = valid_mask(g, w);
        if (mask)
            fluid &= ~(mask[((size_t)y * 2 + 0) * g.P + w] | mask[((size_t)y * 2 + 1) * g.P + w]);
        uint32_t pl[7];
#pragma unroll
        for (int d = 0; d < Q; ++d) pl[d] = planes[((size_t)y * Q + d) * g.P + w];
        nf = __popc(fluid);
        for (uint32_t b = 0; b < 32; ++b) {
            if (!((fluid >> b) & 1u)) continue;
            uint32_t s[7];
#pragma unroll
            for (int d = 0; d < Q; ++d) s[d] = (pl[d] >> b) & 1u;
            int rho, kx, ky;
            if (KIND == kHPP) {
                rho = (int)(s[0] + s[1] + s[2] + s[3]);
                kx = 2 * ((int)s[0] - (int)s[2]);
                ky = (int)s[1] - (int)s[3];
            } else {
                rho = (int)(s[0] + s[1] + s[2] + s[3] + s[4] + s[5]) + (Q == 7 ? (int)s[6] : 0);
                kx = 2 * (int)s[0] + (int)s[1] - (int)s[2] - 2 * (int)s[3] - (int)s[4] + (int)s[5];
                ky = (int)s[1] + (int)s[2] - (int)s[4] - (int)s[5];
            }
            if (rho == 0) continue;
            const int f = 840 / rho;
            sx += (long long)(kx * f) / 2;  // kx is even for HPP; FHP keeps half units
            sy += (long long)(ky * f);
        }
    }
    unsigned long long ux = wave_sum_u64((unsigned long long)sx);
    unsigned long long uy = wave_sum_u64((unsigned long long)sy);
    unsigned long long un = wave_sum_u64((unsigned long long)nf);
    if ((threadIdx.x & 63u) == 0) {
        atomicAdd((unsigned long long*)acc + 0, ux);
        atomicAdd((unsigned long long*)acc + 1, uy);
        atomicAdd((unsigned long long*)acc + 2, un);
    }
}

hipError_t launch_mean_velocity(int kind, const uint32_t* planes, const uint32_t* mask,
                                const Geometry& g, long long* acc3, hipStream_t s)
{
    const dim3 grid(div_up(g.W, 256), g.dim_y);
    switch (kind) {
    case kHPP: hipLaunchKernelGGL(mean_velocity_kernel<kHPP>, grid, dim3(256), 0, s, planes, mask, g, acc3); break;
    case kFHP1: hipLaunchKernelGGL(mean_velocity_kernel<kFHP1>, grid, dim3(256), 0, s, planes, mask, g, acc3); break;
    default: hipLaunchKernelGGL(mean_velocity_kernel<kFHP2>, grid, dim3(256), 0, s, planes, mask, g, acc3); break;
    }
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void fill_u32_kernel(uint32_t* p, uint32_t v, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
        p[i] = v;
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t value, uint64_t n, hipStream_t s)
{
    uint32_t blocks = div_up(n, 256);
    if (blocks > 65535) blocks = 65535;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks), dim3(256), 0, s, p, value, n);
    return hipGetLastError();
}

}  // namespace lgca
