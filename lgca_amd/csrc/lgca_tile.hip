// lgca_tile.hip -- temporally blocked collide-and-propagate on gfx950.
//
// One pass over HBM advances the lattice K steps.  A wave owns 64 consecutive
// 32-cell words of a row (one word per lane) and marches down a tile of TH
// rows; for every input row it loads (7 plane words per lane) it runs the K
// time levels as a register pipeline: level s turns rows j-1, j, j+1 of level
// s-1 into row j of level s, with j lagging the input row by s.  Only level 0
// is read from HBM and only level K is written back, so the state crosses HBM
// once per K steps.
//
// x-halo: a shifted plane needs one bit of the neighbouring word, which the
// neighbouring lane holds (DPP wave_shr / wave_shl).  Lanes 0 and 63 have no
// outer neighbour, so the cells next to the wave edges go wrong, and the wrong
// zone grows by one cell per level.  Read along the lanes, the cells form the
// row in order (a ragged last word contributes only its nb valid cells), so
// after K levels the first and last K cells of the wave are suspect.  The
// host picks a halo of h lanes per side holding at least K cells (h = 1 for
// full words, 2 when an edge lane may hold a short last word, ceil(K / nb)
// when the row is a single word); waves advance by 64 - 2h words and store
// lanes h..63-h.
// y-halo: the tile reads K extra rows above and below and discards them.
//
// Chirality: philox({w, y, t >> 2, 0})[t & 3] serves a word for four steps, so
// with t0 a multiple of 4 one call per word and row covers levels 1-4 (5-8):
// the later levels take their word from a delay line.
#include "lgca_dev.hpp"
#include "lgca_kernels.hpp"

namespace lgca {

namespace {

__device__ __forceinline__ uint32_t from_left(uint32_t v)   // lane i gets lane i-1
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xF, 0xF, true);
}
__device__ __forceinline__ uint32_t from_right(uint32_t v)  // lane i gets lane i+1
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x130, 0xF, 0xF, true);
}

struct LaneCtx {
    uint32_t w;          // word index of this lane (torus, 0..W-1)
    uint32_t first_shl;  // 32 - nb on the lane holding word 0, else 0
    bool last;           // lane holds word W-1
    uint32_t nb;
    uint32_t vmask;      // valid bits of this lane's word
};

// out cell x = in cell x-1, cyclically over dim_x
__device__ __forceinline__ uint32_t shift_m1(uint32_t v, const LaneCtx& c)
{
    const uint32_t L = from_left(v) << c.first_shl;
    return __builtin_amdgcn_alignbit(v, L, 31);
}
// out cell x = in cell x+1, cyclically over dim_x
template <bool RAGGED>
__device__ __forceinline__ uint32_t shift_p1(uint32_t v, const LaneCtx& c)
{
    const uint32_t R = from_right(v);
    const uint32_t o = __builtin_amdgcn_alignbit(R, v, 1);
    if (!RAGGED) return o;
    return c.last ? ((v >> 1) | ((R & 1u) << (c.nb - 1u))) : o;
}

}  // namespace

template <int KIND, int K, bool RAGGED>
__global__ __launch_bounds__(256) void tile_kernel(TileArgs a)
{
    constexpr int Q = num_dir_of_kind(KIND);
    const Geometry& g = a.g;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t strip = blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint32_t stride = 64u - 2u * a.halo;
    const uint32_t nstrips = (g.W + stride - 1) / stride;
    if (strip >= nstrips) return;  // whole wave
    const uint32_t y0 = blockIdx.y * a.th;
    const uint32_t rows = min(a.th, g.dim_y - y0);

    const int wabs = (int)(strip * stride) - (int)a.halo + (int)lane;
    LaneCtx c;
    c.w = (uint32_t)((wabs + (int)g.W) % (int)g.W);
    c.nb = g.nb;
    c.first_shl = (c.w == 0) ? 32u - g.nb : 0u;
    c.last = c.w + 1 == g.W;
    c.vmask = (c.last && g.nb < 32) ? ((1u << g.nb) - 1u) : 0xFFFFFFFFu;
    const bool writer = lane >= a.halo && lane < 64u - a.halo && wabs < (int)g.W;

    // word range of this wave, for the uniform mask / force tests
    const uint32_t* src = a.src;
    uint32_t* dst = a.dst;
    const uint32_t n_iter = rows + 2 * K;
    auto wrap = [&](int r) -> uint32_t {
        r %= (int)g.dim_y;
        return (uint32_t)(r < 0 ? r + (int)g.dim_y : r);
    };

    // pipeline windows between level b and b+1: rows j-1 (planes 1, 2) and j
    uint32_t winA[K][7];  // row j   (all planes kept; the compiler drops unused ones)
    uint32_t winB[K][2];  // row j-1 planes 1 and 2
    // chirality delay line: dl[s][k] = word for level s, k iterations ahead
    uint32_t dl[K + 1][4];
#pragma unroll
    for (int s = 0; s <= K; ++s)
#pragma unroll
        for (int k = 0; k < 4; ++k) dl[s][k] = 0;
#pragma unroll
    for (int b = 0; b < K; ++b) {
#pragma unroll
        for (int d = 0; d < 7; ++d) winA[b][d] = 0;
        winB[b][0] = winB[b][1] = 0;
    }

    uint32_t nxt[7];
    {
        const uint32_t r0 = wrap((int)y0 - K);
        const uint32_t* p = src + (size_t)r0 * Q * g.P + c.w;
#pragma unroll
        for (int d = 0; d < Q; ++d) nxt[d] = p[(size_t)d * g.P];
    }

    for (uint32_t i = 0; i < n_iter; ++i) {
        uint32_t cur[7];
#pragma unroll
        for (int d = 0; d < Q; ++d) cur[d] = nxt[d];
        if (i + 1 < n_iter) {
            const uint32_t rn = wrap((int)(y0 + i + 1) - K);
            const uint32_t* p = src + (size_t)rn * Q * g.P + c.w;
#pragma unroll
            for (int d = 0; d < Q; ++d) nxt[d] = p[(size_t)d * g.P];
        }

        // chirality calls of this iteration, per level (only levels 1, 5 call)
        uint32_t callw[K + 1][4];
#pragma unroll
        for (int s = 0; s <= K; ++s)
#pragma unroll
            for (int k = 0; k < 4; ++k) callw[s][k] = 0;

#pragma unroll
        for (int s = 1; s <= K; ++s) {
            uint32_t out[7];
            const bool active = i >= (uint32_t)(2 * s);      // wave-uniform
            const bool have_prev = i >= (uint32_t)(2 * (s - 1));
            const uint32_t j = wrap((int)(y0 + i) - K - s);  // row of level s
            const uint32_t t = a.t0 + (uint32_t)(s - 1);     // its step
            if (active) {
                // pull stream (src/omp_lattice.cpp:141-180)
                uint32_t in[7];
                const uint32_t* A = winA[s - 1];
                const uint32_t* B = winB[s - 1];
                if (KIND == kHPP) {
                    in[0] = shift_m1(A[0], c);            // from (x-1, j)
                    in[1] = B[0];                         // from (x, j-1), plane 1 kept in B[0]
                    in[2] = shift_p1<RAGGED>(A[2], c);    // from (x+1, j)
                    in[3] = cur[3];                       // from (x, j+1)
                } else {
                    in[0] = shift_m1(A[0], c);
                    in[3] = shift_p1<RAGGED>(A[3], c);
                    if (Q == 7) in[6] = A[6];
                    if (j & 1u) {   // odd row
                        in[1] = B[0];
                        in[2] = shift_p1<RAGGED>(B[1], c);
                        in[4] = shift_p1<RAGGED>(cur[4], c);
                        in[5] = cur[5];
                    } else {
                        in[1] = shift_m1(B[0], c);
                        in[2] = B[1];
                        in[4] = cur[4];
                        in[5] = shift_m1(cur[5], c);
                    }
                }
                uint32_t st[7];
#pragma unroll
                for (int d = 0; d < Q; ++d) st[d] = in[d];
                uint32_t p = 0;
                if (KIND != kHPP) {
                    const int m = (s - 1) & 3;
                    if (K <= 2) {
                        // t0 even: level 1 uses word t0 & 3, level 2 the next one
                        if (s == 1) {
                            const U4 r = philox4x32_10(c.w, j, t >> 2, 0u, a.key0, a.key1);
                            callw[1][0] = r.x; callw[1][1] = r.y; callw[1][2] = r.z; callw[1][3] = r.w;
                            const uint32_t sel = t & 3u;
                            p = sel == 0 ? r.x : (sel == 1 ? r.y : (sel == 2 ? r.z : r.w));
                        } else {
                            p = dl[s][0];
                        }
                    } else if (m == 0) {
                        const U4 r = philox4x32_10(c.w, j, t >> 2, 0u, a.key0, a.key1);
                        callw[s][0] = r.x; callw[s][1] = r.y; callw[s][2] = r.z; callw[s][3] = r.w;
                        p = r.x;
                    } else {
                        p = dl[s][m - 1];
                    }
                }
                collide<KIND, uint32_t>(st, p);
                uint32_t fluid = 0xFFFFFFFFu;
                if (a.mask) {
                    const int lo = a.row_range[j], hi = a.row_range[g.dim_y + j];
                    const bool mine = (int)c.w >= lo && (int)c.w < hi;
                    if (__builtin_amdgcn_ballot_w64(mine) != 0) {
                        const uint32_t ns = a.mask[((size_t)j * 2 + 0) * g.P + c.w];
                        const uint32_t sl = a.mask[((size_t)j * 2 + 1) * g.P + c.w];
                        uint32_t edge = 0;
                        if (c.w == 0) edge |= 1u;
                        if (c.last) edge |= 1u << (g.nb - 1u);
                        apply_solids<KIND, uint32_t>(st, in, ns, sl, edge, j == 0 || j + 1 == g.dim_y);
                        fluid = ~(ns | sl);
                    }
                }
                if (a.fdec && j >= a.fy0 && j < a.fy1) {
                    const bool mine = c.w >= a.fw0 && c.w < a.fw0 + a.fwn;
                    if (__builtin_amdgcn_ballot_w64(mine) != 0) {
                        uint32_t f = 0;
                        if (mine)
                            f = a.fdec[(size_t)(s - 1) * a.fdec_stride + (size_t)(j - a.fy0) * a.fwn +
                                       (c.w - a.fw0)];
                        apply_force<KIND, uint32_t>(st, f & fluid, a.force_dir);
                    }
                }
#pragma unroll
                for (int d = 0; d < Q; ++d) out[d] = RAGGED ? (st[d] & c.vmask) : st[d];
            }
            // slide the window of boundary s-1 (its newest row is cur)
            if (have_prev) {
                if (KIND == kHPP) {
                    winB[s - 1][0] = winA[s - 1][1];
                } else {
                    winB[s - 1][0] = winA[s - 1][1];
                    winB[s - 1][1] = winA[s - 1][2];
                }
#pragma unroll
                for (int d = 0; d < Q; ++d) winA[s - 1][d] = cur[d];
            }
#pragma unroll
            for (int d = 0; d < Q; ++d) cur[d] = active ? out[d] : 0u;
            if (s == K && active) {
                uint32_t* q = dst + (size_t)j * Q * g.P + c.w;
                if (writer) {
#pragma unroll
                    for (int d = 0; d < Q; ++d) q[(size_t)d * g.P] = cur[d];
                }
            }
        }

        // Advance the chirality delay lines.  Level s (m = (s-1) & 3 > 0) uses
        // word m of the call that level s-m made for the same row m iterations
        // earlier: dl[s][0..m-1] is that queue, dl[s][m-1] the oldest entry.
        if (KIND != kHPP) {
#pragma unroll
            for (int s = 2; s <= K; ++s) {
                if (K <= 2) {
                    const uint32_t sel = (a.t0 + 1u) & 3u;
                    dl[s][0] = sel == 0 ? callw[1][0]
                             : (sel == 1 ? callw[1][1] : (sel == 2 ? callw[1][2] : callw[1][3]));
                } else {
                    const int m = (s - 1) & 3;
                    if (m == 0) continue;
#pragma unroll
                    for (int k = m - 1; k >= 1; --k) dl[s][k] = dl[s][k - 1];
                    dl[s][0] = callw[s - m][m];
                }
            }
        }
    }
}

uint32_t tile_halo(const Geometry& g, int k)
{
    if (g.W == 1) return (uint32_t)((k + (int)g.nb - 1) / (int)g.nb);
    return (int)g.nb < k ? 2u : 1u;
}

hipError_t launch_tile(int kind, int k, const TileArgs& a, hipStream_t s)
{
    if (a.halo < tile_halo(a.g, k) || a.halo > 16) return hipErrorInvalidValue;
    const uint32_t nstrips = (a.g.W + (64u - 2u * a.halo) - 1) / (64u - 2u * a.halo);
    const dim3 grid((nstrips + 3) / 4, (a.g.dim_y + a.th - 1) / a.th);
    const bool ragged = a.g.nb != 32;
#define LGCA_TILE_CASE(KD, KK)                                                                 \
    if (kind == KD && k == KK) {                                                             \
        if (ragged) hipLaunchKernelGGL((tile_kernel<KD, KK, true>), grid, dim3(256), 0, s, a);  \
        else hipLaunchKernelGGL((tile_kernel<KD, KK, false>), grid, dim3(256), 0, s, a);        \
        return hipGetLastError();                                                            \
    }
    LGCA_TILE_CASE(kHPP, 1) LGCA_TILE_CASE(kHPP, 2) LGCA_TILE_CASE(kHPP, 4) LGCA_TILE_CASE(kHPP, 8)
    LGCA_TILE_CASE(kFHP1, 1) LGCA_TILE_CASE(kFHP1, 2) LGCA_TILE_CASE(kFHP1, 4) LGCA_TILE_CASE(kFHP1, 8)
    LGCA_TILE_CASE(kFHP2, 1) LGCA_TILE_CASE(kFHP2, 2) LGCA_TILE_CASE(kFHP2, 4) LGCA_TILE_CASE(kFHP2, 8)
#undef LGCA_TILE_CASE
    return hipErrorInvalidValue;
}

}  // namespace lgca
