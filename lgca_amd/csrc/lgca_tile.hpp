// lgca_tile.hpp -- temporally blocked collide-and-propagate on gfx950.
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
// Chirality: philox({w, y, t >> 2, 0})[t & 3] serves a word for the four
// steps of a group.  A pass starting at step t0 is compiled for O = t0 & 3:
// the first level of each group in the pass makes the call for its row, and
// the group's later levels take their word from a delay line, so a pass makes
// one call per word and row per group it touches, whatever its first step.
//
// The kernels are templates; lgca_tile_inst.hip instantiates the single-
// lattice kernels of one (model, K) pair, or the batch kernel of one K, per
// translation unit, and lgca_tile.hip dispatches to them.
#pragma once

#include "lgca_dev.hpp"
#include "lgca_kernels.hpp"

#include <algorithm>
#include <type_traits>

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
    uint32_t wb;         // 4 * w, the lane's byte offset within a plane row
    uint32_t first_shl;  // 32 - nb on the lane holding word 0, else 0
    bool last;           // lane holds word W-1
    uint32_t nb;
    uint32_t vmask;      // valid bits of this lane's word
    uint32_t lsrc, rsrc; // segment mode: byte address of the lanes holding words w-1, w+1
};

// Memory goes through buffer resources: a wave-uniform base in scalar
// registers, a 32-bit per-lane offset and a scalar offset, so no load or
// store needs vector address arithmetic, and the hardware drops accesses past
// `bytes`.  (Word 3 selects raw 32-bit access, as for every gfx9 buffer.)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, 0x00020000);
}
// `soff` is wave-uniform at every call; readfirstlane says so to the
// compiler, which otherwise may reuse a vector copy of the value and wrap the
// access in a loop over its distinct values.
__device__ __forceinline__ uint32_t bload(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    return __builtin_amdgcn_raw_buffer_load_b32(r, voff, __builtin_amdgcn_readfirstlane(soff), 0);
}
__device__ __forceinline__ void bstore(uint32_t v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    __builtin_amdgcn_raw_buffer_store_b32(v, r, voff, __builtin_amdgcn_readfirstlane(soff), 0);
}
// The same at system coherence (cache policy sc0 sc1): the store is written
// through to memory and the load reads the coherent copy, so rows that one
// XCD writes and another reads in the same launch need no L2 maintenance.
constexpr int kCoherent = 17;
__device__ __forceinline__ uint32_t bload_coherent(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    return __builtin_amdgcn_raw_buffer_load_b32(r, voff, __builtin_amdgcn_readfirstlane(soff), kCoherent);
}
__device__ __forceinline__ void bstore_coherent(uint32_t v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    __builtin_amdgcn_raw_buffer_store_b32(v, r, voff, __builtin_amdgcn_readfirstlane(soff), kCoherent);
}

// Segment modes of the walk (template parameter SEG): strip mode, segment
// mode (S = 64 / W rows per wave), and segment mode with rows of exactly 64
// words, one per wave.
constexpr int kStrips = 0, kSegments = 1, kWholeRow = 2;

// Neighbouring words: the adjacent lane (DPP) when a wave holds consecutive
// words of one row; in segment mode, where the row wraps inside its
// segment, the lane holding word w -+ 1 of the same row (ds_bpermute), or,
// when the segment is the whole wave, the rotation of the wave by one lane
// (DPP wave_ror:1 / wave_rol:1).  A permute goes through the LDS crossbar and
// its result is waited for at every shift: walks of small lattices, one wave
// per SIMD, wait it out four times per level.
template <int SEG>
__device__ __forceinline__ uint32_t left_word(uint32_t v, const LaneCtx& c)
{
    if (SEG == kWholeRow) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x13C, 0xF, 0xF, true);
    return SEG ? (uint32_t)__builtin_amdgcn_ds_bpermute((int)c.lsrc, (int)v) : from_left(v);
}
template <int SEG>
__device__ __forceinline__ uint32_t right_word(uint32_t v, const LaneCtx& c)
{
    if (SEG == kWholeRow) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x134, 0xF, 0xF, true);
    return SEG ? (uint32_t)__builtin_amdgcn_ds_bpermute((int)c.rsrc, (int)v) : from_right(v);
}

// Output `sel` (wave-uniform) of a Philox call, by selects on the values
// (a chain over an array would become a dynamically indexed stack array).
__device__ __forceinline__ uint32_t pick4(const U4& r, uint32_t sel)
{
    const uint32_t lo = (sel & 1u) ? r.y : r.x;
    const uint32_t hi = (sel & 1u) ? r.w : r.z;
    return (sel & 2u) ? hi : lo;
}

// Chirality words of a pass whose first step t0 has t0 & 3 == O.  Level s
// (step t0 + s - 1) uses word (O + s - 1) & 3 of its group's call.  The call
// is made by the group's first level in the pass: level s itself when its
// word is 0 or s == 1, else level max(1, s - word), which worked on the same
// row that many iterations earlier.
// O = kRuntimeOffset (batch rounds with K <= 2, whose jobs in one launch may
// start on either even offset) schedules the calls as for O = 0 and picks the
// words at run time, which is exact because such a pass starts on an even
// step: its one or two levels share a group.
constexpr int kRuntimeOffset = -1;
__host__ __device__ constexpr int chir_word(int O, int s) { return ((O < 0 ? 0 : O) + s - 1) & 3; }
__host__ __device__ constexpr bool chir_caller(int O, int s) { return s == 1 || chir_word(O, s) == 0; }
__host__ __device__ constexpr int chir_src(int O, int s)
{
    return chir_caller(O, s) ? s : (s - chir_word(O, s) >= 1 ? s - chir_word(O, s) : 1);
}
__host__ __device__ constexpr int chir_delay(int O, int s) { return s - chir_src(O, s); }

// out cell x = in cell x-1, cyclically over dim_x
template <bool RAGGED, int SEG>
__device__ __forceinline__ uint32_t shift_m1(uint32_t v, const LaneCtx& c)
{
    uint32_t L = left_word<SEG>(v, c);
    if (RAGGED) L <<= c.first_shl;
    return __builtin_amdgcn_alignbit(v, L, 31);
}
// out cell x = in cell x+1, cyclically over dim_x
template <bool RAGGED, int SEG>
__device__ __forceinline__ uint32_t shift_p1(uint32_t v, const LaneCtx& c)
{
    const uint32_t R = right_word<SEG>(v, c);
    const uint32_t o = __builtin_amdgcn_alignbit(R, v, 1);
    if (!RAGGED) return o;
    return c.last ? ((v >> 1) | ((R & 1u) << (c.nb - 1u))) : o;
}

}  // namespace

// K <= 4 fits 128 VGPRs (four waves per SIMD); K = 8 needs more.
template <int K> struct TileOccupancy { static constexpr int waves = K <= 4 ? 4 : 2; };

// One wave's work.  Strip mode (SEG kStrips): x-strip `strip` of row tile
// `tile`, or part `part` of it (tile_part_rows).  Segment mode (SEG
// kSegments or kWholeRow, rows of W <= 64 words): the wave's lanes form
// S = 64 / W segments of W lanes, each holding a whole row; segment s walks
// row tile tile * S + s, and `strip` is unused.  O = a.t0 & 3 (0 when K == 1,
// whose one level picks its word at run time), or kRuntimeOffset.
//
// ALT (strip mode): odd row tiles walk up the lattice, even ones down.  Two
// vertically adjacent tiles then read the K halo rows they share at the same
// point of their walks (a down tile ends where the up tile below it ends, and
// starts where the up tile above it starts), so the second read of each halo
// row finds it in the L2 cache of their XCD instead of HBM.
// PIPE (tile_pipe_kernel, strip mode): the pass is one of several in a
// launch, and the wave's XCD owns the rows [pw.y0, pw.y1) of the lattice.
// Rows of the neighbouring bands (at most K above or below the band) come
// from their owners' exchange rows of the previous pass, after the first
// pass of the launch; the band's own first and last K output rows also go to
// its exchange rows for this pass.
struct PipeWalk {
    uint32_t y0, y1;        // the band's rows
    uint32_t pass;          // pass within the launch
    uint32_t band, prev, next, nband;
    uint32_t* exch;         // [pass & 1][band][0 first / 1 last][K rows][Q planes][P words]
};

template <int KIND, int K, int O, bool RAGGED, int SEG, bool ALT = false, bool PIPE = false>
__device__ __forceinline__ void tile_body(const TileArgs& a, uint32_t strip, uint32_t tile, uint32_t part = 0,
                                          const PipeWalk* pw = nullptr)
{
    constexpr int Q = num_dir_of_kind(KIND);
    const Geometry& g = a.g;
    const uint32_t lane = threadIdx.x & 63u;
    // tiles start on even rows (a.th is even), so the parity of the row a
    // level works on follows from the iteration index: see `odd_row` below
    uint32_t y0, rows;  // per lane in segment mode
    int wabs;
    bool writer;
    LaneCtx c;
    if (SEG) {
        const uint32_t S = 64u / g.W;
        const uint32_t ntiles = (g.dim_y + a.th - 1) / a.th;
        if (tile * S >= ntiles) return;  // whole wave
        const uint32_t seg = lane / g.W;  // S on the 64 - S * W idle lanes
        const uint32_t base = seg * g.W;
        const uint32_t mytile = tile * S + seg;
        const bool live = seg < S && mytile < ntiles;
        y0 = live ? mytile * a.th : 0u;
        rows = live ? min(a.th, g.dim_y - y0) : 0u;
        wabs = (int)(lane - base);  // < W, also on idle lanes
        writer = live;
        // (idle lanes may name a lane past 63; they never feed a live lane)
        c.lsrc = 4u * ((wabs == 0 ? base + g.W - 1u : lane - 1u) & 63u);
        c.rsrc = 4u * (((uint32_t)wabs + 1u == g.W ? base : lane + 1u) & 63u);
    } else {
        const uint32_t stride = 64u - 2u * a.halo;
        const uint32_t nstrips = (g.W + stride - 1) / stride;
        if (strip >= nstrips) return;  // whole wave
        y0 = tile * a.th;
        rows = min(a.th, g.dim_y - y0);
        if (part) {  // the part's rows of the tile (strip-mode orders only)
            uint32_t off, n;
            tile_part_rows(rows, part, off, n);
            y0 += off;
            rows = n;
        }
        wabs = (int)(strip * stride) - (int)a.halo + (int)lane;
        writer = lane >= a.halo && lane < 64u - a.halo && wabs < (int)g.W;
        c.lsrc = c.rsrc = 0;
    }
    c.w = (uint32_t)((wabs + (int)g.W) % (int)g.W);
    c.wb = 4u * c.w;
    c.nb = g.nb;
    c.first_shl = (c.w == 0) ? 32u - g.nb : 0u;
    c.last = c.w + 1 == g.W;
    c.vmask = (c.last && g.nb < 32) ? ((1u << g.nb) - 1u) : 0xFFFFFFFFu;
    uint32_t edge = 0;  // cells of this lane on the east / west columns
    if (c.w == 0) edge |= 1u;
    if (c.last) edge |= 1u << (g.nb - 1u);

    const uint32_t* src = a.src;
    uint32_t* dst = a.dst;
    const uint32_t n_iter = (SEG ? a.th : rows) + 2 * K;

    // Plane d (or mask plane b) of row r at this lane's word.  Strip mode: the
    // row is wave-uniform, so each access builds a descriptor over that row.
    // Segment mode: rows differ per lane, so one descriptor spans the buffer
    // (W <= 64 keeps it under 4 GiB) and the row goes into the lane offset.
    const uint32_t plane_b = g.P * 4u, row_b = (uint32_t)Q * plane_b;
    // PIPE: exchange row k of side `side` (0: the band's first K rows, 1: its
    // last K) of band b, for the pass of parity par
    auto exch_row = [&](uint32_t par, uint32_t b, uint32_t side, uint32_t k) -> uint32_t* {
        return pw->exch + (size_t)((((par * pw->nband + b) * 2u + side) * K + k) * Q) * g.P;
    };
    auto load_plane = [&](uint32_t r, int d) -> uint32_t {
        if (SEG) return bload(make_rsrc(src, g.dim_y * row_b), r * row_b + c.wb, (uint32_t)d * plane_b);
        if (PIPE && pw->pass > 0 && r - pw->y0 >= pw->y1 - pw->y0) {
            // a neighbouring band's row: t rows above the band, or u below it
            uint32_t t = pw->y0 + g.dim_y - r;
            if (t >= g.dim_y) t -= g.dim_y;
            uint32_t u = r + g.dim_y - pw->y1;
            if (u >= g.dim_y) u -= g.dim_y;
            const uint32_t par = (pw->pass - 1u) & 1u;
            const uint32_t* row = (t >= 1 && t <= (uint32_t)K) ? exch_row(par, pw->prev, 1, K - t)
                                                               : exch_row(par, pw->next, 0, u);
            return bload_coherent(make_rsrc(row, row_b), c.wb, (uint32_t)d * plane_b);
        }
        return bload(make_rsrc(src + (size_t)r * Q * g.P, row_b), c.wb, (uint32_t)d * plane_b);
    };
    // Stores and force-decision loads are issued by every lane: a lane that
    // must not access memory gets the offset kOutside, past any descriptor's
    // size, and the hardware drops its store or returns 0 for its load.  An
    // exec-masked branch around them would make the count of memory
    // operations issued after the next row's loads path-dependent, and the
    // waits for those loads would then wait for the stores as well.
    constexpr uint32_t kOutside = 0x80000000u;
    const uint32_t fplane_b = (a.fy1 - a.fy0) * a.fwn * 4u;  // one step's force decisions
    // the pass's K steps of decisions, one after another (a.fdec_stride is
    // one step's words): one descriptor for every level, the level's step in
    // the scalar offset
    const uint32_t fall_b = (uint32_t)K * fplane_b;
    auto store_plane = [&](uint32_t r, int d, uint32_t v, bool wr) {
        if (SEG) bstore(v, make_rsrc(dst, g.dim_y * row_b), wr ? r * row_b + c.wb : kOutside, (uint32_t)d * plane_b);
        else bstore(v, make_rsrc(dst + (size_t)r * Q * g.P, row_b), wr ? c.wb : kOutside, (uint32_t)d * plane_b);
        if (PIPE) {  // the band's first and last K rows also go to its exchange rows
            const uint32_t par = pw->pass & 1u, o0 = r - pw->y0, o1 = pw->y1 - 1u - r;
            if (o0 < (uint32_t)K)
                bstore_coherent(v, make_rsrc(exch_row(par, pw->band, 0, o0), row_b), wr ? c.wb : kOutside,
                                (uint32_t)d * plane_b);
            if (o1 < (uint32_t)K)
                bstore_coherent(v, make_rsrc(exch_row(par, pw->band, 1, K - 1u - o1), row_b),
                                wr ? c.wb : kOutside, (uint32_t)d * plane_b);
        }
    };
    auto load_mask = [&](uint32_t r, int b) -> uint32_t {
        if (SEG) return bload(make_rsrc(a.mask, g.dim_y * 2u * plane_b), r * 2u * plane_b + c.wb, (uint32_t)b * plane_b);
        return bload(make_rsrc(a.mask + (size_t)r * 2 * g.P, 2u * plane_b), c.wb, (uint32_t)b * plane_b);
    };
    const int dy = (int)g.dim_y;
    auto wrap = [&](int r) -> uint32_t {
        if (r < 0) {
            r += dy;
            if (r < 0) r = (r % dy + dy) % dy;
        } else if (r >= dy) {
            r -= dy;
            if (r >= dy) r %= dy;
        }
        return (uint32_t)r;
    };
    // Walk direction (wave-uniform) and the input row of iteration i: down
    // from y0 - K, or up from y0 + rows - 1 + K.
    const bool up = ALT && !SEG && (tile & 1u) != 0;
    auto iter_row = [&](uint32_t i) -> uint32_t {
        return up ? wrap((int)(y0 + rows - 1) + K - (int)i) : wrap((int)(y0 + i) - K);
    };

    // Wave geometry for the uniform mask / force tests: does the wave (words
    // [wlo, wlo + 64) modulo W) touch words [lo, hi)?
    const int wlo = (int)(strip * (64u - 2u * a.halo)) - (int)a.halo;  // strip mode only
    const int whi = wlo + 64;
    auto wave_hits = [&](int lo, int hi) -> bool {
        if (lo >= hi) return false;
        if (g.W < 64) return true;  // the wave wraps the whole row
        const int W = (int)g.W;
        auto hits = [&](int a0, int a1) { return a0 < hi && a1 > lo; };
        if (wlo < 0) return hits(0, whi) || hits(W + wlo, W);
        if (whi > W) return hits(wlo, W) || hits(0, whi - W);
        return hits(wlo, whi);
    };
    // Strip mode: which input rows of the walk hold solids in this wave's
    // words, 64 iterations at a time.  Lane l tests the row of iteration
    // i0 + l against the row's range of non-fluid words (one vector load per
    // bound), and the ballot gives bit l.  A scalar load per row instead would
    // be waited for before each row's branch, and at the start of a launch
    // every wave waits at once.
    auto solid_rows = [&](uint32_t i0) -> uint64_t {
        bool hit = false;
        const uint32_t i = i0 + lane;
        if (i < n_iter) {
            const uint32_t r = iter_row(i);
            hit = wave_hits(a.row_range[r], a.row_range[g.dim_y + r]);
        }
        return __builtin_amdgcn_ballot_w64(hit);
    };
    const bool force_cols = a.fdec && wave_hits((int)a.fw0, (int)(a.fw0 + a.fwn));
    const bool force_mine = a.fdec && c.w >= a.fw0 && c.w < a.fw0 + a.fwn;
    const uint32_t fvoff = force_mine ? (c.w - a.fw0) * 4u : kOutside;  // strip mode
    // K <= 2: the walk's first input row is loaded before the tests below, so
    // that their loads wait out one memory latency together with these (short
    // walks on small lattices feel it).  At K = 4 the registers this holds
    // across the tests push the walks into spilling, so the walk loads it.
    constexpr bool kEarlyRow = K <= 2;
    const uint32_t row_first = iter_row(0);
    uint32_t first[7];
#pragma unroll
    for (int d = 0; d < Q; ++d) first[d] = kEarlyRow ? load_plane(row_first, d) : 0u;

    // Decide once per tile whether the mask and force paths can be needed.
    // srows holds the solid rows of iterations 0..63 (strip mode).
    bool tile_solids = false, tile_force = false;
    uint64_t srows = 0;
    if (SEG) {
        // Segments walk consecutive row tiles, so rows [wy0 - K, wy0 + S th
        // + K) cover every row the live ones read; a wave with no solid row
        // there takes the lean walk.  Forcing: whenever the lattice has it
        // (rows differ per segment).
        tile_force = a.fdec != nullptr;
        if (a.mask) {
            const uint32_t S = 64u / g.W;
            const uint32_t wy0 = tile * S * a.th, span = S * a.th + 2u * (uint32_t)K;
            for (uint32_t i0 = 0; !tile_solids && i0 < span; i0 += 64) {
                const uint32_t i = i0 + lane;
                bool hit = false;
                if (i < span) {
                    const uint32_t r = wrap((int)(wy0 + i) - K);
                    hit = a.row_range[r] < a.row_range[g.dim_y + r];
                }
                tile_solids = __builtin_amdgcn_ballot_w64(hit) != 0;
            }
        }
    } else {
        if (a.mask) {
            srows = solid_rows(0);
            tile_solids = srows != 0;
            for (uint32_t i0 = 64; !tile_solids && i0 < n_iter; i0 += 64) tile_solids = solid_rows(i0) != 0;
        }
        for (uint32_t i0 = 0; force_cols && !tile_force && i0 < n_iter; i0 += 64) {
            const uint32_t i = i0 + lane;
            bool f = false;
            if (i < n_iter) {
                const uint32_t r = iter_row(i);
                f = r >= a.fy0 && r < a.fy1;
            }
            tile_force = __builtin_amdgcn_ballot_w64(f) != 0;
        }
    }

    // MODE 0: all fluid, unforced; 1: forced along x, no solids; 2: the
    // general walk
    auto walk = [&](auto mode_tag, auto up_tag) {
    constexpr int MODE = decltype(mode_tag)::value;
    constexpr bool UP = decltype(up_tag)::value;
    const bool t_solids = MODE == 2 && tile_solids;
    const bool t_force = MODE == 1 || (MODE == 2 && tile_force);
    // pipeline windows between level b and b+1: rows j and the one before it
    // in walk order (j-1 walking down, j+1 walking up), of which only the
    // planes pulled across rows are kept (1, 2 down and 4, 5 up; HPP 1 or 3)
    uint32_t winA[K][7];
    uint32_t winB[K][2];
    uint32_t dl[K + 1][4];            // chirality delay lines
    uint32_t mns[K + 1], msl[K + 1];  // mask delay line: [s] = masks of row i - s
    uint32_t fnext[K + 1];            // force decisions for the next iteration's levels
    uint32_t callw[K + 1][4];         // chirality calls of the current iteration
#pragma unroll
    for (int s = 0; s <= K; ++s) {
        mns[s] = msl[s] = fnext[s] = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) dl[s][k] = callw[s][k] = 0;
    }
#pragma unroll
    for (int b = 0; b < K; ++b) {
#pragma unroll
        for (int d = 0; d < 7; ++d) winA[b][d] = 0;
        winB[b][0] = winB[b][1] = 0;
    }

    // rows: row_i is the input row of the current iteration; lrow[s] the row
    // level s works on (row_i - s on the torus)
    uint32_t row_i = row_first;
    uint32_t lrow[K + 1];
#pragma unroll
    for (int s = 1; s <= K; ++s) lrow[s] = wrap(UP ? (int)row_i + s : (int)row_i - s);

    uint32_t nxt[7];
    uint32_t nns = 0, nsl = 0;  // masks of the prefetched row
    {
        if (t_solids && (SEG || (srows & 1u))) {
            nns = load_mask(row_i, 0);
            nsl = load_mask(row_i, 1);
        }
#pragma unroll
        for (int d = 0; d < Q; ++d) nxt[d] = kEarlyRow ? first[d] : load_plane(row_i, d);
    }

    // One step of the row walk.  STEADY iterations (i >= 2K) have every level
    // active.  PAR is i & 1: with even tile starts, even tile heights and an
    // even dim_y (FHP), the row of level s has parity (K + PAR + s) & 1 walking
    // down and (1 + K + PAR + s) & 1 walking up.
    auto iterate = [&](auto steady_tag, auto par_tag, uint32_t i) {
        constexpr bool STEADY = decltype(steady_tag)::value;
        constexpr int PAR = decltype(par_tag)::value;
        const uint32_t row_next = UP ? (row_i == 0 ? g.dim_y - 1u : row_i - 1u)
                                     : (row_i + 1 == g.dim_y ? 0u : row_i + 1);
        uint32_t cur[7];
#pragma unroll
        for (int d = 0; d < Q; ++d) cur[d] = nxt[d];
        mns[0] = nns;
        msl[0] = nsl;
        uint32_t fcur[K + 1];
#pragma unroll
        for (int s = 1; s <= K; ++s) fcur[s] = fnext[s];

        // side loads for the next iteration, then the next row's planes
        if (i + 1 < n_iter) {
            const uint32_t rn = row_next;
            nns = nsl = 0;
            // the row of iteration i + 1; a new group of 64 is tested on entry
            if (t_solids && !SEG && ((i + 1u) & 63u) == 0) srows = solid_rows(i + 1u);
            if (t_solids && (SEG || ((srows >> ((i + 1u) & 63u)) & 1u))) {
                nns = load_mask(rn, 0);
                nsl = load_mask(rn, 1);
            }
            if (t_force) {
                // one buffer load per level on every lane, reading 0 where
                // there is no decision
#pragma unroll
                for (int s = 1; s <= K; ++s) {
                    // next iteration, level s works on the row level s - 1 works on now
                    const uint32_t jn = s == 1 ? row_i : lrow[s - 1];
                    const bool lev = STEADY || i + 1 >= (uint32_t)(2 * s);
                    const __amdgpu_buffer_rsrc_t fr = make_rsrc(a.fdec, fall_b);
                    const uint32_t fstep = (uint32_t)(s - 1) * fplane_b;
                    if (SEG) {  // rows differ per lane
                        const bool ok = lev && jn >= a.fy0 && jn < a.fy1 && force_mine;
                        const uint32_t voff = ok ? ((jn - a.fy0) * a.fwn + (c.w - a.fw0)) * 4u : kOutside;
                        fnext[s] = bload(fr, voff, fstep);
                    } else {    // the row is uniform: its offset is scalar; an
                                // offset past the buffer reads 0 for rows outside
                        const bool ok = lev && jn >= a.fy0 && jn < a.fy1;
                        fnext[s] = bload(fr, ok ? fvoff : kOutside, fstep + (ok ? (jn - a.fy0) * a.fwn * 4u : 0u));
                    }
                }
            }
#pragma unroll
            for (int d = 0; d < Q; ++d) nxt[d] = load_plane(rn, d);
        }

#pragma unroll
        for (int s = 1; s <= K; ++s) {
            uint32_t out[7];
            const bool active = STEADY || i >= (uint32_t)(2 * s);          // wave-uniform
            const bool have_prev = STEADY || i >= (uint32_t)(2 * (s - 1));
            const uint32_t j = lrow[s];                      // row of level s
            const uint32_t t = a.t0 + (uint32_t)(s - 1);     // its step
            const bool odd_row = KIND != kHPP && (((K + PAR + s + (UP ? 1 : 0)) & 1) != 0);
            if (active) {
                // pull stream (src/omp_lattice.cpp:141-180).  Planes 1 and 2
                // (HPP: 1) come from row j-1, planes 4 and 5 (HPP: 3) from
                // row j+1: the kept older row and the newest row, in the
                // order of the walk's direction.
                uint32_t in[7];
                const uint32_t* A = winA[s - 1];
                const uint32_t* B = winB[s - 1];
                if (KIND == kHPP) {
                    in[0] = shift_m1<RAGGED, SEG>(A[0], c);    // from (x-1, j)
                    in[1] = UP ? cur[1] : B[0];                // from (x, j-1)
                    in[2] = shift_p1<RAGGED, SEG>(A[2], c);    // from (x+1, j)
                    in[3] = UP ? B[0] : cur[3];                // from (x, j+1)
                } else {
                    const uint32_t r1 = UP ? cur[1] : B[0], r2 = UP ? cur[2] : B[1];  // row j-1
                    const uint32_t r4 = UP ? B[0] : cur[4], r5 = UP ? B[1] : cur[5];  // row j+1
                    in[0] = shift_m1<RAGGED, SEG>(A[0], c);
                    in[3] = shift_p1<RAGGED, SEG>(A[3], c);
                    if (Q == 7) in[6] = A[6];
                    if (odd_row) {
                        in[1] = r1;
                        in[2] = shift_p1<RAGGED, SEG>(r2, c);
                        in[4] = shift_p1<RAGGED, SEG>(r4, c);
                        in[5] = r5;
                    } else {
                        in[1] = shift_m1<RAGGED, SEG>(r1, c);
                        in[2] = r2;
                        in[4] = r4;
                        in[5] = shift_m1<RAGGED, SEG>(r5, c);
                    }
                }
                uint32_t st[7];
#pragma unroll
                for (int d = 0; d < Q; ++d) st[d] = in[d];
                uint32_t p = 0;
                if (KIND != kHPP) {
                    if (chir_caller(O, s)) {
                        const U4 r = philox4x32_10(c.w, j, t >> 2, 0u, a.key0, a.key1);
                        callw[s][0] = r.x; callw[s][1] = r.y; callw[s][2] = r.z; callw[s][3] = r.w;
                        p = pick4(r, (K == 1 || O < 0) ? (t & 3u) : (uint32_t)chir_word(O, s));
                    } else {
                        p = dl[s][chir_delay(O, s) - 1];
                    }
                }
                collide<KIND, uint32_t>(st, p);
                if (t_solids) {
                    const uint32_t ns = mns[s], sl = msl[s];
                    if (__builtin_amdgcn_ballot_w64((ns | sl) != 0) != 0) {
                        // rows of NO_SLIP cells only (walls, obstacles) take
                        // the one-select rule
                        if (__builtin_amdgcn_ballot_w64(sl != 0) == 0)
                            apply_no_slip<KIND, uint32_t>(st, in, ns);
                        else
                            apply_solids<KIND, uint32_t>(st, in, ns, sl, edge, j == 0 || j + 1 == g.dim_y);
                    }
                    // fcur is zero outside the forced rows
                    if (t_force && (SEG || (j >= a.fy0 && j < a.fy1)))
                        apply_force<KIND, uint32_t>(st, fcur[s] & ~(ns | sl), a.force_dir);
                } else if (MODE == 1) {
                    // fcur is zero outside the forced rows and columns
                    apply_force<KIND, uint32_t>(st, fcur[s], 1);
                } else if (t_force && (SEG || (j >= a.fy0 && j < a.fy1))) {
                    apply_force<KIND, uint32_t>(st, fcur[s], a.force_dir);
                }
#pragma unroll
                for (int d = 0; d < Q; ++d) out[d] = RAGGED ? (st[d] & c.vmask) : st[d];
            }
            // slide the window of boundary s-1 (its newest row is cur)
            if (have_prev) {
                if (KIND == kHPP) {
                    winB[s - 1][0] = winA[s - 1][UP ? 3 : 1];
                } else {
                    winB[s - 1][0] = winA[s - 1][UP ? 4 : 1];
                    winB[s - 1][1] = winA[s - 1][UP ? 5 : 2];
                }
#pragma unroll
                for (int d = 0; d < Q; ++d) winA[s - 1][d] = cur[d];
            }
            if (active) {
#pragma unroll
                for (int d = 0; d < Q; ++d) cur[d] = out[d];
                // segment mode: rows past a segment's own tile belong to the next
                if (s == K) {
                    const bool wr = writer && (!SEG || i - (uint32_t)(2 * K) < rows);
#pragma unroll
                    for (int d = 0; d < Q; ++d) store_plane(j, d, cur[d], wr);
                }
            }
        }

        // mask delay line: level s next iteration handles the row level s-1 had
        if (t_solids) {
#pragma unroll
            for (int s = K; s >= 1; --s) {
                mns[s] = mns[s - 1];
                msl[s] = msl[s - 1];
            }
        }

        // Advance the chirality delay lines.  A level s that makes no call
        // uses word chir_word of the call level chir_src made for the same row
        // D = chir_delay iterations earlier: dl[s][0..D-1] is that queue,
        // dl[s][D-1] the oldest entry.
        if (KIND != kHPP) {
#pragma unroll
            for (int s = 2; s <= K; ++s) {
                const int D = chir_delay(O, s);
                if (D == 0) continue;
#pragma unroll
                for (int k = D - 1; k >= 1; --k) dl[s][k] = dl[s][k - 1];
                const int c0 = chir_src(O, s);
                if (O < 0)
                    dl[s][0] = pick4(U4{callw[c0][0], callw[c0][1], callw[c0][2], callw[c0][3]},
                                     (a.t0 + (uint32_t)(s - 1)) & 3u);
                else
                    dl[s][0] = callw[c0][chir_word(O, s)];
            }
        }
        // advance the rows
#pragma unroll
        for (int s = K; s >= 2; --s) lrow[s] = lrow[s - 1];
        lrow[1] = row_i;
        row_i = row_next;
    };

    using F = std::false_type;
    using T = std::true_type;
    using P0 = std::integral_constant<int, 0>;
    using P1 = std::integral_constant<int, 1>;
    uint32_t i = 0;
    for (; i < (uint32_t)(2 * K); i += 2) {   // pipeline fill (n_iter > 2K)
        iterate(F{}, P0{}, i);
        iterate(F{}, P1{}, i + 1);
    }
    for (; i + 1 < n_iter; i += 2) {
        iterate(T{}, P0{}, i);
        iterate(T{}, P1{}, i + 1);
    }
    if (i < n_iter) iterate(T{}, P0{}, i);
    };  // walk

    // Most tiles of a large lattice are all fluid and unforced: they take a
    // copy of the walk with that code left out (no tests, no branches).  Tiles
    // forced along x that hold no solids take a copy with the force applied
    // on every level and none of the solid tests (in the general walk they
    // ran 1.75 times as long as a plain walk).  Rows with a short last word,
    // and the multi-pass kernel, keep fewer copies: more would push those
    // kernels past the register budget.
    using M0 = std::integral_constant<int, 0>;
    using M1 = std::integral_constant<int, 1>;
    using M2 = std::integral_constant<int, 2>;
    const int mode = (RAGGED || tile_solids) ? 2 : !tile_force ? 0 : (a.force_dir == 1 && !PIPE) ? 1 : 2;
    auto run = [&](auto up_tag) {
        if (mode == 2) walk(M2{}, up_tag);
        else if (mode == 1) walk(M1{}, up_tag);
        else walk(M0{}, up_tag);
    };
    if constexpr (ALT && !SEG) {
        if (up) {
            run(std::true_type{});
            return;
        }
    }
    run(std::false_type{});
}

// One wave per workgroup.  A CU takes a new workgroup only when every SIMD the
// workgroup would use has room, so with four waves per workgroup the waves
// that end early (strips past the row end, or the light strips beside a
// forced one) kept slots of a CU idle until the longest wave of their
// workgroup ended: on the 32768 x 8192 Karman lattice half of the CUs ran
// with 4 of their 16 wave slots busy for the first quarter of each launch.
//
// Strip-mode wave order (TileOrder): waves over solids or forced words run
// longer, and dispatched in plain order the last of them trail the rest of
// the grid; so they are dispatched first (tile_wave_order; without an order,
// the forced strips of every row tile) and the light ones fill in behind
// them.
template <int KIND, int K, int O, bool RAGGED, int SEG>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(TileOccupancy<K>::waves, 8)))
void tile_kernel(TileArgs a, TileOrder o)
{
    // segment mode: grid (tile groups); strip mode: grid (nstrips * ntiles)
    if (SEG) {
        tile_body<KIND, K, O, RAGGED, SEG>(a, 0, blockIdx.x);
        return;
    }
    // Workgroups go to the 8 XCDs round-robin.  Each XCD takes a contiguous
    // run of the order instead, so the tiles it holds are vertical neighbours
    // whose halo rows meet in its L2.  The heavy and the light waves are
    // mapped apart, so that every XCD gets its share of both (mapped as one
    // range, the forced waves all land on the first XCDs, which then finish
    // last).
    uint32_t strip, tile, part = 0;
    if (o.order) {
        const uint32_t e = o.order[blockIdx.x];
        const uint32_t it = e >> kPartBits;
        part = e & ((1u << kPartBits) - 1u);
        tile = it / o.nstrips;
        strip = it - tile * o.nstrips;
    } else {
        const uint32_t H = o.nheavy * o.ntiles;
        const bool in_heavy = blockIdx.x < H;
        const uint32_t b = in_heavy ? blockIdx.x : blockIdx.x - H, n = in_heavy ? H : gridDim.x - H;
        const uint32_t xcd = b & 7u, q = n >> 3, r = n & 7u;
        const uint32_t L = (in_heavy ? 0u : H) + xcd * q + min(xcd, r) + (b >> 3);
        if (L < H) {
            tile = L / o.nheavy;
            strip = o.heavy[L - tile * o.nheavy];
        } else {
            const uint32_t nl = o.nstrips - o.nheavy, l = L - H;
            tile = l / nl;
            strip = l - tile * nl;  // the strip-th light strip: step over the heavy ones
            for (uint32_t h = 0; h < o.nheavy; ++h)
                if (o.heavy[h] <= strip) ++strip;
        }
    }
    tile_body<KIND, K, O, RAGGED, kStrips, true>(a, strip, tile, part);
}

// Several passes in one launch (strip mode, K = 4 or 8, th >= K), so that the
// end of one pass overlaps the start of the next instead of leaving the GPU
// to drain between launches.  Each XCD owns a band of whole row tiles
// (p.band_y0 / band_y1, its own L2 holding the band's rows) and hands out the
// band's work items pass after pass from its own ticket counter.  A wave
// takes a ticket, waits until every strip of every tile its walk reads (rows
// [y0 - K, y0 + rows + K)) has finished the previous pass, and runs the pass.
// Tickets go out in pass order, so every item a wave waits for is held by a
// running wave, which waits only for earlier passes: the waits always end.
// They are bounded anyway (kPipeSpinLimit polls): a wave that gives up sets
// p.error and leaves, and the host reports the error.
//
// Memory: an item publishes that it has finished pass j by storing j + 1 into
// its own word of p.done with a device-coherent store, after all its other
// stores have completed, and waiters poll those words with device-coherent
// loads, 64 strips per load: flags that one wave stores and others poll,
// both bypassing the non-coherent caches, as the look-back scans of rocPRIM
// do on these GPUs (a read-modify-write count leaves no such guarantee for
// the polls of other XCDs).  Rows written by other CUs of the same XCD meet
// in its L2, so after its waits a wave drops its CU's L1 copies
// (buffer_inv sc0) and reads them as usual.  Rows of the neighbouring bands,
// written on other XCDs, go through the exchange rows with system-coherent
// stores and loads (bload_coherent); the first pass of a launch reads them
// from the state buffer, which the previous launch wrote.  The state buffers
// are written and read by one XCD each during the launch, and each band's
// ticket counter by the waves of one XCD.
constexpr uint32_t kPipeSpinLimit = 1u << 20;
constexpr uint32_t kXccIdReg = (31u << 11) | 20u;  // hwreg(HW_REG_XCC_ID), all 32 bits

// Waits until every strip of tile u has finished `need` passes; lane l polls
// strip s0 + l.
__device__ __forceinline__ bool pipe_wait(const PipeArgs& p, uint32_t u, uint32_t need, uint32_t lane)
{
    for (uint32_t s0 = 0; s0 < p.nstrips; s0 += 64) {
        const bool mine = s0 + lane < p.nstrips;
        const uint32_t* w = p.done + (size_t)u * p.nstrips + s0 + (mine ? lane : 0u);
        for (uint32_t spins = 0;; ++spins) {
            const uint32_t v = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (__builtin_amdgcn_ballot_w64(mine && v < need) == 0) break;
            if (spins >= kPipeSpinLimit) {
                __hip_atomic_store(p.error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return false;
            }
            __builtin_amdgcn_s_sleep(8);
        }
    }
    return true;
}

template <int KIND, int K, int O>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(TileOccupancy<K>::waves, 8)))
void tile_pipe_kernel(TileArgs a, PipeArgs p)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t band = (__builtin_amdgcn_s_getreg(kXccIdReg) & 15u) % p.nband;
    PipeWalk pw;
    pw.y0 = p.band_y0[band];
    pw.y1 = p.band_y1[band];
    pw.band = band;
    pw.prev = band == 0 ? p.nband - 1u : band - 1u;
    pw.next = band + 1u == p.nband ? 0u : band + 1u;
    pw.nband = p.nband;
    pw.exch = p.exch;
    const uint32_t nb = p.band_items[band], dy = a.g.dim_y;
    for (;;) {
        uint32_t tk = 0;
        if (lane == 0) tk = __hip_atomic_fetch_add(p.queue + band, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        tk = __builtin_amdgcn_readfirstlane(tk);
        if (tk >= nb * p.npass) return;
        const uint32_t pass = tk / nb, it = p.order[p.band_first[band] + (tk - pass * nb)];
        const uint32_t tile = it / p.nstrips, strip = it - tile * p.nstrips;
        if (pass > 0) {
            // the tiles holding rows [y0 - K, y0 + rows + K), cyclically
            const uint32_t y0 = tile * a.th, rows = min(a.th, dy - y0);
            const uint32_t lo = y0 >= (uint32_t)K ? y0 - K : y0 + dy - K;
            uint32_t hi = y0 + rows + K - 1u;
            if (hi >= dy) hi -= dy;
            const uint32_t first = lo / a.th, last = hi / a.th;
            for (uint32_t u = first, n = 0;; u = u + 1u == p.ntiles ? 0u : u + 1u, ++n) {
                if (!pipe_wait(p, u, pass, lane)) return;
                if (u == last || n >= 4) break;
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            asm volatile("buffer_inv sc0" ::: "memory");
        }
        TileArgs b = a;
        b.src = p.buf[pass & 1u];
        b.dst = p.buf[(pass + 1u) & 1u];
        b.t0 = a.t0 + pass * (uint32_t)K;
        if (a.fdec) b.fdec = a.fdec + (size_t)pass * p.fdec_pass;
        pw.pass = pass;
        tile_body<KIND, K, O, false, kStrips, true, true>(b, strip, tile, 0, &pw);
        // the pass's stores have all completed before the item says so
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        if (lane == 0)
            __hip_atomic_store(p.done + (size_t)tile * p.nstrips + strip, pass + 1u, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Batched form: every wave takes one work item (job, strip, tile, shape) and
// runs that job's lattice.  A job's step, current buffer and force-window
// origin live in device memory (the job's last wave of a round moves them on),
// so queued rounds need no host-to-device traffic and no launch between them.
// One launch covers all items of one (K, ragged) class, whatever their jobs'
// models: each wave takes its job's model body (the job's Q, wave-uniform), so
// a round of a mixed batch is one launch rather than one per model.  The top
// bit of an item's shape picks segment mode for its job, so strip and segment
// jobs share the launch too.
template <int K, bool RAGGED>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(TileOccupancy<K>::waves, 8)))
void tile_batch_kernel(const BatchJob* __restrict__ jobs, BatchState* state,
                       const WorkItem* __restrict__ items, uint32_t n_items, uint32_t n_heavy)
{
    // One wave (workgroup) per item.  Workgroups go to the 8 XCDs
    // round-robin; the leading (heavy) items and the rest each map to
    // contiguous runs per XCD, as in tile_kernel.
    const bool in_heavy = blockIdx.x < n_heavy;
    const uint32_t b = in_heavy ? blockIdx.x : blockIdx.x - n_heavy;
    const uint32_t n = in_heavy ? n_heavy : gridDim.x - n_heavy;
    const uint32_t xcd = b & 7u, q = n >> 3, r = n & 7u;
    // wave-uniform by construction; readfirstlane lets the compiler keep the
    // job record in scalar registers
    const uint32_t wi = __builtin_amdgcn_readfirstlane((in_heavy ? 0u : n_heavy) + xcd * q + min(xcd, r) + (b >> 3));
    if (wi >= n_items) return;
    const WorkItem it = items[wi];
    const BatchState bs = state[it.job];
    TileArgs a = jobs[it.job].a;
    a.t0 = bs.step;
    a.th = it.shape & 0xFFFFu;
    a.halo = (it.shape >> 16) & 0x7FFFu;
    a.src = bs.cur ? jobs[it.job].buf1 : jobs[it.job].buf0;
    a.dst = bs.cur ? jobs[it.job].buf0 : jobs[it.job].buf1;
    if (a.fdec) a.fdec += (size_t)(bs.step - bs.fdec_step0) * a.fdec_stride;
    // batch rounds start passes of K >= 4 steps on multiples of 4 (O = 0) and
    // 2-step passes on even steps (picked at run time)
    constexpr int O = K <= 2 ? kRuntimeOffset : 0;
    const bool seg = (it.shape >> 31) != 0;
    const uint32_t Q = __builtin_amdgcn_readfirstlane(a.g.Q);
    if (Q == 4) {
        if (seg) tile_body<kHPP, K, O, RAGGED, kSegments>(a, it.strip, it.tile);
        else tile_body<kHPP, K, O, RAGGED, kStrips>(a, it.strip, it.tile);
    } else if (Q == 6) {
        if (seg) tile_body<kFHP1, K, O, RAGGED, kSegments>(a, it.strip, it.tile);
        else tile_body<kFHP1, K, O, RAGGED, kStrips>(a, it.strip, it.tile);
    } else {
        if (seg) tile_body<kFHP2, K, O, RAGGED, kSegments>(a, it.strip, it.tile);
        else tile_body<kFHP2, K, O, RAGGED, kStrips>(a, it.strip, it.tile);
    }
    // The job's last wave of the round advances it.  Every other wave of the
    // job has read the state and counted itself, so nothing in this launch
    // reads the state after the update, and the next launch on the stream
    // sees it.
    if ((threadIdx.x & 63u) == 0) {
        BatchState& st = state[it.job];
        if (atomicAdd(&st.done, 1u) + 1u == it.job_items) {
            st.done = 0;
            st.step = bs.step + (uint32_t)K;
            st.cur = bs.cur ^ 1u;
        }
    }
}

// Launchers of one (model, K) pair; lgca_tile.hip has validated the shape and
// built the grid.  The pass offset O = t0 & 3 selects the instantiation.
template <int KIND, int K>
hipError_t launch_tile_kk(const TileArgs& a, const TileOrder& o, uint32_t nwaves, bool ragged,
                          int seg, hipStream_t s)
{
    const int off = K == 1 ? 0 : (int)(a.t0 & 3u);
    const dim3 grid(nwaves), block(64);
#define LGCA_TILE_LAUNCH(OO, RG, SG) \
    hipLaunchKernelGGL((tile_kernel<KIND, K, OO, RG, SG>), grid, block, 0, s, a, o)
#define LGCA_TILE_OFF(OO)                                                   \
    if (off == OO) {                                                        \
        if (ragged) {                                                       \
            if (seg) LGCA_TILE_LAUNCH(OO, true, kSegments);                 \
            else LGCA_TILE_LAUNCH(OO, true, kStrips);                       \
        } else {                                                            \
            if (seg == kWholeRow) LGCA_TILE_LAUNCH(OO, false, kWholeRow);   \
            else if (seg) LGCA_TILE_LAUNCH(OO, false, kSegments);           \
            else LGCA_TILE_LAUNCH(OO, false, kStrips);                      \
        }                                                                   \
        return hipGetLastError();                                           \
    }
    LGCA_TILE_OFF(0)
    if constexpr (K > 1) {
        LGCA_TILE_OFF(1)
        LGCA_TILE_OFF(2)
        LGCA_TILE_OFF(3)
    }
#undef LGCA_TILE_OFF
#undef LGCA_TILE_LAUNCH
    return hipErrorInvalidValue;
}

// Waves of the strip-mode kernel of one (model, K) pair that a CU holds at
// once (one wave per workgroup), or 0 when the runtime cannot say.
template <int KIND, int K>
int tile_waves_per_cu_kk(bool ragged)
{
    int n = 0;
    const hipError_t e =
        ragged ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, tile_kernel<KIND, K, 0, true, kStrips>, 64, 0)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, tile_kernel<KIND, K, 0, false, kStrips>, 64, 0);
    return e == hipSuccess ? n : 0;
}

// Pipelined launcher (K = 4 or 8, no short last word); lgca_tile.hip has
// built the plan and checked the shape.
template <int KIND, int K>
hipError_t launch_tile_pipe_kk(const TileArgs& a, const PipeArgs& p, uint32_t nwaves, hipStream_t s)
{
    static_assert(K == 4 || K == 8, "pipelined passes keep the chirality offset: K = 4 or 8");
    switch (a.t0 & 3u) {
    case 0: hipLaunchKernelGGL((tile_pipe_kernel<KIND, K, 0>), dim3(nwaves), dim3(64), 0, s, a, p); break;
    case 1: hipLaunchKernelGGL((tile_pipe_kernel<KIND, K, 1>), dim3(nwaves), dim3(64), 0, s, a, p); break;
    case 2: hipLaunchKernelGGL((tile_pipe_kernel<KIND, K, 2>), dim3(nwaves), dim3(64), 0, s, a, p); break;
    default: hipLaunchKernelGGL((tile_pipe_kernel<KIND, K, 3>), dim3(nwaves), dim3(64), 0, s, a, p); break;
    }
    return hipGetLastError();
}

template <int K>
hipError_t launch_tile_batch_kk(const BatchJob* jobs, BatchState* state, const WorkItem* items,
                                uint32_t n_items, uint32_t n_heavy, bool ragged, hipStream_t s)
{
    if (ragged)
        hipLaunchKernelGGL((tile_batch_kernel<K, true>), dim3(n_items), dim3(64), 0, s, jobs, state, items,
                           n_items, n_heavy);
    else
        hipLaunchKernelGGL((tile_batch_kernel<K, false>), dim3(n_items), dim3(64), 0, s, jobs, state, items,
                           n_items, n_heavy);
    return hipGetLastError();
}

}  // namespace lgca
