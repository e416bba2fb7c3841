// lgca_tile.hip -- host side of the temporally blocked step: checks launch
// shapes and grids and dispatches to the per-(model, K) and per-K batch
// launchers that lgca_tile_inst.hip instantiates from lgca_tile.hpp.  The
// plans themselves (halo, strips, wave order, bands) are in lgca_plan.cpp.
#include "lgca_dev.hpp"
#include "lgca_kernels.hpp"

#include <algorithm>
#include <cstdint>

namespace lgca {

hipError_t launch_tile_batch(int k, bool ragged, const BatchJob* jobs, BatchState* state,
                             const WorkItem* items, uint32_t n_items, uint32_t n_heavy, hipStream_t s)
{
    if (!n_items) return hipSuccess;
    n_heavy = std::min(n_heavy, n_items);
    switch (k) {
    case 1: return launch_tile_batch_kk<1>(jobs, state, items, n_items, n_heavy, ragged, s);
    case 2: return launch_tile_batch_kk<2>(jobs, state, items, n_items, n_heavy, ragged, s);
    case 4: return launch_tile_batch_kk<4>(jobs, state, items, n_items, n_heavy, ragged, s);
    case 8: return launch_tile_batch_kk<8>(jobs, state, items, n_items, n_heavy, ragged, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_tile_pipe(int kind, int k, const TileArgs& a, const PipeArgs& p, uint32_t nwaves, hipStream_t s)
{
    if (tile_segments(a.g, k) || a.g.nb != 32 || a.th < (uint32_t)k || (a.th & 1u) || a.th > 65535)
        return hipErrorInvalidValue;
    if (a.halo < tile_halo(a.g, k) || a.halo > 16 || p.nband < 1 || p.nband > kMaxBands || !nwaves)
        return hipErrorInvalidValue;
#define LGCA_PIPE_CASE(KD, KK) \
    if (kind == KD && k == KK) return launch_tile_pipe_kk<KD, KK>(a, p, nwaves, s);
    LGCA_PIPE_CASE(kHPP, 4) LGCA_PIPE_CASE(kHPP, 8)
    LGCA_PIPE_CASE(kFHP1, 4) LGCA_PIPE_CASE(kFHP1, 8)
    LGCA_PIPE_CASE(kFHP2, 4) LGCA_PIPE_CASE(kFHP2, 8)
#undef LGCA_PIPE_CASE
    return hipErrorInvalidValue;
}

int tile_waves_per_cu(int kind, int k, bool ragged)
{
#define LGCA_OCC_CASE(KD, KK) \
    if (kind == KD && k == KK) return tile_waves_per_cu_kk<KD, KK>(ragged);
    LGCA_OCC_CASE(kHPP, 1) LGCA_OCC_CASE(kHPP, 2) LGCA_OCC_CASE(kHPP, 4) LGCA_OCC_CASE(kHPP, 8)
    LGCA_OCC_CASE(kFHP1, 1) LGCA_OCC_CASE(kFHP1, 2) LGCA_OCC_CASE(kFHP1, 4) LGCA_OCC_CASE(kFHP1, 8)
    LGCA_OCC_CASE(kFHP2, 1) LGCA_OCC_CASE(kFHP2, 2) LGCA_OCC_CASE(kFHP2, 4) LGCA_OCC_CASE(kFHP2, 8)
#undef LGCA_OCC_CASE
    return 0;
}

hipError_t launch_tile(int kind, int k, const TileArgs& a, hipStream_t s, const uint32_t* order, uint32_t n_order)
{
    const bool seg = tile_segments(a.g, k);
    if (!seg && (a.halo < tile_halo(a.g, k) || a.halo > 16)) return hipErrorInvalidValue;
    if (a.th < 2 || (a.th & 1u) || a.th > 65535) return hipErrorInvalidValue;
    // the pass's k steps of force decisions go through one buffer descriptor,
    // with room past its end for the offset that reads 0
    if (a.fdec && (uint64_t)k * a.fdec_stride * 4u >= ((uint64_t)1 << 31)) return hipErrorInvalidValue;
    const uint32_t ntiles = (a.g.dim_y + a.th - 1) / a.th;
    uint32_t nwaves = 0;  // one wave per workgroup
    TileOrder o{};
    if (seg) {
        nwaves = (ntiles + 64 / a.g.W - 1) / (64 / a.g.W);
    } else {
        const uint32_t stride = 64u - 2u * a.halo;
        const uint32_t nstrips = (a.g.W + stride - 1) / stride;
        o.nstrips = nstrips;
        o.ntiles = ntiles;
        if ((uint64_t)nstrips * ntiles >= ((uint64_t)1 << 31)) return hipErrorInvalidValue;
        nwaves = nstrips * ntiles;
        o.order = order;
        if (order) {
            // an order lists every tile and strip once, whole or in up to 4 parts
            if ((uint64_t)n_order < nwaves || (uint64_t)n_order > 4ull * nwaves) return hipErrorInvalidValue;
            nwaves = n_order;
        }
        // without an order: the heavy strips are those over a forced word
        if (!order && a.fdec && a.fwn && nstrips <= 16384) {
            for (uint32_t st = 0; st < nstrips; ++st) {
                if (!tile_strip_hits(a.g, stride, a.halo, st, a.fw0, (int64_t)a.fw0 + a.fwn)) continue;
                if (o.nheavy == kMaxHeavy) { o.nheavy = 0; break; }  // many: plain order
                o.heavy[o.nheavy++] = st;
            }
            if (o.nheavy == nstrips) o.nheavy = 0;
        }
    }
    const bool ragged = a.g.nb != 32;
    // segment mode with one row of exactly 64 words per wave: DPP rotations
    // instead of permutes for the neighbouring words
    const int seg_mode = !seg ? 0 : (a.g.W == 64 && !ragged) ? 2 : 1;
#define LGCA_TILE_CASE(KD, KK) \
    if (kind == KD && k == KK) return launch_tile_kk<KD, KK>(a, o, nwaves, ragged, seg_mode, s);
    LGCA_TILE_CASE(kHPP, 1) LGCA_TILE_CASE(kHPP, 2) LGCA_TILE_CASE(kHPP, 4) LGCA_TILE_CASE(kHPP, 8)
    LGCA_TILE_CASE(kFHP1, 1) LGCA_TILE_CASE(kFHP1, 2) LGCA_TILE_CASE(kFHP1, 4) LGCA_TILE_CASE(kFHP1, 8)
    LGCA_TILE_CASE(kFHP2, 1) LGCA_TILE_CASE(kFHP2, 2) LGCA_TILE_CASE(kFHP2, 4) LGCA_TILE_CASE(kFHP2, 8)
#undef LGCA_TILE_CASE
    return hipErrorInvalidValue;
}

}  // namespace lgca
