// lgca_tile.hip -- host side of the temporally blocked step: launch shapes,
// grids and the strip-mode block order, dispatched to the per-(model, K) and
// per-K batch launchers that lgca_tile_inst.hip instantiates from lgca_tile.hpp.
#include "lgca_dev.hpp"
#include "lgca_kernels.hpp"

#include <algorithm>

namespace lgca {

hipError_t launch_tile_batch(int k, bool ragged, const BatchJob* jobs, BatchState* state,
                             const WorkItem* items, uint32_t n_items, uint32_t n_heavy, hipStream_t s)
{
    if (!n_items) return hipSuccess;
    const uint32_t nblk = (n_items + 3) / 4, hblk = std::min<uint32_t>((n_heavy + 3) / 4, nblk);
    switch (k) {
    case 1: return launch_tile_batch_kk<1>(jobs, state, items, n_items, nblk, hblk, ragged, s);
    case 2: return launch_tile_batch_kk<2>(jobs, state, items, n_items, nblk, hblk, ragged, s);
    case 4: return launch_tile_batch_kk<4>(jobs, state, items, n_items, nblk, hblk, ragged, s);
    case 8: return launch_tile_batch_kk<8>(jobs, state, items, n_items, nblk, hblk, ragged, s);
    default: return hipErrorInvalidValue;
    }
}

uint32_t tile_halo(const Geometry& g, int k)
{
    if (g.W == 1) return (uint32_t)((k + (int)g.nb - 1) / (int)g.nb);
    return (int)g.nb < k ? 2u : 1u;
}

uint32_t tile_strips(const Geometry& g, int k)
{
    const uint32_t stride = 64u - 2u * tile_halo(g, k);
    return (g.W + stride - 1) / stride;
}

bool tile_segments(const Geometry& g, int k)
{
    return g.W <= 64 && (64 / g.W >= 2 || tile_strips(g, k) >= 2);
}

hipError_t launch_tile(int kind, int k, const TileArgs& a, hipStream_t s)
{
    const bool seg = tile_segments(a.g, k);
    if (!seg && (a.halo < tile_halo(a.g, k) || a.halo > 16)) return hipErrorInvalidValue;
    if (a.th < 2 || (a.th & 1u) || a.th > 65535) return hipErrorInvalidValue;
    const uint32_t ntiles = (a.g.dim_y + a.th - 1) / a.th;
    dim3 grid;
    TileOrder o{};
    if (seg) {
        const uint32_t groups = (ntiles + 64 / a.g.W - 1) / (64 / a.g.W);
        grid = dim3((groups + 3) / 4, 1);
    } else {
        const uint32_t stride = 64u - 2u * a.halo;
        const uint32_t nstrips = (a.g.W + stride - 1) / stride;
        o.nbx = (nstrips + 3) / 4;
        o.ntiles = ntiles;
        if ((uint64_t)o.nbx * ntiles >= ((uint64_t)1 << 31)) return hipErrorInvalidValue;
        grid = dim3(o.nbx * ntiles);
        // heavy block columns: those with a strip over a forced word; strip
        // st spans words [st * stride - halo, + 64) modulo W
        if (a.fdec && a.fwn && nstrips <= 4096) {
            const int64_t W = a.g.W;
            auto hits_force = [&](uint32_t st) {
                if (W <= 64) return true;  // the strip wraps the whole row
                const int64_t a0 = (((int64_t)st * stride - a.halo) % W + W) % W;
                for (int64_t k = -1; k <= 1; ++k) {
                    const int64_t b0 = (int64_t)a.fw0 + k * W, b1 = b0 + a.fwn;
                    if (a0 < b1 && a0 + 64 > b0) return true;
                }
                return false;
            };
            for (uint32_t bx = 0; bx < o.nbx; ++bx) {
                bool heavy = false;
                for (uint32_t st = 4 * bx; st < 4 * bx + 4 && st < nstrips; ++st) heavy |= hits_force(st);
                if (!heavy) continue;
                if (o.nheavy == kMaxHeavy) { o.nheavy = 0; break; }  // many: plain order
                o.heavy[o.nheavy++] = bx;
            }
            if (o.nheavy == o.nbx) o.nheavy = 0;
        }
    }
    const bool ragged = a.g.nb != 32;
    const uint32_t nblocks = grid.x;
#define LGCA_TILE_CASE(KD, KK) \
    if (kind == KD && k == KK) return launch_tile_kk<KD, KK>(a, o, nblocks, ragged, seg, s);
    LGCA_TILE_CASE(kHPP, 1) LGCA_TILE_CASE(kHPP, 2) LGCA_TILE_CASE(kHPP, 4) LGCA_TILE_CASE(kHPP, 8)
    LGCA_TILE_CASE(kFHP1, 1) LGCA_TILE_CASE(kFHP1, 2) LGCA_TILE_CASE(kFHP1, 4) LGCA_TILE_CASE(kFHP1, 8)
    LGCA_TILE_CASE(kFHP2, 1) LGCA_TILE_CASE(kFHP2, 2) LGCA_TILE_CASE(kFHP2, 4) LGCA_TILE_CASE(kFHP2, 8)
#undef LGCA_TILE_CASE
    return hipErrorInvalidValue;
}

}  // namespace lgca
