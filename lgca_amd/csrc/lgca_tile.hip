// lgca_tile.hip -- host side of the temporally blocked step: launch shapes,
// grids and the strip-mode wave order, dispatched to the per-(model, K) and
// per-K batch launchers that lgca_tile_inst.hip instantiates from lgca_tile.hpp.
#include "lgca_dev.hpp"
#include "lgca_kernels.hpp"

#include <algorithm>
#include <climits>
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

namespace {

// Does strip st (words [st * stride - halo, + 64) modulo W) touch words
// [lo, hi) of its rows?
bool strip_hits(const Geometry& g, uint32_t stride, uint32_t halo, uint32_t st, int64_t lo, int64_t hi)
{
    if (lo >= hi) return false;
    const int64_t W = g.W;
    if (W <= 64) return true;  // the strip wraps the whole row
    const int64_t a0 = (((int64_t)st * stride - halo) % W + W) % W;
    for (int64_t k = -1; k <= 1; ++k)
        if (a0 < hi + k * W && a0 + 64 > lo + k * W) return true;
    return false;
}

// Appends v in the order that hands each XCD a contiguous run of it:
// workgroups go to the 8 XCDs round-robin, so entry b of the appended range
// runs on XCD (start + b) mod 8.
void append_xcd_runs(std::vector<uint32_t>& order, const std::vector<uint32_t>& v)
{
    const uint32_t n = (uint32_t)v.size(), q = n >> 3, r = n & 7u;
    for (uint32_t b = 0; b < n; ++b) {
        const uint32_t xcd = b & 7u;
        order.push_back(v[xcd * q + std::min(xcd, r) + (b >> 3)]);
    }
}

// The cost class of every strip-mode wave of a pass (tile * nstrips + strip):
// 2 when its walk crosses solids (about twice a plain walk), 1 when it crosses
// forced words (about a sixth longer), else 0.  False when the pass is not in
// strip mode or has too many waves to list.
bool wave_classes(const TileArgs& a, int k, const int32_t* row_lo, const int32_t* row_hi,
                  uint32_t& nstrips, uint32_t& ntiles, std::vector<uint8_t>& cls)
{
    if (tile_segments(a.g, k) || a.th < 2) return false;
    const uint32_t stride = 64u - 2u * a.halo;
    nstrips = (a.g.W + stride - 1) / stride;
    ntiles = (a.g.dim_y + a.th - 1) / a.th;
    if ((uint64_t)nstrips * ntiles > ((uint64_t)1 << 24)) return false;
    const bool forced = a.fdec && a.fwn;
    std::vector<uint8_t> fstrip(nstrips, 0);
    if (forced)
        for (uint32_t st = 0; st < nstrips; ++st)
            fstrip[st] = strip_hits(a.g, stride, a.halo, st, a.fw0, (int64_t)a.fw0 + a.fwn);
    cls.assign((size_t)nstrips * ntiles, 0);
    const int64_t dy = a.g.dim_y;
    for (uint32_t t = 0; t < ntiles; ++t) {
        const int64_t y0 = (int64_t)t * a.th, rows = std::min<int64_t>(a.th, dy - y0);
        int64_t lo = INT64_MAX, hi = INT64_MIN;  // hull of the walk rows' solid words
        bool frow = false;
        for (int64_t y = y0 - k; y < y0 + rows + k; ++y) {
            const int64_t r = ((y % dy) + dy) % dy;
            if (row_lo && row_lo[r] < row_hi[r]) {
                lo = std::min<int64_t>(lo, row_lo[r]);
                hi = std::max<int64_t>(hi, row_hi[r]);
            }
            frow |= r >= (int64_t)a.fy0 && r < (int64_t)a.fy1;
        }
        for (uint32_t st = 0; st < nstrips; ++st) {
            uint8_t c = 0;
            if (lo < hi && strip_hits(a.g, stride, a.halo, st, lo, hi)) c = 2;
            else if (forced && frow && fstrip[st]) c = 1;
            cls[(size_t)t * nstrips + st] = c;
        }
    }
    return true;
}

}  // namespace

bool tile_wave_order(const TileArgs& a, int k, const int32_t* row_lo, const int32_t* row_hi,
                     std::vector<uint32_t>& order)
{
    order.clear();
    uint32_t nstrips = 0, ntiles = 0;
    std::vector<uint8_t> cls;
    if (!wave_classes(a, k, row_lo, row_hi, nstrips, ntiles, cls)) return false;
    std::vector<uint32_t> solid, force, plain;
    for (uint32_t it = 0; it < (uint32_t)cls.size(); ++it)
        (cls[it] == 2 ? solid : cls[it] == 1 ? force : plain).push_back(it);
    // The few long solid walks go first, spread over the XCDs; then the
    // forced and the plain ones, each in contiguous runs per XCD, so that
    // the tiles an XCD holds are vertical neighbours whose halo rows meet in
    // its L2.
    order.reserve((size_t)nstrips * ntiles);
    order.insert(order.end(), solid.begin(), solid.end());
    append_xcd_runs(order, force);
    append_xcd_runs(order, plain);
    return true;
}

bool tile_pipe_plan(const TileArgs& a, int k, uint32_t nband, const int32_t* row_lo, const int32_t* row_hi,
                    PipePlan& plan)
{
    plan = PipePlan{};
    if ((k != 4 && k != 8) || nband < 1 || nband > kMaxBands) return false;
    if (a.g.nb != 32 || a.th < (uint32_t)k || (a.th & 1u)) return false;
    uint32_t nstrips = 0, ntiles = 0;
    std::vector<uint8_t> cls;
    if (!wave_classes(a, k, row_lo, row_hi, nstrips, ntiles, cls)) return false;
    if (ntiles < 2 * nband) return false;
    // bands of whole tiles with about equal work, at least two tiles each
    std::vector<double> cost(ntiles + 1, 0.0);  // prefix sums
    for (uint32_t t = 0; t < ntiles; ++t) {
        double c = 0.0;
        for (uint32_t st = 0; st < nstrips; ++st) c += 1.0 + (double)cls[(size_t)t * nstrips + st] * 0.5;
        cost[t + 1] = cost[t] + c;
    }
    std::vector<uint32_t> cut(nband + 1, 0);
    cut[nband] = ntiles;
    for (uint32_t b = 1; b < nband; ++b) {
        const double target = cost[ntiles] * b / nband;
        uint32_t t = cut[b - 1] + 2;
        while (t < ntiles && cost[t] + 0.5 * (cost[t + 1] - cost[t]) < target) ++t;
        cut[b] = std::min(std::max(t, cut[b - 1] + 2), ntiles - 2 * (nband - b));
    }
    for (uint32_t b = 0; b < nband; ++b) {
        const uint32_t y0 = cut[b] * a.th, y1 = std::min<uint64_t>((uint64_t)cut[b + 1] * a.th, a.g.dim_y);
        if (y1 - y0 < 2u * (uint32_t)k) return false;
        plan.band_y0[b] = y0;
        plan.band_y1[b] = y1;
        plan.band_first[b] = (uint32_t)plan.order.size();
        for (uint32_t t = cut[b]; t < cut[b + 1]; ++t)
            for (uint32_t st = 0; st < nstrips; ++st) plan.order.push_back(t * nstrips + st);
        plan.band_items[b] = (uint32_t)plan.order.size() - plan.band_first[b];
    }
    plan.nband = nband;
    plan.nstrips = nstrips;
    plan.ntiles = ntiles;
    return true;
}

size_t tile_pipe_exch_words(const Geometry& g, int k, uint32_t nband)
{
    return (size_t)2 * nband * 2 * (size_t)k * g.Q * g.P;
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

hipError_t launch_tile(int kind, int k, const TileArgs& a, hipStream_t s, const uint32_t* order)
{
    const bool seg = tile_segments(a.g, k);
    if (!seg && (a.halo < tile_halo(a.g, k) || a.halo > 16)) return hipErrorInvalidValue;
    if (a.th < 2 || (a.th & 1u) || a.th > 65535) return hipErrorInvalidValue;
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
        // without an order: the heavy strips are those over a forced word
        if (!order && a.fdec && a.fwn && nstrips <= 16384) {
            for (uint32_t st = 0; st < nstrips; ++st) {
                if (!strip_hits(a.g, stride, a.halo, st, a.fw0, (int64_t)a.fw0 + a.fwn)) continue;
                if (o.nheavy == kMaxHeavy) { o.nheavy = 0; break; }  // many: plain order
                o.heavy[o.nheavy++] = st;
            }
            if (o.nheavy == nstrips) o.nheavy = 0;
        }
    }
    const bool ragged = a.g.nb != 32;
#define LGCA_TILE_CASE(KD, KK) \
    if (kind == KD && k == KK) return launch_tile_kk<KD, KK>(a, o, nwaves, ragged, seg, s);
    LGCA_TILE_CASE(kHPP, 1) LGCA_TILE_CASE(kHPP, 2) LGCA_TILE_CASE(kHPP, 4) LGCA_TILE_CASE(kHPP, 8)
    LGCA_TILE_CASE(kFHP1, 1) LGCA_TILE_CASE(kFHP1, 2) LGCA_TILE_CASE(kFHP1, 4) LGCA_TILE_CASE(kFHP1, 8)
    LGCA_TILE_CASE(kFHP2, 1) LGCA_TILE_CASE(kFHP2, 2) LGCA_TILE_CASE(kFHP2, 4) LGCA_TILE_CASE(kFHP2, 8)
#undef LGCA_TILE_CASE
    return hipErrorInvalidValue;
}

}  // namespace lgca
