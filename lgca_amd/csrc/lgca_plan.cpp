// lgca_plan.cpp -- host-only launch planning: lattice geometry, the tiled
// step's halo and strip counts, the strip-mode wave dispatch order and the
// band plan of multi-pass launches.  Plain C++ (no device code), so that test
// programs can build it with the host compiler and check the plans without a
// GPU.
#include "lgca_kernels.hpp"

#include <algorithm>
#include <climits>
#include <cstdint>

namespace lgca {

Geometry make_geometry(uint32_t dim_x, uint32_t dim_y, uint32_t Q)
{
    Geometry g{};
    g.dim_x = dim_x;
    g.dim_y = dim_y;
    g.W = (dim_x + 31) / 32;
    g.P = (g.W + 3) & ~3u;
    g.Wq = g.P / 4;
    g.nb = dim_x - 32 * (g.W - 1);
    g.Q = Q;
    uint32_t l = 0;
    while ((1u << l) < g.Wq && l < 6) ++l;
    g.qb_log = l;
    return g;
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

// Does strip st (words [st * stride - halo, + 64) modulo W) touch words
// [lo, hi) of its rows?
bool tile_strip_hits(const Geometry& g, uint32_t stride, uint32_t halo, uint32_t st, int64_t lo, int64_t hi)
{
    if (lo >= hi) return false;
    const int64_t W = g.W;
    if (W <= 64) return true;  // the strip wraps the whole row
    const int64_t a0 = (((int64_t)st * stride - halo) % W + W) % W;
    for (int64_t k = -1; k <= 1; ++k)
        if (a0 < hi + k * W && a0 + 64 > lo + k * W) return true;
    return false;
}

namespace {

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
            fstrip[st] = tile_strip_hits(a.g, stride, a.halo, st, a.fw0, (int64_t)a.fw0 + a.fwn);
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
            if (lo < hi && tile_strip_hits(a.g, stride, a.halo, st, lo, hi)) c = 2;
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
    for (uint32_t it = 0; it < (uint32_t)cls.size(); ++it) {
        const uint32_t e = it << kPartBits;  // whole tile
        if (cls[it] != 2) {
            (cls[it] == 1 ? force : plain).push_back(e);
            continue;
        }
        // A walk over solids runs up to 2.5 times as long as a plain one,
        // long enough to outlast the pass: it is cut into quarters of the
        // tile (halves on short tiles), each walking its rows plus the
        // K-row halo, which costs more in all but ends sooner.
        const uint32_t t = it / nstrips, rows = std::min(a.th, a.g.dim_y - t * a.th);
        if (rows >= 16)
            for (uint32_t p = 3; p <= 6; ++p) solid.push_back(e | p);
        else if (rows >= 8)
            for (uint32_t p = 1; p <= 2; ++p) solid.push_back(e | p);
        else
            solid.push_back(e);
    }
    // The solid walks go first, spread over the XCDs; then the forced and
    // the plain ones, each in contiguous runs per XCD, so that the tiles an
    // XCD holds are vertical neighbours whose halo rows meet in its L2.
    order.reserve(solid.size() + force.size() + plain.size());
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

}  // namespace lgca
