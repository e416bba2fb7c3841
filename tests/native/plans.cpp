// Host-side launch planning of the tiled step (lgca_amd/csrc/lgca_plan.cpp),
// checked without a GPU:
//   * tile_wave_order covers every row of every row tile and strip of a pass
//     exactly once, whole tiles or parts of an even number of rows starting
//     on even rows; the walks over solid rows come before all others (cut
//     into parts on tiles of 8 rows or more), the forced strips before the
//     unforced ones, and a lattice without solids or forcing is handed out
//     in contiguous ascending runs, one per XCD (entries b, b + 8, ... of the
//     order run on one XCD);
//   * tile_pipe_plan cuts the row tiles into contiguous bands of whole tiles
//     (at least two tiles and 2K rows each) that cover the lattice once, and
//     lists every wave once, band after band; it refuses the shapes its
//     kernel cannot run.
// Prints "PLANS PASSED" and exits 0 when every check holds.
#include "lgca_kernels.hpp"

#include <algorithm>
#include <cstdio>
#include <vector>

using namespace lgca;

static int failures = 0;

#define CHECK(cond, ...)                         \
    do {                                         \
        if (!(cond)) {                           \
            std::printf("FAILED: " __VA_ARGS__); \
            std::printf("\n");                   \
            ++failures;                          \
        }                                        \
    } while (0)

enum Shape { kOpen, kWallsAndBlock };

// Non-fluid word ranges per row: none, or walls on the first and last row
// plus a 3-word block over the middle fifth of the rows.
struct Solids {
    std::vector<int32_t> lo, hi;
    std::vector<uint8_t> row;  // row holds a solid
    uint32_t y0 = 0, y1 = 0;   // block rows
};

static Solids solids(const Geometry& g, Shape shape)
{
    Solids s;
    s.lo.assign(g.dim_y, 0x7FFFFFFF);
    s.hi.assign(g.dim_y, 0);
    s.row.assign(g.dim_y, 0);
    if (shape == kOpen) return s;
    for (uint32_t y : {0u, g.dim_y - 1}) {
        s.lo[y] = 0;
        s.hi[y] = (int32_t)g.W;
        s.row[y] = 1;
    }
    s.y0 = g.dim_y / 2 - g.dim_y / 10;
    s.y1 = g.dim_y / 2 + g.dim_y / 10;
    for (uint32_t y = s.y0; y < s.y1; ++y) {
        s.lo[y] = (int32_t)(g.W / 8);
        s.hi[y] = (int32_t)(g.W / 8 + 3);
        s.row[y] = 1;
    }
    return s;
}

// Does the walk of tile t (its rows and k rows either side, periodic) cross
// a row that holds solids?
static bool walk_crosses(const Solids& s, uint32_t dim_y, uint32_t th, uint32_t t, int k)
{
    const int64_t y0 = (int64_t)t * th, y1 = std::min<int64_t>(y0 + th, dim_y);
    for (int64_t y = y0 - k; y < y1 + k; ++y)
        if (s.row[(size_t)(((y % dim_y) + dim_y) % dim_y)]) return true;
    return false;
}

static void check_order(const Geometry& g, Shape shape, int k, uint32_t th, bool forced)
{
    const Solids s = solids(g, shape);
    TileArgs a{};
    a.g = g;
    a.th = th;
    a.halo = tile_halo(g, k);
    uint32_t dummy = 0;
    if (forced) {  // the first two words of every row
        a.fdec = &dummy;
        a.fw0 = 0;
        a.fwn = 2;
        a.fy0 = 0;
        a.fy1 = g.dim_y;
    }
    std::vector<uint32_t> order;
    const bool ok = tile_wave_order(a, k, shape == kOpen ? nullptr : s.lo.data(),
                                    shape == kOpen ? nullptr : s.hi.data(), order);
    CHECK(ok == !tile_segments(g, k), "order given iff strip mode (%u x %u, k %d, th %u)", g.dim_x, g.dim_y, k, th);
    if (!ok) return;

    const uint32_t ns = tile_strips(g, k), nt = (g.dim_y + th - 1) / th;
    // rows covered per wave (tile * ns + strip), and the positions of its
    // first and last entries
    std::vector<uint32_t> pos((size_t)ns * nt, UINT32_MAX), last((size_t)ns * nt, 0), covered((size_t)ns * nt, 0);
    std::vector<uint8_t> parts((size_t)ns * nt, 0);
    bool ok_parts = order.size() >= pos.size() && order.size() <= 4 * pos.size();
    for (size_t i = 0; ok_parts && i < order.size(); ++i) {
        const uint32_t w = order[i] >> kPartBits, part = order[i] & ((1u << kPartBits) - 1u);
        ok_parts = w < pos.size() && part <= 6;
        if (!ok_parts) break;
        const uint32_t t = w / ns, rows = std::min(th, g.dim_y - t * th);
        uint32_t off = 0, n = 0;
        tile_part_rows(rows, part, off, n);
        // parts of one split (halves or quarters) only, each even-sized,
        // starting on an even row, and covering the tile once in all
        const uint8_t kind = part == 0 ? 1 : part <= 2 ? 2 : 4;
        ok_parts = (parts[w] == 0 || parts[w] == kind) && n > 0 && off % 2 == 0 &&
                   (rows % 2 != 0 || n % 2 == 0) && off + n <= rows;
        parts[w] = kind;
        covered[w] += n;
        pos[w] = std::min(pos[w], (uint32_t)i);
        last[w] = std::max(last[w], (uint32_t)i);
    }
    for (size_t w = 0; ok_parts && w < pos.size(); ++w)
        ok_parts = covered[w] == std::min(th, g.dim_y - (uint32_t)(w / ns) * th);
    CHECK(ok_parts, "order covers every row once (%u x %u, k %d, th %u, forced %d)", g.dim_x, g.dim_y, k, th,
          (int)forced);
    if (!ok_parts) return;
    // every part of a split lies in the tile and the parts do not overlap:
    // each part's rows are counted once above and their sum is the tile's
    for (uint32_t rows = 2; rows <= 64; ++rows)
        for (uint32_t f : {2u, 4u}) {
            uint32_t next = 0;
            for (uint32_t p = (f == 2 ? 1u : 3u); p < (f == 2 ? 3u : 7u); ++p) {
                uint32_t off = 0, n = 0;
                tile_part_rows(rows, p, off, n);
                if (rows >= 4 * f) CHECK(off == next && n >= 2, "parts tile the rows (%u rows, part %u)", rows, p);
                next = off + n;
            }
            if (rows >= 4 * f) CHECK(next == rows, "parts end at the tile end (%u rows)", rows);
        }

    // solid walks first: every wave of a tile whose walk crosses a wall row
    // (all its strips cross solids) before every wave of a tile that crosses
    // no solid row
    uint32_t last_wall = 0, first_clear = UINT32_MAX;
    // forced strips first among the clear tiles: strip 0 covers words 0, 1;
    // a middle strip covers neither
    const uint32_t mid = ns / 2;
    const uint32_t stride = 64u - 2u * a.halo;
    const bool mid_clear = ns >= 3 && mid * stride >= a.halo + 2 && mid * stride + 64 <= g.W + a.halo;
    uint32_t last_forced = 0, first_mid = UINT32_MAX;
    for (uint32_t t = 0; t < nt; ++t) {
        const bool wall = shape != kOpen && (walk_crosses(s, g.dim_y, th, t, k)) &&
                          ((int64_t)t * th - k < 1 || (int64_t)(t + 1) * th + k > (int64_t)g.dim_y - 1);
        const bool clear = shape == kOpen || !walk_crosses(s, g.dim_y, th, t, k);
        for (uint32_t st = 0; st < ns; ++st) {
            const uint32_t p = pos[(size_t)t * ns + st], q = last[(size_t)t * ns + st];
            if (wall) last_wall = std::max(last_wall, q);
            if (clear) first_clear = std::min(first_clear, p);
            if (clear && forced && st == 0) last_forced = std::max(last_forced, q);
            if (clear && forced && mid_clear && st == mid) first_mid = std::min(first_mid, p);
        }
    }
    if (shape != kOpen && first_clear != UINT32_MAX)
        CHECK(last_wall < first_clear, "solid walks lead (%u x %u, k %d, th %u)", g.dim_x, g.dim_y, k, th);
    if (forced && mid_clear && first_mid != UINT32_MAX)
        CHECK(last_forced < first_mid, "forced strips lead the clear tiles (%u x %u, k %d, th %u)", g.dim_x,
              g.dim_y, k, th);

    // solid walks are cut into parts on tiles of 8 rows or more
    if (shape != kOpen)
        for (uint32_t st = 0; st < ns; ++st)
            CHECK(parts[st] == (th >= 16 ? 4 : th >= 8 ? 2 : 1), "solid walk of tile 0 cut (%u x %u, k %d, th %u)",
                  g.dim_x, g.dim_y, k, th);
    // one class: entries b, b + 8, b + 16, ... form an ascending run of
    // consecutive whole waves, and the runs of XCDs 0..7 follow each other
    if (shape == kOpen && !forced) {
        uint32_t next = 0;
        for (uint32_t x = 0; x < 8 && x < order.size(); ++x)
            for (size_t b = x; b < order.size(); b += 8) {
                CHECK(order[b] == next << kPartBits, "contiguous XCD runs (%u x %u, k %d, th %u, entry %zu)",
                      g.dim_x, g.dim_y, k, th, b);
                next = (order[b] >> kPartBits) + 1;
            }
    }
}

static void check_plan(const Geometry& g, Shape shape, int k, uint32_t th)
{
    const Solids s = solids(g, shape);
    TileArgs a{};
    a.g = g;
    a.th = th;
    a.halo = tile_halo(g, k);
    const uint32_t nband = 8;
    PipePlan plan;
    const bool ok = tile_pipe_plan(a, k, nband, shape == kOpen ? nullptr : s.lo.data(),
                                   shape == kOpen ? nullptr : s.hi.data(), plan);
    const uint32_t nt = (g.dim_y + th - 1) / th;
    const bool allowed = (k == 4 || k == 8) && !tile_segments(g, k) && g.nb == 32 && th >= (uint32_t)k &&
                         !(th & 1u) && nt >= 2 * nband;
    if (!allowed) {
        CHECK(!ok, "plan refused (%u x %u, k %d, th %u)", g.dim_x, g.dim_y, k, th);
        return;
    }
    if (g.dim_y % th == 0) CHECK(ok, "plan given (%u x %u, k %d, th %u)", g.dim_x, g.dim_y, k, th);
    if (!ok) return;
    CHECK(plan.nband == nband && plan.ntiles == nt && plan.nstrips == tile_strips(g, k), "plan sizes");
    std::vector<uint8_t> seen((size_t)plan.nstrips * plan.ntiles, 0);
    uint32_t next_row = 0, next_item = 0;
    for (uint32_t b = 0; b < plan.nband; ++b) {
        CHECK(plan.band_y0[b] == next_row, "bands are contiguous (band %u)", b);
        CHECK(plan.band_y0[b] % th == 0, "band %u starts on a tile", b);
        CHECK(plan.band_y1[b] - plan.band_y0[b] >= 2u * (uint32_t)k, "band %u has 2K rows", b);
        CHECK(plan.band_first[b] == next_item, "items band after band (band %u)", b);
        CHECK(plan.band_items[b] >= 2 * plan.nstrips, "band %u has two tiles", b);
        CHECK(plan.band_items[b] % plan.nstrips == 0, "band %u has whole tiles", b);
        for (uint32_t i = 0; i < plan.band_items[b]; ++i) {
            const uint32_t it = plan.order[plan.band_first[b] + i];
            const uint32_t y = it / plan.nstrips * th;
            CHECK(y >= plan.band_y0[b] && y < plan.band_y1[b], "item %u in band %u", it, b);
            CHECK(it < seen.size() && !seen[it], "item %u listed once", it);
            if (it < seen.size()) seen[it] = 1;
        }
        next_row = plan.band_y1[b];
        next_item += plan.band_items[b];
    }
    CHECK(next_row == g.dim_y, "bands cover the rows (%u x %u, k %d, th %u)", g.dim_x, g.dim_y, k, th);
    CHECK(plan.order.size() == seen.size() &&
              std::count(seen.begin(), seen.end(), (uint8_t)1) == (long)seen.size(),
          "every item listed (%u x %u, k %d, th %u)", g.dim_x, g.dim_y, k, th);
}

int main()
{
    // strip mode with whole and short last words, segment mode, one word
    for (uint32_t dim_x : {32768u, 4096u, 3000u, 2048u, 100000u, 20u}) {
        for (uint32_t dim_y : {8192u, 1024u, 512u, 1000u}) {
            const Geometry g = make_geometry(dim_x, dim_y, 7);
            for (int k : {1, 2, 4, 8})
                for (uint32_t th : {2u, 8u, 12u, 16u, 40u, 64u, 1000u})
                    for (Shape shape : {kOpen, kWallsAndBlock}) {
                        for (bool forced : {false, true}) check_order(g, shape, k, th, forced);
                        check_plan(g, shape, k, th);
                    }
        }
    }
    if (failures) {
        std::printf("%d checks failed\n", failures);
        return 1;
    }
    std::printf("PLANS PASSED\n");
    return 0;
}
