// lgca_kernels.hpp -- launch interface of the lattice-gas kernels.
//
// HBM layout of one state buffer ("row-interleaved bit-planes"):
//   word[(y * Q + d) * P + w]   bit b = cell (32*w + b, y), direction d
// P (row pitch) is the 32-cell word count W rounded up to a multiple of 4 so
// every thread moves whole 16-byte quads; bits past dim_x are kept zero.
// Cell-type masks use the same addressing with two planes per row:
//   mask[(y * 2 + 0) * P + w]  NO_SLIP,   mask[(y * 2 + 1) * P + w]  SLIP.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace lgca {

struct Geometry {
    uint32_t dim_x, dim_y;
    uint32_t W;    // 32-cell words per row
    uint32_t P;    // row pitch in words, multiple of 4
    uint32_t Wq;   // 4-word quads per row (P / 4)
    uint32_t nb;   // valid bits in word W-1, 1..32
    uint32_t Q;    // directions
    uint32_t qb_log;  // log2 of quads per workgroup row segment (<= 6)
};

Geometry make_geometry(uint32_t dim_x, uint32_t dim_y, uint32_t Q);

struct StepArgs {
    Geometry g;
    const uint32_t* src;
    uint32_t* dst;
    const uint32_t* mask;       // cell-type masks, or null when all FLUID
    const int32_t* row_range;   // lo[dim_y] then hi[dim_y]: words [lo, hi) of row y hold non-fluid cells
    uint32_t key0, key1;        // hash key = seed
    uint32_t step;              // hash step (absolute step mod 2^32)
    const uint32_t* fdec;       // force decisions of this step, or null
    uint32_t fy0, fy1;          // rows with decisions
    uint32_t fw0, fwn;          // first word and word count with decisions
    int32_t force_dir;          // 1 x, 2 y
};

hipError_t launch_step(int kind, const StepArgs& a, hipStream_t s);

// Temporally blocked step: K steps per pass, tiles of `th` rows.
struct TileArgs {
    Geometry g;
    const uint32_t* src;
    uint32_t* dst;
    const uint32_t* mask;
    const int32_t* row_range;
    uint32_t key0, key1;
    uint32_t t0;                // first step of the pass (mod 2^32), any value
    uint32_t th;                // rows per tile
    uint32_t halo;              // lanes per wave edge that are computed but not stored
    const uint32_t* fdec;       // decisions of step t0, or null
    uint64_t fdec_stride;       // words between consecutive steps' decisions
    uint32_t fy0, fy1, fw0, fwn;
    int32_t force_dir;
};
uint32_t tile_halo(const Geometry& g, int k);    // smallest valid TileArgs::halo
uint32_t tile_strips(const Geometry& g, int k);  // x-strips per row tile in strip mode
// Segment mode: rows of W <= 64 words, S = 64 / W whole rows per wave (one per
// row tile), used when it beats strips (S >= 2, or a row needing two strips).
bool tile_segments(const Geometry& g, int k);
// Does strip st of a pass (words [st * stride - halo, + 64) of its rows, taken
// modulo W) touch words [lo, hi) of those rows?
bool tile_strip_hits(const Geometry& g, uint32_t stride, uint32_t halo, uint32_t st, int64_t lo, int64_t hi);
// A pass of k steps; `order` is a dispatch order from tile_wave_order for the
// same arguments and k, in device memory, with n_order entries, or null.
hipError_t launch_tile(int kind, int k, const TileArgs& a, hipStream_t s, const uint32_t* order = nullptr,
                       uint32_t n_order = 0);
// Strip-mode dispatch order of a pass, from the rows' non-fluid word ranges
// row_lo / row_hi on the host (null when the lattice has no solids): one
// entry per wave, (tile * strips + strip) << kPartBits | part, where part 0 is
// the whole row tile and parts 1-2 and 3-6 its halves and quarters
// (tile_part_rows).  Every row of every tile and strip is in exactly one
// entry.  False in segment mode, where the order is fixed.
constexpr uint32_t kPartBits = 3;
bool tile_wave_order(const TileArgs& a, int k, const int32_t* row_lo, const int32_t* row_hi,
                     std::vector<uint32_t>& order);
// Rows [off, off + n) of a tile of `rows` rows that part `part` covers: parts
// of an even number of rows each (the last takes the rest), so that a part of
// a tile that starts on an even row starts on one too.
__host__ __device__ inline void tile_part_rows(uint32_t rows, uint32_t part, uint32_t& off, uint32_t& n)
{
    const uint32_t f = part == 0 ? 1u : part <= 2 ? 2u : 4u;  // parts of the tile
    const uint32_t idx = part == 0 ? 0u : part <= 2 ? part - 1u : part - 3u;
    const uint32_t q = (rows / f) & ~1u;
    off = idx * q;
    n = idx + 1 == f ? rows - off : q;
}

// Strip-mode wave order of tile_kernel (one wave per workgroup, one per row
// tile and strip): entry blockIdx.x of `order` when there is one, otherwise
// the heavy strips (those over forced words, at most kMaxHeavy, ascending) of
// every row tile before the others (nheavy 0: plain order).
constexpr uint32_t kMaxHeavy = 32;
struct TileOrder {
    uint32_t nstrips;           // strips per row tile
    uint32_t ntiles;            // row tiles
    uint32_t nheavy;            // heavy strips
    uint32_t heavy[kMaxHeavy];  // their indices, ascending
    const uint32_t* order;      // device order (tile_wave_order), or null
};

// Several K-step passes in one launch (tile_pipe_kernel): one row band per
// XCD, each band's work items handed out pass after pass, each item waiting
// for the previous pass of the tiles it reads.
constexpr uint32_t kMaxBands = 8;
struct PipeArgs {
    uint32_t* buf[2];            // pass p reads buf[p & 1] and writes buf[(p + 1) & 1]
    uint32_t npass;              // passes in the launch
    uint32_t nband;              // row bands, one per XCD
    uint32_t nstrips, ntiles;
    uint32_t* queue;             // [nband] tickets handed out (0 at launch)
    uint32_t* done;              // [ntiles][nstrips] passes each item has finished (0 at launch)
    uint32_t* error;             // set when a wait gives up (0 at launch)
    const uint32_t* order;       // the bands' items (tile * nstrips + strip), band after band
    uint32_t band_first[kMaxBands], band_items[kMaxBands];
    uint32_t band_y0[kMaxBands], band_y1[kMaxBands];
    uint32_t* exch;              // [2][nband][2][K][Q][P] words: each band's first and last K rows
    uint64_t fdec_pass;          // decision words per pass (K steps)
};
// Band partition and item order of a pipelined launch: nband bands of whole
// row tiles with about equal work (walks over solids counted twice), items
// tile-major within each band.  False when the shape does not allow one
// (segment mode, a short last word, K not 4 or 8, th < K, fewer than two
// tiles or 2K rows per band).
struct PipePlan {
    std::vector<uint32_t> order;
    uint32_t nband = 0, nstrips = 0, ntiles = 0;
    uint32_t band_first[kMaxBands] = {}, band_items[kMaxBands] = {};
    uint32_t band_y0[kMaxBands] = {}, band_y1[kMaxBands] = {};
};
bool tile_pipe_plan(const TileArgs& a, int k, uint32_t nband, const int32_t* row_lo, const int32_t* row_hi,
                    PipePlan& plan);
size_t tile_pipe_exch_words(const Geometry& g, int k, uint32_t nband);
// Launches p.npass passes of k steps from a.t0 (a.src / a.dst unused; p.buf
// holds the buffers) on nwaves waves; p's counters must be zero.
hipError_t launch_tile_pipe(int kind, int k, const TileArgs& a, const PipeArgs& p, uint32_t nwaves,
                            hipStream_t s);

// Batched ensembles.  BatchJob is a job's static descriptor (its TileArgs
// without src/dst/t0/th/halo, which come from BatchState and WorkItem);
// BatchState is the job's dynamic state on the device; a WorkItem is one
// wave's tile.
struct BatchJob {
    TileArgs a;            // geometry, masks, keys, force region and window base
    uint32_t* buf0;
    uint32_t* buf1;
    uint32_t force_q;      // per-cell force sampling parameter (force_q)
    uint32_t fx0, fx1;     // forced columns [fx0, fx1)
};
struct BatchState {
    uint32_t step;         // absolute step (mod 2^32)
    uint32_t cur;          // 0: state in buf0, 1: in buf1
    uint32_t fdec_step0;   // step of the first decision plane at a.fdec
    uint32_t done;         // waves of the current round that have finished
};
struct WorkItem {
    uint32_t job, strip, tile;  // segment mode: strip 0, tile = group of 64 / W row tiles
    uint32_t shape;        // rows per tile | halo << 16 | segment mode << 31
    uint32_t job_items;    // items of this job in the round
};
// One class launch of a round: every item (all of one K and raggedness, any
// models) advances its job K steps reading its BatchState.  The wave that
// finishes a job's last item of the round moves the job K steps on and swaps
// its buffers, for the next round's launches to read.
// n_heavy: the class's leading items that can take the general walk.
hipError_t launch_tile_batch(int k, bool ragged, const BatchJob* jobs, BatchState* state,
                             const WorkItem* items, uint32_t n_items, uint32_t n_heavy, hipStream_t s);

// Launchers of one (model, K) pair and of one K's batch kernel, instantiated
// by lgca_tile_inst.hip; grid and shapes come checked from launch_tile /
// launch_tile_batch.
// seg: 0 strip mode, 1 segment mode, 2 segment mode with rows of 64 words
template <int KIND, int K>
hipError_t launch_tile_kk(const TileArgs& a, const TileOrder& o, uint32_t nwaves, bool ragged,
                          int seg, hipStream_t s);
template <int K>
hipError_t launch_tile_batch_kk(const BatchJob* jobs, BatchState* state, const WorkItem* items,
                                uint32_t n_items, uint32_t n_heavy, bool ragged, hipStream_t s);
template <int KIND, int K>
hipError_t launch_tile_pipe_kk(const TileArgs& a, const PipeArgs& p, uint32_t nwaves, hipStream_t s);
template <int KIND, int K>
int tile_waves_per_cu_kk(bool ragged);
// Waves of the strip-mode tiled kernel of model `kind` and k steps per pass
// that a CU holds at once, or 0 when unknown.
int tile_waves_per_cu(int kind, int k, bool ragged);

struct ForceDrawArgs {
    uint32_t* dec;              // [steps][fy1 - fy0][fwn]
    uint32_t nsteps;
    uint32_t step0;
    uint32_t key0, key1;
    uint32_t q;                 // first-fire sampling parameter, see force_q
    uint32_t fx0, fx1, fy0, fy1, fw0, fwn;
    uint32_t dim_x;
};
hipError_t launch_force_draw(const ForceDrawArgs& a, hipStream_t s);

// Hash threshold of probability p: draws (uniform 32-bit) below it fire.
inline uint64_t hash_threshold(double p)
{
    if (!(p > 0.0)) return 0;
    if (p >= 1.0) return (uint64_t)1 << 32;
    return (uint64_t)(p * 4294967296.0);  // p * 2^32 < 2^32: truncation is floor
}

// First-fire parameter of the per-cell force for threshold thr in [1, 2^32]:
// q = 2^32 - thr (see force_word in lgca_dev.hpp).
inline uint32_t force_q(uint64_t thr) { return (uint32_t)(((uint64_t)1 << 32) - thr); }

hipError_t launch_pack(const uint8_t* state, uint32_t* planes, const Geometry& g,
                       uint32_t* err, hipStream_t s);
hipError_t launch_unpack(const uint32_t* planes, uint8_t* state, const Geometry& g, hipStream_t s);
hipError_t launch_types(const uint8_t* types, uint32_t* mask, int32_t* row_range,
                        const Geometry& g, uint32_t* err, hipStream_t s);
hipError_t launch_types_unpack(const uint32_t* mask, uint8_t* types, const Geometry& g,
                               hipStream_t s);
hipError_t launch_init_fill(uint32_t* planes, const uint32_t* mask, const uint8_t* region,
                            const Geometry& g, uint32_t key0, uint32_t key1, uint64_t thr,
                            hipStream_t s);
hipError_t launch_cell_fields(int kind, const uint32_t* planes, const Geometry& g,
                              float* density, float* momentum, hipStream_t s);
hipError_t launch_coarse(int kind, const uint32_t* planes, const Geometry& g, uint32_t r,
                         float* mean_density, float* mean_momentum, float* velocity,
                         hipStream_t s);
hipError_t launch_count(const uint32_t* words, uint64_t nwords, unsigned long long* out,
                        hipStream_t s);
hipError_t launch_mean_velocity(int kind, const uint32_t* planes, const uint32_t* mask,
                                const Geometry& g, long long* acc3, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint32_t value, uint64_t n, hipStream_t s);

}  // namespace lgca
