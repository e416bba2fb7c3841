// lgca_batch.hip -- batched ensembles and serving: the lgca_batch_* entry
// points of include/lgca.h.
//
// A batch keeps every admitted job's bit-planes resident in one device arena
// and advances all of them together.  A round gives each running job k steps
// (k = the tuned steps per pass, less at a job's end) and is one launch of
// tile_batch_kernel per (model, k, ragged width) class; the wave that finishes
// a job's last work item of the round moves the job's step and buffer parity
// on the device.
// Work items (one per wave: job, x-strip, row tile, tile shape) are rebuilt
// only when the set of running jobs or the tuning changes, so a steady stretch
// of rounds is launches alone.  Per-cell force decisions are drawn for windows
// of kWindow steps by one launch over all forced jobs that need a new window.
//
// A worker thread owns the GPU stream: it admits queued jobs in order while
// their memory fits the HBM budget (and their cells the active-cell cap),
// queues rounds at most kInflight ahead of the GPU, post-processes finished
// jobs (coarse fields and state bytes) and marks them done when the GPU has
// passed their completion event.  Client calls touch only host-side tables,
// except lgca_batch_result, which copies on a stream of its own.
//
// Tiles: every job in a round uses the same tile height th, chosen so the
// round has about `wave_target` waves; the autotuner times rounds with each
// (steps per pass, wave target) candidate on the live mix and keeps the
// fastest per site update, and tunes again when the mix has changed by 2x.
// Results never depend on the tuning: the hash is keyed by (seed, step, cell).
#include "lgca.h"

#include "lgca_dev.hpp"
#include "lgca_kernels.hpp"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

using namespace lgca;

namespace lgca {

// One admitted job: its job-table entry and how to set it up on the device.
struct JobInstall {
    BatchJob job;
    uint32_t slot;
    int32_t geometry;      // LGCA_GEOM_*
    int32_t cx, cy;        // disc centre
    uint32_t halfw_off;    // disc half-widths at halfw[halfw_off .. + nh)
    uint32_t nh;
    uint32_t generated;    // 1: cell types from the geometry and state from the hash fill
    uint32_t pad;
    uint64_t fill_thr;
};

__global__ __launch_bounds__(256) void install_kernel(BatchJob* jobs, BatchState* state,
                                                      const JobInstall* __restrict__ rec, uint32_t n)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    jobs[rec[i].slot] = rec[i].job;
    state[rec[i].slot] = BatchState{0u, 0u, 0u, 0u};
}

// Bits of word w whose cells x = 32w + b lie in [xa, xb].
__device__ __forceinline__ uint32_t bits_in_range(uint32_t w, int64_t xa, int64_t xb)
{
    const int64_t lo = max(xa - 32 * (int64_t)w, (int64_t)0);
    const int64_t hi = min(xb - 32 * (int64_t)w, (int64_t)31);
    if (lo > hi) return 0u;
    const uint32_t n = (uint32_t)(hi - lo + 1);
    return (n == 32 ? 0xFFFFFFFFu : ((1u << n) - 1u)) << (uint32_t)lo;
}

// Set-up of the generated jobs of one admission, blockIdx.y = record: NO_SLIP
// walls on rows 0 and dim_y - 1 (pipe, cylinder) and the disc |y - cy| < nh,
// |x - cx| <= halfw[|y - cy|] (cylinder) as mask planes with their per-row
// word ranges, the hash fill of the FLUID cells in buffer 0 and zeros in
// buffer 1 (padding words included).
__global__ __launch_bounds__(256) void setup_kernel(const JobInstall* __restrict__ rec,
                                                    const int32_t* __restrict__ halfw_all)
{
    const JobInstall& r = rec[blockIdx.y];
    if (!r.generated) return;
    const Geometry g = r.job.a.g;
    uint32_t* mask = const_cast<uint32_t*>(r.job.a.mask);
    int32_t* rr = const_cast<int32_t*>(r.job.a.row_range);
    const int32_t* hw = halfw_all + r.halfw_off;
    const uint64_t n = (uint64_t)g.dim_y * g.P;
    for (uint64_t idx = (uint64_t)blockIdx.x * 256 + threadIdx.x; idx < n; idx += (uint64_t)gridDim.x * 256) {
        const uint32_t y = (uint32_t)(idx / g.P), w = (uint32_t)(idx % g.P);
        const uint32_t valid = w + 1 < g.W ? 0xFFFFFFFFu
                             : (w + 1 == g.W ? (g.nb == 32 ? 0xFFFFFFFFu : (1u << g.nb) - 1u) : 0u);
        uint32_t ns = 0;
        int32_t lo = 0x7FFFFFFF, hi = 0;
        if (r.geometry != LGCA_GEOM_PERIODIC && (y == 0 || y + 1 == g.dim_y)) {
            ns = valid;
            lo = 0;
            hi = (int32_t)g.W;
        }
        if (r.geometry == LGCA_GEOM_CYLINDER) {
            const int64_t dy = abs((int64_t)y - r.cy);
            if (dy < (int64_t)r.nh) {
                const int64_t xa = max((int64_t)r.cx - hw[dy], (int64_t)0);
                const int64_t xb = min((int64_t)r.cx + hw[dy], (int64_t)g.dim_x - 1);
                if (xa <= xb) {
                    ns |= bits_in_range(w, xa, xb) & valid;
                    lo = min(lo, (int32_t)(xa >> 5));
                    hi = max(hi, (int32_t)(xb >> 5) + 1);
                }
            }
        }
        if (mask) {
            mask[((size_t)y * 2 + 0) * g.P + w] = ns;
            mask[((size_t)y * 2 + 1) * g.P + w] = 0u;
            if (w == 0) {
                rr[y] = lo;
                rr[g.dim_y + y] = hi;
            }
        }
        const uint32_t fluid = valid & ~ns;
        for (uint32_t d = 0; d < g.Q; ++d) {
            const size_t at = ((size_t)y * g.Q + d) * g.P + w;
            r.job.buf0[at] = fluid ? (fill_word(w, y, d, g.dim_x, r.job.a.key0, r.job.a.key1, r.fill_thr) & fluid) : 0u;
            r.job.buf1[at] = 0u;
        }
    }
}

// Force windows of many jobs: blockIdx.y picks the job, threads stride over
// its force region's (row, word) pairs and draw every four-step group of the
// window for theirs.  The window starts at the job's current device step,
// which becomes its fdec_step0.  (A thread per (group, row, word) triple paid
// two 32-bit divisions per triple to find its row and word.)
__global__ __launch_bounds__(256) void batch_draw_kernel(const BatchJob* __restrict__ jobs,
                                                         BatchState* state,
                                                         const uint32_t* __restrict__ slots,
                                                         uint32_t nsteps)
{
    __shared__ uint32_t Qt[32];
    const uint32_t slot = slots[blockIdx.y];
    const BatchJob& j = jobs[slot];
    force_table(Qt, j.force_q);
    const uint32_t step0 = state[slot].step;
    const uint32_t rows = j.a.fy1 - j.a.fy0, fwn = j.a.fwn;
    const uint32_t groups = ((step0 & 3u) + nsteps + 3u) / 4u;
    const uint64_t plane = (uint64_t)rows * fwn;
    uint32_t* dec = const_cast<uint32_t*>(j.a.fdec);
    for (uint64_t it = (uint64_t)blockIdx.x * 256 + threadIdx.x; it < plane; it += (uint64_t)gridDim.x * 256) {
        uint32_t yy, ww;
        if (plane < ((uint64_t)1 << 32)) {  // 32-bit division, the usual case
            yy = (uint32_t)it / fwn;
            ww = (uint32_t)it - yy * fwn;
        } else {
            yy = (uint32_t)(it / fwn);
            ww = (uint32_t)(it - (uint64_t)yy * fwn);
        }
        const uint32_t w = j.a.fw0 + ww;
        const uint32_t cols = force_columns(w, j.fx0, j.fx1, j.a.g.dim_x);
        for (uint32_t g = 0; g < groups; ++g)
            draw_group(dec + it, plane, nsteps, step0, g, w, j.a.fy0 + yy, j.a.key0, j.a.key1, Qt, cols);
    }
    // fdec_step0 is not the field the other blocks read
    if (blockIdx.x == 0 && threadIdx.x == 0) state[slot].fdec_step0 = step0;
}

}  // namespace lgca

namespace {

constexpr uint32_t kWindow = 32;          // steps per force window
constexpr int kInflight = 4;              // rounds queued ahead of the GPU
constexpr size_t kAlign = 256;
constexpr size_t kStagingHalf = 32u << 20;  // pinned upload staging, two halves
constexpr uint32_t kMinTileRows = 8, kMaxTileRows = 512;
constexpr int kTuneRounds = 3;            // rounds timed per candidate

size_t align_up(size_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int kind_of_model(int model) { return model == LGCA_HPP ? kHPP : (model == LGCA_FHP_I ? kFHP1 : kFHP2); }

// The reference's disc test (src/lattice.cpp:264-279): dist is a float
// rounded from the double sqrt, compared in double with diameter / 2.0.
bool inside_disc(int64_t dx, int64_t dy, float diameter)
{
    const float dist = (float)std::sqrt((double)(dx * dx + dy * dy));
    return (double)dist < (double)diameter / 2.0;
}

// Half-widths of the disc per |dy|: halfw[dy] = the largest dx >= 0 with
// (dx, dy) inside; the table stops at the first row with none.
std::vector<int32_t> disc_halfwidths(float diameter)
{
    std::vector<int32_t> hw;
    if (!(diameter > 0.0f)) return hw;
    int64_t dx = (int64_t)std::ceil((double)diameter / 2.0) + 1;
    for (int64_t dy = 0;; ++dy) {
        while (dx >= 0 && !inside_disc(dx, dy, diameter)) --dx;
        if (dx < 0) break;
        hw.push_back((int32_t)dx);
    }
    return hw;
}

// Device memory of one job, as offsets into one arena block.
struct JobLayout {
    size_t buf, mask, row_range, bytes, md, mm, vel, fdec, halfw, total;
};

struct JobParams {
    int kind = 0;
    Geometry g{};
    size_t state_words = 0;
    bool has_mask = false;
    bool forced = false;
    uint64_t force_thr = 0;
    uint32_t fx0 = 0, fx1 = 0, fy0 = 0, fy1 = 0, fw0 = 0, fwn = 0;
    uint64_t nc = 0;
};

JobParams job_params(const lgca_job_desc& d)
{
    JobParams p;
    p.kind = kind_of_model(d.model);
    p.g = make_geometry(d.dim_x, d.dim_y, (uint32_t)num_dir_of_kind(p.kind));
    p.state_words = (size_t)d.dim_y * p.g.Q * p.g.P;
    p.has_mask = d.geometry != LGCA_GEOM_PERIODIC;
    p.force_thr = hash_threshold(d.force_p);
    p.fx0 = d.fx0;
    p.fx1 = std::min(d.fx1, d.dim_x);
    p.fy0 = d.fy0;
    p.fy1 = std::min(d.fy1, d.dim_y);
    p.forced = d.force_dir != LGCA_FORCE_NONE && p.force_thr != 0 && p.fx0 < p.fx1 && p.fy0 < p.fy1;
    if (p.forced) {
        p.fw0 = p.fx0 / 32;
        p.fwn = (p.fx1 + 31) / 32 - p.fw0;
    }
    p.nc = (uint64_t)(d.dim_x / (2 * d.coarse_radius)) * (d.dim_y / (2 * d.coarse_radius));
    return p;
}

JobLayout job_layout(const lgca_job_desc& d, const JobParams& p, size_t n_halfw)
{
    JobLayout L{};
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off += align_up(bytes ? bytes : 1);
        return at;
    };
    L.buf = take(2 * p.state_words * 4);
    L.mask = p.has_mask ? take((size_t)d.dim_y * 2 * p.g.P * 4) : SIZE_MAX;
    L.row_range = p.has_mask ? take((size_t)d.dim_y * 2 * 4) : SIZE_MAX;
    L.bytes = take((size_t)d.dim_x * d.dim_y);
    L.md = take(p.nc * 4);
    L.mm = take(p.nc * 8);
    L.vel = take(p.nc * 8);
    L.fdec = p.forced ? take((size_t)kWindow * (p.fy1 - p.fy0) * p.fwn * 4) : SIZE_MAX;
    L.halfw = n_halfw ? take(n_halfw * 4) : SIZE_MAX;
    L.total = off;
    return L;
}

struct Job {
    uint64_t id = 0;
    uint32_t readers = 0;  // lgca_batch_result copies in flight (under the batch's mutex)
    lgca_job_desc d{};
    std::vector<uint8_t> types_in, state_in;
    std::vector<int32_t> halfw;
    JobParams p;
    JobLayout L{};
    int status = LGCA_JOB_QUEUED;
    double t_submit = 0, t_admit = 0, t_done = 0;
    // worker-side
    char* mem = nullptr;
    int slot = -1;
    uint64_t step = 0;       // steps queued
    uint32_t cur = 0;        // buffer holding the state after the queued steps
    uint64_t window_end = 0; // force decisions exist for steps < window_end
    hipEvent_t done_ev = nullptr;
    hipEvent_t last_ev = nullptr;  // end of the job's last round on the main stream
};

// First-fit allocator over a few large hipMalloc blocks.
struct Arena {
    struct Block {
        char* base;
        size_t size;
        std::map<size_t, size_t> free_;  // offset -> length
    };
    std::vector<Block> blocks;
    size_t reserved = 0;
    size_t block_bytes = (size_t)1 << 30;

    hipError_t alloc(size_t bytes, char** out)
    {
        for (Block& b : blocks) {
            for (auto it = b.free_.begin(); it != b.free_.end(); ++it) {
                if (it->second < bytes) continue;
                const size_t off = it->first, rem = it->second - bytes;
                b.free_.erase(it);
                if (rem) b.free_[off + bytes] = rem;
                *out = b.base + off;
                return hipSuccess;
            }
        }
        const size_t sz = std::max(bytes, block_bytes);
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, sz);
        if (e != hipSuccess) return e;
        blocks.push_back(Block{(char*)p, sz, {}});
        if (sz > bytes) blocks.back().free_[bytes] = sz - bytes;
        reserved += sz;
        *out = (char*)p;
        return hipSuccess;
    }

    void free(char* p, size_t bytes)
    {
        for (Block& b : blocks) {
            if (p < b.base || p >= b.base + b.size) continue;
            size_t off = (size_t)(p - b.base), len = bytes;
            auto next = b.free_.lower_bound(off);
            if (next != b.free_.end() && next->first == off + len) {
                len += next->second;
                next = b.free_.erase(next);
            }
            if (next != b.free_.begin()) {
                auto prev = std::prev(next);
                if (prev->first + prev->second == off) {
                    off = prev->first;
                    len += prev->second;
                    b.free_.erase(prev);
                }
            }
            b.free_[off] = len;
            return;
        }
    }

    void release_all()
    {
        for (Block& b : blocks) (void)hipFree(b.base);
        blocks.clear();
        reserved = 0;
    }
};

struct ClassRange {
    int k;
    bool ragged;
    uint32_t offset, count;
    uint32_t heavy;  // leading items that can take the general walk
};

struct Candidate {
    int k;
    uint32_t waves;
};

}  // namespace

struct lgca_batch {
    int device = 0;
    uint64_t budget = 0;
    hipStream_t stream = nullptr;       // worker: set-up and rounds
    hipStream_t post_stream = nullptr;  // worker: post-processing of finished jobs
    hipStream_t copy_stream = nullptr;  // lgca_batch_result
    std::thread worker;

    std::mutex mu;
    std::condition_variable cv_work;    // submissions, releases, stop
    std::condition_variable cv_done;    // completions
    std::condition_variable cv_readers; // a job's result copies finished
    bool stop = false;
    std::string err;                    // sticky worker failure
    std::string call_err;               // last failed client call
    uint64_t next_id = 1;
    std::unordered_map<uint64_t, std::unique_ptr<Job>> jobs;
    std::deque<Job*> queue;             // submitted, not admitted
    std::deque<uint64_t> done_ids;      // for lgca_batch_poll
    uint64_t bytes_in_use = 0;
    uint64_t n_submitted = 0, n_done = 0;
    Arena arena;

    // tuning (under mu)
    int user_k = 0;
    uint32_t user_waves = 0;
    uint64_t max_active_cells = 0;
    bool tuning_changed = false;

    // worker-private
    std::vector<Job*> active;
    uint64_t active_cells = 0;
    std::vector<Job*> slot_job;
    std::vector<int> free_slots;
    BatchJob* d_jobs = nullptr;
    BatchState* d_state = nullptr;
    BatchJob* h_jobs = nullptr;         // pinned mirror of the job table
    BatchState* h_state = nullptr;
    uint32_t cap = 0;
    WorkItem* d_items[2] = {nullptr, nullptr};
    size_t items_cap[2] = {0, 0};
    int items_buf = 0;
    uint32_t* d_draw = nullptr;
    size_t draw_cap = 0;
    char* d_setup = nullptr;            // admission records and disc tables
    size_t setup_cap = 0;
    uint32_t* d_err = nullptr;          // sink for the codec kernels' flags
    char* staging = nullptr;            // pinned, two halves
    size_t staging_used = 0;
    int staging_half = 0;
    hipEvent_t staging_ev[2] = {nullptr, nullptr};
    std::vector<ClassRange> classes;
    std::vector<uint64_t> sig, last_sig;
    std::deque<hipEvent_t> inflight;
    std::vector<hipEvent_t> event_pool;
    std::vector<Job*> finishing;        // post-processed, waiting for their event

    // autotuner (worker-private)
    int k_cur = 4;
    uint32_t waves_cur = 4096;
    std::vector<Candidate> cands;
    size_t cand_i = 0;
    int cand_rounds = 0;
    std::vector<hipEvent_t> tune_ev;    // start/end per candidate
    std::vector<uint64_t> tune_su;
    bool tuning = false, evaluating = false;
    uint64_t tuned_wave_rows = 0;       // sum of strips * rows when last tuned
    uint64_t rounds_since_tune = 0;
    uint64_t n_tunings = 0;

    // stats (under mu)
    uint64_t rounds = 0, launches = 0, site_updates = 0;
    int stat_k = 0;
    uint32_t stat_waves = 0;
    uint64_t stat_tunings = 0;
};

namespace {

lgca_status set_call_error(lgca_batch* b, lgca_status code, const std::string& msg)
{
    if (b) {
        std::lock_guard<std::mutex> lk(b->mu);
        b->call_err = msg;
    }
    return code;
}

thread_local std::string g_batch_create_error;

struct DevGuard {
    int prev = -1;
    explicit DevGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DevGuard()
    {
        int now = -1;
        if (prev >= 0 && hipGetDevice(&now) == hipSuccess && now != prev) (void)hipSetDevice(prev);
    }
};

std::string validate(const lgca_job_desc& d)
{
    if (d.model < LGCA_HPP || d.model > LGCA_FHP_III) return "unknown model";
    if (d.dim_x == 0 || d.dim_y == 0 || d.coarse_radius == 0) return "zero size or radius";
    if (d.model != LGCA_HPP && (d.dim_y % 2) != 0) return "FHP models need an even dim_y";
    if (d.dim_x < 2 * d.coarse_radius || d.dim_y < 2 * d.coarse_radius)
        return "dim_x and dim_y must be at least 2 * coarse_radius";
    if ((uint64_t)d.dim_x * d.dim_y >= (1ull << 40) || d.dim_y > 65535 || d.dim_x >= (1u << 30))
        return "lattice too large (dim_y <= 65535, dim_x < 2^30)";
    if (d.steps >= (1ull << 32)) return "steps must be below 2^32";
    if (!d.state && !(d.fill >= 0.0 && d.fill <= 1.0)) return "fill must lie in [0, 1]";
    if (d.geometry < LGCA_GEOM_PERIODIC || d.geometry > LGCA_GEOM_TYPES) return "unknown geometry";
    if (d.geometry == LGCA_GEOM_TYPES && !d.cell_types) return "LGCA_GEOM_TYPES needs cell_types";
    if (d.geometry == LGCA_GEOM_CYLINDER && !(d.cyl_diameter >= 0.0 && d.cyl_diameter < 1e9))
        return "cylinder diameter must lie in [0, 1e9)";
    if (d.force_dir != LGCA_FORCE_NONE && d.force_dir != LGCA_FORCE_X && d.force_dir != LGCA_FORCE_Y)
        return "force_dir must be NONE, X or Y";
    if (d.force_dir != LGCA_FORCE_NONE && !(d.force_p >= 0.0 && d.force_p <= 1.0))
        return "force_p must lie in [0, 1]";
    if (d.force_dir != LGCA_FORCE_NONE) {
        // a pass reads its steps' decisions through one buffer descriptor
        const uint64_t x1 = std::min(d.fx1, d.dim_x), y1 = std::min(d.fy1, d.dim_y);
        if (d.fx0 < x1 && d.fy0 < y1 &&
            (y1 - d.fy0) * ((x1 + 31) / 32 - d.fx0 / 32) * 8u * 4u >= (1ull << 31))
            return "force region too large (the decisions of an 8-step pass must stay under 2 GiB)";
    }
    return "";
}

// ---------------------------------------------------------------- worker

struct WorkerError {
    std::string msg;
};

void hipchk(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw WorkerError{std::string(what) + ": " + hipGetErrorString(e)};
}

hipEvent_t take_event(lgca_batch* b)
{
    if (!b->event_pool.empty()) {
        hipEvent_t e = b->event_pool.back();
        b->event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    hipchk(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    return e;
}

// Copies host bytes to the device through the pinned staging halves.
void upload(lgca_batch* b, void* dst, const void* src, size_t bytes)
{
    const char* s = (const char*)src;
    char* d = (char*)dst;
    while (bytes) {
        if (b->staging_used == kStagingHalf) {
            // switch halves; the other one is free once its last copy ran
            hipchk(hipEventRecord(b->staging_ev[b->staging_half], b->stream), "hipEventRecord");
            b->staging_half ^= 1;
            hipchk(hipEventSynchronize(b->staging_ev[b->staging_half]), "hipEventSynchronize");
            b->staging_used = 0;
        }
        const size_t n = std::min(bytes, kStagingHalf - b->staging_used);
        char* st = b->staging + (size_t)b->staging_half * kStagingHalf + b->staging_used;
        std::memcpy(st, s, n);
        hipchk(hipMemcpyAsync(d, st, n, hipMemcpyHostToDevice, b->stream), "hipMemcpyAsync(upload)");
        // keep 256-byte alignment of later staging regions
        b->staging_used = std::min(kStagingHalf, b->staging_used + align_up(n));
        s += n;
        d += n;
        bytes -= n;
    }
}

void grow_tables(lgca_batch* b, uint32_t need)
{
    if (need <= b->cap) return;
    uint32_t cap = std::max<uint32_t>(b->cap ? 2 * b->cap : 1024, need);
    hipchk(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    BatchJob* dj = nullptr;
    BatchState* ds = nullptr;
    BatchJob* hj = nullptr;
    BatchState* hs = nullptr;
    hipchk(hipMalloc((void**)&dj, (size_t)cap * sizeof(BatchJob)), "hipMalloc(job table)");
    hipchk(hipMalloc((void**)&ds, (size_t)cap * sizeof(BatchState)), "hipMalloc(state table)");
    hipchk(hipHostMalloc((void**)&hj, (size_t)cap * sizeof(BatchJob), hipHostMallocDefault), "hipHostMalloc");
    hipchk(hipHostMalloc((void**)&hs, (size_t)cap * sizeof(BatchState), hipHostMallocDefault), "hipHostMalloc");
    if (b->cap) {
        hipchk(hipMemcpy(dj, b->d_jobs, (size_t)b->cap * sizeof(BatchJob), hipMemcpyDeviceToDevice), "copy");
        hipchk(hipMemcpy(ds, b->d_state, (size_t)b->cap * sizeof(BatchState), hipMemcpyDeviceToDevice), "copy");
        std::memcpy(hj, b->h_jobs, (size_t)b->cap * sizeof(BatchJob));
        std::memcpy(hs, b->h_state, (size_t)b->cap * sizeof(BatchState));
        (void)hipFree(b->d_jobs);
        (void)hipFree(b->d_state);
        (void)hipHostFree(b->h_jobs);
        (void)hipHostFree(b->h_state);
    }
    for (uint32_t i = cap; i-- > b->cap;) b->free_slots.push_back((int)i);
    b->slot_job.resize(cap, nullptr);
    b->d_jobs = dj;
    b->d_state = ds;
    b->h_jobs = hj;
    b->h_state = hs;
    b->cap = cap;
}

// Grows a worker-owned device buffer; the stream is drained first because
// queued work may still read the old one.
void grow_buffer(lgca_batch* b, char** buf, size_t* cap, size_t need, const char* what)
{
    if (need <= *cap) return;
    hipchk(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    if (*buf) (void)hipFree(*buf);
    *buf = nullptr;
    *cap = 0;
    const size_t c = std::max(need + need / 2, (size_t)1 << 20);
    hipchk(hipMalloc((void**)buf, c), what);
    *cap = c;
}

// Sets up newly admitted jobs: one install launch writes their job-table
// entries and zero device states, one set-up launch builds the masks and hash
// fill of every job with a built-in geometry; jobs with given cell types or
// state take the upload and codec path.
void admit_all(lgca_batch* b, const std::vector<Job*>& jobs)
{
    if (jobs.empty()) return;
    if (b->free_slots.size() < jobs.size()) grow_tables(b, b->cap + (uint32_t)jobs.size());
    std::vector<JobInstall> rec(jobs.size());
    std::vector<int32_t> halfw;
    for (size_t i = 0; i < jobs.size(); ++i) {
        Job* j = jobs[i];
        j->slot = b->free_slots.back();
        b->free_slots.pop_back();
        b->slot_job[j->slot] = j;
        const JobParams& p = j->p;
        char* m = j->mem;
        JobInstall& r = rec[i];
        std::memset(&r, 0, sizeof r);
        BatchJob& e = r.job;
        e.a.g = p.g;
        e.a.mask = p.has_mask ? (const uint32_t*)(m + j->L.mask) : nullptr;
        e.a.row_range = p.has_mask ? (const int32_t*)(m + j->L.row_range) : nullptr;
        e.a.key0 = (uint32_t)j->d.seed;
        e.a.key1 = (uint32_t)(j->d.seed >> 32);
        if (p.forced) {
            e.a.fdec = (const uint32_t*)(m + j->L.fdec);
            e.a.fdec_stride = (uint64_t)(p.fy1 - p.fy0) * p.fwn;
            e.a.fy0 = p.fy0;
            e.a.fy1 = p.fy1;
            e.a.fw0 = p.fw0;
            e.a.fwn = p.fwn;
            e.a.force_dir = j->d.force_dir;
            e.force_q = force_q(p.force_thr);
            e.fx0 = p.fx0;
            e.fx1 = p.fx1;
        }
        e.buf0 = (uint32_t*)(m + j->L.buf);
        e.buf1 = e.buf0 + p.state_words;
        r.slot = (uint32_t)j->slot;
        r.geometry = j->d.geometry;
        r.cx = j->d.cyl_x;
        r.cy = j->d.cyl_y;
        r.halfw_off = (uint32_t)halfw.size();
        r.nh = (uint32_t)j->halfw.size();
        halfw.insert(halfw.end(), j->halfw.begin(), j->halfw.end());
        r.generated = j->d.geometry != LGCA_GEOM_TYPES && j->state_in.empty();
        r.fill_thr = hash_threshold(j->d.fill);
        b->h_jobs[j->slot] = e;
        j->step = 0;
        j->cur = 0;
        j->window_end = 0;
    }
    const size_t rec_bytes = align_up(rec.size() * sizeof(JobInstall));
    grow_buffer(b, &b->d_setup, &b->setup_cap, rec_bytes + halfw.size() * 4 + 16, "hipMalloc(set-up)");
    const JobInstall* d_rec = (const JobInstall*)b->d_setup;
    const int32_t* d_halfw = (const int32_t*)(b->d_setup + rec_bytes);
    upload(b, b->d_setup, rec.data(), rec.size() * sizeof(JobInstall));
    if (!halfw.empty()) upload(b, b->d_setup + rec_bytes, halfw.data(), halfw.size() * 4);
    const uint32_t n = (uint32_t)rec.size();
    hipLaunchKernelGGL(install_kernel, dim3((n + 255) / 256), dim3(256), 0, b->stream, b->d_jobs, b->d_state,
                       d_rec, n);
    hipchk(hipGetLastError(), "install_kernel");
    uint64_t most = 0;
    for (const Job* j : jobs) most = std::max<uint64_t>(most, (uint64_t)j->d.dim_y * j->p.g.P);
    const uint32_t bx = (uint32_t)std::min<uint64_t>((most + 255) / 256, 256);
    for (uint32_t o = 0; o < n; o += 65535) {
        const uint32_t ny = std::min<uint32_t>(65535, n - o);
        hipLaunchKernelGGL(setup_kernel, dim3(bx, ny), dim3(256), 0, b->stream, d_rec + o, d_halfw);
        hipchk(hipGetLastError(), "setup_kernel");
    }
    // given cell types or state
    for (size_t i = 0; i < jobs.size(); ++i) {
        if (rec[i].generated) continue;
        Job* j = jobs[i];
        const JobParams& p = j->p;
        char* m = j->mem;
        uint32_t* buf0 = (uint32_t*)(m + j->L.buf);
        uint32_t* mask = p.has_mask ? (uint32_t*)(m + j->L.mask) : nullptr;
        int32_t* rr = p.has_mask ? (int32_t*)(m + j->L.row_range) : nullptr;
        uint8_t* bytes = (uint8_t*)(m + j->L.bytes);
        hipchk(launch_fill_u32(buf0, 0u, 2 * (uint64_t)p.state_words, b->stream), "fill");
        if (p.has_mask) {
            if (j->d.geometry == LGCA_GEOM_TYPES) {
                upload(b, bytes, j->types_in.data(), j->types_in.size());
            } else {
                // built-in geometry with a given state: the set-up kernel's
                // masks, made here from bytes as for given types
                std::vector<uint8_t> t((size_t)j->d.dim_x * j->d.dim_y, (uint8_t)LGCA_FLUID);
                for (uint32_t x = 0; x < j->d.dim_x; ++x) {
                    t[x] = LGCA_SOLID_NO_SLIP;
                    t[(size_t)(j->d.dim_y - 1) * j->d.dim_x + x] = LGCA_SOLID_NO_SLIP;
                }
                for (size_t dy = 0; dy < j->halfw.size(); ++dy)
                    for (int sgn = -1; sgn <= 1; sgn += 2) {
                        const int64_t y = (int64_t)j->d.cyl_y + sgn * (int64_t)dy;
                        if (y < 0 || y >= (int64_t)j->d.dim_y || (sgn == 1 && dy == 0)) continue;
                        const int64_t xa = std::max<int64_t>((int64_t)j->d.cyl_x - j->halfw[dy], 0);
                        const int64_t xb = std::min<int64_t>((int64_t)j->d.cyl_x + j->halfw[dy], (int64_t)j->d.dim_x - 1);
                        for (int64_t x = xa; x <= xb; ++x) t[(size_t)y * j->d.dim_x + x] = LGCA_SOLID_NO_SLIP;
                    }
                upload(b, bytes, t.data(), t.size());
            }
            hipchk(launch_fill_u32((uint32_t*)rr, 0x7FFFFFFFu, j->d.dim_y, b->stream), "fill");
            hipchk(launch_fill_u32((uint32_t*)rr + j->d.dim_y, 0u, j->d.dim_y, b->stream), "fill");
            hipchk(launch_types(bytes, mask, rr, p.g, b->d_err, b->stream), "types");
        }
        if (!j->state_in.empty()) {
            upload(b, bytes, j->state_in.data(), j->state_in.size());
            hipchk(launch_pack(bytes, buf0, p.g, b->d_err, b->stream), "pack");
        } else {
            hipchk(launch_init_fill(buf0, mask, nullptr, p.g, (uint32_t)j->d.seed, (uint32_t)(j->d.seed >> 32),
                                    hash_threshold(j->d.fill), b->stream),
                   "init_fill");
        }
    }
}

// Coarse fields and state bytes of a finished job, then its completion event.
// They run on the post-processing stream, after the job's last round on the
// main stream: no later round touches the job's buffers, so the next rounds
// overlap them.
void finish(lgca_batch* b, Job* j)
{
    const JobParams& p = j->p;
    char* m = j->mem;
    const uint32_t* st = (const uint32_t*)(m + j->L.buf) + (j->cur ? p.state_words : 0);
    j->last_ev = take_event(b);
    hipchk(hipEventRecord(j->last_ev, b->stream), "hipEventRecord");
    hipchk(hipStreamWaitEvent(b->post_stream, j->last_ev, 0), "hipStreamWaitEvent");
    hipchk(launch_coarse(p.kind, st, p.g, j->d.coarse_radius, (float*)(m + j->L.md), (float*)(m + j->L.mm),
                         (float*)(m + j->L.vel), b->post_stream),
           "coarse");
    hipchk(launch_unpack(st, (uint8_t*)(m + j->L.bytes), p.g, b->post_stream), "unpack");
    j->done_ev = take_event(b);
    hipchk(hipEventRecord(j->done_ev, b->post_stream), "hipEventRecord");
    b->finishing.push_back(j);
}

// Marks finished jobs whose event has passed; `block` waits for all of them.
void reap(lgca_batch* b, bool block)
{
    if (b->finishing.empty()) return;
    std::vector<Job*> still;
    std::vector<Job*> done;
    for (Job* j : b->finishing) {
        hipError_t q = block ? hipEventSynchronize(j->done_ev) : hipEventQuery(j->done_ev);
        if (q == hipErrorNotReady) {
            still.push_back(j);
            continue;
        }
        hipchk(q, "hipEventQuery");
        done.push_back(j);
    }
    b->finishing.swap(still);
    if (done.empty()) return;
    const double t = now_s();
    {
        std::lock_guard<std::mutex> lk(b->mu);
        for (Job* j : done) {
            b->event_pool.push_back(j->done_ev);
            b->event_pool.push_back(j->last_ev);
            j->done_ev = nullptr;
            j->last_ev = nullptr;
            b->slot_job[j->slot] = nullptr;
            b->free_slots.push_back(j->slot);
            j->slot = -1;
            j->status = LGCA_JOB_DONE;
            j->t_done = t;
            b->done_ids.push_back(j->id);
            ++b->n_done;
        }
    }
    b->cv_done.notify_all();
}

int pick_k(int K, uint64_t left, uint64_t step)
{
    if (K >= 8 && left >= 8 && (step & 3) == 0) return 8;
    if (K >= 4 && left >= 4 && (step & 3) == 0) return 4;
    if (K >= 2 && left >= 2 && (step & 1) == 0) return 2;
    return 1;
}

// Waves a job's round would have with one-row tiles: strips x rows, or rows /
// (rows per wave) in segment mode.
uint64_t job_wave_rows(const Geometry& g, int k)
{
    if (tile_segments(g, k)) return ((uint64_t)g.dim_y + 64 / g.W - 1) / (64 / g.W);
    return (uint64_t)tile_strips(g, k) * g.dim_y;
}

uint64_t wave_rows(const lgca_batch* b, int k)
{
    uint64_t s = 0;
    for (const Job* j : b->active) s += job_wave_rows(j->p.g, k);
    return s;
}

uint32_t tile_rows_for(uint64_t wave_rows_total, uint32_t waves)
{
    uint64_t th = wave_rows_total / std::max<uint32_t>(waves, 1);
    th = std::max<uint64_t>(th, kMinTileRows);
    th = std::min<uint64_t>(th, kMaxTileRows);
    th += th & 1u;
    return (uint32_t)th;
}

// Whether a work item reading rows [a, b) of a job (on the torus) can take
// the general walk: rows with walls, the disc, given cell types or forcing.
// It only orders the items of a class (those first, so that waves that run
// longer start first and the rest fill in behind them); results never depend
// on it.
bool rows_heavy(const Job* j, int64_t a, int64_t b)
{
    const int64_t H = (int64_t)j->d.dim_y;
    if (b - a >= H) {
        a = 0;
        b = H;
    }
    auto hits = [&](int64_t lo, int64_t hi) {  // [lo, hi) inside [0, H)
        auto span = [&](int64_t s0, int64_t s1) { return s0 < hi && s1 > lo; };
        if (j->p.forced && span(j->p.fy0, j->p.fy1)) return true;
        if (j->d.geometry == LGCA_GEOM_PERIODIC) return false;
        if (j->d.geometry == LGCA_GEOM_TYPES) return true;
        if (span(0, 1) || span(H - 1, H)) return true;
        if (j->d.geometry == LGCA_GEOM_CYLINDER && !j->halfw.empty()) {
            const int64_t nh = (int64_t)j->halfw.size(), cy = j->d.cyl_y;
            if (span(cy - nh + 1, cy + nh)) return true;
        }
        return false;
    };
    const int64_t a0 = ((a % H) + H) % H, len = b - a;
    if (a0 + len <= H) return hits(a0, a0 + len);
    return hits(a0, H) || hits(0, a0 + len - H);
}

// Builds this round's work items if the mix changed.
void build_round(lgca_batch* b, const std::vector<int>& ks, uint32_t th)
{
    const size_t n = b->active.size();
    b->sig.clear();
    // The items depend on each job's geometry, slot, k and the tile height.
    // A job keeps its slot while it runs, but a finished job's slot goes to a
    // later job of another size: the signature names jobs by their ids, never
    // reused, so that a new job in an old slot always rebuilds the items.
    b->sig.push_back(th);
    for (size_t i = 0; i < n; ++i) {
        b->sig.push_back(b->active[i]->id);
        b->sig.push_back(((uint64_t)b->active[i]->slot << 8) | (uint64_t)ks[i]);
    }
    if (b->sig == b->last_sig) return;
    b->last_sig = b->sig;
    // classes in a fixed order: k, ragged (every model, segment and strip jobs
    // share a launch)
    std::map<uint64_t, std::vector<WorkItem>> cls, light;  // light: appended after
    for (size_t i = 0; i < n; ++i) {
        const Job* j = b->active[i];
        const Geometry& g = j->p.g;
        const int k = ks[i];
        const bool seg = tile_segments(g, k);
        const uint32_t halo = seg ? 0u : tile_halo(g, k);
        const uint32_t jth = std::min<uint32_t>(th, g.dim_y + (g.dim_y & 1u));
        const uint32_t tiles = (g.dim_y + jth - 1) / jth;
        const uint64_t key = ((uint64_t)k << 8) | (g.nb != 32 ? 1u : 0u);
        std::vector<WorkItem>& vh = cls[key];
        std::vector<WorkItem>& vl = light[key];
        if (seg) {
            const uint32_t S = 64 / g.W, groups = (tiles + S - 1) / S;
            for (uint32_t t = 0; t < groups; ++t) {
                const int64_t y0 = (int64_t)t * S * jth;
                std::vector<WorkItem>& v = rows_heavy(j, y0 - k, y0 + (int64_t)S * jth + k) ? vh : vl;
                v.push_back(WorkItem{(uint32_t)j->slot, 0u, t, jth | (1u << 31), groups});
            }
        } else {
            const uint32_t strips = tile_strips(g, k);
            for (uint32_t t = 0; t < tiles; ++t) {
                const int64_t y0 = (int64_t)t * jth, rows = std::min<int64_t>(jth, g.dim_y - y0);
                std::vector<WorkItem>& v = rows_heavy(j, y0 - k, y0 + rows + k) ? vh : vl;
                for (uint32_t s = 0; s < strips; ++s)
                    v.push_back(WorkItem{(uint32_t)j->slot, s, t, jth | (halo << 16), tiles * strips});
            }
        }
    }
    std::map<uint64_t, uint32_t> nheavy;
    for (auto& kv : light) {
        std::vector<WorkItem>& v = cls[kv.first];
        nheavy[kv.first] = (uint32_t)v.size();
        v.insert(v.end(), kv.second.begin(), kv.second.end());
    }
    size_t total = 0;
    for (auto& kv : cls) total += kv.second.size();
    const size_t bytes = total * sizeof(WorkItem) + 16;
    const int ib = b->items_buf ^= 1;
    // at least 16 MiB, so a growing mix rarely drains the stream to grow it
    grow_buffer(b, (char**)&b->d_items[ib], &b->items_cap[ib], std::max<size_t>(bytes, (size_t)16 << 20),
                "hipMalloc(work items)");
    std::vector<WorkItem> all;
    all.reserve(total);
    b->classes.clear();
    for (auto& kv : cls) {
        ClassRange c;
        c.k = (int)((kv.first >> 8) & 0xFF);
        c.ragged = (kv.first & 1u) != 0;
        c.offset = (uint32_t)all.size();
        c.count = (uint32_t)kv.second.size();
        c.heavy = nheavy[kv.first];
        b->classes.push_back(c);
        all.insert(all.end(), kv.second.begin(), kv.second.end());
    }
    upload(b, b->d_items[ib], all.data(), all.size() * sizeof(WorkItem));
}

// New force windows for forced jobs whose window ends before step + k, and
// for those close to that, in one launch.
void draw_windows(lgca_batch* b, const std::vector<int>& ks)
{
    bool need = false;
    for (size_t i = 0; i < b->active.size(); ++i) {
        const Job* j = b->active[i];
        if (j->p.forced && j->step + ks[i] > j->window_end) need = true;
    }
    if (!need) return;
    std::vector<uint32_t> slots;
    for (size_t i = 0; i < b->active.size(); ++i) {
        Job* j = b->active[i];
        if (!j->p.forced || j->step + ks[i] + kWindow / 2 <= j->window_end) continue;
        slots.push_back((uint32_t)j->slot);
        j->window_end = j->step + kWindow;
    }
    if (b->draw_cap < slots.size()) {
        hipchk(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        if (b->d_draw) (void)hipFree(b->d_draw);
        b->d_draw = nullptr;
        b->draw_cap = std::max<size_t>(slots.size() * 2, 256);
        hipchk(hipMalloc((void**)&b->d_draw, b->draw_cap * 4), "hipMalloc(draw list)");
    }
    upload(b, b->d_draw, slots.data(), slots.size() * 4);
    uint64_t most = 0;
    for (uint32_t s : slots) {
        const Job* j = b->slot_job[s];
        most = std::max<uint64_t>(most, (uint64_t)(j->p.fy1 - j->p.fy0) * j->p.fwn);
    }
    const uint32_t bx = (uint32_t)std::min<uint64_t>((most + 255) / 256, 128);
    // blockIdx.y covers at most 65535 jobs per launch
    for (size_t o = 0; o < slots.size(); o += 65535) {
        const uint32_t ny = (uint32_t)std::min<size_t>(65535, slots.size() - o);
        hipLaunchKernelGGL(batch_draw_kernel, dim3(bx, ny), dim3(256), 0, b->stream, b->d_jobs, b->d_state,
                           b->d_draw + o, kWindow);
        hipchk(hipGetLastError(), "batch_draw_kernel");
    }
}

void start_tuning(lgca_batch* b, int user_k, uint32_t user_waves, uint64_t wr)
{
    b->evaluating = false;  // a restart abandons the previous measurement
    b->cands.clear();
    const int ks[2] = {4, 8};
    const uint32_t ws[4] = {2048, 4096, 8192, 16384};
    for (int k : ks) {
        if (user_k && k != user_k) continue;
        for (uint32_t w : ws) {
            if (user_waves && w != user_waves) continue;
            b->cands.push_back(Candidate{k, w});
        }
    }
    if (user_k && b->cands.empty())
        for (uint32_t w : ws)
            if (!user_waves || w == user_waves) b->cands.push_back(Candidate{user_k, w});
    if (user_waves && b->cands.empty()) b->cands.push_back(Candidate{user_k ? user_k : 4, user_waves});
    b->tuned_wave_rows = wr;
    b->rounds_since_tune = 0;
    if (b->cands.size() <= 1) {
        if (!b->cands.empty()) {
            b->k_cur = b->cands[0].k;
            b->waves_cur = b->cands[0].waves;
        }
        b->tuning = false;
        return;
    }
    b->tuning = true;
    b->cand_i = 0;
    b->cand_rounds = 0;
    for (hipEvent_t e : b->tune_ev) (void)hipEventDestroy(e);
    b->tune_ev.assign(2 * b->cands.size(), nullptr);
    for (auto& e : b->tune_ev) hipchk(hipEventCreate(&e), "hipEventCreate");
    b->tune_su.assign(b->cands.size(), 0);
    b->k_cur = b->cands[0].k;
    b->waves_cur = b->cands[0].waves;
}

// Picks the fastest candidate once the GPU has passed the last timing event.
void evaluate_tuning(lgca_batch* b, bool block)
{
    hipEvent_t last = b->tune_ev.back();
    hipError_t q = block ? hipEventSynchronize(last) : hipEventQuery(last);
    if (q == hipErrorNotReady) return;
    hipchk(q, "hipEventQuery");
    double best = -1.0;
    for (size_t c = 0; c < b->cands.size(); ++c) {
        float ms = 0.0f;
        hipchk(hipEventElapsedTime(&ms, b->tune_ev[2 * c], b->tune_ev[2 * c + 1]), "hipEventElapsedTime");
        const double rate = ms > 0.0f ? (double)b->tune_su[c] / ms : 0.0;
        if (rate > best) {
            best = rate;
            b->k_cur = b->cands[c].k;
            b->waves_cur = b->cands[c].waves;
        }
    }
    b->evaluating = false;
    ++b->n_tunings;
}

void run_round(lgca_batch* b)
{
    int user_k;
    uint32_t user_waves;
    bool retune;
    {
        std::lock_guard<std::mutex> lk(b->mu);
        user_k = b->user_k;
        user_waves = b->user_waves;
        retune = b->tuning_changed;
        b->tuning_changed = false;
    }
    if (b->evaluating) evaluate_tuning(b, false);
    const uint64_t wr_now = wave_rows(b, 4);
    // A mix that has doubled since the last tuning (jobs still arriving) is
    // tuned again at once, a shrinking one after 64 rounds; a tuning under way
    // restarts only when the mix has grown 4x since it began.
    if (!b->tuning && !b->evaluating) {
        const bool grown = b->tuned_wave_rows != 0 && wr_now > 2 * b->tuned_wave_rows;
        const bool shrunk = b->rounds_since_tune >= 64 && wr_now * 2 < b->tuned_wave_rows;
        if (retune || b->tuned_wave_rows == 0 || shrunk || grown) start_tuning(b, user_k, user_waves, wr_now);
    } else if (retune || wr_now > 4 * b->tuned_wave_rows) {
        start_tuning(b, user_k, user_waves, wr_now);
    }
    ++b->rounds_since_tune;

    const size_t n = b->active.size();
    std::vector<int> ks(n);
    uint64_t su = 0;
    for (size_t i = 0; i < n; ++i) {
        const Job* j = b->active[i];
        ks[i] = pick_k(b->k_cur, j->d.steps - j->step, j->step);
        su += (uint64_t)j->d.dim_x * j->d.dim_y * (uint64_t)ks[i];
    }
    const uint32_t th = tile_rows_for(wave_rows(b, b->k_cur), b->waves_cur);
    build_round(b, ks, th);
    draw_windows(b, ks);

    if (b->tuning && b->cand_rounds == 0)
        hipchk(hipEventRecord(b->tune_ev[2 * b->cand_i], b->stream), "hipEventRecord");
    // One launch per (k, ragged) class: the jobs of every model share it (each
    // wave takes its job's model body), so a round is usually one launch, and
    // only jobs at their last steps add a short one.
    uint64_t nl = 0;
    for (const ClassRange& c : b->classes) {
        hipchk(launch_tile_batch(c.k, c.ragged, b->d_jobs, b->d_state, b->d_items[b->items_buf] + c.offset,
                                 c.count, c.heavy, b->stream),
               "tile_batch_kernel");
        ++nl;
    }
    if (b->tuning) {
        b->tune_su[b->cand_i] += su;
        if (++b->cand_rounds == kTuneRounds) {
            hipchk(hipEventRecord(b->tune_ev[2 * b->cand_i + 1], b->stream), "hipEventRecord");
            b->cand_rounds = 0;
            if (++b->cand_i == b->cands.size()) {
                b->tuning = false;
                b->evaluating = true;
            } else {
                b->k_cur = b->cands[b->cand_i].k;
                b->waves_cur = b->cands[b->cand_i].waves;
            }
        }
    }

    // host-side bookkeeping of the queued steps
    std::vector<Job*> still;
    still.reserve(n);
    for (size_t i = 0; i < n; ++i) {
        Job* j = b->active[i];
        j->step += (uint64_t)ks[i];
        j->cur ^= 1u;
        if (j->step == j->d.steps) {
            finish(b, j);
            b->active_cells -= (uint64_t)j->d.dim_x * j->d.dim_y;
        } else {
            still.push_back(j);
        }
    }
    b->active.swap(still);
    hipEvent_t ev = take_event(b);
    hipchk(hipEventRecord(ev, b->stream), "hipEventRecord");
    b->inflight.push_back(ev);
    {
        std::lock_guard<std::mutex> lk(b->mu);
        ++b->rounds;
        b->launches += nl;
        b->site_updates += su;
        b->stat_k = b->k_cur;
        b->stat_waves = b->waves_cur;
        b->stat_tunings = b->n_tunings;
    }
    while ((int)b->inflight.size() > kInflight) {
        hipchk(hipEventSynchronize(b->inflight.front()), "hipEventSynchronize");
        b->event_pool.push_back(b->inflight.front());
        b->inflight.pop_front();
    }
}

// A worker failure is sticky: unfinished jobs fail and every later call reports it.
void fail_all(lgca_batch* b, const std::string& msg)
{
    std::lock_guard<std::mutex> lk(b->mu);
    b->err = msg;
    for (auto& kv : b->jobs)
        if (kv.second->status == LGCA_JOB_QUEUED || kv.second->status == LGCA_JOB_RUNNING)
            kv.second->status = LGCA_JOB_FAILED;
    b->cv_done.notify_all();
}

void worker_main(lgca_batch* b)
{
    DevGuard dg(b->device);
    try {
        for (;;) {
            // admit what fits, in order
            std::vector<Job*> admitted;
            {
                std::unique_lock<std::mutex> lk(b->mu);
                for (;;) {
                    if (b->stop) return;
                    while (!b->queue.empty()) {
                        Job* j = b->queue.front();
                        if (b->bytes_in_use + j->L.total > b->budget) break;
                        const uint64_t cells = (uint64_t)j->d.dim_x * j->d.dim_y;
                        if (b->max_active_cells && !b->active.empty() &&
                            b->active_cells + cells > b->max_active_cells)
                            break;
                        hipError_t e = b->arena.alloc(j->L.total, &j->mem);
                        if (e != hipSuccess) throw WorkerError{std::string("arena: ") + hipGetErrorString(e)};
                        b->bytes_in_use += j->L.total;
                        b->queue.pop_front();
                        j->status = LGCA_JOB_RUNNING;
                        j->t_admit = now_s();
                        b->active_cells += cells;
                        b->active.push_back(j);
                        admitted.push_back(j);
                    }
                    if (!b->active.empty() || !b->finishing.empty() || !admitted.empty()) break;
                    b->cv_work.wait(lk);
                }
            }
            admit_all(b, admitted);
            for (Job* j : admitted) {
                if (j->d.steps == 0) {
                    finish(b, j);
                    b->active_cells -= (uint64_t)j->d.dim_x * j->d.dim_y;
                    b->active.erase(std::find(b->active.begin(), b->active.end(), j));
                }
            }
            if (!b->active.empty()) {
                run_round(b);
                reap(b, false);
            } else {
                // nothing left to queue: wait for the GPU to finish what is queued
                if (b->evaluating) evaluate_tuning(b, true);
                reap(b, true);
                for (hipEvent_t e : b->inflight) b->event_pool.push_back(e);
                b->inflight.clear();
            }
        }
    } catch (const WorkerError& e) {
        fail_all(b, e.msg);
    } catch (const std::exception& e) {
        fail_all(b, std::string("host error: ") + e.what());
    }
}

lgca_status worker_failed(lgca_batch* b)
{
    // caller holds b->mu
    b->call_err = "batch worker failed: " + b->err;
    return LGCA_ERR_HIP;
}

}  // namespace

// ----------------------------------------------------------- entry points

extern "C" lgca_status lgca_batch_create(lgca_batch** out, int device, uint64_t hbm_budget_bytes)
{
    if (!out) {
        g_batch_create_error = "lgca_batch_create: out is NULL";
        return LGCA_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
        g_batch_create_error = "lgca_batch_create: no such GPU";
        return LGCA_ERR_INVALID_ARGUMENT;
    }
    if (hbm_budget_bytes == 0) {
        g_batch_create_error = "lgca_batch_create: zero budget";
        return LGCA_ERR_INVALID_ARGUMENT;
    }
    auto* b = new lgca_batch();
    b->device = device;
    b->budget = hbm_budget_bytes;
    b->arena.block_bytes = (size_t)std::min<uint64_t>((uint64_t)1 << 30, align_up(hbm_budget_bytes));
    DevGuard dg(device);
    auto fail = [&](hipError_t e, const char* what) {
        g_batch_create_error = std::string(what) + ": " + hipGetErrorString(e);
        lgca_batch_destroy(b);
        return e == hipErrorOutOfMemory ? LGCA_ERR_OUT_OF_MEMORY : LGCA_ERR_HIP;
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess) return fail(e, "hipStreamCreate");
    if ((e = hipStreamCreateWithFlags(&b->copy_stream, hipStreamNonBlocking)) != hipSuccess) return fail(e, "hipStreamCreate");
    if ((e = hipStreamCreateWithFlags(&b->post_stream, hipStreamNonBlocking)) != hipSuccess) return fail(e, "hipStreamCreate");
    if ((e = hipHostMalloc((void**)&b->staging, 2 * kStagingHalf, hipHostMallocDefault)) != hipSuccess)
        return fail(e, "hipHostMalloc(staging)");
    for (auto& ev : b->staging_ev)
        if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) return fail(e, "hipEventCreate");
    // both halves start free
    for (auto& ev : b->staging_ev)
        if ((e = hipEventRecord(ev, b->stream)) != hipSuccess) return fail(e, "hipEventRecord");
    if ((e = hipMalloc((void**)&b->d_err, 64)) != hipSuccess) return fail(e, "hipMalloc");
    try {
        grow_tables(b, 1024);
    } catch (const WorkerError& we) {
        g_batch_create_error = we.msg;
        lgca_batch_destroy(b);
        return LGCA_ERR_HIP;
    }
    b->worker = std::thread(worker_main, b);
    *out = b;
    return LGCA_OK;
}

extern "C" lgca_status lgca_batch_destroy(lgca_batch* b)
{
    if (!b) return LGCA_OK;
    {
        std::lock_guard<std::mutex> lk(b->mu);
        b->stop = true;
    }
    b->cv_work.notify_all();
    if (b->worker.joinable()) b->worker.join();
    DevGuard dg(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->post_stream) (void)hipStreamSynchronize(b->post_stream);
    if (b->copy_stream) (void)hipStreamSynchronize(b->copy_stream);
    for (auto& kv : b->jobs) {
        if (kv.second->done_ev) (void)hipEventDestroy(kv.second->done_ev);
        if (kv.second->last_ev) (void)hipEventDestroy(kv.second->last_ev);
    }
    for (hipEvent_t e : b->inflight) (void)hipEventDestroy(e);
    for (hipEvent_t e : b->event_pool) (void)hipEventDestroy(e);
    for (hipEvent_t e : b->tune_ev) (void)hipEventDestroy(e);
    for (hipEvent_t e : b->staging_ev)
        if (e) (void)hipEventDestroy(e);
    b->arena.release_all();
    void* dptrs[] = {b->d_jobs, b->d_state, b->d_items[0], b->d_items[1], b->d_draw, b->d_err, b->d_setup};
    for (void* p : dptrs)
        if (p) (void)hipFree(p);
    void* hptrs[] = {b->h_jobs, b->h_state, b->staging};
    for (void* p : hptrs)
        if (p) (void)hipHostFree(p);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->copy_stream) (void)hipStreamDestroy(b->copy_stream);
    if (b->post_stream) (void)hipStreamDestroy(b->post_stream);
    delete b;
    return LGCA_OK;
}

extern "C" const char* lgca_batch_last_error(const lgca_batch* b)
{
    if (!b) return g_batch_create_error.c_str();
    return b->call_err.c_str();
}

extern "C" lgca_status lgca_batch_job_bytes(const lgca_job_desc* desc, uint64_t* bytes)
{
    if (!desc || !bytes) return LGCA_ERR_INVALID_ARGUMENT;
    if (!validate(*desc).empty()) return LGCA_ERR_INVALID_ARGUMENT;
    const JobParams p = job_params(*desc);
    size_t nh = 0;
    if (desc->geometry == LGCA_GEOM_CYLINDER) nh = disc_halfwidths((float)desc->cyl_diameter).size();
    *bytes = job_layout(*desc, p, nh).total;
    return LGCA_OK;
}

extern "C" lgca_status lgca_batch_submit(lgca_batch* b, const lgca_job_desc* desc, uint64_t* job_id)
{
    if (!b || !desc || !job_id) return set_call_error(b, LGCA_ERR_INVALID_ARGUMENT, "lgca_batch_submit: NULL argument");
    const std::string why = validate(*desc);
    if (!why.empty()) return set_call_error(b, LGCA_ERR_INVALID_ARGUMENT, "lgca_batch_submit: " + why);
    auto j = std::make_unique<Job>();
    j->d = *desc;
    j->d.cell_types = nullptr;
    j->d.state = nullptr;
    j->p = job_params(*desc);
    const size_t n = (size_t)desc->dim_x * desc->dim_y;
    if (desc->geometry == LGCA_GEOM_TYPES) {
        j->types_in.assign(desc->cell_types, desc->cell_types + n);
        for (uint8_t t : j->types_in)
            if (t > 2) return set_call_error(b, LGCA_ERR_INVALID_STATE, "lgca_batch_submit: cell types must be 0, 1 or 2");
    }
    if (desc->state) {
        j->state_in.assign(desc->state, desc->state + n);
        const uint32_t Q = j->p.g.Q;
        for (uint8_t v : j->state_in)
            if (v >> Q) return set_call_error(b, LGCA_ERR_INVALID_STATE, "lgca_batch_submit: state bytes carry bits at or above num_dir");
    }
    if (desc->geometry == LGCA_GEOM_CYLINDER) j->halfw = disc_halfwidths((float)desc->cyl_diameter);
    j->L = job_layout(j->d, j->p, j->halfw.size());
    std::lock_guard<std::mutex> lk(b->mu);
    if (!b->err.empty()) return worker_failed(b);
    if (j->L.total > b->budget) {
        b->call_err = "lgca_batch_submit: the job needs " + std::to_string(j->L.total) +
                      " device bytes, more than the batch budget";
        return LGCA_ERR_BUDGET;
    }
    j->id = b->next_id++;
    j->t_submit = now_s();
    j->status = LGCA_JOB_QUEUED;
    *job_id = j->id;
    b->queue.push_back(j.get());
    b->jobs.emplace(j->id, std::move(j));
    ++b->n_submitted;
    b->cv_work.notify_all();
    return LGCA_OK;
}

extern "C" lgca_status lgca_batch_wait(lgca_batch* b, uint64_t job_id, double timeout_s)
{
    if (!b) return LGCA_ERR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lk(b->mu);
    Job* j = nullptr;
    if (job_id != LGCA_ALL_JOBS) {
        auto it = b->jobs.find(job_id);
        if (it == b->jobs.end()) {
            b->call_err = "lgca_batch_wait: unknown job";
            return LGCA_ERR_INVALID_ARGUMENT;
        }
        j = it->second.get();
    }
    auto ready = [&] {
        if (!b->err.empty()) return true;
        if (j) return j->status >= LGCA_JOB_DONE;
        return b->n_done == b->n_submitted;
    };
    if (timeout_s < 0) {
        b->cv_done.wait(lk, ready);
    } else if (!b->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) {
        b->call_err = "lgca_batch_wait: timed out";
        return LGCA_ERR_INVALID_STATE;
    }
    if (!b->err.empty()) return worker_failed(b);
    return LGCA_OK;
}

extern "C" lgca_status lgca_batch_poll(lgca_batch* b, uint64_t* ids, uint32_t max_ids, uint32_t* n)
{
    if (!b || !n || (max_ids && !ids)) return LGCA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(b->mu);
    uint32_t c = 0;
    while (c < max_ids && !b->done_ids.empty()) {
        ids[c++] = b->done_ids.front();
        b->done_ids.pop_front();
    }
    *n = c;
    if (!b->err.empty()) return worker_failed(b);
    return LGCA_OK;
}

extern "C" lgca_status lgca_batch_job_info(lgca_batch* b, uint64_t job_id, lgca_job_info* info)
{
    if (!b || !info) return LGCA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(b->mu);
    auto it = b->jobs.find(job_id);
    if (it == b->jobs.end()) {
        b->call_err = "lgca_batch_job_info: unknown job";
        return LGCA_ERR_INVALID_ARGUMENT;
    }
    const Job* j = it->second.get();
    std::memset(info, 0, sizeof *info);
    info->status = j->status;
    info->step = j->status == LGCA_JOB_DONE || j->status == LGCA_JOB_RELEASED ? j->d.steps : 0;
    info->device_bytes = j->L.total;
    info->submit_s = j->t_submit;
    info->admit_s = j->t_admit;
    info->done_s = j->t_done;
    return LGCA_OK;
}

extern "C" lgca_status lgca_batch_result(lgca_batch* b, uint64_t job_id, uint8_t* state,
                                         float* mean_density, float* mean_momentum)
{
    if (!b) return LGCA_ERR_INVALID_ARGUMENT;
    char* m = nullptr;
    JobLayout L{};
    size_t n = 0, nc = 0;
    Job* j = nullptr;
    {
        std::lock_guard<std::mutex> lk(b->mu);
        auto it = b->jobs.find(job_id);
        if (it == b->jobs.end() || it->second->status != LGCA_JOB_DONE) {
            b->call_err = "lgca_batch_result: job unknown, not done or released";
            return LGCA_ERR_INVALID_ARGUMENT;
        }
        j = it->second.get();
        // lgca_batch_release waits for this count to drop to zero before the
        // job's memory can go to another job
        ++j->readers;
        m = j->mem;
        L = j->L;
        n = (size_t)j->d.dim_x * j->d.dim_y;
        nc = j->p.nc;
    }
    DevGuard dg(b->device);
    hipError_t e = hipSuccess;
    if (state && e == hipSuccess) e = hipMemcpyAsync(state, m + L.bytes, n, hipMemcpyDeviceToHost, b->copy_stream);
    if (mean_density && e == hipSuccess)
        e = hipMemcpyAsync(mean_density, m + L.md, nc * 4, hipMemcpyDeviceToHost, b->copy_stream);
    if (mean_momentum && e == hipSuccess)
        e = hipMemcpyAsync(mean_momentum, m + L.mm, nc * 8, hipMemcpyDeviceToHost, b->copy_stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->copy_stream);
    {
        std::lock_guard<std::mutex> lk(b->mu);
        if (--j->readers == 0) b->cv_readers.notify_all();
    }
    if (e != hipSuccess)
        return set_call_error(b, LGCA_ERR_HIP, std::string("lgca_batch_result: ") + hipGetErrorString(e));
    return LGCA_OK;
}

extern "C" lgca_status lgca_batch_release(lgca_batch* b, uint64_t job_id)
{
    if (!b) return LGCA_ERR_INVALID_ARGUMENT;
    {
        std::unique_lock<std::mutex> lk(b->mu);
        auto it = b->jobs.find(job_id);
        if (it == b->jobs.end() || it->second->status != LGCA_JOB_DONE) {
            b->call_err = "lgca_batch_release: job unknown, not done or already released";
            return LGCA_ERR_INVALID_ARGUMENT;
        }
        Job* j = it->second.get();
        // result copies of this job still in flight finish first (jobs stay
        // in the table for the batch's lifetime, so j stays valid)
        b->cv_readers.wait(lk, [&] { return j->readers == 0; });
        if (j->status != LGCA_JOB_DONE) {
            b->call_err = "lgca_batch_release: job already released";
            return LGCA_ERR_INVALID_ARGUMENT;
        }
        b->arena.free(j->mem, j->L.total);
        b->bytes_in_use -= j->L.total;
        j->mem = nullptr;
        j->status = LGCA_JOB_RELEASED;
        j->types_in.clear();
        j->types_in.shrink_to_fit();
        j->state_in.clear();
        j->state_in.shrink_to_fit();
    }
    b->cv_work.notify_all();
    return LGCA_OK;
}

extern "C" lgca_status lgca_batch_set_tuning(lgca_batch* b, int steps_per_pass, uint32_t wave_target,
                                             uint64_t max_active_cells)
{
    if (!b) return LGCA_ERR_INVALID_ARGUMENT;
    if (steps_per_pass != 0 && steps_per_pass != 2 && steps_per_pass != 4 && steps_per_pass != 8)
        return set_call_error(b, LGCA_ERR_INVALID_ARGUMENT,
                              "lgca_batch_set_tuning: steps_per_pass must be 0, 2, 4 or 8");
    if (wave_target > (1u << 24))
        return set_call_error(b, LGCA_ERR_INVALID_ARGUMENT, "lgca_batch_set_tuning: wave_target too large");
    std::lock_guard<std::mutex> lk(b->mu);
    b->user_k = steps_per_pass;
    b->user_waves = wave_target;
    b->max_active_cells = max_active_cells;
    b->tuning_changed = true;
    return LGCA_OK;
}

extern "C" lgca_status lgca_batch_get_stats(lgca_batch* b, lgca_batch_stats* s)
{
    if (!b || !s) return LGCA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(b->mu);
    std::memset(s, 0, sizeof *s);
    s->jobs_submitted = b->n_submitted;
    s->jobs_done = b->n_done;
    uint64_t active = 0;
    for (auto& kv : b->jobs) active += kv.second->status == LGCA_JOB_RUNNING;
    s->jobs_active = active;
    s->jobs_queued = b->queue.size();
    s->rounds = b->rounds;
    s->launches = b->launches;
    s->site_updates = b->site_updates;
    s->bytes_in_use = b->bytes_in_use;
    s->arena_bytes = b->arena.reserved;
    s->steps_per_pass = b->stat_k;
    s->wave_target = b->stat_waves;
    s->tunings = b->stat_tunings;
    return LGCA_OK;
}
