// lgca_capi.hip -- the C ABI of include/lgca.h over the gfx950 kernels.
//
// A handle owns, on its GPU: two state buffers (ping-pong bit-planes), the
// output snapshot, the cell-type masks with a per-row range of non-fluid words,
// force-decision planes for a window of upcoming steps, post-processing
// outputs and a small block for counters and validation flags.
#include "lgca.h"

#include "lgca_dev.hpp"
#include "lgca_kernels.hpp"

#include <algorithm>
#include <climits>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace lgca;

namespace lgca {
// Device-resident state of the viewers' forcing controller.
struct Controller {
    long long acc[3];   // mean-velocity sums of the snapshot: x, y (units of 1/840), FLUID cells
    long long forcing;  // the controller's current forcing amount
    long long amount;   // this tick: amount to apply, or -1
    long long reverted; // particles reverted by the last applied force
    uint32_t calls;     // count-mode force call counter (hash stream 3)
    float v[2];         // mean velocity the last decision saw
};

// Steps per pass and rows per tile of one tiled launch.
struct LaunchShape {
    int k;
    uint32_t th;
};
// A strip-mode dispatch order in device memory (tile_wave_order), for passes
// of k steps, th-row tiles and halo `halo` under the cell types and forcing
// that `key` names.
struct WaveOrder {
    int k;
    uint32_t th, halo;
    uint64_t key;
    uint32_t* dev;
    uint32_t n;  // entries
};
// A dispatch order for launch_tile: entries in device memory and their count
// (null and 0: the kernel's own order).
struct OrderRef {
    const uint32_t* dev = nullptr;
    uint32_t n = 0;
};
// A pipelined launch's band plan (tile_pipe_plan) and its item order in
// device memory, cached like WaveOrder.
struct PipeCache {
    int k;
    uint32_t th, halo;
    uint64_t key;
    PipePlan plan;
    uint32_t* order;
};
// Counter words of a pipelined launch: the bands' ticket counters, the error
// word, then one finished-pass count per work item (row tile x strip).
constexpr uint32_t kPipeQueue = 0, kPipeError = kMaxBands, kPipeDone = 16;
// steps_per_pass value that lets the trial choose it
constexpr int kAutoSteps = -1;
}  // namespace lgca

struct lgca_lattice {
    int model = 0;
    int kind = 0;
    uint32_t Q = 0;
    uint32_t dim_x = 0, dim_y = 0, r = 1;
    int device = 0;
    Geometry g{};
    size_t state_words = 0;
    uint32_t* state[2] = {nullptr, nullptr};
    int cur = 0;
    uint32_t* snap = nullptr;
    uint32_t* mask = nullptr;       // allocated once any non-fluid cell is set
    int32_t* row_range = nullptr;
    std::vector<int32_t> row_lo, row_hi;  // host copy of row_range (empty without solids)
    uint64_t types_version = 0;     // counts cell-type uploads
    std::vector<WaveOrder> orders;  // cached dispatch orders
    // several passes per launch (lgca_set_pipeline)
    int pipe = 0;                   // 1: on where the launch shape allows it
    uint32_t nxcc = 1, ncu = 0;     // XCDs and CUs of the device
    std::vector<PipeCache> pipes;   // cached band plans
    uint32_t* pipe_exch = nullptr;  // exchange rows
    size_t pipe_exch_cap = 0;       // words
    uint32_t* pipe_ctr = nullptr;   // counter words (kPipe*)
    size_t pipe_ctr_cap = 0;        // words
    bool pipe_check = false;        // a pipelined launch ran since its error word was read
    uint32_t* dec = nullptr;        // force decisions, [k][rows][fwn]
    size_t dec_cap = 0;             // words
    uint64_t dec_step0 = 0;
    uint32_t dec_count = 0;
    float* cell_density = nullptr;
    float* cell_momentum = nullptr;
    float* mean_density = nullptr;
    float* mean_momentum = nullptr;
    float* velocity = nullptr;
    uint8_t* bytes = nullptr;       // N-byte staging for u8 state / types
    uint32_t* visited = nullptr;    // count-mode force: first-visit bitmap
    unsigned long long* counters = nullptr;  // 16 words: counters and flags
    uint64_t seed = 0, step = 0, force_calls = 0;
    int force_dir = LGCA_FORCE_NONE;
    uint64_t force_thr = 0;
    uint32_t fx0 = 0, fx1 = 0, fy0 = 0, fy1 = 0;
    int tile_k = kAutoSteps;        // steps per pass of the tiled kernel (0: quad kernel)
    uint32_t tile_th = 0;           // rows per tile, 0 = automatic
    // Automatic launch shape: a timed trial of a few (steps per pass, rows per
    // tile) pairs on the first launches, then the fastest (plan_launch)
    std::vector<LaunchShape> trial; // candidates, the rule's own first
    std::vector<hipEvent_t> trial_ev;  // start / end of every timed group of launches
    std::vector<uint32_t> trial_steps; // steps launched inside each group
    uint32_t trial_issued = 0;      // candidate launches queued so far
    uint32_t trial_group = 1;       // launches per timed group
    int trial_open = -1;            // group whose end event is still to come
    LaunchShape tuned{0, 0};        // the chosen shape, th 0 while undecided
    uint32_t tuned_rem4 = 0;        // rows of the fastest 4-step candidate, for the
                                    // 4-step remainders of an 8-step choice (0: rule)
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    // forcing controller of the viewers (lgca_controller_*)
    Controller* ctrl = nullptr;
    int ctrl_dir = LGCA_FORCE_NONE;
    float ctrl_u = 0.0f;
    int64_t ctrl_eq = 0;
    bool calls_on_device = false;   // the count-mode call counter lives in ctrl
    std::string err;
    std::mutex mu;
};

namespace {

thread_local std::string g_create_error;

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard()
    {
        int now = -1;
        if (prev >= 0 && hipGetDevice(&now) == hipSuccess && now != prev) (void)hipSetDevice(prev);
    }
};

lgca_status fail(lgca_lattice* h, lgca_status code, const std::string& msg)
{
    if (h) h->err = msg;
    else g_create_error = msg;
    return code;
}

lgca_status hip_fail(lgca_lattice* h, hipError_t e, const char* what)
{
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return fail(h, e == hipErrorOutOfMemory ? LGCA_ERR_OUT_OF_MEMORY : LGCA_ERR_HIP, m);
}

#define HIPCHK(h, expr)                                             \
    do {                                                            \
        hipError_t e_ = (expr);                                     \
        if (e_ != hipSuccess) return hip_fail((h), e_, #expr);      \
    } while (0)

// Starts the launch-shape trial afresh (the workload or the stream changed).
void reset_trial(lgca_lattice* h)
{
    h->trial.clear();
    h->trial_issued = 0;
    h->trial_open = -1;
    h->tuned = LaunchShape{0, 0};
    h->tuned_rem4 = 0;
}

template <typename T>
hipError_t dalloc(T** p, size_t bytes)
{
    void* v = nullptr;
    hipError_t e = hipMalloc(&v, bytes ? bytes : 16);
    *p = (T*)v;
    return e;
}

lgca_status ensure_bytes(lgca_lattice* h)
{
    if (!h->bytes) HIPCHK(h, dalloc(&h->bytes, (size_t)h->dim_x * h->dim_y));
    return LGCA_OK;
}

lgca_status ensure_mask(lgca_lattice* h)
{
    if (!h->mask) {
        HIPCHK(h, dalloc(&h->mask, (size_t)h->dim_y * 2 * h->g.P * 4));
        HIPCHK(h, dalloc(&h->row_range, (size_t)h->dim_y * 2 * 4));
    }
    return LGCA_OK;
}

lgca_status read_counters(lgca_lattice* h, unsigned long long* out, int n)
{
    HIPCHK(h, hipMemcpyAsync(out, h->counters, n * sizeof(unsigned long long),
                             hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return LGCA_OK;
}

lgca_status clear_counters(lgca_lattice* h)
{
    HIPCHK(h, launch_fill_u32((uint32_t*)h->counters, 0u, 16, h->stream));
    return LGCA_OK;
}

// Brings the count-mode force call counter back to the host after device
// controller ticks advanced it.
lgca_status sync_force_calls(lgca_lattice* h)
{
    if (!h->calls_on_device) return LGCA_OK;
    uint32_t calls = 0;
    HIPCHK(h, hipMemcpyAsync(&calls, &h->ctrl->calls, 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->force_calls = (h->force_calls & ~(uint64_t)0xFFFFFFFFu) | calls;
    h->calls_on_device = false;
    return LGCA_OK;
}

}  // namespace

// -------------------------------------------------------------- count force
// apply_body_force(int forcing), src/omp_lattice.cpp:254-346.  One wave walks
// the candidate sequence 64 tries at a time.  A revisit of a cell can never
// flip (after a visit each flip rule of the cell is infeasible), so only first
// visits count: the wave drops candidates that repeat an earlier lane's cell,
// tests-and-sets a per-call visited bitmap, sums the reverts of its lanes in
// order and stops at the first try whose running total reaches `forcing`
// (always after at least one try), exactly as the sequential loop would.
namespace lgca {

template <int KIND>
__device__ __forceinline__ void count_force_body(uint32_t* planes, const uint32_t* mask, const Geometry& g,
                                                 uint32_t key0, uint32_t key1, uint32_t call, int dir,
                                                 long long forcing, uint32_t* visited, long long* result)
{
    constexpr int Q = num_dir_of_kind(KIND);
    const uint32_t lane = threadIdx.x;
    const uint64_t n = (uint64_t)g.dim_x * g.dim_y;
    const uint64_t it_max = 2 * n;
    long long reverted = 0;
    for (uint64_t base = 0; base < it_max; base += 64) {
        const uint64_t i = base + lane;
        const bool active = i < it_max;
        // flips of the previous batch (L2 atomics) must be visible to this one
        __threadfence();
        const U4 rr = philox4x32_10((uint32_t)i, (uint32_t)(i >> 32), call, 3u, key0, key1);
        const uint64_t cell = ((((uint64_t)rr.y) << 32) | rr.x) % n;
        bool dup = false;
        for (uint32_t j = 0; j < 64; ++j) {
            const uint64_t cj = __shfl(cell, (int)j, 64);
            if (j < lane && base + j < it_max && cj == cell) dup = true;
        }
        bool first = false;
        const uint32_t y = (uint32_t)(cell / g.dim_x), x = (uint32_t)(cell % g.dim_x);
        const uint32_t w = x >> 5, bit = 1u << (x & 31u);
        if (active && !dup) {
            const uint32_t old = atomicOr(visited + (cell >> 5), 1u << (cell & 31u));
            first = (old & (1u << (cell & 31u))) == 0;
        }
        uint32_t flip[7] = {0, 0, 0, 0, 0, 0, 0};
        int k = 0;
        if (first) {
            bool fluid = true;
            if (mask)
                fluid = ((mask[((size_t)y * 2 + 0) * g.P + w] | mask[((size_t)y * 2 + 1) * g.P + w]) & bit) == 0;
            if (fluid) {
                uint32_t s[7], o[7];
#pragma unroll
                for (int d = 0; d < Q; ++d)
                    s[d] = o[d] = __hip_atomic_load(planes + ((size_t)y * Q + d) * g.P + w,
                                                    __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & bit;
                apply_force<KIND>(o, bit, dir);
#pragma unroll
                for (int d = 0; d < Q; ++d) {
                    flip[d] = s[d] ^ o[d];
                    k += flip[d] ? 1 : 0;
                }
                k /= 2;
            }
        }
        // inclusive prefix of k over lanes
        int incl = k;
        for (int off = 1; off < 64; off <<= 1) {
            const int t = __shfl_up(incl, off, 64);
            if ((int)lane >= off) incl += t;
        }
        const bool hit = active && (reverted + incl >= forcing);
        const unsigned long long hits = __ballot(hit);
        const uint32_t limit = hits ? (uint32_t)__ffsll((long long)hits) - 1u : 63u;
        if (lane <= limit && first && k > 0) {
#pragma unroll
            for (int d = 0; d < Q; ++d)
                if (flip[d]) atomicXor(planes + ((size_t)y * Q + d) * g.P + w, flip[d]);
        }
        if (lane > limit && first) atomicAnd(visited + (cell >> 5), ~(1u << (cell & 31u)));
        reverted += __shfl(incl, (int)limit, 64);
        if (hits) break;
    }
    if (lane == 0) *result = reverted;
}

template <int KIND>
__global__ __launch_bounds__(64) void body_force_count_kernel(uint32_t* planes, const uint32_t* mask,
                                                              Geometry g, uint32_t key0, uint32_t key1,
                                                              uint32_t call, int dir, long long forcing,
                                                              uint32_t* visited, long long* result)
{
    count_force_body<KIND>(planes, mask, g, key0, key1, call, dir, forcing, visited, result);
}

// The same, with the amount and call number the controller left on the GPU;
// nothing happens (and no call is counted) when the amount is negative.
template <int KIND>
__global__ __launch_bounds__(64) void controlled_force_kernel(uint32_t* planes, const uint32_t* mask,
                                                              Geometry g, uint32_t key0, uint32_t key1,
                                                              int dir, Controller* c, uint32_t* visited)
{
    const long long amount = c->amount;
    if (amount < 0) return;
    const uint32_t call = c->calls;
    count_force_body<KIND>(planes, mask, g, key0, key1, call, dir, amount, visited, &c->reverted);
    if (threadIdx.x == 0) c->calls = call + 1;
}

// The viewers' decision (apps/pipe/pipe_viewer.cpp:107-125): the mean x
// velocity of the snapshot, in the FP32 value lgca_mean_velocity returns; if
// it is below u, the forcing becomes `eq` once it exceeds 0.9 u (compared in
// double, as the viewer's float * double promotes), and the forcing applies.
__global__ void controller_decide_kernel(Controller* c, float u, long long eq, double unit_y)
{
    const double nf = (double)c->acc[2];
    const float v0 = nf > 0 ? (float)((double)c->acc[0] / (840.0 * nf)) : 0.0f;
    const float v1 = nf > 0 ? (float)((double)c->acc[1] * unit_y / (840.0 * nf)) : 0.0f;
    c->v[0] = v0;
    c->v[1] = v1;
    long long amount = -1;
    if (v0 < u) {
        if ((double)v0 > 0.9 * (double)u) c->forcing = eq;
        amount = c->forcing;
    }
    c->amount = amount;
}

}  // namespace lgca

// --------------------------------------------------------------- lifecycle

extern "C" lgca_status lgca_create(lgca_lattice** out, int model, uint32_t dim_x, uint32_t dim_y,
                                   uint32_t coarse_radius, int device)
{
    if (!out) return fail(nullptr, LGCA_ERR_INVALID_ARGUMENT, "lgca_create: out is NULL");
    *out = nullptr;
    if (model < LGCA_HPP || model > LGCA_FHP_III)
        return fail(nullptr, LGCA_ERR_INVALID_ARGUMENT, "lgca_create: unknown model");
    if (dim_x == 0 || dim_y == 0 || coarse_radius == 0)
        return fail(nullptr, LGCA_ERR_INVALID_ARGUMENT, "lgca_create: zero size or radius");
    if (model != LGCA_HPP && (dim_y % 2) != 0)
        return fail(nullptr, LGCA_ERR_INVALID_ARGUMENT,
                    "lgca_create: FHP models need an even dim_y (hexagonal row parity)");
    if (dim_x < 2 * coarse_radius || dim_y < 2 * coarse_radius)
        return fail(nullptr, LGCA_ERR_INVALID_ARGUMENT,
                    "lgca_create: dim_x and dim_y must be at least 2 * coarse_radius");
    if ((uint64_t)dim_x * dim_y >= (1ull << 40) || dim_y > 65535 || dim_x >= (1u << 30))
        return fail(nullptr, LGCA_ERR_INVALID_ARGUMENT,
                    "lgca_create: lattice too large (dim_y <= 65535, dim_x < 2^30)");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return fail(nullptr, LGCA_ERR_INVALID_ARGUMENT, "lgca_create: no such GPU");

    auto* h = new lgca_lattice();
    h->model = model;
    h->kind = model == LGCA_HPP ? kHPP : (model == LGCA_FHP_I ? kFHP1 : kFHP2);
    h->Q = (uint32_t)num_dir_of_kind(h->kind);
    h->dim_x = dim_x;
    h->dim_y = dim_y;
    h->r = coarse_radius;
    h->device = device;
    h->g = make_geometry(dim_x, dim_y, h->Q);
    h->state_words = (size_t)dim_y * h->Q * h->g.P;
    DeviceGuard dg(device);
    auto bail = [&](lgca_status st) {
        g_create_error = h->err;
        lgca_destroy(h);
        return st;
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&h->own_stream, hipStreamNonBlocking)) != hipSuccess)
        return bail(hip_fail(h, e, "hipStreamCreate"));
    h->stream = h->own_stream;
    {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeNumberOfXccs, device) == hipSuccess && n >= 1)
            h->nxcc = (uint32_t)std::min<int>(n, (int)kMaxBands);
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && n >= 1)
            h->ncu = (uint32_t)n;
    }
    const size_t sb = h->state_words * 4;
    if ((e = dalloc(&h->state[0], sb)) != hipSuccess) return bail(hip_fail(h, e, "hipMalloc(state)"));
    if ((e = dalloc(&h->state[1], sb)) != hipSuccess) return bail(hip_fail(h, e, "hipMalloc(state)"));
    if ((e = dalloc(&h->snap, sb)) != hipSuccess) return bail(hip_fail(h, e, "hipMalloc(snapshot)"));
    if ((e = dalloc(&h->counters, 16 * sizeof(unsigned long long))) != hipSuccess)
        return bail(hip_fail(h, e, "hipMalloc(counters)"));
    const uint64_t nw = (uint64_t)h->state_words;
    if ((e = launch_fill_u32(h->state[0], 0u, nw, h->stream)) != hipSuccess ||
        (e = launch_fill_u32(h->state[1], 0u, nw, h->stream)) != hipSuccess ||
        (e = launch_fill_u32(h->snap, 0u, nw, h->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(h->stream)) != hipSuccess)
        return bail(hip_fail(h, e, "initial fill"));
    *out = h;
    return LGCA_OK;
}

extern "C" lgca_status lgca_destroy(lgca_lattice* h)
{
    if (!h) return LGCA_OK;
    {
        DeviceGuard dg(h->device);
        if (h->own_stream) (void)hipStreamSynchronize(h->own_stream);
        if (h->stream && h->stream != h->own_stream) (void)hipStreamSynchronize(h->stream);
        void* ptrs[] = {h->state[0], h->state[1], h->snap, h->mask, h->row_range, h->dec,
                        h->cell_density, h->cell_momentum, h->mean_density, h->mean_momentum,
                        h->velocity, h->bytes, h->visited, h->counters, h->ctrl};
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
        for (const WaveOrder& o : h->orders) (void)hipFree(o.dev);
        for (const PipeCache& c : h->pipes) (void)hipFree(c.order);
        if (h->pipe_exch) (void)hipFree(h->pipe_exch);
        if (h->pipe_ctr) (void)hipFree(h->pipe_ctr);
        for (hipEvent_t ev : h->trial_ev) (void)hipEventDestroy(ev);
        if (h->own_stream) (void)hipStreamDestroy(h->own_stream);
    }
    delete h;
    return LGCA_OK;
}

extern "C" const char* lgca_last_error(const lgca_lattice* h)
{
    return h ? h->err.c_str() : g_create_error.c_str();
}

extern "C" lgca_status lgca_get_info(const lgca_lattice* h, lgca_info* info)
{
    if (!h || !info) return LGCA_ERR_INVALID_ARGUMENT;
    std::memset(info, 0, sizeof *info);
    info->model = h->model;
    info->num_dir = h->Q;
    info->dim_x = h->dim_x;
    info->dim_y = h->dim_y;
    info->coarse_radius = h->r;
    info->coarse_dim_x = h->dim_x / (2 * h->r);
    info->coarse_dim_y = h->dim_y / (2 * h->r);
    info->words_per_row = h->g.W;
    info->row_pitch = h->g.P;
    info->tile_rows = h->tile_th ? h->tile_th + (h->tile_th & 1u) : h->tuned.th;
    info->num_cells = (uint64_t)h->dim_x * h->dim_y;
    info->num_coarse_cells = (uint64_t)info->coarse_dim_x * info->coarse_dim_y;
    info->state_bytes = (uint64_t)h->state_words * 4;
    uint64_t dev = 3 * info->state_bytes + 16 * 8;
    if (h->mask) dev += (uint64_t)h->dim_y * 2 * h->g.P * 4 + (uint64_t)h->dim_y * 8;
    dev += h->dec_cap * 4;
    if (h->cell_density) dev += info->num_cells * 4;
    if (h->cell_momentum) dev += info->num_cells * 8;
    if (h->mean_density) dev += info->num_coarse_cells * 20;
    if (h->bytes) dev += info->num_cells;
    if (h->visited) dev += (info->num_cells + 31) / 32 * 4;
    info->device_bytes = dev;
    info->step = h->step;
    info->seed = h->seed;
    if (h->tile_k != kAutoSteps) info->steps_per_pass = h->tile_k;
    else info->steps_per_pass = h->tile_th ? 4 : (h->tuned.th ? h->tuned.k : -1);
    info->pass_pipeline = (uint32_t)h->pipe;
    return LGCA_OK;
}

extern "C" lgca_status lgca_set_stream(lgca_lattice* h, void* hip_stream)
{
    if (!h) return LGCA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(h->mu);
    DeviceGuard dg(h->device);
    hipStream_t next = hip_stream ? (hipStream_t)hip_stream : h->own_stream;
    if (next != h->stream) {
        // order the new stream after everything queued so far
        hipEvent_t ev;
        HIPCHK(h, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        HIPCHK(h, hipEventRecord(ev, h->stream));
        HIPCHK(h, hipStreamWaitEvent(next, ev, 0));
        HIPCHK(h, hipEventDestroy(ev));
        h->stream = next;
        reset_trial(h);  // a trial's events must share one stream
    }
    return LGCA_OK;
}

// After a sync: the error word of the pipelined launches queued since the
// last read (a pass that gave up waiting has left the state wrong).
static lgca_status check_pipe(lgca_lattice* h)
{
    if (!h->pipe_check) return LGCA_OK;
    h->pipe_check = false;
    uint32_t e = 0;
    HIPCHK(h, hipMemcpy(&e, h->pipe_ctr + kPipeError, 4, hipMemcpyDeviceToHost));
    if (e) return fail(h, LGCA_ERR_HIP, "lgca_step: a pipelined pass gave up waiting for its neighbours");
    return LGCA_OK;
}

extern "C" lgca_status lgca_synchronize(lgca_lattice* h)
{
    if (!h) return LGCA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(h->mu);
    DeviceGuard dg(h->device);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return check_pipe(h);
}

extern "C" lgca_status lgca_set_seed(lgca_lattice* h, uint64_t seed)
{
    if (!h) return LGCA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(h->mu);
    h->seed = seed;
    h->dec_count = 0;
    return LGCA_OK;
}

extern "C" lgca_status lgca_set_step(lgca_lattice* h, uint64_t step)
{
    if (!h) return LGCA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(h->mu);
    h->step = step;
    h->dec_count = 0;
    return LGCA_OK;
}

// ------------------------------------------------------------ cell types

static lgca_status upload_types(lgca_lattice* h, const uint8_t* types, hipMemcpyKind kind)
{
    const size_t n = (size_t)h->dim_x * h->dim_y;
    reset_trial(h);
    lgca_status st = ensure_bytes(h);
    if (st) return st;
    HIPCHK(h, hipMemcpyAsync(h->bytes, types, n, kind, h->stream));
    if ((st = ensure_mask(h))) return st;
    if ((st = clear_counters(h))) return st;
    // row ranges start empty: lo[y] = INT_MAX, hi[y] = 0
    HIPCHK(h, launch_fill_u32((uint32_t*)h->row_range, 0x7FFFFFFFu, h->dim_y, h->stream));
    HIPCHK(h, launch_fill_u32((uint32_t*)h->row_range + h->dim_y, 0u, h->dim_y, h->stream));
    HIPCHK(h, launch_types(h->bytes, h->mask, h->row_range, h->g, (uint32_t*)h->counters, h->stream));
    unsigned long long flag = 0;
    if ((st = read_counters(h, &flag, 1))) return st;
    if (flag & 1ull) return fail(h, LGCA_ERR_INVALID_STATE, "cell types must be 0, 1 or 2");
    // the rows' solid word ranges on the host, for the dispatch order; the
    // stream is idle here, so the cached orders can go
    ++h->types_version;
    for (const WaveOrder& o : h->orders) (void)hipFree(o.dev);
    h->orders.clear();
    for (const PipeCache& c : h->pipes) (void)hipFree(c.order);
    h->pipes.clear();
    h->row_lo.resize(h->dim_y);
    h->row_hi.resize(h->dim_y);
    HIPCHK(h, hipMemcpy(h->row_lo.data(), h->row_range, (size_t)h->dim_y * 4, hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(h->row_hi.data(), h->row_range + h->dim_y, (size_t)h->dim_y * 4, hipMemcpyDeviceToHost));
    return LGCA_OK;
}

extern "C" lgca_status lgca_set_cell_types(lgca_lattice* h, const uint8_t* host_types)
{
    if (!h || !host_types) return LGCA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(h->mu);
    DeviceGuard dg(h->device);
    return upload_types(h, host_types, hipMemcpyHostToDevice);
}

extern "C" lgca_status lgca_get_cell_types(lgca_lattice* h, uint8_t* host_types)
{
    if (!h || !host_types) return LGCA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(h->mu);
    DeviceGuard dg(h->device);
    const size_t n = (size_t)h->dim_x * h->dim_y;
    lgca_status st = ensure_bytes(h);
    if (st) return st;
    HIPCHK(h, launch_types_unpack(h->mask, h->bytes, h->g, h->stream));
    HIPCHK(h, hipMemcpyAsync(host_types, h->bytes, n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return LGCA_OK;
}

// ----------------------------------------------------------------- state

static lgca_status upload_state(lgca_lattice* h, const uint8_t* src, hipMemcpyKind kind)
{
    const size_t n = (size_t)h->dim_x * h->dim_y;
    lgca_status st = ensure_bytes(h);
    if (st) return st;
    HIPCHK(h, hipMemcpyAsync(h->bytes, src, n, kind, h->stream));
    if ((st = clear_counters(h))) return st;
    HIPCHK(h, launch_fill_u32(h->state[h->cur], 0u, h->state_words, h->stream));
    HIPCHK(h, launch_pack(h->bytes, h->state[h->cur], h->g, (uint32_t*)h->counters, h->stream));
    unsigned long long flag = 0;
    if ((st = read_counters(h, &flag, 1))) return st;
    if (flag & 1ull) {
        HIPCHK(h, launch_fill_u32(h->state[h->cur], 0u, h->state_words, h->stream));
        return fail(h, LGCA_ERR_INVALID_STATE, "state bytes carry bits at or above num_dir");
    }
    return LGCA_OK;
}

extern "C" lgca_status lgca_set_state_u8(lgca_lattice* h, const uint8_t* host_state)
{
    if (!h || !host_state) return LGCA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(h->mu);
    DeviceGuard dg(h->device);
    return upload_state(h, host_state, hipMemcpyHostToDevice);
}

extern "C" lgca_status lgca_set_state_u8_device(lgca_lattice* h, const uint8_t* dev_state)
{
    if (!h || !dev_state) return LGCA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(h->mu);
    DeviceGuard dg(h->device);
    return upload_state(h, dev_state, hipMemcpyDeviceToDevice);
}

static lgca_status download_state(lgca_lattice* h, uint8_t* dst, hipMemcpyKind kind)
{
    const size_t n = (size_t)h->dim_x * h->dim_y;
    if (kind == hipMemcpyDeviceToDevice) {
        HIPCHK(h, launch_unpack(h->state[h->cur], dst, h->g, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        return check_pipe(h);
    }
    lgca_status st = ensure_bytes(h);
    if (st) return st;
    HIPCHK(h, launch_unpack(h->state[h->cur], h->bytes, h->g, h->stream));
    HIPCHK(h, hipMemcpyAsync(dst, h->bytes, n, kind, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return check_pipe(h);
}

extern "C" lgca_status lgca_get_state_u8(lgca_lattice* h, uint8_t* host_state)
{
    if (!h || !host_state) return LGCA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(h->mu);
    DeviceGuard dg(h->device);
    return download_state(h, host_state, hipMemcpyDeviceToHost);
}

extern "C" lgca_status lgca_get_state_u8_device(lgca_lattice* h, uint8_t* dev_state)
{
    if (!h || !dev_state) return LGCA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(h->mu);
    DeviceGuard dg(h->device);
    return download_state(h, dev_state, hipMemcpyDeviceToDevice);
}

extern "C" lgca_status lgca_init_random(lgca_lattice* h, double p_fill)
{
    if (!h || !(p_fill >= 0.0) || p_fill > 1.0) return LGCA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(h->mu);
    DeviceGuard dg(h->device);
    HIPCHK(h, launch_init_fill(h->state[h->cur], h->mask, nullptr, h->g, (uint32_t)h->seed,
                               (uint32_t)(h->seed >> 32), hash_threshold(p_fill), h->stream));
    return LGCA_OK;
}

extern "C" lgca_status lgca_init_random_masked(lgca_lattice* h, double p_fill,
                                               const uint8_t* host_region)
{
    if (!h || !host_region || !(p_fill >= 0.0) || p_fill > 1.0) return LGCA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(h->mu);
    DeviceGuard dg(h->device);
    lgca_status st = ensure_bytes(h);
    if (st) return st;
    HIPCHK(h, hipMemcpyAsync(h->bytes, host_region, (size_t)h->dim_x * h->dim_y,
                             hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, launch_init_fill(h->state[h->cur], h->mask, h->bytes, h->g, (uint32_t)h->seed,
                               (uint32_t)(h->seed >> 32), hash_threshold(p_fill), h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return LGCA_OK;
}

// ----------------------------------------------------------------- step

extern "C" lgca_status lgca_set_forcing(lgca_lattice* h, int dir, double p, uint32_t x0, uint32_t x1,
                                        uint32_t y0, uint32_t y1)
{
    if (!h) return LGCA_ERR_INVALID_ARGUMENT;
    if (dir != LGCA_FORCE_NONE && dir != LGCA_FORCE_X && dir != LGCA_FORCE_Y)
        return fail(h, LGCA_ERR_INVALID_ARGUMENT, "lgca_set_forcing: dir must be NONE, X or Y");
    if (dir != LGCA_FORCE_NONE && !(p >= 0.0 && p <= 1.0))
        return fail(h, LGCA_ERR_INVALID_ARGUMENT, "lgca_set_forcing: p must lie in [0, 1]");
    std::lock_guard<std::mutex> lk(h->mu);
    reset_trial(h);
    h->force_dir = dir;
    h->force_thr = hash_threshold(p);
    h->fx0 = x0;
    h->fx1 = x1 < h->dim_x ? x1 : h->dim_x;
    h->fy0 = y0;
    h->fy1 = y1 < h->dim_y ? y1 : h->dim_y;
    if (h->fx0 >= h->fx1 || h->fy0 >= h->fy1 || h->force_thr == 0) h->force_dir = LGCA_FORCE_NONE;
    h->dec_count = 0;
    return LGCA_OK;
}

// Decision words per step of the force region: its rows x its words.
static size_t decision_words_per_step(const lgca_lattice* h)
{
    const uint32_t fw0 = h->fx0 / 32, fw1 = (h->fx1 + 31) / 32;
    return (size_t)(fw1 - fw0) * (h->fy1 - h->fy0);
}

// Queues the force draw of steps [step, step + nsteps) into dec.
static lgca_status draw_decisions(lgca_lattice* h, uint32_t* dec, uint32_t nsteps)
{
    const uint32_t fw0 = h->fx0 / 32, fw1 = (h->fx1 + 31) / 32;
    ForceDrawArgs fa{};
    fa.dec = dec;
    fa.nsteps = nsteps;
    fa.step0 = (uint32_t)h->step;
    fa.key0 = (uint32_t)h->seed;
    fa.key1 = (uint32_t)(h->seed >> 32);
    fa.q = force_q(h->force_thr);
    fa.fx0 = h->fx0;
    fa.fx1 = h->fx1;
    fa.fy0 = h->fy0;
    fa.fy1 = h->fy1;
    fa.fw0 = fw0;
    fa.fwn = fw1 - fw0;
    fa.dim_x = h->dim_x;
    HIPCHK(h, launch_force_draw(fa, h->stream));
    return LGCA_OK;
}

// Force decisions are drawn for a window of upcoming steps at once; a window
// always starts at the current step and covers at least `need` steps.
static lgca_status prepare_decisions(lgca_lattice* h, uint64_t remaining, uint32_t need)
{
    const size_t per_step = decision_words_per_step(h);
    // up to 64 steps, capped at 64 MiB of decision words, never fewer than `need`
    size_t k = (size_t)(64ull << 20) / 4 / per_step;
    if (k > 64) k = 64;
    if (k > remaining) k = (size_t)remaining;
    if (k < need) k = need;
    if (h->dec_cap < k * per_step) {
        if (h->dec) {
            HIPCHK(h, hipStreamSynchronize(h->stream));
            HIPCHK(h, hipFree(h->dec));
        }
        h->dec = nullptr;
        h->dec_cap = 0;
        HIPCHK(h, dalloc(&h->dec, k * per_step * 4));
        h->dec_cap = k * per_step;
    }
    lgca_status st = draw_decisions(h, h->dec, (uint32_t)k);
    if (st) return st;
    h->dec_step0 = h->step;
    h->dec_count = (uint32_t)k;
    return LGCA_OK;
}

// Steps of force decisions a pipelined launch may cover: up to 256, as far
// as 64 MiB of decision words allow, at least one pass.
static uint64_t pipe_window_steps(const lgca_lattice* h, int k)
{
    const uint64_t fit = (64ull << 20) / 4 / std::max<size_t>(decision_words_per_step(h), 1);
    return std::max<uint64_t>(std::min<uint64_t>(fit, 256), (uint64_t)k);
}

// Rows per tile: tall enough to amortise the 2K halo rows, short enough to
// give the GPU a few thousand waves.
static uint32_t auto_tile_rows(const lgca_lattice* h, int k)
{
    if (h->tile_th) return h->tile_th + (h->tile_th & 1u);
    // rows x waves per row: strips per row, or 1 / (rows per wave) in segment mode
    const uint64_t wave_rows = tile_segments(h->g, k) ? (uint64_t)h->dim_y / (64 / h->g.W)
                                                      : (uint64_t)h->dim_y * tile_strips(h->g, k);
    // about 2.1 rounds of resident waves (four per SIMD on 1024 SIMDs); on the
    // 32768 x 8192 lattice this gives 16-row tiles, the fastest 4-step shape
    // there.  Small lattices get short tiles: a wave's walk of th + 2K rows is
    // then their critical path (2-row tiles run the 2048 x 512 pipe 1.6x
    // faster than 8)
    uint64_t th = wave_rows / 8704;
    if (th < 2) th = 2;
    if (th > 256) th = 256;
    if (th > h->dim_y) th = h->dim_y;
    th += th & 1u;  // tiles start on even rows
    return (uint32_t)th;
}

// Keeps candidate `pick` of a finished trial, and the fastest 4-step
// candidate for the 4-step passes that finish a call of an 8-step choice.
static void keep_trial_choice(lgca_lattice* h, const std::vector<LaunchShape>& cand,
                              const std::vector<float>& best, size_t pick)
{
    h->tuned = cand[pick];
    h->tuned_rem4 = 0;
    float b4 = 1e30f;
    for (size_t i = 0; i < cand.size(); ++i)
        if (cand[i].k == 4 && best[i] < b4) {
            b4 = best[i];
            h->tuned_rem4 = cand[i].th;
        }
}

// Rows per tile of a k-step pass that is not of the kept steps per pass.
static uint32_t other_tile_rows(const lgca_lattice* h, int k)
{
    if (k == 4 && !h->tile_th && h->tuned.th && h->tuned_rem4) return h->tuned_rem4;
    return auto_tile_rows(h, k);
}

// The largest steps per launch, at most `want` and `left`.  A pass may start
// at any step: the tiled kernels are compiled for each offset of the first
// step within its four-step chirality group.
static int steps_for_launch(int want, uint64_t left)
{
    if (want >= 8 && left >= 8) return 8;
    if (want >= 4 && left >= 4) return 4;
    if (want >= 2 && left >= 2) return 2;
    return 1;
}

// Shape of the next tiled launch (steps k, rows per tile th), and the trial
// group whose start event goes before it and whose end event goes after it
// (-1: none).  With automatic rows, the first launches run candidate shapes,
// each twice in a group of consecutive launches between two events on the
// lattice's stream: at the configured steps per pass (4 when automatic) the
// rule's height, then 4, 8, 12, 16, 20, 24, 32 rows; with automatic steps on
// a lattice that fills the GPU (the rule's height is 4 rows or more), 8-step
// passes of 16 to 64 rows, which move the state through HBM half as often
// (the 32768 x 8192 Karman lattice runs 8-step passes of 40 rows 1.4x faster
// than 4-step passes of 16), and on a lattice the rule gives 2-row tiles,
// 2-step passes of 2 and 4 rows, whose shorter walks suit lattices too
// small to fill the GPU (the 2048 x 512 pipe runs 10 % faster so).  Groups are single launches, or four
// when 2-step passes compete: then the gaps between launches must count as
// they do in a long run, and timing each launch alone penalised the shorter
// 2-step launches.  Once the last group has finished (queried, never waited
// for) the fastest per step is kept, the rule's unless another beats it by
// more than 2 %.  Results do not depend on the shape.
constexpr uint32_t kTrialReps = 2;

// The trial's candidate shapes, the rule's own first.
static TileArgs tile_args(const lgca_lattice* h, uint32_t th, const uint32_t* fdec);
static TileArgs with_k(TileArgs a, int k);

// Waves of a strip-mode pass of k steps and th-row tiles (tile_wave_order's
// entries: solid walks count once per part), 0 when it has no such order.
static uint64_t pass_waves(const lgca_lattice* h, int k, uint32_t th)
{
    static const uint32_t any_decision = 0;  // only its presence is read
    const TileArgs a = with_k(tile_args(h, th, h->force_thr ? &any_decision : nullptr), k);
    const bool solids = h->mask && h->row_lo.size() == h->dim_y;
    std::vector<uint32_t> v;
    if (!tile_wave_order(a, k, solids ? h->row_lo.data() : nullptr, solids ? h->row_hi.data() : nullptr, v))
        return 0;
    return v.size();
}

// The shortest even tile height (8 rows or more) whose k-step passes have at
// most `slots` waves, or 0 when none has.  The wave count falls as tiles grow
// (fewer tiles, each cut the same way), so a bisection finds it.
static uint32_t fill_tile_rows(const lgca_lattice* h, int k, uint64_t slots)
{
    uint32_t lo = 4, hi = std::min<uint32_t>(h->dim_y, 65534) / 2;  // heights 2 lo .. 2 hi
    if (hi < lo) return 0;
    const uint64_t top = pass_waves(h, k, 2 * hi);
    if (!top || top > slots) return 0;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        const uint64_t n = pass_waves(h, k, 2 * mid);
        if (n && n <= slots) hi = mid;
        else lo = mid + 1;
    }
    return 2 * lo;
}

static std::vector<LaunchShape> trial_candidates(const lgca_lattice* h)
{
    const int base_k = h->tile_k == kAutoSteps ? 4 : h->tile_k;
    auto too_big = [&](int ck) {  // see lgca_step: decisions of a pass over 2 GiB
        return h->force_dir != LGCA_FORCE_NONE &&
               (uint64_t)ck * decision_words_per_step(h) * 4u >= ((uint64_t)1 << 31);
    };
    int k0 = base_k;
    while (k0 > 1 && too_big(k0)) k0 /= 2;
    std::vector<LaunchShape> c(1, LaunchShape{k0, auto_tile_rows(h, k0)});
    auto add = [&](int ck, uint32_t cth) {
        if (cth > h->dim_y || too_big(ck)) return;  // too big: cut by lgca_step, not timed
        for (const LaunchShape& s : c)
            if (s.k == ck && s.th == cth) return;
        c.push_back(LaunchShape{ck, cth});
    };
    for (uint32_t r : {4u, 8u, 12u, 16u, 20u, 24u, 32u}) add(base_k, r);
    if (h->tile_k == kAutoSteps && c[0].th == 2) {
        // small lattices: their walks are short chains of dependent steps, and
        // fewer or more steps per launch can win (the 2048 x 512 pipe keeps
        // 2-step passes, the 256 x 256 HPP lattice 8-step ones: 0.84 against
        // 1.14 us per step at 4)
        add(2, 2);
        add(2, 4);
        add(8, 2);
    }
    if (h->tile_k == kAutoSteps && c[0].th >= 4) {
        for (uint32_t r : {16u, 24u, 32u, 40u, 48u, 64u}) add(8, r);
        // Tiles tall enough that a pass's waves fill the GPU's wave slots
        // once (or twice) and no more: a pass with a few waves past a whole
        // number of rounds ends with a tail of them on an otherwise idle GPU
        // (32768 x 8192 Karman lattice, 8-step passes, 2048 slots: 76-row
        // tiles, 1986 waves, 21.1 us per step; 72 rows, 2094 waves, 23.2 us)
        const uint32_t slots = h->ncu * (uint32_t)std::max(tile_waves_per_cu(h->kind, 8, h->g.nb != 32), 0);
        for (uint32_t m = 1; m <= 2 && slots && !tile_segments(h->g, 8); ++m) {
            const uint32_t th = fill_tile_rows(h, 8, (uint64_t)m * slots);
            if (th) {
                add(8, th);
                add(8, th + 4);
            }
        }
    }
    return c;
}

static lgca_status plan_launch(lgca_lattice* h, uint64_t left, int* k, uint32_t* th, int* ev_start,
                               int* ev_end)
{
    *ev_start = *ev_end = -1;
    const int base_k = h->tile_k == kAutoSteps ? 4 : h->tile_k;
    if (h->tile_th) {
        *k = steps_for_launch(base_k, left);
        *th = auto_tile_rows(h, *k);
        return LGCA_OK;
    }
    if (!h->tuned.th && h->trial.empty()) {
        h->trial = trial_candidates(h);
        h->trial_group = 1;
        for (const LaunchShape& s : h->trial)
            if (s.k == 2 && base_k != 2) h->trial_group = 4;
        h->trial_issued = 0;
        h->trial_open = -1;
        const size_t groups = kTrialReps * h->trial.size();
        while (h->trial_ev.size() < 2 * groups) {
            hipEvent_t ev;
            HIPCHK(h, hipEventCreate(&ev));
            h->trial_ev.push_back(ev);
        }
        h->trial_steps.assign(groups, 0);
    }
    const uint32_t groups = kTrialReps * (uint32_t)h->trial.size();
    const uint32_t n = h->trial_group * groups;  // candidate launches in all
    if (!h->tuned.th && h->trial_issued == n && h->trial_open < 0) {
        const hipError_t q = hipEventQuery(h->trial_ev[2 * (groups - 1) + 1]);
        if (q != hipErrorNotReady) {
            HIPCHK(h, q);
            std::vector<float> best(h->trial.size(), 1e30f);
            for (uint32_t e = 0; e < groups; ++e) {
                float ms = 0.0f;
                HIPCHK(h, hipEventElapsedTime(&ms, h->trial_ev[2 * e], h->trial_ev[2 * e + 1]));
                const size_t c = e / kTrialReps;
                best[c] = std::min(best[c], ms / (float)std::max<uint32_t>(h->trial_steps[e], 1));
            }
            size_t pick = 0;
            for (size_t c = 1; c < best.size(); ++c)
                if (best[c] < best[pick]) pick = c;
            if (best[0] <= 1.02f * best[pick]) pick = 0;
            keep_trial_choice(h, h->trial, best, pick);
        }
    }
    if (h->tuned.th) {
        *k = steps_for_launch(h->tuned.k, left);
        *th = *k == h->tuned.k ? h->tuned.th : other_tile_rows(h, *k);
        return LGCA_OK;
    }
    if (h->trial_issued < n) {
        const uint32_t e = h->trial_issued / h->trial_group, pos = h->trial_issued % h->trial_group;
        const LaunchShape& c = h->trial[e / kTrialReps];
        *k = steps_for_launch(c.k, left);
        if (*k == c.k) {
            // a launch that the steps left force to another k inside a
            // group is timed with it, and its steps are counted
            *th = c.th;
            if (pos == 0) *ev_start = (int)e;
            if (pos + 1 == h->trial_group) *ev_end = (int)e;
            ++h->trial_issued;
            return LGCA_OK;
        }
    } else {
        *k = steps_for_launch(base_k, left);  // the rule's shape meanwhile
    }
    *th = auto_tile_rows(h, *k);
    return LGCA_OK;
}

extern "C" lgca_status lgca_set_tuning(lgca_lattice* h, int steps_per_pass, uint32_t tile_rows)
{
    if (!h) return LGCA_ERR_INVALID_ARGUMENT;
    if (steps_per_pass != kAutoSteps && steps_per_pass != 0 && steps_per_pass != 1 && steps_per_pass != 2 &&
        steps_per_pass != 4 && steps_per_pass != 8)
        return fail(h, LGCA_ERR_INVALID_ARGUMENT,
                    "lgca_set_tuning: steps_per_pass must be -1, 0, 1, 2, 4 or 8");
    if (tile_rows > 65534)
        return fail(h, LGCA_ERR_INVALID_ARGUMENT, "lgca_set_tuning: tile_rows must be at most 65534");
    std::lock_guard<std::mutex> lk(h->mu);
    h->tile_k = steps_per_pass;
    h->tile_th = tile_rows;
    reset_trial(h);
    return LGCA_OK;
}

// Multi-pass launches stay switched off: on the 4096 x 1024 periodic lattice
// at 8 steps per pass and 16-row tiles, their waits did not complete (the
// first launch ran past 100 s, with completion published first by a
// read-modify-write count, then by device-coherent stores), and the cause is
// not known.  lgca_set_pipeline(h, 1) fails until it is.
constexpr bool kPipelineAvailable = false;

extern "C" lgca_status lgca_set_pipeline(lgca_lattice* h, int on)
{
    if (!h) return LGCA_ERR_INVALID_ARGUMENT;
    if (on != 0 && on != 1) return fail(h, LGCA_ERR_INVALID_ARGUMENT, "lgca_set_pipeline: on must be 0 or 1");
    if (on && !kPipelineAvailable)
        return fail(h, LGCA_ERR_INVALID_STATE,
                    "lgca_set_pipeline: multi-pass launches are switched off in this build (their waits did "
                    "not complete on the GPU)");
    std::lock_guard<std::mutex> lk(h->mu);
    h->pipe = on;
    return LGCA_OK;
}

// Arguments of one tiled pass from the live state into the other buffer, at
// the current step, with rows per tile th and the force decisions of that
// step (or null).
static TileArgs tile_args(const lgca_lattice* h, uint32_t th, const uint32_t* fdec)
{
    const uint32_t fw0 = h->fx0 / 32, fw1 = (h->fx1 + 31) / 32;
    TileArgs a{};
    a.g = h->g;
    a.src = h->state[h->cur];
    a.dst = h->state[h->cur ^ 1];
    a.mask = h->mask;
    a.row_range = h->row_range;
    a.key0 = (uint32_t)h->seed;
    a.key1 = (uint32_t)(h->seed >> 32);
    a.t0 = (uint32_t)h->step;
    a.th = th;
    a.fdec = fdec;
    a.fdec_stride = (uint64_t)(fw1 - fw0) * (h->fy1 - h->fy0);
    a.fy0 = h->fy0;
    a.fy1 = h->fy1;
    a.fw0 = fw0;
    a.fwn = fw1 - fw0;
    a.force_dir = h->force_dir;
    return a;
}

// The halo of a pass of k steps (a TileArgs from tile_args with it set).
static TileArgs with_k(TileArgs a, int k)
{
    a.halo = tile_halo(a.g, k);
    return a;
}

// The dispatch order of a strip-mode pass of k steps with arguments a, built
// and cached on first use; null in segment mode, when it cannot be built, or
// when it is not cached and may_alloc is false (the kernel then uses its own
// fixed order; results never depend on the order).
static OrderRef wave_order(lgca_lattice* h, int k, const TileArgs& a, bool may_alloc)
{
    if (tile_segments(h->g, k)) return {};
    uint64_t key = h->types_version;
    for (uint64_t v : {(uint64_t)(a.fdec != nullptr), (uint64_t)a.fw0, (uint64_t)a.fwn, (uint64_t)a.fy0,
                       (uint64_t)a.fy1})
        key = key * 0x100000001B3ull ^ v;
    for (const WaveOrder& o : h->orders)
        if (o.k == k && o.th == a.th && o.halo == a.halo && o.key == key) return {o.dev, o.n};
    if (!may_alloc) return {};
    std::vector<uint32_t> v;
    const bool solids = h->mask && h->row_lo.size() == h->dim_y;
    if (!tile_wave_order(a, k, solids ? h->row_lo.data() : nullptr, solids ? h->row_hi.data() : nullptr, v))
        return {};
    if (h->orders.size() >= 16) {
        // the oldest may still be read by queued passes
        if (hipStreamSynchronize(h->stream) != hipSuccess) return {};
        (void)hipFree(h->orders.front().dev);
        h->orders.erase(h->orders.begin());
    }
    uint32_t* dev = nullptr;
    if (dalloc(&dev, v.size() * 4) != hipSuccess) return {};
    if (hipMemcpy(dev, v.data(), v.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(dev);
        return {};
    }
    h->orders.push_back(WaveOrder{k, a.th, a.halo, key, dev, (uint32_t)v.size()});
    return {dev, (uint32_t)v.size()};
}

// The band plan of a pipelined launch of k-step passes with arguments a, built
// and cached on first use (with the exchange rows and counters sized for
// it); null when the shape does not allow one or an allocation fails (the
// caller then launches pass by pass).
static const PipeCache* pipe_plan(lgca_lattice* h, int k, const TileArgs& a)
{
    uint64_t key = h->types_version;
    for (uint64_t v : {(uint64_t)(a.fdec != nullptr), (uint64_t)a.fw0, (uint64_t)a.fwn, (uint64_t)a.fy0,
                       (uint64_t)a.fy1})
        key = key * 0x100000001B3ull ^ v;
    for (const PipeCache& c : h->pipes)
        if (c.k == k && c.th == a.th && c.halo == a.halo && c.key == key) return &c;
    PipeCache c{k, a.th, a.halo, key, PipePlan{}, nullptr};
    const bool solids = h->mask && h->row_lo.size() == h->dim_y;
    if (!tile_pipe_plan(a, k, h->nxcc, solids ? h->row_lo.data() : nullptr, solids ? h->row_hi.data() : nullptr,
                        c.plan))
        return nullptr;
    const size_t xw = tile_pipe_exch_words(h->g, k, c.plan.nband);
    const size_t cw = kPipeDone + (size_t)c.plan.ntiles * c.plan.nstrips;
    // the buffers grow only between launches: wait for queued ones first
    if ((xw > h->pipe_exch_cap || cw > h->pipe_ctr_cap || h->pipes.size() >= 16) &&
        hipStreamSynchronize(h->stream) != hipSuccess)
        return nullptr;
    if (xw > h->pipe_exch_cap) {
        if (h->pipe_exch) (void)hipFree(h->pipe_exch);
        h->pipe_exch = nullptr;
        h->pipe_exch_cap = 0;
        if (dalloc(&h->pipe_exch, xw * 4) != hipSuccess) return nullptr;
        h->pipe_exch_cap = xw;
    }
    if (cw > h->pipe_ctr_cap) {
        if (h->pipe_ctr) (void)hipFree(h->pipe_ctr);
        h->pipe_ctr = nullptr;
        h->pipe_ctr_cap = 0;
        if (dalloc(&h->pipe_ctr, cw * 4) != hipSuccess) return nullptr;
        h->pipe_ctr_cap = cw;
    }
    if (dalloc(&c.order, c.plan.order.size() * 4) != hipSuccess) return nullptr;
    if (hipMemcpy(c.order, c.plan.order.data(), c.plan.order.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(c.order);
        return nullptr;
    }
    if (h->pipes.size() >= 16) {
        (void)hipFree(h->pipes.front().order);
        h->pipes.erase(h->pipes.begin());
    }
    h->pipes.push_back(std::move(c));
    return &h->pipes.back();
}

static lgca_status step_locked(lgca_lattice* h, uint64_t nsteps, uint32_t* ext_dec = nullptr,
                               size_t ext_words = 0);

// The launch-shape trial of plan_launch, run at once and waited for: each
// candidate shape is timed over groups of consecutive passes covering eight
// steps (three interleaved rounds, the best of each kept), and the fastest
// per step is kept, the rule's unless another beats it by more than 2 %; the
// kept shape then makes one untimed pass at each other step offset within a
// four-step group (the kernels are compiled per offset), so that the steps
// that follow run warm.  Every pass reads the live state and writes the other
// buffer, which the next step overwrites anyway, so neither the state nor the
// step changes.
constexpr uint32_t kTuneReps = 3;
extern "C" lgca_status lgca_tune(lgca_lattice* h)
{
    if (!h) return LGCA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(h->mu);
    DeviceGuard dg(h->device);
    if (h->tile_k == 0 || h->tile_th) return LGCA_OK;  // no automatic choice to make
    const std::vector<LaunchShape> cand = trial_candidates(h);
    int kmax = 1;
    for (const LaunchShape& c : cand) kmax = std::max(kmax, c.k);
    // decisions for the candidates' passes and for the warm-up passes below,
    // which start up to three steps later
    const uint32_t* fdec = nullptr;
    const size_t per_step = decision_words_per_step(h);
    const uint32_t need = (uint32_t)kmax + 3;
    if (h->force_dir != LGCA_FORCE_NONE) {
        if (h->dec_count == 0 || h->step < h->dec_step0 || h->step + need > h->dec_step0 + h->dec_count) {
            lgca_status st = prepare_decisions(h, 64, need);
            if (st) return st;
        }
        fdec = h->dec + (size_t)(h->step - h->dec_step0) * per_step;
    }
    hipEvent_t ev[2] = {nullptr, nullptr};
    lgca_status st = LGCA_OK;
    auto fail_hip = [&](hipError_t e, const char* what) {
        if (st == LGCA_OK && e != hipSuccess) st = hip_fail(h, e, what);
        return e != hipSuccess;
    };
    if (!fail_hip(hipEventCreate(&ev[0]), "hipEventCreate") && !fail_hip(hipEventCreate(&ev[1]), "hipEventCreate")) {
        std::vector<float> best(cand.size(), 1e30f);
        for (uint32_t rep = 0; rep < kTuneReps && st == LGCA_OK; ++rep) {
            for (size_t i = 0; i < cand.size() && st == LGCA_OK; ++i) {
                const TileArgs a = with_k(tile_args(h, cand[i].th, fdec), cand[i].k);
                const int n = 8 / cand[i].k;
                // the first pass of a shape loads its code: untimed
                const OrderRef ord = wave_order(h, cand[i].k, a, true);
                if (rep == 0 && fail_hip(launch_tile(h->kind, cand[i].k, a, h->stream, ord.dev, ord.n), "launch_tile"))
                    break;
                if (fail_hip(hipEventRecord(ev[0], h->stream), "hipEventRecord")) break;
                for (int l = 0; l < n && st == LGCA_OK; ++l)
                    fail_hip(launch_tile(h->kind, cand[i].k, a, h->stream, ord.dev, ord.n), "launch_tile");
                if (st || fail_hip(hipEventRecord(ev[1], h->stream), "hipEventRecord") ||
                    fail_hip(hipEventSynchronize(ev[1]), "hipEventSynchronize"))
                    break;
                float ms = 0.0f;
                if (fail_hip(hipEventElapsedTime(&ms, ev[0], ev[1]), "hipEventElapsedTime")) break;
                best[i] = std::min(best[i], ms / 8.0f);
            }
        }
        if (st == LGCA_OK) {
            size_t pick = 0;
            for (size_t i = 1; i < best.size(); ++i)
                if (best[i] < best[pick]) pick = i;
            if (best[0] <= 1.02f * best[pick]) pick = 0;
            h->trial.clear();
            h->trial_issued = 0;
            h->trial_open = -1;
            keep_trial_choice(h, cand, best, pick);
            // one untimed pass of the kept shape at each later offset within
            // the four-step group, and of the shapes that finish calls whose
            // steps are not a multiple of it (fewer steps per pass) at every
            // offset, so that each kernel the following steps use has run once
            // (a first launch loads code and takes longer)
            for (int k = h->tuned.k; k >= 1 && st == LGCA_OK; k /= 2) {
                const uint32_t th = k == h->tuned.k ? h->tuned.th : other_tile_rows(h, k);
                for (uint32_t o = k == h->tuned.k ? 1 : 0; o < 4 && (k > 1 || o == 0) && st == LGCA_OK; ++o) {
                    TileArgs a = with_k(tile_args(h, th, fdec ? fdec + o * per_step : nullptr), k);
                    a.t0 += o;
                    const OrderRef ord = wave_order(h, k, a, true);
                    fail_hip(launch_tile(h->kind, k, a, h->stream, ord.dev, ord.n), "launch_tile");
                }
            }
            if (st == LGCA_OK) fail_hip(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
        }
    }
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    return st;
}

extern "C" lgca_status lgca_step(lgca_lattice* h, uint64_t nsteps)
{
    if (!h) return LGCA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(h->mu);
    DeviceGuard dg(h->device);
    return step_locked(h, nsteps);
}

// ------------------------------------------------- step without allocating
// lgca_step_noalloc draws its force decisions into `scratch`, a fixed array
// in device memory, instead of the handle's decision buffer, which lgca_step
// allocates (and grows) on demand.  Each GPU the library runs on holds its
// own copy of the array until the process ends.  Host code never passes the
// array's name to a kernel or copy: on the host it is only a stand-in for
// the device symbol.  The array's address on each GPU is looked up once, at
// that GPU's first call, with the GPU current, and kept in g_scratch_dev.
constexpr size_t kNoAllocScratchWords = (size_t)2 << 20;  // 8 MiB of decision words
static __device__ __align__(256) uint32_t scratch[kNoAllocScratchWords];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_dev[kMaxScratchDevices];  // per GPU; null until looked up
}  // namespace

// The device address of `scratch` on the handle's GPU (current when called).
static lgca_status scratch_address(lgca_lattice* h, uint32_t** out)
{
    if (h->device < 0 || h->device >= kMaxScratchDevices)
        return fail(h, LGCA_ERR_INVALID_ARGUMENT, "lgca_step_noalloc: device index beyond 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[h->device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(h, e, "hipGetSymbolAddress(scratch)");
        g_scratch_dev[h->device] = (uint32_t*)p;
    }
    *out = g_scratch_dev[h->device];
    return LGCA_OK;
}

// The launch shape kept for lgca_step: the fixed rows, else the trial's
// choice once made, else the rule's.
static LaunchShape kept_shape(const lgca_lattice* h)
{
    const int base_k = h->tile_k == kAutoSteps ? 4 : h->tile_k;
    if (!h->tile_th && h->tuned.th) return h->tuned;
    return LaunchShape{base_k, auto_tile_rows(h, base_k)};
}

extern "C" lgca_status lgca_step_noalloc(lgca_lattice* h, uint64_t nsteps)
{
    if (!h) return LGCA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(h->mu);
    DeviceGuard dg(h->device);
    uint32_t* dev_scratch = nullptr;
    lgca_status st = scratch_address(h, &dev_scratch);
    if (st) return st;
    if (h->force_dir != LGCA_FORCE_NONE) {
        const int k = h->tile_k == 0 ? 1 : kept_shape(h).k;
        const size_t need = decision_words_per_step(h) * (size_t)k;
        if (need > kNoAllocScratchWords)
            return fail(h, LGCA_ERR_INVALID_ARGUMENT,
                        "lgca_step_noalloc: a " + std::to_string(k) + "-step pass needs " +
                            std::to_string(need) + " force-decision words; the fixed scratch holds " +
                            std::to_string(kNoAllocScratchWords));
    }
    return step_locked(h, nsteps, dev_scratch, kNoAllocScratchWords);
}

// The body of lgca_step; the caller holds h->mu and the device.  With
// ext_dec, force decisions are drawn into those ext_words words (windows of
// at least one pass, checked by the caller) instead of the handle's own
// buffer, and launches take the kept shape without trial launches, so that
// nothing is allocated.
static lgca_status step_locked(lgca_lattice* h, uint64_t nsteps, uint32_t* ext_dec, size_t ext_words)
{
    constexpr uint64_t kMaxPassDecisionBytes = (uint64_t)1 << 31;
    const bool forcing = h->force_dir != LGCA_FORCE_NONE;
    const uint32_t fw0 = h->fx0 / 32, fw1 = (h->fx1 + 31) / 32;
    const size_t per_step = decision_words_per_step(h);
    uint64_t ext_step0 = 0;  // window held in ext_dec
    uint32_t ext_count = 0;
    uint64_t left = nsteps;
    while (left > 0) {
        int k = 1, ev_start = -1, ev_end = -1;
        uint32_t th = 0;
        if (h->tile_k != 0 && ext_dec) {
            const LaunchShape sh = kept_shape(h);
            k = steps_for_launch(sh.k, left);
            th = k == sh.k ? sh.th : other_tile_rows(h, k);
        } else if (h->tile_k != 0) {
            lgca_status st = plan_launch(h, left, &k, &th, &ev_start, &ev_end);
            if (st) return st;
        }
        // A pass reads its k steps of force decisions through one buffer
        // descriptor of under 2 GiB: lattices whose forced region is larger
        // than that per pass (some 2 G forced cells at 8 steps) take fewer
        // steps per pass (a trial launch so cut is timed with its group).
        if (forcing && h->tile_k != 0 && (uint64_t)k * per_step * 4u >= kMaxPassDecisionBytes) {
            while (k > 1 && (uint64_t)k * per_step * 4u >= kMaxPassDecisionBytes) k /= 2;
            if ((uint64_t)k * per_step * 4u >= kMaxPassDecisionBytes)
                return fail(h, LGCA_ERR_INVALID_ARGUMENT, "lgca_step: the force region's decisions of one step exceed 2 GiB");
            th = other_tile_rows(h, k);
        }
        // Several passes of the kept shape in one launch: all the steps left, as
        // far as one window of force decisions reaches (not while the trial
        // times single launches, nor without allocating)
        uint64_t npass = 1;
        const PipeCache* pc = nullptr;
        if (h->pipe && !ext_dec && h->tile_k != 0 && ev_start < 0 && ev_end < 0 && h->trial_open < 0 &&
            (k == 4 || k == 8) && left >= 2 * (uint64_t)k) {
            uint64_t np = left / (uint64_t)k;
            if (forcing) np = std::min<uint64_t>(np, pipe_window_steps(h, k) / (uint64_t)k);
            if (np >= 2) {
                // the plan reads only whether decisions exist, not where
                pc = pipe_plan(h, k, with_k(tile_args(h, th, forcing ? h->state[0] : nullptr), k));
                if (pc) npass = np;
            }
        }
        const uint64_t span = npass * (uint64_t)k;  // steps of this launch
        const uint32_t* fdec = nullptr;
        if (forcing && ext_dec) {
            if (ext_count == 0 || h->step + k > ext_step0 + ext_count) {
                uint64_t kw = std::min<uint64_t>(ext_words / per_step, 64);
                kw = std::max<uint64_t>(std::min<uint64_t>(kw, left), (uint64_t)k);
                lgca_status st = draw_decisions(h, ext_dec, (uint32_t)kw);
                if (st) return st;
                ext_step0 = h->step;
                ext_count = (uint32_t)kw;
            }
            fdec = ext_dec + (size_t)(h->step - ext_step0) * per_step;
        } else if (forcing) {
            if (h->dec_count == 0 || h->step < h->dec_step0 ||
                h->step + span > h->dec_step0 + h->dec_count) {
                lgca_status st = prepare_decisions(h, left, (uint32_t)span);
                if (st) return st;
            }
            fdec = h->dec + (size_t)(h->step - h->dec_step0) * per_step;
        }
        if (pc) {
            const PipePlan& pl = pc->plan;
            const TileArgs a = with_k(tile_args(h, th, fdec), k);
            PipeArgs p{};
            p.buf[0] = h->state[h->cur];
            p.buf[1] = h->state[h->cur ^ 1];
            p.npass = (uint32_t)npass;
            p.nband = pl.nband;
            p.nstrips = pl.nstrips;
            p.ntiles = pl.ntiles;
            p.queue = h->pipe_ctr + kPipeQueue;
            p.error = h->pipe_ctr + kPipeError;
            p.done = h->pipe_ctr + kPipeDone;
            p.order = pc->order;
            for (uint32_t b = 0; b < pl.nband; ++b) {
                p.band_first[b] = pl.band_first[b];
                p.band_items[b] = pl.band_items[b];
                p.band_y0[b] = pl.band_y0[b];
                p.band_y1[b] = pl.band_y1[b];
            }
            p.exch = h->pipe_exch;
            p.fdec_pass = (uint64_t)k * per_step;
            // every wave slot of the GPU (two waves per SIMD at 8 steps, four at 4)
            uint32_t nwaves = std::max<uint32_t>(h->ncu, 1) * 4u * (k == 8 ? 2u : 4u);
            nwaves = (nwaves + pl.nband - 1) / pl.nband * pl.nband;
            HIPCHK(h, launch_fill_u32(h->pipe_ctr, 0u, kPipeDone + (size_t)pl.ntiles * pl.nstrips, h->stream));
            HIPCHK(h, launch_tile_pipe(h->kind, k, a, p, nwaves, h->stream));
            h->pipe_check = true;
            h->cur ^= (int)(npass & 1u);
            h->step += span;
            left -= span;
            continue;
        }
        if (h->tile_k == 0) {
            StepArgs a{};
            a.g = h->g;
            a.mask = h->mask;
            a.row_range = h->row_range;
            a.key0 = (uint32_t)h->seed;
            a.key1 = (uint32_t)(h->seed >> 32);
            a.force_dir = h->force_dir;
            a.fdec = fdec;
            a.fy0 = h->fy0;
            a.fy1 = h->fy1;
            a.fw0 = fw0;
            a.fwn = fw1 - fw0;
            a.src = h->state[h->cur];
            a.dst = h->state[h->cur ^ 1];
            a.step = (uint32_t)h->step;
            HIPCHK(h, launch_step(h->kind, a, h->stream));
        } else {
            if (ev_start >= 0) {
                HIPCHK(h, hipEventRecord(h->trial_ev[2 * ev_start], h->stream));
                h->trial_open = ev_start;
                h->trial_steps[ev_start] = 0;
            }
            const TileArgs a = with_k(tile_args(h, th, fdec), k);
            const OrderRef ord = wave_order(h, k, a, ext_dec == nullptr);
            HIPCHK(h, launch_tile(h->kind, k, a, h->stream, ord.dev, ord.n));
            if (h->trial_open >= 0) h->trial_steps[h->trial_open] += (uint32_t)k;
            if (ev_end >= 0) {
                HIPCHK(h, hipEventRecord(h->trial_ev[2 * ev_end + 1], h->stream));
                h->trial_open = -1;
            }
        }
        h->cur ^= 1;
        h->step += (uint64_t)k;
        left -= (uint64_t)k;
    }
    return LGCA_OK;
}

extern "C" lgca_status lgca_apply_body_force_count(lgca_lattice* h, int dir, int64_t forcing,
                                                   int64_t* reverted)
{
    if (!h || (dir != LGCA_FORCE_X && dir != LGCA_FORCE_Y)) return LGCA_ERR_INVALID_ARGUMENT;
    if (forcing < INT32_MIN || forcing > INT32_MAX)
        return fail(h, LGCA_ERR_INVALID_ARGUMENT,
                    "lgca_apply_body_force_count: forcing must fit an int, as apply_body_force(int) takes it");
    // The reference compares its unsigned count of reverted particles with the
    // int amount (src/omp_lattice.cpp:263, :345), so the amount counts modulo
    // 2^32: a negative one runs all 2N candidates.
    const long long target = forcing < 0 ? (long long)forcing + 4294967296LL : (long long)forcing;
    std::lock_guard<std::mutex> lk(h->mu);
    DeviceGuard dg(h->device);
    lgca_status st0 = sync_force_calls(h);
    if (st0) return st0;
    const uint64_t n = (uint64_t)h->dim_x * h->dim_y;
    const uint64_t vw = (n + 31) / 32;
    if (!h->visited) HIPCHK(h, dalloc(&h->visited, vw * 4));
    HIPCHK(h, launch_fill_u32(h->visited, 0u, vw, h->stream));
    lgca_status st = clear_counters(h);
    if (st) return st;
    const uint32_t call = (uint32_t)h->force_calls++;
    const uint32_t k0 = (uint32_t)h->seed, k1 = (uint32_t)(h->seed >> 32);
    long long* res = (long long*)h->counters;
    switch (h->kind) {
    case kHPP: hipLaunchKernelGGL(body_force_count_kernel<kHPP>, dim3(1), dim3(64), 0, h->stream, h->state[h->cur], h->mask, h->g, k0, k1, call, dir, target, h->visited, res); break;
    case kFHP1: hipLaunchKernelGGL(body_force_count_kernel<kFHP1>, dim3(1), dim3(64), 0, h->stream, h->state[h->cur], h->mask, h->g, k0, k1, call, dir, target, h->visited, res); break;
    default: hipLaunchKernelGGL(body_force_count_kernel<kFHP2>, dim3(1), dim3(64), 0, h->stream, h->state[h->cur], h->mask, h->g, k0, k1, call, dir, target, h->visited, res); break;
    }
    HIPCHK(h, hipGetLastError());
    unsigned long long v = 0;
    if ((st = read_counters(h, &v, 1))) return st;
    if (reverted) *reverted = (int64_t)v;
    return LGCA_OK;
}

// ----------------------------------------------------------- post-process

extern "C" lgca_status lgca_snapshot(lgca_lattice* h)
{
    if (!h) return LGCA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(h->mu);
    DeviceGuard dg(h->device);
    HIPCHK(h, hipMemcpyAsync(h->snap, h->state[h->cur], h->state_words * 4, hipMemcpyDeviceToDevice,
                             h->stream));
    return LGCA_OK;
}

static lgca_status run_coarse(lgca_lattice* h)
{
    const size_t nc = (size_t)(h->dim_x / (2 * h->r)) * (h->dim_y / (2 * h->r));
    if (!h->mean_density) {
        HIPCHK(h, dalloc(&h->mean_density, nc * 4));
        HIPCHK(h, dalloc(&h->mean_momentum, nc * 8));
        HIPCHK(h, dalloc(&h->velocity, nc * 8));
    }
    HIPCHK(h, launch_coarse(h->kind, h->snap, h->g, h->r, h->mean_density, h->mean_momentum,
                            h->velocity, h->stream));
    return LGCA_OK;
}

extern "C" lgca_status lgca_post_process(lgca_lattice* h, float* cell_density, float* cell_momentum,
                                         float* mean_density, float* mean_momentum)
{
    if (!h) return LGCA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(h->mu);
    DeviceGuard dg(h->device);
    const size_t n = (size_t)h->dim_x * h->dim_y;
    const size_t nc = (size_t)(h->dim_x / (2 * h->r)) * (h->dim_y / (2 * h->r));
    if (cell_density || cell_momentum) {
        if (cell_density && !h->cell_density) HIPCHK(h, dalloc(&h->cell_density, n * 4));
        if (cell_momentum && !h->cell_momentum) HIPCHK(h, dalloc(&h->cell_momentum, n * 8));
        HIPCHK(h, launch_cell_fields(h->kind, h->snap, h->g, cell_density ? h->cell_density : nullptr,
                                     cell_momentum ? h->cell_momentum : nullptr, h->stream));
        if (cell_density)
            HIPCHK(h, hipMemcpyAsync(cell_density, h->cell_density, n * 4, hipMemcpyDeviceToHost, h->stream));
        if (cell_momentum)
            HIPCHK(h, hipMemcpyAsync(cell_momentum, h->cell_momentum, n * 8, hipMemcpyDeviceToHost, h->stream));
    }
    if (mean_density || mean_momentum) {
        lgca_status st = run_coarse(h);
        if (st) return st;
        if (mean_density)
            HIPCHK(h, hipMemcpyAsync(mean_density, h->mean_density, nc * 4, hipMemcpyDeviceToHost, h->stream));
        if (mean_momentum)
            HIPCHK(h, hipMemcpyAsync(mean_momentum, h->mean_momentum, nc * 8, hipMemcpyDeviceToHost, h->stream));
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return LGCA_OK;
}

extern "C" lgca_status lgca_post_process_device(lgca_lattice* h, const float** mean_density,
                                                const float** mean_momentum, const float** velocity)
{
    if (!h) return LGCA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(h->mu);
    DeviceGuard dg(h->device);
    lgca_status st = run_coarse(h);
    if (st) return st;
    if (mean_density) *mean_density = h->mean_density;
    if (mean_momentum) *mean_momentum = h->mean_momentum;
    if (velocity) *velocity = h->velocity;
    return LGCA_OK;
}

extern "C" lgca_status lgca_count_particles(lgca_lattice* h, uint64_t* count)
{
    if (!h || !count) return LGCA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(h->mu);
    DeviceGuard dg(h->device);
    lgca_status st = clear_counters(h);
    if (st) return st;
    HIPCHK(h, launch_count(h->state[h->cur], h->state_words, h->counters, h->stream));
    unsigned long long v = 0;
    if ((st = read_counters(h, &v, 1))) return st;
    *count = v;
    return LGCA_OK;
}

extern "C" lgca_status lgca_mean_velocity(lgca_lattice* h, float out[2])
{
    if (!h || !out) return LGCA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(h->mu);
    DeviceGuard dg(h->device);
    lgca_status st = clear_counters(h);
    if (st) return st;
    HIPCHK(h, launch_mean_velocity(h->kind, h->snap, h->mask, h->g, (long long*)h->counters, h->stream));
    unsigned long long v[3] = {0, 0, 0};
    if ((st = read_counters(h, v, 3))) return st;
    const double sx = (double)(long long)v[0], sy = (double)(long long)v[1];
    const double nf = (double)v[2];
    const double unit_y = h->kind == kHPP ? 1.0 : (double)0.866025388240814208984375f;
    out[0] = nf > 0 ? (float)(sx / (840.0 * nf)) : 0.0f;
    out[1] = nf > 0 ? (float)(sy * unit_y / (840.0 * nf)) : 0.0f;
    return LGCA_OK;
}

// ------------------------------------------------------------ checkpoints
// File layout (little-endian), all fixed-size fields first:
//   "LGCACKP1"  8-byte magic
//   u32 version (2; files of version 1 are still read), i32 model,
//   u32 dim_x, dim_y, coarse_radius, flags (bit 0: controller state below),
//   u64 seed, step, force_calls,
//   i32 force_dir, u32 0, u64 force_thr, u32 fx0, fx1, fy0, fy1,
//   u64 checksum: FNV-1a-64 of the header with this field zero, then of the
//       payload (version 1: of the payload only),
//   version 2 only: i32 controller direction, f32 controller u,
//   i64 equilibrium forcing, current forcing, amount of the last tick, its
//   reverted particles, f32 mean velocity the last tick saw (x, y);
//   payload: num_cells cell-type bytes, num_cells state bytes (bit d =
//   direction d, the Bitset layout of src/lgca_bitset.h:229-231), and in
//   version 2 num_cells bytes of the output snapshot.
namespace {

struct CheckpointHeader {
    char magic[8];
    uint32_t version;
    int32_t model;
    uint32_t dim_x, dim_y, coarse_radius, flags;
    uint64_t seed, step, force_calls;
    int32_t force_dir;
    uint32_t reserved1;
    uint64_t force_thr;
    uint32_t fx0, fx1, fy0, fy1;
    uint64_t checksum;
    // version 2
    int32_t ctrl_dir;
    float ctrl_u;
    int64_t ctrl_eq, ctrl_forcing, ctrl_amount, ctrl_reverted;
    float ctrl_v[2];
};
constexpr size_t kHeaderV1 = 96;
static_assert(offsetof(CheckpointHeader, ctrl_dir) == kHeaderV1, "checkpoint header layout");
static_assert(sizeof(CheckpointHeader) == 144, "checkpoint header layout");
constexpr char kCheckpointMagic[8] = {'L', 'G', 'C', 'A', 'C', 'K', 'P', '1'};
constexpr uint32_t kCheckpointFlagController = 1;

uint64_t fnv1a64(const uint8_t* p, size_t n, uint64_t h = 0xcbf29ce484222325ull)
{
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 0x100000001b3ull;
    }
    return h;
}

// Version 2 checksum: the header with its checksum field zero, then the payload.
uint64_t checkpoint_checksum(const CheckpointHeader& hd, const std::string& payload)
{
    CheckpointHeader z = hd;
    z.checksum = 0;
    const uint64_t h = fnv1a64((const uint8_t*)&z, sizeof z);
    return fnv1a64((const uint8_t*)payload.data(), payload.size(), h);
}

}  // namespace

extern "C" lgca_status lgca_save_checkpoint(lgca_lattice* h, const char* path)
{
    if (!h || !path) return LGCA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(h->mu);
    DeviceGuard dg(h->device);
    lgca_status st0 = sync_force_calls(h);
    if (st0) return st0;
    const size_t n = (size_t)h->dim_x * h->dim_y;
    std::string buf(3 * n, '\0');
    uint8_t* types = (uint8_t*)&buf[0];
    uint8_t* state = types + n;
    uint8_t* snap = state + n;
    lgca_status st = ensure_bytes(h);
    if (st) return st;
    HIPCHK(h, launch_types_unpack(h->mask, h->bytes, h->g, h->stream));
    HIPCHK(h, hipMemcpyAsync(types, h->bytes, n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, launch_unpack(h->snap, h->bytes, h->g, h->stream));
    HIPCHK(h, hipMemcpyAsync(snap, h->bytes, n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if ((st = download_state(h, state, hipMemcpyDeviceToHost))) return st;
    CheckpointHeader hd{};
    std::memcpy(hd.magic, kCheckpointMagic, 8);
    hd.version = 2;
    hd.model = h->model;
    hd.dim_x = h->dim_x;
    hd.dim_y = h->dim_y;
    hd.coarse_radius = h->r;
    hd.seed = h->seed;
    hd.step = h->step;
    hd.force_calls = h->force_calls;
    hd.force_dir = h->force_dir;
    hd.force_thr = h->force_thr;
    hd.fx0 = h->fx0;
    hd.fx1 = h->fx1;
    hd.fy0 = h->fy0;
    hd.fy1 = h->fy1;
    if (h->ctrl) {
        Controller c;
        HIPCHK(h, hipMemcpyAsync(&c, h->ctrl, sizeof c, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        hd.flags |= kCheckpointFlagController;
        hd.ctrl_dir = h->ctrl_dir;
        hd.ctrl_u = h->ctrl_u;
        hd.ctrl_eq = h->ctrl_eq;
        hd.ctrl_forcing = c.forcing;
        hd.ctrl_amount = c.amount;
        hd.ctrl_reverted = c.reverted;
        hd.ctrl_v[0] = c.v[0];
        hd.ctrl_v[1] = c.v[1];
    }
    hd.checksum = checkpoint_checksum(hd, buf);
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(h, LGCA_ERR_INVALID_ARGUMENT, std::string("cannot open ") + path);
    const bool ok = std::fwrite(&hd, sizeof hd, 1, f) == 1 &&
                    std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    const bool closed = std::fclose(f) == 0;
    if (!ok || !closed) return fail(h, LGCA_ERR_INVALID_ARGUMENT, std::string("cannot write ") + path);
    return LGCA_OK;
}

extern "C" lgca_status lgca_load_checkpoint(lgca_lattice** out, const char* path, int device)
{
    if (!out || !path) return fail(nullptr, LGCA_ERR_INVALID_ARGUMENT, "lgca_load_checkpoint: NULL argument");
    *out = nullptr;
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(nullptr, LGCA_ERR_INVALID_ARGUMENT, std::string("cannot open ") + path);
    CheckpointHeader hd{};
    std::string buf;
    bool ok = std::fread(&hd, kHeaderV1, 1, f) == 1 && std::memcmp(hd.magic, kCheckpointMagic, 8) == 0 &&
              (hd.version == 1 || hd.version == 2);
    if (ok && hd.version == 2) ok = std::fread((char*)&hd + kHeaderV1, sizeof hd - kHeaderV1, 1, f) == 1;
    // sizes and ranges before anything is allocated
    ok = ok && hd.dim_x > 0 && hd.dim_y > 0 && (uint64_t)hd.dim_x * hd.dim_y < (1ull << 40) &&
         hd.force_dir >= LGCA_FORCE_NONE && hd.force_dir <= LGCA_FORCE_Y && hd.force_thr <= (1ull << 32);
    if (ok && hd.version == 1) hd.flags = 0;
    const bool has_ctrl = ok && (hd.flags & kCheckpointFlagController);
    if (has_ctrl)
        ok = (hd.ctrl_dir == LGCA_FORCE_X || hd.ctrl_dir == LGCA_FORCE_Y) && hd.ctrl_eq >= 0 &&
             hd.ctrl_forcing >= 0 && hd.ctrl_amount >= -1;
    if (ok) {
        const size_t n = (size_t)hd.dim_x * hd.dim_y;
        buf.resize((hd.version == 2 ? 3 : 2) * n);
        ok = std::fread(&buf[0], 1, buf.size(), f) == buf.size() && std::fgetc(f) == EOF &&
             (hd.version == 2 ? checkpoint_checksum(hd, buf)
                              : fnv1a64((const uint8_t*)buf.data(), buf.size())) == hd.checksum;
    }
    std::fclose(f);
    if (!ok) return fail(nullptr, LGCA_ERR_INVALID_STATE, std::string("not a valid checkpoint: ") + path);
    lgca_lattice* h = nullptr;
    lgca_status st = lgca_create(&h, hd.model, hd.dim_x, hd.dim_y, hd.coarse_radius, device);
    if (st) return st;
    const size_t n = (size_t)hd.dim_x * hd.dim_y;
    Controller c{};  // outlives the stream synchronisation below
    {
        std::lock_guard<std::mutex> lk(h->mu);
        DeviceGuard dg(h->device);
        st = upload_types(h, (const uint8_t*)buf.data(), hipMemcpyHostToDevice);
        // the snapshot first (version 1 files hold none: the saved state stands in)
        if (!st && hd.version == 2) {
            st = upload_state(h, (const uint8_t*)buf.data() + 2 * n, hipMemcpyHostToDevice);
            if (!st) {
                const hipError_t e =
                    hipMemcpyAsync(h->snap, h->state[h->cur], h->state_words * 4, hipMemcpyDeviceToDevice, h->stream);
                if (e != hipSuccess) st = hip_fail(h, e, "hipMemcpyAsync(snapshot)");
            }
        }
        if (!st) st = upload_state(h, (const uint8_t*)buf.data() + n, hipMemcpyHostToDevice);
        if (!st && hd.version == 1) {
            const hipError_t e =
                hipMemcpyAsync(h->snap, h->state[h->cur], h->state_words * 4, hipMemcpyDeviceToDevice, h->stream);
            if (e != hipSuccess) st = hip_fail(h, e, "hipMemcpyAsync(snapshot)");
        }
        if (!st) {
            h->seed = hd.seed;
            h->step = hd.step;
            h->force_calls = hd.force_calls;
            // the normalisation of lgca_set_forcing: a window clipped to the
            // lattice, and no forcing where it is empty or p is 0
            h->force_dir = hd.force_dir;
            h->force_thr = hd.force_thr;
            h->fx0 = hd.fx0;
            h->fx1 = std::min(hd.fx1, hd.dim_x);
            h->fy0 = hd.fy0;
            h->fy1 = std::min(hd.fy1, hd.dim_y);
            if (h->fx0 >= h->fx1 || h->fy0 >= h->fy1 || h->force_thr == 0) h->force_dir = LGCA_FORCE_NONE;
            h->dec_count = 0;
        }
        if (!st && has_ctrl) {
            c.forcing = hd.ctrl_forcing;
            c.amount = hd.ctrl_amount;
            c.reverted = hd.ctrl_reverted;
            c.calls = (uint32_t)hd.force_calls;
            c.v[0] = hd.ctrl_v[0];
            c.v[1] = hd.ctrl_v[1];
            hipError_t e = dalloc(&h->ctrl, sizeof(Controller));
            if (e == hipSuccess) e = hipMemcpyAsync(h->ctrl, &c, sizeof c, hipMemcpyHostToDevice, h->stream);
            if (e != hipSuccess) st = hip_fail(h, e, "controller state");
            h->ctrl_dir = hd.ctrl_dir;
            h->ctrl_u = hd.ctrl_u;
            h->ctrl_eq = hd.ctrl_eq;
            h->calls_on_device = false;
        }
        if (!st) {
            const hipError_t e = hipStreamSynchronize(h->stream);
            if (e != hipSuccess) st = hip_fail(h, e, "hipStreamSynchronize");
        }
    }
    if (st) {
        g_create_error = h->err;
        lgca_destroy(h);
        return st;
    }
    *out = h;
    return LGCA_OK;
}

// ------------------------------------------------------- forcing controller

extern "C" lgca_status lgca_controller_set(lgca_lattice* h, int dir, float u, int64_t forcing,
                                           int64_t equilibrium_forcing)
{
    if (!h) return LGCA_ERR_INVALID_ARGUMENT;
    if (dir != LGCA_FORCE_X && dir != LGCA_FORCE_Y)
        return fail(h, LGCA_ERR_INVALID_ARGUMENT, "lgca_controller_set: dir must be X or Y");
    // the viewers' amounts come from get_initial_forcing / get_equilibrium_forcing
    // (size_t, src/lattice.cpp:445-461) and pass through an int parameter
    if (forcing < 0 || forcing > INT32_MAX || equilibrium_forcing < 0 || equilibrium_forcing > INT32_MAX)
        return fail(h, LGCA_ERR_INVALID_ARGUMENT, "lgca_controller_set: amounts must lie in [0, 2^31)");
    std::lock_guard<std::mutex> lk(h->mu);
    DeviceGuard dg(h->device);
    lgca_status st = sync_force_calls(h);
    if (st) return st;
    if (!h->ctrl) HIPCHK(h, dalloc(&h->ctrl, sizeof(Controller)));
    HIPCHK(h, launch_fill_u32((uint32_t*)h->ctrl, 0u, sizeof(Controller) / 4, h->stream));
    // forcing (64-bit) as two words, written by fill kernels: no host buffer
    // has to outlive the call
    uint32_t* f = (uint32_t*)&h->ctrl->forcing;
    HIPCHK(h, launch_fill_u32(f, (uint32_t)(uint64_t)forcing, 1, h->stream));
    HIPCHK(h, launch_fill_u32(f + 1, (uint32_t)((uint64_t)forcing >> 32), 1, h->stream));
    h->ctrl_dir = dir;
    h->ctrl_u = u;
    h->ctrl_eq = equilibrium_forcing;
    return LGCA_OK;
}

extern "C" lgca_status lgca_controller_tick(lgca_lattice* h, uint64_t steps)
{
    if (!h) return LGCA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(h->mu);
    if (!h->ctrl) return fail(h, LGCA_ERR_INVALID_STATE, "lgca_controller_tick: no lgca_controller_set yet");
    DeviceGuard dg(h->device);
    if (!h->calls_on_device) {
        HIPCHK(h, launch_fill_u32(&h->ctrl->calls, (uint32_t)h->force_calls, 1, h->stream));
        h->calls_on_device = true;
    }
    // decision on the snapshot's mean velocity
    HIPCHK(h, launch_fill_u32((uint32_t*)h->ctrl->acc, 0u, 6, h->stream));
    HIPCHK(h, launch_mean_velocity(h->kind, h->snap, h->mask, h->g, h->ctrl->acc, h->stream));
    const double unit_y = h->kind == kHPP ? 1.0 : (double)0.866025388240814208984375f;
    hipLaunchKernelGGL(controller_decide_kernel, dim3(1), dim3(1), 0, h->stream, h->ctrl, h->ctrl_u,
                       (long long)h->ctrl_eq, unit_y);
    HIPCHK(h, hipGetLastError());
    // count-mode force with the decided amount
    const uint64_t n = (uint64_t)h->dim_x * h->dim_y;
    const uint64_t vw = (n + 31) / 32;
    if (!h->visited) HIPCHK(h, dalloc(&h->visited, vw * 4));
    HIPCHK(h, launch_fill_u32(h->visited, 0u, vw, h->stream));
    const uint32_t k0 = (uint32_t)h->seed, k1 = (uint32_t)(h->seed >> 32);
    switch (h->kind) {
    case kHPP: hipLaunchKernelGGL(controlled_force_kernel<kHPP>, dim3(1), dim3(64), 0, h->stream, h->state[h->cur], h->mask, h->g, k0, k1, h->ctrl_dir, h->ctrl, h->visited); break;
    case kFHP1: hipLaunchKernelGGL(controlled_force_kernel<kFHP1>, dim3(1), dim3(64), 0, h->stream, h->state[h->cur], h->mask, h->g, k0, k1, h->ctrl_dir, h->ctrl, h->visited); break;
    default: hipLaunchKernelGGL(controlled_force_kernel<kFHP2>, dim3(1), dim3(64), 0, h->stream, h->state[h->cur], h->mask, h->g, k0, k1, h->ctrl_dir, h->ctrl, h->visited); break;
    }
    HIPCHK(h, hipGetLastError());
    // the steps, then the snapshot the next tick decides on
    lgca_status st = step_locked(h, steps);
    if (st) return st;
    HIPCHK(h, hipMemcpyAsync(h->snap, h->state[h->cur], h->state_words * 4, hipMemcpyDeviceToDevice,
                             h->stream));
    return LGCA_OK;
}

extern "C" lgca_status lgca_controller_get(lgca_lattice* h, int64_t* forcing, int64_t* applied,
                                           float velocity[2])
{
    if (!h) return LGCA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(h->mu);
    if (!h->ctrl) return fail(h, LGCA_ERR_INVALID_STATE, "lgca_controller_get: no lgca_controller_set yet");
    DeviceGuard dg(h->device);
    Controller c;
    HIPCHK(h, hipMemcpyAsync(&c, h->ctrl, sizeof c, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (forcing) *forcing = c.forcing;
    if (applied) *applied = c.amount;
    if (velocity) {
        velocity[0] = c.v[0];
        velocity[1] = c.v[1];
    }
    return LGCA_OK;
}
