/*
 * lgca.h -- C ABI of the MI355X lattice-gas engine (liblgca_hip.so).
 *
 * Plain C types only: an opaque handle, plain pointers and sizes.  Every entry
 * point returns an lgca_status (0 = success); the text of the last failure on
 * a handle is available from lgca_last_error().  Host pointers are ordinary
 * host memory; entry points that name a "device" pointer take memory on the
 * handle's GPU.
 *
 * One handle is one lattice of one model resident on one GPU: the state
 * (Q bit-planes, ping-pong), the cell-type masks, the output snapshot and the
 * coarse fields.  Work is queued on the handle's HIP stream; entry points that
 * return data to the host wait for it.  A handle may be driven by two host
 * threads at once only as the reference's viewers do (one thread stepping and
 * snapshotting, another post-processing the snapshot); otherwise a handle is
 * not re-entrant.
 *
 * Reference interface replaced: template<Model> class Lattice, src/lattice.h:32-238,
 * whose only live implementation is OMP_Lattice, src/omp_lattice.h:27-78.
 */
#ifndef LGCA_H_
#define LGCA_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LGCA_API __attribute__((visibility("default")))

typedef int lgca_status;
enum {
    LGCA_OK = 0,
    LGCA_ERR_INVALID_ARGUMENT = 1, /* bad size, model, range or pointer     */
    LGCA_ERR_HIP = 2,              /* a HIP runtime call failed             */
    LGCA_ERR_OUT_OF_MEMORY = 3,    /* device allocation failed              */
    LGCA_ERR_INVALID_STATE = 4,    /* state bytes carry bits >= Q, or cell
                                      types outside 0..2                    */
    LGCA_ERR_BUDGET = 5            /* request exceeds an HBM budget         */
};

/* Models, src/lgca_common.h:49-54. */
enum { LGCA_HPP = 0, LGCA_FHP_I = 1, LGCA_FHP_II = 2, LGCA_FHP_III = 3 };
/* Cell types, src/lgca_common.h:56-60 (one byte per cell at this ABI). */
enum { LGCA_FLUID = 0, LGCA_SOLID_NO_SLIP = 1, LGCA_SOLID_SLIP = 2 };
/* Body-force direction, m_bf_dir of src/lattice.h:68 ('x' or 'y'). */
enum { LGCA_FORCE_NONE = 0, LGCA_FORCE_X = 1, LGCA_FORCE_Y = 2 };

typedef struct lgca_lattice lgca_lattice;

typedef struct lgca_info {
    int32_t model;
    uint32_t num_dir;        /* Q: 4, 6, 7, 7                                   */
    uint32_t dim_x, dim_y;
    uint32_t coarse_radius;  /* r                                              */
    uint32_t coarse_dim_x, coarse_dim_y;
    uint32_t words_per_row;  /* 32-cell words per row and plane                */
    uint32_t row_pitch;      /* words between consecutive planes (>= words)   */
    uint32_t tile_rows;      /* rows per tile of the tiled kernel's steps_per_pass
                                launches: lgca_set_tuning's value, else the height
                                the timed trial of the first launches chose (0
                                until it has)                                   */
    uint64_t num_cells;
    uint64_t num_coarse_cells;
    uint64_t state_bytes;    /* one state buffer: dim_y * Q * row_pitch * 4   */
    uint64_t device_bytes;   /* everything the handle holds on the GPU        */
    uint64_t step;           /* absolute step counter (hash input)            */
    uint64_t seed;
    int32_t steps_per_pass;  /* steps per pass of lgca_step: lgca_set_tuning's
                                value, else the one the trial chose (2, 4 or 8;
                                -1 until it has)                                */
    uint32_t pass_pipeline;  /* 1 when lgca_set_pipeline has turned on launches
                                of several passes                               */
} lgca_info;

/* Creates a lattice of dim_x x dim_y cells on GPU `device`, all FLUID and
 * empty.  Replaces the constructors Lattice(...) / OMP_Lattice(...)
 * (src/lattice.cpp:29-157, src/omp_lattice.cpp:74-88), which derive the size
 * from (test case, Re, Ma, r); the size is explicit here and the derivation
 * lives in the C++ wrapper.  FHP models need an even dim_y, as the reference
 * asserts (src/lattice.cpp:141).  The coarse grid has floor(dim / 2r) cells per
 * axis, which is what the reference computes when its divisibility asserts
 * (src/lattice.cpp:152-153) are compiled out; both sizes must be >= 2r. */
LGCA_API lgca_status lgca_create(lgca_lattice** out, int model, uint32_t dim_x, uint32_t dim_y,
                        uint32_t coarse_radius, int device);
/* Frees the handle; replaces ~OMP_Lattice / free_memory (src/omp_lattice.cpp:91-97,473-489). */
LGCA_API lgca_status lgca_destroy(lgca_lattice* h);
/* Text of the last failure on h (or of the last failed lgca_create when h is NULL). */
LGCA_API const char* lgca_last_error(const lgca_lattice* h);
LGCA_API lgca_status lgca_get_info(const lgca_lattice* h, lgca_info* info);

/* Queue work on an external HIP stream (a hipStream_t; NULL restores the
 * handle's own stream). */
LGCA_API lgca_status lgca_set_stream(lgca_lattice* h, void* hip_stream);
LGCA_API lgca_status lgca_synchronize(lgca_lattice* h);

/* Seed and absolute step of the counter-based hash that stands in for the
 * reference's rand() (src/lgca_bitset.h:284-288, src/omp_lattice.cpp:84,269). */
LGCA_API lgca_status lgca_set_seed(lgca_lattice* h, uint64_t seed);
LGCA_API lgca_status lgca_set_step(lgca_lattice* h, uint64_t step);

/* Cell types, one byte per cell (0 FLUID, 1 NO_SLIP, 2 SLIP), row-major.
 * Replaces writes to m_cell_type_cpu by apply_cell_type_all / apply_bc_* /
 * apply_boundary_cell_type_* (src/lattice.cpp:221-395). */
LGCA_API lgca_status lgca_set_cell_types(lgca_lattice* h, const uint8_t* host_types);
LGCA_API lgca_status lgca_get_cell_types(lgca_lattice* h, uint8_t* host_types);

/* State, one byte per cell, bit d = particle in direction d (the Bitset
 * layout, src/lgca_bitset.h:229-231).  Bytes with bits >= Q are rejected.
 * Replaces copy_data_to_device / copy_data_from_device (src/lattice.h:206-209). */
LGCA_API lgca_status lgca_set_state_u8(lgca_lattice* h, const uint8_t* host_state);
LGCA_API lgca_status lgca_get_state_u8(lgca_lattice* h, uint8_t* host_state);
/* Device-pointer forms of the same (num_cells bytes on the handle's GPU). */
LGCA_API lgca_status lgca_set_state_u8_device(lgca_lattice* h, const uint8_t* dev_state);
LGCA_API lgca_status lgca_get_state_u8_device(lgca_lattice* h, uint8_t* dev_state);

/* Fills every FLUID cell, every direction, with probability p_fill from the
 * hash (stream 2); other cells keep their state.  Replaces init_random
 * (src/lattice.cpp:198-217), whose threshold is p = 1/Q. */
LGCA_API lgca_status lgca_init_random(lgca_lattice* h, double p_fill);
/* The same restricted to cells whose byte in host_region (num_cells bytes) is
 * non-zero; replaces init_diffusion (src/lattice.cpp:465-496), whose disc the
 * caller rasterises. */
LGCA_API lgca_status lgca_init_random_masked(lgca_lattice* h, double p_fill,
                                             const uint8_t* host_region);

/* Per-cell body force fused into the step: after collision, each FLUID cell in
 * [x0, x1) x [y0, y1) is forced with probability p (hash stream 1) using the
 * flip rule of apply_body_force (src/omp_lattice.cpp:293-337).  dir
 * LGCA_FORCE_NONE disables it. */
LGCA_API lgca_status lgca_set_forcing(lgca_lattice* h, int dir, double p, uint32_t x0, uint32_t x1,
                             uint32_t y0, uint32_t y1);

/* Advances nsteps collide-and-propagate steps (src/omp_lattice.cpp:100-249),
 * each fused with the per-cell force when set.  Queued; returns at once. */
LGCA_API lgca_status lgca_step(lgca_lattice* h, uint64_t nsteps);

/* lgca_step without allocating: the same steps, bits and launch shapes, for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  No counterpart in
 * the reference.  lgca_step allocates the handle's force-decision buffer on
 * demand; this entry point draws the decisions into a fixed array of 2^21
 * words (8 MiB) that the library holds on each GPU, and runs no launch-shape
 * trial (it takes the shape a trial or lgca_tune kept, else the rule's).  It
 * fails with LGCA_ERR_INVALID_ARGUMENT when one pass needs more decision words
 * than that: (force rows) x (32-cell words of the force columns) x (steps per
 * pass).  The array is shared by every call on a GPU and keeps the last call's
 * decisions: calls whose queued work overlaps on one GPU (several handles on
 * different streams) race; calls on one handle's stream do not. */
LGCA_API lgca_status lgca_step_noalloc(lgca_lattice* h, uint64_t nsteps);

/* Launch shape of lgca_step: steps fused per pass over HBM (1, 2, 4 or 8; 0
 * selects the one-step kernel that keeps a 128-cell quad per thread; -1, the
 * default, lets the trial choose: 4, 8 on a lattice that fills the GPU where
 * 8-step passes are faster, or 2 where 2-step passes are faster on a lattice
 * too small to fill it) and rows per tile (0 = automatic: a timed trial
 * of a few shapes on the first launches, kept thereafter and restarted by this
 * call, lgca_set_forcing, new cell types or a new stream).  Results do not
 * depend on it.  No counterpart in the reference; used by the autotuner and
 * benchmarks. */
LGCA_API lgca_status lgca_set_tuning(lgca_lattice* h, int steps_per_pass, uint32_t tile_rows);
/* Several passes of 4 or 8 steps per launch (on = 1; 0, the default, launches
 * pass by pass).  Each XCD of the GPU takes a band of row tiles and runs its
 * tiles pass after pass, each tile starting as soon as its neighbours have
 * finished the previous pass, so one pass's last waves overlap the next
 * pass's first instead of leaving the GPU to drain between launches.  Used
 * once the launch shape is fixed (lgca_set_tuning) or kept (the trial, or
 * lgca_tune), on lattices of full 32-cell words with at least two row tiles
 * and 2K rows per XCD, as far as one window of force decisions reaches (256
 * steps, fewer when the decisions of that many steps exceed 64 MiB); other
 * steps run pass by pass.  Results do not depend on it.  lgca_synchronize
 * and the state downloads report LGCA_ERR_HIP if a pass ever gave up waiting
 * for its neighbours (a bounded wait; it cannot hang the GPU).  No
 * counterpart in the reference.  In this build the mode is switched off:
 * on = 1 fails with LGCA_ERR_INVALID_STATE (its waits did not complete on
 * the GPU, for a reason not yet found), and on = 0 is accepted. */
LGCA_API lgca_status lgca_set_pipeline(lgca_lattice* h, int on);
/* Runs that trial now and waits for it: every candidate shape is timed on
 * passes that read the live state and write the other state buffer, and the
 * fastest per step is kept (lgca_get_info reports it), with the fastest
 * 4-step shape for the shorter passes that finish calls of an 8-step choice.
 * Candidates include tile heights whose passes fill the GPU's wave slots
 * once or twice.  Each kernel the kept shapes use then runs once untimed, at
 * every offset of the first step within its four-step group.  The state and
 * the step do not change.  Does nothing when lgca_set_tuning fixed the rows
 * per tile or selected the one-step kernel.  No counterpart in the
 * reference; benchmarks call it before their warm-up, so that timed steps
 * run the kept shapes from the first. */
LGCA_API lgca_status lgca_tune(lgca_lattice* h);

/* Count-mode body force: replaces apply_body_force(int forcing)
 * (src/omp_lattice.cpp:254-346).  Candidate cells come from the hash (stream
 * 3, keyed by a per-handle call counter) instead of rand() % N; the loop stops
 * once `forcing` particles are reverted or 2N candidates are spent, and always
 * makes at least one try.  `forcing` must fit an int (else
 * LGCA_ERR_INVALID_ARGUMENT); as in the reference, which compares an unsigned
 * count with it, a negative amount counts modulo 2^32 and so runs all 2N
 * candidates.  *reverted receives the count (may be NULL). */
LGCA_API lgca_status lgca_apply_body_force_count(lgca_lattice* h, int dir, int64_t forcing,
                                        int64_t* reverted);

/* Copies the live state into the output snapshot that post-processing reads;
 * replaces copy_data_to_output_buffer (src/lattice.cpp:437-441). */
LGCA_API lgca_status lgca_snapshot(lgca_lattice* h);

/* Post-processes the snapshot (src/omp_lattice.cpp:349-454): per-cell density
 * (N floats) and momentum (2N floats, x/y interleaved), coarse density (Nc)
 * and coarse momentum (2Nc).  Host pointers; any may be NULL to skip it. */
LGCA_API lgca_status lgca_post_process(lgca_lattice* h, float* cell_density, float* cell_momentum,
                              float* mean_density, float* mean_momentum);
/* Coarse fields left on the GPU: device pointers valid until the next call
 * or lgca_destroy, filled once the handle's stream reaches this point (see
 * lgca_synchronize).  velocity = mean momentum / mean density (0 where empty). */
LGCA_API lgca_status lgca_post_process_device(lgca_lattice* h, const float** mean_density,
                                     const float** mean_momentum, const float** velocity);

/* Particles in the live state; replaces get_n_particles (src/lattice.cpp:180-195). */
LGCA_API lgca_status lgca_count_particles(lgca_lattice* h, uint64_t* count);
/* Mean over FLUID cells of momentum / density of the snapshot; replaces
 * get_mean_velocity (src/omp_lattice.cpp:508-557).  Summed exactly. */
LGCA_API lgca_status lgca_mean_velocity(lgca_lattice* h, float out[2]);

/* The pipe and karman viewers' forcing controller (apps/pipe/pipe_viewer.cpp:107-125,
 * apps/karman/karman_viewer.cpp:108-124) on the GPU.  lgca_controller_set takes the
 * force direction, the target velocity u (Lattice::u(), src/lattice.cpp:64-67) and the
 * initial and equilibrium forcing amounts (get_initial_forcing /
 * get_equilibrium_forcing, src/lattice.cpp:445-461).  lgca_controller_tick is one
 * viewer tick: the mean x velocity of the output snapshot, as lgca_mean_velocity
 * computes it; if it is below u, the forcing becomes the equilibrium amount once the
 * velocity exceeds 0.9 u and is applied as lgca_apply_body_force_count does (same
 * call counter); then `steps` steps and a new snapshot.  Nothing is read back, so
 * ticks queue on the stream while the host post-processes earlier snapshots.
 * Both amounts must lie in [0, 2^31). */
LGCA_API lgca_status lgca_controller_set(lgca_lattice* h, int dir, float u, int64_t forcing,
                                         int64_t equilibrium_forcing);
LGCA_API lgca_status lgca_controller_tick(lgca_lattice* h, uint64_t steps);
/* Waits for the stream, then reports the controller's current forcing, the amount the
 * last tick applied (-1: none) and the mean velocity that tick saw. */
LGCA_API lgca_status lgca_controller_get(lgca_lattice* h, int64_t* forcing, int64_t* applied,
                                         float velocity[2]);

/* Checkpoint of the live state: model, sizes, seed, step, per-cell force
 * settings, the count-mode force call counter, cell types, state bytes, the
 * output snapshot and the forcing controller's state, in one file whose
 * checksum covers header and payload.  A run resumed from it continues bit for
 * bit as the saved run would have, post-processing and controller ticks
 * included.  No counterpart in the reference, whose state lives only in the
 * process (src/lgca_bitset.h:31-245). */
LGCA_API lgca_status lgca_save_checkpoint(lgca_lattice* h, const char* path);
/* Creates a handle on GPU `device` from a checkpoint file.  Damaged files
 * (LGCA_ERR_INVALID_STATE) are refused; force settings are normalised as
 * lgca_set_forcing does.  Files of the earlier layout, without a snapshot or
 * controller state, are read too: their snapshot is the saved state. */
LGCA_API lgca_status lgca_load_checkpoint(lgca_lattice** out, const char* path, int device);

/* ------------------------------------------------------------------------
 * Batched ensembles and serving.  No counterpart in the reference, which runs
 * one lattice per process: a job here is one run of that lattice, set up as
 * apply_bc_periodic / apply_bc_pipe / apply_bc_karman_vortex_street
 * (src/lattice.cpp:221-280) or from given cell types, filled as init_random
 * (src/lattice.cpp:198-217), advanced `steps` collide-and-propagate steps
 * (src/omp_lattice.cpp:100-249) with the optional per-cell force, and
 * post-processed (src/omp_lattice.cpp:349-454).  Every job's bits are the ones
 * the single-lattice entry points above give for the same parameters.
 *
 * A batch owns one GPU stream and a worker thread.  Jobs are submitted at any
 * time; the worker admits them in order while their device memory fits the
 * batch's HBM budget (and their cells the active-cell cap), and advances
 * all admitted jobs together, a few steps per round, in one kernel launch per
 * (model, steps per pass, ragged width) class.  A finished job's outputs stay
 * on the GPU until released.  All batch entry points are thread-safe.
 * ---------------------------------------------------------------------- */
typedef struct lgca_batch lgca_batch;

enum { LGCA_GEOM_PERIODIC = 0, LGCA_GEOM_PIPE = 1, LGCA_GEOM_CYLINDER = 2, LGCA_GEOM_TYPES = 3 };
enum { LGCA_JOB_QUEUED = 0, LGCA_JOB_RUNNING = 1, LGCA_JOB_DONE = 2, LGCA_JOB_RELEASED = 3,
       LGCA_JOB_FAILED = 4 };
#define LGCA_ALL_JOBS UINT64_MAX

typedef struct lgca_job_desc {
    int32_t model;             /* LGCA_HPP .. LGCA_FHP_III                             */
    uint32_t dim_x, dim_y;     /* as lgca_create                                       */
    uint32_t coarse_radius;    /* r of the coarse fields (window 2r)                   */
    uint64_t seed;             /* hash key of this job                                 */
    uint64_t steps;            /* steps to run                                         */
    double fill;               /* per-bit fill of FLUID cells (ignored if `state`)     */
    int32_t geometry;          /* LGCA_GEOM_*: all FLUID; NO_SLIP rows 0 and dim_y-1;
                                  those plus a NO_SLIP disc; or `cell_types`          */
    int32_t cyl_x, cyl_y;      /* disc centre                                           */
    double cyl_diameter;       /* disc: (float)sqrt(dx^2 + dy^2) < diameter / 2,
                                  as src/lattice.cpp:264-279                          */
    int32_t force_dir;         /* per-cell force as lgca_set_forcing                   */
    double force_p;
    uint32_t fx0, fx1, fy0, fy1;
    const uint8_t* cell_types; /* LGCA_GEOM_TYPES: num_cells bytes, copied at submit   */
    const uint8_t* state;      /* initial state bytes (copied at submit) or NULL       */
} lgca_job_desc;

typedef struct lgca_job_info {
    int32_t status;            /* LGCA_JOB_*                                           */
    uint32_t reserved;
    uint64_t step;             /* steps done (exact once DONE)                         */
    uint64_t device_bytes;     /* what the job holds against the budget               */
    double submit_s, admit_s, done_s;  /* host clock (seconds, monotonic); 0 = not yet */
} lgca_job_info;

typedef struct lgca_batch_stats {
    uint64_t jobs_submitted, jobs_done, jobs_active, jobs_queued;
    uint64_t rounds;           /* rounds queued                                        */
    uint64_t launches;         /* step-kernel launches queued                          */
    uint64_t site_updates;     /* cells x steps queued                                 */
    uint64_t bytes_in_use;     /* device bytes of admitted, unreleased jobs            */
    uint64_t arena_bytes;      /* device memory reserved by the batch                  */
    int32_t steps_per_pass;    /* current tuning                                       */
    uint32_t wave_target;      /* waves per round the tile height aims at              */
    uint64_t tunings;          /* autotuner decisions taken                            */
} lgca_batch_stats;

/* Creates a batch on GPU `device` whose jobs may hold at most
 * hbm_budget_bytes of device memory together. */
LGCA_API lgca_status lgca_batch_create(lgca_batch** out, int device, uint64_t hbm_budget_bytes);
/* Stops the worker (unfinished jobs are abandoned) and frees everything. */
LGCA_API lgca_status lgca_batch_destroy(lgca_batch* b);
LGCA_API const char* lgca_batch_last_error(const lgca_batch* b);
/* Device bytes the job would hold (what admission checks against the budget). */
LGCA_API lgca_status lgca_batch_job_bytes(const lgca_job_desc* desc, uint64_t* bytes);
/* Validates and queues a job; *job_id identifies it from then on.  Fails with
 * LGCA_ERR_BUDGET if the job alone exceeds the budget. */
LGCA_API lgca_status lgca_batch_submit(lgca_batch* b, const lgca_job_desc* desc, uint64_t* job_id);
/* Waits until the job (or every submitted job, LGCA_ALL_JOBS) is DONE, at
 * most timeout_s seconds (< 0: no limit).  LGCA_ERR_INVALID_STATE on time-out. */
LGCA_API lgca_status lgca_batch_wait(lgca_batch* b, uint64_t job_id, double timeout_s);
/* Ids of jobs that became DONE since the last poll, in completion order. */
LGCA_API lgca_status lgca_batch_poll(lgca_batch* b, uint64_t* ids, uint32_t max_ids, uint32_t* n);
LGCA_API lgca_status lgca_batch_job_info(lgca_batch* b, uint64_t job_id, lgca_job_info* info);
/* Outputs of a DONE job into host memory (any may be NULL): the state bytes
 * (num_cells), coarse density (Nc) and coarse momentum (2Nc, x/y
 * interleaved) of its final state. */
LGCA_API lgca_status lgca_batch_result(lgca_batch* b, uint64_t job_id, uint8_t* state,
                                       float* mean_density, float* mean_momentum);
/* Returns a DONE job's device memory to the budget. */
LGCA_API lgca_status lgca_batch_release(lgca_batch* b, uint64_t job_id);
/* Launch shape: steps per pass (2, 4 or 8; 0 = autotune), waves per round the
 * tile height aims at (0 = autotune), and the cap on cells of running jobs
 * (0 = none).  Results do not depend on it. */
LGCA_API lgca_status lgca_batch_set_tuning(lgca_batch* b, int steps_per_pass, uint32_t wave_target,
                                           uint64_t max_active_cells);
LGCA_API lgca_status lgca_batch_get_stats(lgca_batch* b, lgca_batch_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* LGCA_H_ */
