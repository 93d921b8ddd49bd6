/*
 * pcs_abi.h -- C ABI of the MI355X simulation hot path for ABC-PMC particle
 * concentration estimation (walk -> detect -> bin -> cumulative-histogram
 * distance).
 *
 * Plain C types only: caller-owned host buffers, library-owned device
 * buffers, int return codes (PCS_OK == 0, negative on error), no exceptions
 * across the boundary. One pcs_ctx / pcs_batch / pcs_server must not be used
 * by two threads at once.
 *
 * The reference has no plugin or FFI layer; its hot path sits behind two
 * internal seams that these entry points replace (citations relative to the
 * reference repository):
 *   - per-job launch pair K_simulate_system_unr + K_distance issued by
 *     dev_worker_ (src/gpu/detail/simulation_impl.inl:1064-1097), with
 *     SimulateSetup / SimulateRun / HistogramRecorder arguments
 *     (src/gpu/cusim/kernel-simulate.cuh:13-63, src/gpu/cusim/histogram2d.cuh:42-60);
 *   - the reference histogram + SAT upload
 *     (src/gpu/detail/simulation_impl.inl:605-629, src/gpu/cusim/histogram2d.inl:6-18);
 *   - the multi-GPU queue pool and completion callback
 *     (src/gpu/detail/simulation_impl.inl:631-730, 889-901, 1121-1136),
 *     replaced by resident batches and a serving loop on one GPU.
 *
 * Histogram layout (differs from the reference by one row): row r, column b
 * of an n_k x n_de row-major uint32 matrix counts tracks of length k = r + 1
 * whose diffusion estimate falls in bin b = trunc(DE / de_bin_width). The
 * reference keeps an always-empty row 0 (kmax = 1 + max frames,
 * src/gpu/detail/simulation_impl.inl:513); the distance is unchanged by it.
 */
#ifndef PCS_ABI_H
#define PCS_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PCS_ABI_VERSION 1
#define PCS_MAX_SPECIES 8
#define PCS_MAX_NK 4096
#define PCS_MAX_NDE 4096

/* Return codes. */
#define PCS_OK 0
#define PCS_EINVAL (-1)    /* invalid argument */
#define PCS_ENOMEM (-2)    /* device or host allocation failed */
#define PCS_EHIP (-3)      /* HIP runtime error (see pcs_last_error) */
#define PCS_EAGAIN (-4)    /* serving: no room now, try again */
#define PCS_EOVERFLOW (-5) /* a record arena overflowed and could not be rerun */
#define PCS_ENODEV (-6)    /* no usable GPU */
#define PCS_ESTATE (-7)    /* call not valid in the object's current state */

/* Boundary condition in z (x and y are always periodic). */
#define PCS_Z_PERIODIC 0  /* reference: src/gpu/cusim/kernel-simulate-unroll.inl:176 */
#define PCS_Z_REFLECT 1   /* model extension */

/* Detection model. */
#define PCS_DETECT_SLAB 0     /* |z| <= az/2 (src/gpu/cusim/utility.inl:102-153) */
#define PCS_DETECT_GAUSS_Z 1  /* probability exp(-z^2 / (2 sigma^2)); extension */

/* Distance kind; both return log10 of an exact integer sum over the 2D
 * cumulative histograms S (simulated) and R (observed). */
#define PCS_DIST_L2SQ 0  /* log10 sum (S-R)^2: src/cpu/distance.jl:1-8 */
#define PCS_DIST_L1 1    /* log10 sum |S-R|: extension */

/* Execution forms of pcs_simulate / pcs_batch_run. */
#define PCS_MODE_PLAIN 0    /* one launch pair per simulation */
#define PCS_MODE_BATCHED 1  /* one persistent launch per batch in a hipGraph */
#define PCS_MODE_BATCHED_DIRECT 2 /* the batched kernels launched without the graph,
                                     with events between them (kernel timing) */
#define PCS_MODE_PLAIN_STREAMS 3  /* the plain form's launch pairs spread round-robin
                                     over PCS_PLAIN_STREAMS streams of one priority:
                                     the reference's job pool keeps one pair in flight
                                     per queue (src/gpu/main.cpp:45 "1/30",
                                     src/gpu/detail/simulation_impl.inl:686-710, 889-901) */
#define PCS_PLAIN_STREAMS 4       /* hardware queues a process gets by default */

/* Model shared by all simulations of one ABC run
 * (SimulateSetup, src/gpu/cusim/kernel-simulate.cuh:13-38, filled at
 * src/gpu/detail/simulation_impl.inl:568-579). */
typedef struct pcs_system {
    uint32_t n_species;    /* C in [1, PCS_MAX_SPECIES] */
    uint32_t kmin;         /* shortest recorded track, >= 2 */
    uint32_t n_k;          /* histogram rows, [1, PCS_MAX_NK] */
    uint32_t n_de;         /* histogram columns, [1, PCS_MAX_NDE] */
    float de_bin_width;    /* ub_de / n_de */
    float four_dt;         /* 4 * deltat */
    float half_l[3];       /* half box extents Lx/2, Ly/2, Lz/2 (um) */
    float half_a[2];       /* half detection extents ax/2, ay/2 (um) */
    uint32_t z_boundary;   /* PCS_Z_* */
    uint32_t detect_model; /* PCS_DETECT_* */
    float detect_sigma;    /* sigma of PCS_DETECT_GAUSS_Z (um) */
    uint32_t dist_kind;    /* PCS_DIST_* */
} pcs_system;

/* One simulation (SimulateRun, src/gpu/cusim/kernel-simulate.cuh:46-63,
 * precomputed at src/gpu/detail/simulation_impl.inl:306-341). Videos are
 * given as a range into the video arrays passed alongside. */
typedef struct pcs_proposal {
    uint32_t sim_id;       /* Philox counter word; unique per simulation and seed */
    uint32_t first_video;  /* index of the first video in the video arrays */
    uint32_t n_videos;     /* number of videos */
    uint32_t reserved;     /* must be 0 */
    float half_az[PCS_MAX_SPECIES];  /* az_i / 2 per species (fixed depth: all equal) */
    float cum_prob[PCS_MAX_SPECIES]; /* cumulative abundance; entry C-1 >= 1 */
    float sigma[PCS_MAX_SPECIES];    /* random-walk step sqrt(2 m_i dt) per species */
} pcs_proposal;

typedef struct pcs_ctx pcs_ctx;
typedef struct pcs_batch pcs_batch;
typedef struct pcs_server pcs_server;

/* ----------------------------------------------------------------- basics */

int pcs_abi_version(void);
const char* pcs_strerror(int code);
/* Last error message recorded on this thread (HIP error text, argument). */
const char* pcs_last_error(void);
/* Number of visible GPUs (0 when none). */
int pcs_device_count(void);

/* Observed histogram from track lists, as built by prepare_
 * (src/gpu/detail/simulation_impl.inl:605-616) with row K-1: DE >= ub_de is
 * skipped, bin = trunc(DE / (ub_de / n_de)). hist (n_k*n_de) is overwritten.
 * Tracks outside the matrix are counted in *dropped (nullable). Host only. */
int pcs_reference_hist(const uint64_t* K, const double* DE, uint64_t n, double ub_de,
                       uint32_t n_k, uint32_t n_de, uint32_t* hist, uint64_t* dropped);

/* ---------------------------------------------------------------- context */

/* Open a context on GPU `device` for one model and its observed histogram
 * (plain counts, n_k*n_de; the cumulative table is built internally, replacing
 * the host SAT + upload of src/gpu/detail/simulation_impl.inl:617-629). */
int pcs_open(const pcs_system* sys, const uint32_t* obs_hist, int device, pcs_ctx** out);
/* Releases the caller's handle. Batches created from the context keep it
 * alive until they are destroyed; no other call may use ctx afterwards. */
int pcs_close(pcs_ctx* ctx);

/* One-shot: simulate n proposals and write log10 distances (n doubles).
 * hist_out (nullable) receives n*n_k*n_de histograms; tracks_out (nullable)
 * the number of binned tracks per proposal. Blocks until done. */
int pcs_simulate(pcs_ctx* ctx, const pcs_proposal* props, uint32_t n,
                 const uint32_t* video_frames, const uint32_t* video_particles, uint32_t n_videos,
                 uint64_t seed, int mode, double* dist_out, uint32_t* hist_out,
                 uint32_t* tracks_out);

/* Allocation-free variant of pcs_simulate (batched form, direct launches).
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (tight serving loops, harnesses that forbid hipMalloc). It
 * calls no hipMalloc/hipFree: proposal records, tables, counters, results,
 * the record arena and the sort scratch of a run live in one fixed device
 * array of PCS_STATIC_SCRATCH_BYTES per GPU, reserved when the library is
 * loaded and held until the process ends. Limits: n <= PCS_STATIC_MAX_SIMS,
 * n_videos <= PCS_STATIC_MAX_VIDEOS, proposals with consecutive video ranges
 * (props[i].first_video = sum of the earlier n_videos), and the planned
 * record slices plus exact reruns of overflowed ones within the array
 * (about 16 million records); beyond them it fails with PCS_EINVAL or
 * PCS_EOVERFLOW. Every call on one GPU shares that GPU's array and leaves
 * its contents behind: calls that overlap on one GPU (any context) race.
 * Results equal pcs_simulate's. Blocks until done. tracks_out and exact_out
 * (n uint64 exact distance sums) are nullable. */
#define PCS_STATIC_SCRATCH_BYTES (128u << 20)
#define PCS_STATIC_MAX_SIMS 8192u
#define PCS_STATIC_MAX_VIDEOS 65536u
int pcs_simulate_static(pcs_ctx* ctx, const pcs_proposal* props, uint32_t n,
                        const uint32_t* video_frames, const uint32_t* video_particles, uint32_t n_videos,
                        uint64_t seed, double* dist_out, uint32_t* tracks_out, uint64_t* exact_out);

/* ---------------------------------------------------------------- batches */

/* A batch holds proposals resident in HBM; load uploads and plans it,
 * run enqueues the walk + distance asynchronously on the batch stream,
 * sync waits, read copies results back. After a load (or pcs_batch_propose)
 * that failed, the batch holds nothing runnable: run and read return
 * PCS_ESTATE until a load or propose succeeds. */
int pcs_batch_create(pcs_ctx* ctx, pcs_batch** out);
int pcs_batch_load(pcs_batch* b, const pcs_proposal* props, uint32_t n,
                   const uint32_t* video_frames, const uint32_t* video_particles, uint32_t n_videos);
int pcs_batch_run(pcs_batch* b, uint64_t seed, int mode);
int pcs_batch_sync(pcs_batch* b);
int pcs_batch_read(pcs_batch* b, double* dist_out, uint32_t* hist_out, uint32_t* tracks_out);
/* Exact integer distance sums of the last run (n uint64: sum (S-R)^2 or
 * sum |S-R|), of which the log10 distances are taken. Call after
 * pcs_batch_read. */
int pcs_batch_read_exact(pcs_batch* b, uint64_t* exact_out);
/* Device time of the last run (HIP events on the batch stream), ms:
 * out[0] whole run, out[1] walk kernel(s), out[2] distance kernel(s). */
int pcs_batch_timing(pcs_batch* b, float out[3]);
/* Number of proposals whose record arena overflowed in the last run and
 * were rerun (informational). */
int pcs_batch_overflow_reruns(pcs_batch* b, uint32_t* n);
/* Record arena of the loaded batch: total slice capacity in records and the
 * device bytes held by the record and sort-scratch arenas. */
int pcs_batch_arena(pcs_batch* b, uint64_t* capacity_records, uint64_t* bytes);
int pcs_batch_destroy(pcs_batch* b);

/* Prior and model of the device proposal step. Parameters are rows of
 * P = 2C + Z values per member: m_1..m_C (um^2/s), log10 c_1..log10 c_C
 * (c per mL), az_1..az_Z (um), Z = 1 (fixed depth) or C. */
typedef struct pcs_prior {
    uint32_t n_components;       /* C, the context's n_species */
    uint32_t n_depths;           /* Z */
    double m_lo, m_hi;           /* prior interval of m */
    double log10c_lo, log10c_hi; /* of log10 c */
    double az_lo, az_hi;         /* of the detection depth az */
    double deltat;               /* frame interval (s) */
    double volume;               /* Lx Ly Lz / 1e12: particles per video per (1/mL) */
} pcs_prior;

/* Proposal step of one ABC-PMC round on the GPU, then the batch's plan (tile
 * tables, record slices) on the GPU: the batch is left ready for
 * pcs_batch_run, as after pcs_batch_load. Replaces the reference's host loop
 * over waiting samples (src/gpu/detail/simulation_impl.inl:256-341;
 * src/cpu/estimate.jl:126-150). For trial t < n_trials: a parent i with
 * probability w_i (the first member whose cumulative weight cum_w[i] reaches
 * a uniform u, wrand_idx_), every parameter p displaced by N(0, tau_p^2) and
 * reflected into its closed prior interval (displace_), components sorted by
 * m, c = 10^log10 c, cumulative abundances (last = 2), sigma = sqrt(2 m dt),
 * half depths, and per video a Poisson(sum c * volume) particle count. The
 * random words are Philox4x32-10 keyed by `seed` with counters (word, trial,
 * slot, 0x70726F70), apart from the walk's; the normals are FP32 Box-Muller
 * pairs of the walk's own math, the rest is FP64 in a fixed operation order,
 * so a CPU restatement reproduces every result bit for bit.
 * theta: P x n_pop (row-major), cum_w: n_pop non-decreasing, tau: P; or
 * theta = cum_w = NULL and n_pop = 0: the resident population of
 * pcs_batch_set_population;
 * video_frames: n_trials x videos_per_trial frame counts (trial-major);
 * trial t gets sim_id first_sim_id + t and videos [t J, (t + 1) J).
 * theta_out (nullable) receives the proposed parameters, P x n_trials. */
int pcs_batch_propose(pcs_batch* b, const pcs_prior* prior, const double* theta, const double* cum_w,
                      const double* tau, uint32_t n_pop, uint32_t n_trials, const uint32_t* video_frames,
                      uint32_t videos_per_trial, uint32_t first_sim_id, uint64_t seed, double* theta_out);
/* Population of an ABC-PMC generation kept resident in the batch: theta
 * (n_params x n_pop, the rows of pcs_batch_propose) and normalised weights w
 * (n_pop). Their cumulative sums are formed on the device in member order
 * (the same sums as a sequential cumsum). pcs_batch_propose then takes
 * theta = cum_w = NULL (and n_pop = 0) to draw from this population, and
 * pcs_batch_pmc_weights weighs the proposed trials against it. */
int pcs_batch_set_population(pcs_batch* b, const double* theta, const double* w, uint32_t n_pop, uint32_t n_params);
/* pmc-standard importance weights of the batch's proposed trials (the last
 * pcs_batch_propose, parameters resident on the device) against its resident
 * population (pcs_batch_set_population), as pcs_pmc_weights on the same
 * values: w_out[j] proportional to 1 / sum_i w_i prod_p N(theta_new[p][j] -
 * theta[p][i]; 0, tau[p]) over the rows with tau[p] > 0 (rows with tau 0 have
 * no spread and give every pair the same factor); normalised to sum 1
 * (n_trials doubles). */
int pcs_batch_pmc_weights(pcs_batch* b, const double* tau, double* w_out);
/* The same in two halves, so that the weights overlap the walk: _async
 * enqueues them on the batch's own weights stream, ordered after the last
 * pcs_batch_propose only (they need the proposed parameters, not the
 * distances: called after pcs_batch_run they run beside the walk), and
 * returns; _wait blocks until they are done and writes w_out as
 * pcs_batch_pmc_weights does. A load, propose or set_population in between
 * waits for them and drops them (_wait then returns PCS_ESTATE). */
int pcs_batch_pmc_weights_async(pcs_batch* b, const double* tau);
int pcs_batch_pmc_weights_wait(pcs_batch* b, double* w_out);
/* The loaded or proposed batch's proposal records (n) and video particle
 * counts (n_videos); each nullable. */
int pcs_batch_read_proposals(pcs_batch* b, pcs_proposal* props_out, uint32_t* video_particles_out);

/* Launch shape of the batched walk kernel. Zero fields mean "default".
 * Results do not depend on the shape (the random stream is counter based). */
typedef struct pcs_tuning {
    uint32_t particles_per_lane; /* 1 or 2 */
    uint32_t waves_per_group;    /* 1, 2, 4 */
    uint32_t groups_per_cu;      /* persistent grid = CUs * groups_per_cu */
    uint32_t stage_records;      /* LDS record staging per wave: 128, 256, 512 */
} pcs_tuning;
int pcs_set_tuning(pcs_ctx* ctx, const pcs_tuning* t);
int pcs_get_tuning(pcs_ctx* ctx, pcs_tuning* t);
/* Autotuner: times the batched walk kernel (HIP events)
 * on the given calibration proposals for every candidate
 * launch shape - particles per lane {1,2} x waves per group {2,4} x groups
 * per CU {2,4,6,8,12,16} (those the hardware can co-schedule) - and keeps
 * the fastest in the context. best/ms (nullable) receive the winner and its
 * walk time (the faster of 2 runs; one untimed run warms each kernel first).
 * Results never depend on the shape. */
int pcs_autotune(pcs_ctx* ctx, const pcs_proposal* props, uint32_t n, const uint32_t* video_frames,
                 const uint32_t* video_particles, uint32_t n_videos, pcs_tuning* best, float* ms);

/* Scale of the per-simulation record slice estimate (default 0.5; the bound
 * is tracks <= detected particle-frames / kmin). Slices that overflow are
 * rerun exactly, so this only trades memory for reruns. Takes effect at the
 * next pcs_batch_load. */
int pcs_set_record_capacity_factor(pcs_ctx* ctx, double factor);

/* Evaluates the walk kernel's FP32 math on the GPU for n 32-bit words:
 * radius[i] = sqrt(-2 ln u) for the walk's radius uniform u = (float)(w | 255) 2^-32,
 * (s[i], c[i]) = sin, cos of the walk's angle 2 pi (a + 1/2) 2^-16, a = w >> 16.
 * Test hook for bit-exactness of the device math; no reference counterpart. */
int pcs_selftest_math(int device, const uint32_t* w, float* radius, float* s, float* c, uint64_t n);

/* Counts, in *n_wrong, the floats whose bit patterns lie in [lo_bits, hi_bits)
 * (positive normal floats) on which the walk kernel's square root (the radius
 * of its Box-Muller pairs) differs from the correctly rounded one; each result
 * is checked on the GPU against the midpoints to its neighbours, squared
 * exactly in FP64. Test hook; no reference counterpart. */
int pcs_selftest_sqrt(int device, uint32_t lo_bits, uint32_t hi_bits, uint64_t* n_wrong);

/* Copies `bytes` (a multiple of 16) between two fresh device buffers with a
 * plain streaming kernel and reports its time in *ms (nullable). Calibration
 * hook for hardware byte counters (known traffic: bytes read + bytes
 * written); no reference counterpart. */
int pcs_selftest_copy(int device, uint64_t bytes, float* ms);

/* ------------------------------------------------------- beyond the walk */

/* Track lists of one simulation: K and DE of every track with k >= kmin, in
 * no particular order, with the walk of pcs_simulate (same stream, same FP32
 * arithmetic), without binning - the synthetic-experiment generator of
 * src/cpu/generate_experiment.jl:1-115 on the GPU. Writes up to cap pairs;
 * *n_out receives the number of tracks (may exceed cap). */
int pcs_generate_tracks(pcs_ctx* ctx, const pcs_proposal* prop, const uint32_t* video_frames,
                        const uint32_t* video_particles, uint32_t n_videos, uint64_t seed, uint32_t* K, float* DE,
                        uint64_t cap, uint64_t* n_out);

/* pmc-standard importance weights on the GPU
 * (src/gpu/detail/simulation_impl.inl:832-871, src/cpu/estimate.jl:193-208):
 * w[j] proportional to 1 / sum_i w_old[i] prod_p N(theta_new[p][j] - theta_old[p][i]; 0, tau[p]),
 * normalised to sum 1 (the Gaussian normalisation constants cancel).
 * theta_* are n_params rows (row-major, row p = parameter p) of n_new /
 * n_old members; n_params <= 24. FP64 throughout. */
int pcs_pmc_weights(int device, const double* theta_new, uint32_t n_new, const double* theta_old,
                    const double* w_old, uint32_t n_old, const double* tau, uint32_t n_params, double* w_out);

/* ---------------------------------------------------------------- serving */

/* Serving forms (pcs_server_config.flags):
 *   0                      host-driven: a scheduler thread packs the requests
 *                          admitted in one pass into a launch group (walk +
 *                          distance kernels) on the next of 4 streams.
 *   PCS_SERVE_DEVICE_LOOP  device-resident: one persistent kernel takes
 *                          requests from a ring in pinned host memory, claims
 *                          their tiles in arrival order, computes each distance
 *                          in the group that finishes the last tile and writes
 *                          the result back to host memory; the host thread
 *                          only admits (record arena, ring slot) and collects.
 *                          The kernel leaves after loop_idle_us without work
 *                          and is launched again when requests arrive.
 *                          A request of at most one tile (64 particles) per
 *                          CU, admitted while at most 12 earlier requests are
 *                          unfinished, is claimed a tile at a time and each
 *                          tile walked by a whole group of 4 waves, three
 *                          producing its random normals, one stepping (a lone
 *                          wave's instruction stream is its latency); other
 *                          requests have a wave walk each tile.
 *   PCS_SERVE_LOOP_ONE_WAVE_TILES  with PCS_SERVE_DEVICE_LOOP: every tile is
 *                          walked by one wave, as in a large request.
 * All give the same result for the same request. In either form the host
 * thread keeps polling (yielding its core) for 2 ms after its last work before
 * it sleeps, so that a request arriving meanwhile is published at once. */
#define PCS_SERVE_DEVICE_LOOP 1u
#define PCS_SERVE_LOOP_ONE_WAVE_TILES 2u

typedef struct pcs_server_config {
    uint64_t arena_bytes;    /* HBM budget for track records (admission), 0 = 1 GiB */
    uint32_t max_inflight;   /* simulations on the GPU at once, 0 = 4096 */
    uint32_t max_videos;     /* videos per request, 0 = 64 */
    uint32_t flags;          /* 0, or PCS_SERVE_DEVICE_LOOP (| PCS_SERVE_LOOP_ONE_WAVE_TILES) */
    uint32_t loop_idle_us;   /* device loop: idle time before it leaves, 0 = 20000 */
} pcs_server_config;

typedef struct pcs_request {
    uint64_t tag;            /* caller's id, echoed in the completion */
    uint64_t seed;           /* Philox key of this simulation */
    uint32_t sys_id;         /* from pcs_server_add_system */
    uint32_t n_videos;
    const uint32_t* frames;     /* [n_videos], copied at submit */
    const uint32_t* particles;  /* [n_videos], copied at submit */
    pcs_proposal prop;       /* first_video is ignored */
} pcs_request;

typedef struct pcs_completion {
    uint64_t tag;
    double dist;             /* log10 distance */
    uint32_t tracks;         /* binned tracks */
    int32_t status;          /* PCS_OK or error */
    double submit_us;        /* host clock at submit (us since server start) */
    double done_us;          /* host clock when the result was observed */
    uint64_t exact;          /* exact integer distance sum (saturated at 2^64-1) */
} pcs_completion;

int pcs_server_create(int device, const pcs_server_config* cfg, pcs_server** out);
int pcs_server_add_system(pcs_server* s, const pcs_system* sys, const uint32_t* obs_hist, uint32_t* sys_id);
/* Non-blocking; PCS_EAGAIN when the request queue is full. */
int pcs_server_submit(pcs_server* s, const pcs_request* req);
/* Collect up to max completions; waits up to wait_us microseconds for at
 * least one (0 = do not wait). */
int pcs_server_poll(pcs_server* s, pcs_completion* out, uint32_t max, uint32_t* n_out, uint32_t wait_us);
/* Requests submitted and not yet returned by poll. */
int pcs_server_pending(pcs_server* s, uint32_t* n);
/* Scale of the record slice estimate of requests submitted from now on
 * (default 1.0; see pcs_set_record_capacity_factor). A request whose slice
 * overflows is rerun with a slice of exactly its size before it completes. */
int pcs_server_set_record_capacity_factor(pcs_server* s, double factor);
/* Requests rerun so far after a record slice overflow (informational). */
int pcs_server_overflow_reruns(pcs_server* s, uint64_t* n);
/* Device loop: requests completed so far whose tiles were each walked by a
 * whole group (informational; 0 in the host-driven form). No counterpart in
 * the reference (its runs have no request loop): for a serving caller's
 * monitoring. */
int pcs_server_group_walked(pcs_server* s, uint64_t* n);
/* The server's clock (us since server start; CLOCK_MONOTONIC), the clock of
 * submit_us and done_us: lets a load generator stamp its own scheduled
 * arrival times in the same frame. */
int pcs_server_now_us(pcs_server* s, double* us);
int pcs_server_destroy(pcs_server* s);

#ifdef __cplusplus
}
#endif

#endif
