// pcs_internal.h -- device-side layouts and kernel launchers shared by the
// C-ABI implementation (pcs_abi.hip) and the kernels (pcs_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pcs_abi.h"

namespace pcs {

constexpr int kWave = 64;

// One model, resident in HBM (one entry per system of a context or server).
struct DevSys {
    uint32_t n_species, kmin, n_k, n_de;
    float de_bin_width, four_dt;
    float half_l[3];
    float half_a[2];
    uint32_t z_boundary, detect_model;
    float detect_sigma;
    uint32_t dist_kind;
    uint32_t maxrow_obs;      // last row holding observed tracks (0 if none)
    const uint32_t* obs_sat;  // observed summed-area table, n_k * n_de uint32
    uint64_t pad1;
};
static_assert(sizeof(DevSys) == 80, "DevSys layout");

// One simulation, resident in HBM. Its record slice rec[0, rec_cap) holds
// packed (row << 16 | bin) track records; srt is a scratch slice of the same
// size for the distance kernel's counting sort. Its Philox key is the launch
// seed XOR (key0, key1): batches use key 0 and one seed per run, served
// requests carry their own key.
struct DevProp {
    uint32_t sim_id, first_video, n_videos, sys_id;
    uint32_t* rec;
    uint32_t* srt;
    uint32_t rec_cap;
    uint32_t key0, key1;
    uint32_t pad;
    float half_az[PCS_MAX_SPECIES];
    float cum_prob[PCS_MAX_SPECIES];
    float sigma[PCS_MAX_SPECIES];
};
static_assert(sizeof(DevProp) == 144, "DevProp layout");

// Record slice size for one simulation (host and device): every recorded
// track spans at least kmin detected frames, so tracks <= detected
// particle-frames / kmin. The expected detected fraction is the detection
// volume over the box volume; `factor` scales that bound and a small floor
// absorbs fluctuations. A slice that still overflows is rerun exactly.
// Sys is pcs_system or DevSys (same field names).
template <typename Sys>
__host__ __device__ inline uint32_t slice_capacity(const Sys& s, const float* half_az, uint32_t n_videos,
                                                   const uint32_t* frames, const uint32_t* particles, double factor)
{
    const double rx = (double)s.half_a[0] / s.half_l[0], ry = (double)s.half_a[1] / s.half_l[1];
    const double fxy = (rx < 1.0 ? rx : 1.0) * (ry < 1.0 ? ry : 1.0);
    double fz = 0.0;
    if (s.detect_model == PCS_DETECT_GAUSS_Z) {
        const double g = (double)s.detect_sigma * 2.5066282746310002 / (2.0 * s.half_l[2]);  // sigma sqrt(2 pi) / Lz
        fz = g < 1.0 ? g : 1.0;
    } else {
        for (uint32_t c = 0; c < s.n_species; ++c) {
            const double r = (double)half_az[c] / s.half_l[2];
            const double rz = r < 1.0 ? r : 1.0;
            fz = rz > fz ? rz : fz;
        }
    }
    double pf = 0.0, maxbound = 0.0;
    for (uint32_t v = 0; v < n_videos; ++v) {
        pf += (double)particles[v] * (double)frames[v];
        maxbound += (double)particles[v] * (double)((frames[v] + 1u) / (s.kmin + 1u));
    }
    const double e = pf * fxy * fz / (double)s.kmin;
    double cap = factor * e + 4.0 * sqrt(e) + 64.0;
    cap = cap < maxbound + 1.0 ? cap : maxbound + 1.0;
    cap = cap < 4.0e9 ? cap : 4.0e9;
    return (uint32_t)ceil(cap);
}

// Arguments of the walk kernel. Videos are the unit of the work plan: video v
// belongs to proposal vprop[v], has vframes[v] frames and vparticles[v]
// particles, and owns tiles [vtile[v], vtile[v+1]) of 64 * PPL particles.
struct WalkArgs {
    const DevSys* sys;
    const DevProp* props;
    const uint32_t* vframes;
    const uint32_t* vparticles;
    const uint32_t* vprop;
    const uint32_t* vtile;   // n_videos + 1 entries
    uint32_t n_videos;
    uint32_t tile_lo, tile_hi;  // tiles processed by this launch
    uint32_t stage_cap;         // LDS records per wave
    uint32_t* rec_count;        // per proposal, atomically bumped
    const uint64_t* seed;       // device-resident Philox key
    uint32_t* tile_counter;     // persistent form: next tile (relative to tile_lo)
    const uint64_t* tile_total; // persistent form, nullable: tiles planned on the device (caps tile_hi)
};

struct DistArgs {
    const DevSys* sys;
    const DevProp* props;
    uint32_t prop_lo;           // proposal of block 0
    uint32_t n_props;
    const uint32_t* rec_count;
    double* dist;               // log10 distance per proposal
    uint64_t* exact;            // integer sum per proposal
    uint32_t* tracks;           // binned tracks per proposal
    int32_t* status;            // PCS_OK or PCS_EOVERFLOW per proposal
    uint32_t colh_off;          // LDS word offset of the column histogram (set by launch_distance)
    uint32_t srtl_off;          // LDS word offset of the sorted bins (set by launch_distance)
    uint32_t srtl_cap;          // sorted 16-bit bins that fit in LDS (set by launch_distance)
};

// ------------------------------------------------ device-resident request loop
//
// Requests are numbered r = 0, 1, 2, ... in the order the host publishes them
// and live in ring slot r % R. The host writes a LoopReq (+ the slot's video
// table) into pinned, coherent host memory and publishes it by storing
// seq = r + 1 last. Group 0 of the loop kernel is the admitter: in request
// order it copies each published request into its slot once the slot is free.
// A slot's state moves through
//   free (0, or 4(r-R)+3)  ->  4r+1 (being copied in)  ->  4r+2 (tiles open)
//   ->  4r+3 (result written; slot free for r + R).
// The other groups are workers: a cursor names the oldest request that may
// still have unclaimed tiles, and a worker claims up to 4 tiles of it with a
// fetch-add on the slot's claim word (low 32 bits of r << 32 | next tile).
// The group that finishes a request's last tile computes its distance and
// writes a LoopDone into host memory, seq = r + 1 last.
constexpr uint32_t kLoopMaxTiles = (1u << 30);
constexpr int kLoopWaves = 4;  // waves per group; a group claims up to 4 tiles at a time

struct LoopReq {
    uint64_t seq;
    uint32_t n_videos;
    uint32_t tiles;  // >= 1 (a request without particles has one empty tile)
    DevProp prop;    // first_video = 0; rec/srt point at the admitted arena range
};
static_assert(sizeof(LoopReq) == 160, "LoopReq layout");

// PCS_TRACE = 1 (the tracing build, `make trace`): every request's completion
// carries the device wall clock (100 MHz) at its admission, its first tile
// claim, the start of its distance, the start of the distance's row sweep
// (after the record passes) and the end, and the host prints one line
// per completed request to stderr (see latency_anatomy.py).
#ifndef PCS_TRACE
#define PCS_TRACE 0
#endif

struct LoopDone {
    uint64_t seq;
    double dist;
    uint64_t exact;
    uint32_t tracks;
    int32_t status;
    uint32_t group_walked;  // the request's tiles were claimed a tile per group
    uint32_t pad;
#if PCS_TRACE
    uint64_t t_admit, t_claim, t_dist, t_sweep, t_done;
#endif
};
static_assert(sizeof(LoopDone) == (PCS_TRACE ? 80 : 40), "LoopDone layout");

struct LoopSlot {
    uint64_t state;
    uint32_t tiles, coop;  // one 8-byte word: a worker reads both at once; coop: a tile per group
    uint32_t tiles_done, rec_count, n_videos, pad;
    uint64_t claim;  // (r mod 2^32) << 32 | next unclaimed tile
    DevProp prop;
#if PCS_TRACE
    uint64_t t_admit, t_claim;
#endif
};
static_assert(sizeof(LoopSlot) == (PCS_TRACE ? 200 : 184), "LoopSlot layout");

// CU index of a group: XCC id << 8 | SE/SH/CU id of HW_ID, below 2048.
constexpr uint32_t kLoopCuSlots = 2048;

struct LoopCtl {
    uint64_t cursor;      // oldest request that may have unclaimed tiles
    uint64_t last_claim;  // wall clock of the latest tile claim, any group
    uint64_t admit_next;  // next request the admitter copies in (kept across launches)
    uint64_t finished;    // requests whose result is written (kept across launches)
    uint32_t quit;        // set when the loop should end: every group leaves
    uint32_t error;       // non-zero: a claim could not be resolved (the loop ended)
};

// Per-slot video table, (maxv + 1) entries of 3 words at slot * 3 * (maxv + 1)
// in host memory (ring_vid) and HBM (slot_vid): entry v = {frames, particles,
// first tile} of video v, entry n_videos = {0, 0, total tiles}.

struct LoopArgs {
    const DevSys* sys;
    LoopSlot* slots;
    uint32_t* slot_vid;
    const LoopReq* ring;
    const uint32_t* ring_vid;
    LoopDone* done;
    const uint32_t* stop;   // host memory: non-zero = leave now
    LoopCtl* ctl;
    uint32_t* cu_token;     // kLoopCuSlots poller tokens, one per CU (zero between launches)
    uint32_t R, maxv;
    uint32_t stage_cap;     // LDS records per wave
    uint32_t nk_lds, nde_lds;
    uint32_t srtl_cap;      // sorted 16-bit bins of a distance that fit in LDS (after the column histogram)
    uint32_t coop_max_tiles;  // requests of at most this many tiles are claimed a tile at a time and each
                              // tile walked by a whole group (0: never) when admitted while at most
    uint32_t coop_max_open;   // this many earlier requests are unfinished
    uint64_t idle_ticks;    // wall-clock ticks without work before the loop ends
};

// Tables of a batch built on the device from the caller's proposal records
// (prepare_kernel): per proposal its DevProp with a record slice placed in
// [rec, rec + arena_cap) (slices past the arena's end are cut short and
// overflow), per video its proposal and tile prefix. Proposals hold
// consecutive video ranges: props[i].first_video = sum of earlier n_videos.
struct PrepArgs {
    const DevSys* sys;
    const pcs_proposal* in;   // n records
    const uint32_t* vframes;
    const uint32_t* vparticles;
    uint32_t n, n_videos;
    uint32_t tile_particles;  // 64 * particles per lane
    uint32_t pad;
    double cap_factor;
    DevProp* props;
    uint32_t* vprop;
    uint32_t* vtile;          // n_videos + 1
    uint32_t* rec;
    uint32_t* srt;
    uint64_t arena_cap;       // records in rec (and in srt)
    uint32_t* rec_count;      // zeroed
    uint32_t* tile_counter;   // zeroed
    uint64_t* totals;         // [0] planned records, [1] tiles
};

// Proposal step of an ABC-PMC round (propose_kernel), one wave per trial:
// parent by cumulative weight, displaced parameters reflected into the prior,
// components sorted by m, the trial's pcs_proposal record and Poisson particle
// counts of its videos. Parameters are rows of P = 2C + Z: m_1..m_C,
// log10 c_1..log10 c_C, az_1..az_Z.
struct ProposeArgs {
    const double* theta;       // P x n_pop
    const double* cum_w;       // n_pop
    const uint32_t* frames_in; // n_trials x J, trial-major
    double* theta_out;         // P x n_trials
    pcs_proposal* props;       // n_trials records
    uint32_t* vframes;         // n_trials x J
    uint32_t* vparticles;      // n_trials x J
    uint32_t n_pop, n_trials, J, C, Z, first_sim_id;
    uint64_t seed;
    double lo[3], hi[3];       // prior bounds of m, log10 c, az
    double deltat, volume;
    double tau[3 * PCS_MAX_SPECIES];
};

// Launchers (pcs_kernels.hip). waves_per_group in {1,2,4}; ppl in {1,2}.
hipError_t launch_reset(uint32_t* rec_count, uint32_t n, uint32_t* tile_counter, hipStream_t s);
hipError_t launch_walk(const WalkArgs& a, int ppl, int waves_per_group, bool persistent,
                       uint32_t grid_groups, bool gauss, bool zreflect, hipStream_t s);
hipError_t launch_distance(const DistArgs& a, uint32_t n_k_max, uint32_t n_de_max, hipStream_t s);
// Track-list form (k, DE) of the walk, one wave per tile, tiles of 64 particles.
hipError_t launch_tracks(const WalkArgs& a, bool gauss, bool zref, hipStream_t s);
// pmc kernel sums: `part` holds pmc_chunks(n_old) x n_new partial sums.
uint32_t pmc_chunks(uint32_t n_old);
hipError_t launch_pmc_sums(const double* xn, uint32_t n_new, const double* xo, const double* w_old, uint32_t n_old,
                           uint32_t P, double* part, double* wsum, hipStream_t s);
hipError_t launch_selftest_math(const uint32_t* w, float* radius, float* s, float* c, uint64_t n, hipStream_t st);
hipError_t launch_selftest_copy(const void* src, void* dst, uint64_t bytes, hipStream_t st);
hipError_t launch_selftest_sqrt(uint32_t lo, uint32_t n, unsigned long long* bad, hipStream_t st);
// Request loop: `grid` groups of kLoopWaves waves; n_de_max selects the
// distance segment width. Returns once launched.
hipError_t launch_serve_loop(const LoopArgs& a, uint32_t grid, uint32_t n_de_max, hipStream_t s);
size_t serve_loop_lds(const LoopArgs& a);
hipError_t serve_loop_occupancy(const LoopArgs& a, uint32_t n_de_max, int* groups_per_cu);
// Sets a.srtl_cap to the sorted bins that fit in the LDS each of the loop's
// groups leaves on a CU without costing a group per CU, then reports the
// groups per CU at that size.
hipError_t serve_loop_fit(LoopArgs& a, uint32_t n_de_max, int* groups_per_cu);
// Whether a group's LDS past the angle table holds the cooperative walk's
// exchange buffers (the distance's working set shares that room).
bool serve_loop_coop_fits(const LoopArgs& a);
hipError_t launch_prepare(const PrepArgs& a, hipStream_t s);
hipError_t launch_propose(const ProposeArgs& a, hipStream_t s);
// Sequential prefix sum of w (one lane, member order) into cum.
hipError_t launch_cumsum(const double* w, uint32_t n, double* cum, hipStream_t s);
// dst[k][j] = src[rows[k]][j] / tau[k] for k < n_rows (rows of n doubles).
struct ScaleRows {
    uint32_t rows[3 * PCS_MAX_SPECIES];
    double tau[3 * PCS_MAX_SPECIES];
    uint32_t n_rows;
};
hipError_t launch_scale_rows(const double* src, uint32_t n, const ScaleRows& sr, double* dst, hipStream_t s);
hipError_t launch_dense_hist(const DevProp* props, uint32_t prop_lo, uint32_t n_props,
                             const uint32_t* rec_count, uint32_t max_count,
                             uint32_t n_k, uint32_t n_de, uint32_t* hist, hipStream_t s);

}  // namespace pcs
