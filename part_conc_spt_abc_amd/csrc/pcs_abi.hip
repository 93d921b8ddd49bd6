// pcs_abi.hip -- C ABI of the hot path: contexts, resident batches (plain and
// batched forms), overflow reruns. The serving loop lives in pcs_server.hip.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "pcs_common.h"

using namespace pcs;

// ------------------------------------------------------------ error state

namespace pcs {

static thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int hip_error(hipError_t e, const char* what)
{
    return set_error(PCS_EHIP, "%s: %s", what, hipGetErrorString(e));
}

int validate_system(const pcs_system* s)
{
    if (!s) return set_error(PCS_EINVAL, "system is NULL");
    if (s->n_species < 1 || s->n_species > PCS_MAX_SPECIES)
        return set_error(PCS_EINVAL, "n_species %u outside [1, %d]", s->n_species, PCS_MAX_SPECIES);
    if (s->kmin < 2) return set_error(PCS_EINVAL, "kmin %u < 2", s->kmin);
    if (s->n_k < 1 || s->n_k > PCS_MAX_NK) return set_error(PCS_EINVAL, "n_k %u outside [1, %d]", s->n_k, PCS_MAX_NK);
    if (s->n_de < 1 || s->n_de > PCS_MAX_NDE)
        return set_error(PCS_EINVAL, "n_de %u outside [1, %d]", s->n_de, PCS_MAX_NDE);
    if (!(s->de_bin_width > 0.0f) || !std::isfinite(s->de_bin_width))
        return set_error(PCS_EINVAL, "de_bin_width must be positive");
    if (!(s->four_dt > 0.0f) || !std::isfinite(s->four_dt)) return set_error(PCS_EINVAL, "four_dt must be positive");
    for (int i = 0; i < 3; ++i)
        if (!(s->half_l[i] > 0.0f) || !std::isfinite(s->half_l[i]))
            return set_error(PCS_EINVAL, "half_l[%d] must be positive", i);
    for (int i = 0; i < 2; ++i)
        if (!(s->half_a[i] >= 0.0f) || !std::isfinite(s->half_a[i]))
            return set_error(PCS_EINVAL, "half_a[%d] must be >= 0", i);
    if (s->z_boundary > PCS_Z_REFLECT) return set_error(PCS_EINVAL, "z_boundary %u unknown", s->z_boundary);
    if (s->detect_model > PCS_DETECT_GAUSS_Z) return set_error(PCS_EINVAL, "detect_model %u unknown", s->detect_model);
    if (s->detect_model == PCS_DETECT_GAUSS_Z && !(s->detect_sigma > 0.0f))
        return set_error(PCS_EINVAL, "detect_sigma must be positive for gaussian detection");
    if (s->dist_kind > PCS_DIST_L1) return set_error(PCS_EINVAL, "dist_kind %u unknown", s->dist_kind);
    return PCS_OK;
}

int validate_proposal(const pcs_system& s, const pcs_proposal& p)
{
    for (uint32_t c = 0; c < s.n_species; ++c) {
        if (!(p.sigma[c] >= 0.0f) || !std::isfinite(p.sigma[c]))
            return set_error(PCS_EINVAL, "proposal %u: sigma[%u] invalid", p.sim_id, c);
        if (!(p.half_az[c] >= 0.0f) || !std::isfinite(p.half_az[c]))
            return set_error(PCS_EINVAL, "proposal %u: half_az[%u] invalid", p.sim_id, c);
    }
    if (!(p.cum_prob[s.n_species - 1] >= 1.0f))
        return set_error(PCS_EINVAL, "proposal %u: cum_prob[C-1] must be >= 1", p.sim_id);
    return PCS_OK;
}

// Summed-area table of the observed histogram (uint32 like the reference's
// host SAT, src/gpu/detail/simulation_impl.inl:617-628) and its deepest row.
void build_sat(const uint32_t* H, uint32_t n_k, uint32_t n_de, std::vector<uint32_t>& S, uint32_t& maxrow)
{
    S.assign((size_t)n_k * n_de, 0u);
    maxrow = 0;
    for (uint32_t i = 0; i < n_k; ++i) {
        uint32_t run = 0;
        for (uint32_t j = 0; j < n_de; ++j) {
            const uint32_t h = H ? H[(size_t)i * n_de + j] : 0u;
            if (h) maxrow = i;
            run += h;
            S[(size_t)i * n_de + j] = run + (i ? S[(size_t)(i - 1) * n_de + j] : 0u);
        }
    }
}

DevSys make_devsys(const pcs_system& s, uint32_t maxrow, const uint32_t* dsat)
{
    DevSys d{};
    d.n_species = s.n_species;
    d.kmin = s.kmin;
    d.n_k = s.n_k;
    d.n_de = s.n_de;
    d.de_bin_width = s.de_bin_width;
    d.four_dt = s.four_dt;
    for (int i = 0; i < 3; ++i) d.half_l[i] = s.half_l[i];
    d.half_a[0] = s.half_a[0];
    d.half_a[1] = s.half_a[1];
    d.z_boundary = s.z_boundary;
    d.detect_model = s.detect_model;
    d.detect_sigma = s.detect_sigma;
    d.dist_kind = s.dist_kind;
    d.maxrow_obs = maxrow;
    d.obs_sat = dsat;
    return d;
}

uint32_t record_capacity(const pcs_system& s, const pcs_proposal& p, const uint32_t* frames,
                         const uint32_t* particles, double factor)
{
    return slice_capacity(s, p.half_az, p.n_videos, frames, particles, factor);
}

}  // namespace pcs

// --------------------------------------------------------------- context

struct pcs_ctx {
    int device = 0;
    int n_cu = 0;
    pcs_system sys{};
    DevSys dsys{};
    DevSys* d_sys = nullptr;
    uint32_t* d_sat = nullptr;
    hipStream_t stream = nullptr;
    pcs_tuning tune{1, 4, 8, 256};
    double cap_factor = 0.5;
    // The caller's handle plus one per live batch: pcs_close drops the
    // caller's, and the context is freed when its last batch is destroyed.
    std::atomic<int> refs{1};
};

struct pcs_batch {
    pcs_ctx* ctx = nullptr;
    hipStream_t stream = nullptr;
    // host copies
    std::vector<pcs_proposal> props;
    std::vector<DevProp> hprops;
    std::vector<uint32_t> vframes, vparticles, vprop, vtile;
    uint32_t n_props = 0, n_videos = 0, n_tiles = 0;
    int planned_ppl = 0;
    uint64_t arena_records = 0;
    uint32_t* d_rec2 = nullptr;  // rerun arena of the last run (overflowed slices)
    uint32_t* d_srt2 = nullptr;
    // device
    DevProp* d_props = nullptr;
    uint32_t *d_vframes = nullptr, *d_vparticles = nullptr, *d_vprop = nullptr, *d_vtile = nullptr;
    uint32_t *d_rec = nullptr, *d_srt = nullptr, *d_count = nullptr, *d_tracks = nullptr, *d_tile_counter = nullptr;
    double* d_dist = nullptr;
    uint64_t* d_exact = nullptr;
    int32_t* d_status = nullptr;
    uint64_t* d_seed = nullptr;
    uint64_t* h_seed = nullptr;  // pinned
    size_t cap_props = 0, cap_videos = 0, cap_arena = 0;
    // device-planned batch (pcs_batch_propose): the proposal records, the
    // population it was drawn from and the proposed parameters stay on the
    // device; the plan (tiles, slices) is made there, host copies of vtile
    // are fetched only when a plain launch or a rerun needs them
    bool dev_planned = false;
    bool host_vtile = false;     // vtile (host) holds the device plan
    uint32_t J = 0;              // videos per trial
    pcs_proposal* d_in = nullptr;
    double *d_theta_pop = nullptr, *d_cumw = nullptr, *d_theta_new = nullptr;
    uint64_t* d_totals = nullptr;  // [0] planned records, [1] tiles
    size_t cap_in = 0, cap_pop = 0, cap_new = 0;
    uint64_t last_need = 0;      // records the last device plan asked for
    // resident population (pcs_batch_set_population): d_theta_pop, d_cumw
    // and its weights d_wpop; the trials' parameters of the last proposal
    // step stay in d_theta_new (P rows of n_props)
    double* d_wpop = nullptr;
    bool pop_resident = false;
    uint32_t pop_n = 0, pop_P = 0, new_P = 0;
    // The last pcs_batch_load / pcs_batch_propose completed: the tables,
    // plan and buffers describe n_props simulations. Cleared as soon as
    // either starts changing them, so a failed call leaves nothing runnable.
    bool ready = false;
    // pcs_batch_pmc_weights workspace: scaled rows, kernel sums, chunk sums
    double *d_xn = nullptr, *d_xo = nullptr, *d_wsum = nullptr, *d_part = nullptr;
    size_t cap_xn = 0, cap_xo = 0, cap_wsum = 0, cap_part = 0;
    // graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    pcs_tuning graph_tune{};
    bool graph_events = false;
    // plain form over several streams (PCS_MODE_PLAIN_STREAMS), made on first use
    hipStream_t pstream[PCS_PLAIN_STREAMS] = {};
    hipEvent_t pfork = nullptr;
    hipEvent_t pjoin[PCS_PLAIN_STREAMS] = {};
    // pmc weights enqueued by pcs_batch_pmc_weights_async on their own stream
    // (after the proposals on the batch stream), collected by ..._wait
    hipStream_t wstream = nullptr;
    hipEvent_t wfork = nullptr;
    bool w_pending = false, w_uniform = false;
    uint32_t w_n = 0;
    // timing
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    int last_mode = -1;
    bool ran = false;
    uint32_t reruns = 0;
};

namespace {

template <typename T>
int dev_alloc(T** p, size_t n)
{
    if (*p) {
        hipFree(*p);
        *p = nullptr;
    }
    if (n == 0) n = 1;
    const hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return set_error(PCS_ENOMEM, "hipMalloc(%zu bytes): %s", n * sizeof(T), hipGetErrorString(e));
    }
    return PCS_OK;
}

#define PCS_HIP(call)                                         \
    do {                                                      \
        const hipError_t e_ = (call);                         \
        if (e_ != hipSuccess) return hip_error(e_, #call);    \
    } while (0)

#define PCS_TRY(call)                   \
    do {                                \
        const int r_ = (call);          \
        if (r_ != PCS_OK) return r_;    \
    } while (0)

void destroy_graph(pcs_batch* b)
{
    if (b->gexec) hipGraphExecDestroy(b->gexec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
}

int check_tuning(const pcs_tuning& t)
{
    if (t.particles_per_lane != 1 && t.particles_per_lane != 2)
        return set_error(PCS_EINVAL, "particles_per_lane must be 1 or 2");
    if (t.waves_per_group != 1 && t.waves_per_group != 2 && t.waves_per_group != 4)
        return set_error(PCS_EINVAL, "waves_per_group must be 1, 2 or 4");
    if (t.groups_per_cu < 1 || t.groups_per_cu > 32) return set_error(PCS_EINVAL, "groups_per_cu outside [1, 32]");
    if (t.stage_records != 128 && t.stage_records != 256 && t.stage_records != 512)
        return set_error(PCS_EINVAL, "stage_records must be 128, 256 or 512");
    return PCS_OK;
}

// Tiles of 64*ppl particles per video, prefix-summed.
void plan_tiles(pcs_batch* b, int ppl)
{
    const uint32_t tp = 64u * (uint32_t)ppl;
    b->vtile.resize(b->n_videos + 1);
    uint32_t acc = 0;
    for (uint32_t v = 0; v < b->n_videos; ++v) {
        b->vtile[v] = acc;
        acc += (b->vparticles[v] + tp - 1u) / tp;
    }
    b->vtile[b->n_videos] = acc;
    b->n_tiles = acc;
    b->planned_ppl = ppl;
}

int upload_plan(pcs_batch* b)
{
    PCS_HIP(hipMemcpyAsync(b->d_vtile, b->vtile.data(), b->vtile.size() * 4, hipMemcpyHostToDevice, b->stream));
    return PCS_OK;
}

WalkArgs walk_args(pcs_batch* b, uint32_t tile_lo, uint32_t tile_hi)
{
    WalkArgs a{};
    a.sys = b->ctx->d_sys;
    a.props = b->d_props;
    a.vframes = b->d_vframes;
    a.vparticles = b->d_vparticles;
    a.vprop = b->d_vprop;
    a.vtile = b->d_vtile;
    a.n_videos = b->n_videos;
    a.tile_lo = tile_lo;
    a.tile_hi = tile_hi;
    a.stage_cap = b->ctx->tune.stage_records;
    a.rec_count = b->d_count;
    a.seed = b->d_seed;
    a.tile_counter = b->d_tile_counter;
    a.tile_total = b->dev_planned ? b->d_totals + 1 : nullptr;
    return a;
}

// Plan of a device-planned batch, made on the device from its resident
// proposal records: DevProp table with record slices, video map, tile prefix.
int device_plan(pcs_batch* b)
{
    pcs_ctx* c = b->ctx;
    PrepArgs pa{};
    pa.sys = c->d_sys;
    pa.in = b->d_in;
    pa.vframes = b->d_vframes;
    pa.vparticles = b->d_vparticles;
    pa.n = b->n_props;
    pa.n_videos = b->n_videos;
    pa.tile_particles = 64u * c->tune.particles_per_lane;
    pa.cap_factor = c->cap_factor;
    pa.props = b->d_props;
    pa.vprop = b->d_vprop;
    pa.vtile = b->d_vtile;
    pa.rec = b->d_rec;
    pa.srt = b->d_srt;
    pa.arena_cap = b->d_rec ? b->cap_arena : 0u;
    pa.rec_count = b->d_count;
    pa.tile_counter = b->d_tile_counter;
    pa.totals = b->d_totals;
    PCS_HIP(launch_prepare(pa, b->stream));
    b->planned_ppl = (int)c->tune.particles_per_lane;
    b->host_vtile = false;
    return PCS_OK;
}

// Host copy of a device plan's tile prefix (plain launches and reruns
// address tiles from the host).
int fetch_plan(pcs_batch* b)
{
    if (!b->dev_planned || b->host_vtile) return PCS_OK;
    b->vtile.resize((size_t)b->n_videos + 1u);
    PCS_HIP(hipMemcpyAsync(b->vtile.data(), b->d_vtile, ((size_t)b->n_videos + 1u) * 4u, hipMemcpyDeviceToHost,
                           b->stream));
    PCS_HIP(hipStreamSynchronize(b->stream));
    b->n_tiles = b->vtile[b->n_videos];
    b->host_vtile = true;
    return PCS_OK;
}

DistArgs dist_args(pcs_batch* b, uint32_t lo, uint32_t n)
{
    DistArgs d{};
    d.sys = b->ctx->d_sys;
    d.props = b->d_props;
    d.prop_lo = lo;
    d.n_props = n;
    d.rec_count = b->d_count;
    d.dist = b->d_dist;
    d.exact = b->d_exact;
    d.tracks = b->d_tracks;
    d.status = b->d_status;
    return d;
}

// Enqueue the three batched kernels (reset, persistent walk, distance).
int enqueue_batched(pcs_batch* b, bool with_events)
{
    pcs_ctx* c = b->ctx;
    const pcs_tuning& t = c->tune;
    const bool gauss = c->sys.detect_model == PCS_DETECT_GAUSS_Z;
    const bool zref = c->sys.z_boundary == PCS_Z_REFLECT;
    if (with_events) PCS_HIP(hipEventRecord(b->ev[0], b->stream));
    PCS_HIP(launch_reset(b->d_count, b->n_props, b->d_tile_counter, b->stream));
    if (with_events) PCS_HIP(hipEventRecord(b->ev[1], b->stream));
    const uint32_t grid = (uint32_t)c->n_cu * t.groups_per_cu;
    PCS_HIP(launch_walk(walk_args(b, 0, b->dev_planned ? 0xffffffffu : b->n_tiles), (int)t.particles_per_lane,
                        (int)t.waves_per_group, true, grid, gauss, zref, b->stream));
    if (with_events) PCS_HIP(hipEventRecord(b->ev[2], b->stream));
    PCS_HIP(launch_distance(dist_args(b, 0, b->n_props), c->sys.n_k, c->sys.n_de, b->stream));
    if (with_events) PCS_HIP(hipEventRecord(b->ev[3], b->stream));
    return PCS_OK;
}

// Streams (same priority as the batch stream) and fork/join events of the
// multi-stream plain form.
int plain_streams(pcs_batch* b)
{
    if (b->pfork) return PCS_OK;
    for (int k = 0; k < PCS_PLAIN_STREAMS; ++k) {
        PCS_HIP(hipStreamCreateWithFlags(&b->pstream[k], hipStreamNonBlocking));
        PCS_HIP(hipEventCreateWithFlags(&b->pjoin[k], hipEventDisableTiming));
    }
    PCS_HIP(hipEventCreateWithFlags(&b->pfork, hipEventDisableTiming));
    return PCS_OK;
}

// Plain form: per simulation, one walk launch over its tiles and one
// distance launch (the reference's per-job pair, src/gpu/detail/simulation_impl.inl:1064-1097),
// on the batch stream or, with `streams`, round-robin over PCS_PLAIN_STREAMS
// streams that fork from the batch stream and join it again (the reference's
// job pool keeps one pair in flight per queue, simulation_impl.inl:686-710).
int enqueue_plain(pcs_batch* b, const std::vector<uint32_t>* subset, bool streams = false)
{
    pcs_ctx* c = b->ctx;
    const pcs_tuning& t = c->tune;
    const bool gauss = c->sys.detect_model == PCS_DETECT_GAUSS_Z;
    const bool zref = c->sys.z_boundary == PCS_Z_REFLECT;
    const uint32_t n = subset ? (uint32_t)subset->size() : b->n_props;
    PCS_TRY(fetch_plan(b));
    if (streams) {
        PCS_TRY(plain_streams(b));
        PCS_HIP(hipEventRecord(b->pfork, b->stream));
        for (int k = 0; k < PCS_PLAIN_STREAMS; ++k) PCS_HIP(hipStreamWaitEvent(b->pstream[k], b->pfork, 0));
    }
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t P = subset ? (*subset)[i] : i;
        const pcs_proposal& p = b->props[P];
        const uint32_t lo = b->vtile[p.first_video], hi = b->vtile[p.first_video + p.n_videos];
        hipStream_t s = streams ? b->pstream[i % PCS_PLAIN_STREAMS] : b->stream;
        PCS_HIP(launch_walk(walk_args(b, lo, hi), (int)t.particles_per_lane, (int)t.waves_per_group, false, 0, gauss,
                            zref, s));
        PCS_HIP(launch_distance(dist_args(b, P, 1), c->sys.n_k, c->sys.n_de, s));
    }
    if (streams) {
        for (int k = 0; k < PCS_PLAIN_STREAMS; ++k) {
            PCS_HIP(hipEventRecord(b->pjoin[k], b->pstream[k]));
            PCS_HIP(hipStreamWaitEvent(b->stream, b->pjoin[k], 0));
        }
    }
    return PCS_OK;
}

// Capture reset + walk + distance into one graph. Events recorded during
// capture do not yield elapsed times on this runtime, so the graph is
// captured without them; PCS_MODE_BATCHED_DIRECT launches the same kernels
// with events between them for kernel timing.
int capture_graph_once(pcs_batch* b, bool events)
{
    destroy_graph(b);
    PCS_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int r = enqueue_batched(b, events);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (r != PCS_OK) {
        if (g) hipGraphDestroy(g);
        return r;
    }
    if (e != hipSuccess) return hip_error(e, "hipStreamEndCapture");
    b->graph = g;
    PCS_HIP(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
    b->graph_tune = b->ctx->tune;
    b->graph_events = events;
    return PCS_OK;
}

int capture_graph(pcs_batch* b) { return capture_graph_once(b, false); }

bool same_tune(const pcs_tuning& a, const pcs_tuning& b) { return std::memcmp(&a, &b, sizeof(a)) == 0; }

// Proposals whose slice overflowed are rerun (plain form) with a slice of
// exactly the size the first pass counted (the counter keeps counting past
// the slice end), placed in a second arena; the counter-based stream makes
// the rerun reproduce the same tracks. Their table entries point there until
// the next run restores them.
int rerun_overflows(pcs_batch* b)
{
    std::vector<int32_t> st(b->n_props);
    std::vector<uint32_t> cnt(b->n_props);
    PCS_HIP(hipMemcpyAsync(st.data(), b->d_status, b->n_props * 4, hipMemcpyDeviceToHost, b->stream));
    PCS_HIP(hipMemcpyAsync(cnt.data(), b->d_count, b->n_props * 4, hipMemcpyDeviceToHost, b->stream));
    PCS_HIP(hipStreamSynchronize(b->stream));
    std::vector<uint32_t> over;
    uint64_t need = 0;
    for (uint32_t i = 0; i < b->n_props; ++i)
        if (st[i] == PCS_EOVERFLOW) {
            over.push_back(i);
            need += cnt[i];
        }
    b->reruns = (uint32_t)over.size();
    if (over.empty()) return PCS_OK;
    PCS_TRY(dev_alloc(&b->d_rec2, need));
    PCS_TRY(dev_alloc(&b->d_srt2, need));
    // the overflowed proposals' table entries, patched to their exact slices
    std::vector<DevProp> dp(over.size());
    if (b->dev_planned) {
        for (size_t k = 0; k < over.size(); ++k)
            PCS_HIP(hipMemcpyAsync(&dp[k], b->d_props + over[k], sizeof(DevProp), hipMemcpyDeviceToHost, b->stream));
        PCS_HIP(hipStreamSynchronize(b->stream));
    } else {
        for (size_t k = 0; k < over.size(); ++k) dp[k] = b->hprops[over[k]];
    }
    uint64_t off = 0;
    for (size_t k = 0; k < over.size(); ++k) {
        const uint32_t P = over[k];
        dp[k].rec = b->d_rec2 + off;
        dp[k].srt = b->d_srt2 + off;
        dp[k].rec_cap = cnt[P];
        off += cnt[P];
        PCS_HIP(hipMemcpyAsync(b->d_props + P, &dp[k], sizeof(DevProp), hipMemcpyHostToDevice, b->stream));
        PCS_HIP(hipMemsetAsync(b->d_count + P, 0, 4, b->stream));
    }
    PCS_TRY(enqueue_plain(b, &over));
    PCS_HIP(hipStreamSynchronize(b->stream));
    return PCS_OK;
}

// Undo rerun_overflows before a new run: the table again points at the
// planned slices (re-planned on the device for a device-planned batch,
// unless `replan` is false because a new plan follows anyway).
int restore_slices(pcs_batch* b, bool replan = true)
{
    if (!b->d_rec2) return PCS_OK;
    if (b->dev_planned) {
        if (replan) PCS_TRY(device_plan(b));
    } else if (b->n_props) {
        PCS_HIP(hipMemcpyAsync(b->d_props, b->hprops.data(), b->n_props * sizeof(DevProp), hipMemcpyHostToDevice,
                               b->stream));
    }
    PCS_HIP(hipStreamSynchronize(b->stream));
    hipFree(b->d_rec2);
    hipFree(b->d_srt2);
    b->d_rec2 = b->d_srt2 = nullptr;
    return PCS_OK;
}

}  // namespace

// ------------------------------------------------- allocation-free variant
//
// pcs_simulate_static keeps every device buffer of a run (proposal records,
// tables, counters, results, record arena and sort scratch) in `scratch`, a
// fixed array of PCS_STATIC_SCRATCH_BYTES on each GPU, held until the
// process ends. The array's name is a host-side stand-in, not a device
// address: its address on a GPU is looked up once, when the first context
// opens on that GPU (pcs_open, with the GPU current), and kept per device in
// g_scratch; host code uses only that pointer.
static __device__ __align__(256) unsigned char scratch[PCS_STATIC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;

// Workspace of pcs_pmc_weights on one GPU (grown on demand, kept until exit).
struct PmcWork {
    std::mutex mu;
    double *xn = nullptr, *xo = nullptr, *w = nullptr, *s = nullptr, *part = nullptr;
    size_t cap[5] = {0, 0, 0, 0, 0};
    hipStream_t stream = nullptr;
};
PmcWork g_pmc[kMaxDevices];
unsigned char* g_scratch[kMaxDevices] = {};

// With `device` current.
int scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return set_error(PCS_EINVAL, "device %d above %d", device, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return PCS_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_error(e, "hipGetSymbolAddress(scratch)");
    if (!ptr) return set_error(PCS_EHIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch[device] = static_cast<unsigned char*>(ptr);
    return PCS_OK;
}

unsigned char* scratch_ptr(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return (device >= 0 && device < kMaxDevices) ? g_scratch[device] : nullptr;
}

// Byte offsets of one run's buffers in the scratch array (256-byte aligned);
// the rest holds the record arena and, after it, the sort scratch.
struct StaticLayout {
    size_t props, in, vframes, vparticles, vprop, vtile, count, tracks, status, dist, exact, misc, rec, srt;
    uint64_t arena_cap;  // records
};

bool static_layout(uint32_t n, uint32_t nv, StaticLayout& L)
{
    size_t o = 0;
    auto take = [&o](size_t bytes) {
        const size_t at = o;
        o += (bytes + 255u) & ~(size_t)255u;
        return at;
    };
    L.props = take((size_t)n * sizeof(DevProp));
    L.in = take((size_t)n * sizeof(pcs_proposal));
    L.vframes = take((size_t)nv * 4u);
    L.vparticles = take((size_t)nv * 4u);
    L.vprop = take((size_t)nv * 4u);
    L.vtile = take(((size_t)nv + 1u) * 4u);
    L.count = take((size_t)n * 4u);
    L.tracks = take((size_t)n * 4u);
    L.status = take((size_t)n * 4u);
    L.dist = take((size_t)n * 8u);
    L.exact = take((size_t)n * 8u);
    L.misc = take(64);  // seed (u64), tile counter (u32), totals (2 x u64)
    if (o >= PCS_STATIC_SCRATCH_BYTES) return false;
    L.arena_cap = (PCS_STATIC_SCRATCH_BYTES - o) / 8u;
    L.rec = o;
    L.srt = o + (size_t)L.arena_cap * 4u;
    return true;
}

}  // namespace

// ------------------------------------------------------------------- ABI

extern "C" {

int pcs_abi_version(void) { return PCS_ABI_VERSION; }

const char* pcs_strerror(int code)
{
    switch (code) {
        case PCS_OK: return "ok";
        case PCS_EINVAL: return "invalid argument";
        case PCS_ENOMEM: return "out of memory";
        case PCS_EHIP: return "HIP runtime error";
        case PCS_EAGAIN: return "resource temporarily unavailable";
        case PCS_EOVERFLOW: return "record arena overflow";
        case PCS_ENODEV: return "no usable GPU";
        case PCS_ESTATE: return "invalid state";
        default: return "unknown error";
    }
}

const char* pcs_last_error(void) { return g_last_error.c_str(); }

int pcs_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int pcs_reference_hist(const uint64_t* K, const double* DE, uint64_t n, double ub_de, uint32_t n_k, uint32_t n_de,
                       uint32_t* hist, uint64_t* dropped)
{
    if ((!K || !DE) && n) return set_error(PCS_EINVAL, "K/DE NULL");
    if (!hist || n_k == 0 || n_de == 0 || !(ub_de > 0.0)) return set_error(PCS_EINVAL, "bad histogram shape");
    std::memset(hist, 0, (size_t)n_k * n_de * 4);
    const double w = ub_de / (double)n_de;
    uint64_t drop = 0;
    for (uint64_t i = 0; i < n; ++i) {
        if (!(DE[i] < ub_de)) continue;
        const long idx = (long)(DE[i] / w);
        if (K[i] < 1 || K[i] > n_k || idx < 0 || idx >= (long)n_de) {
            ++drop;
            continue;
        }
        hist[(size_t)(K[i] - 1) * n_de + (size_t)idx] += 1u;
    }
    if (dropped) *dropped = drop;
    return PCS_OK;
}

int pcs_open(const pcs_system* sys, const uint32_t* obs_hist, int device, pcs_ctx** out)
{
    if (!out) return set_error(PCS_EINVAL, "out is NULL");
    *out = nullptr;
    PCS_TRY(validate_system(sys));
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(PCS_ENODEV, "no GPU visible");
    if (device < 0 || device >= ndev) return set_error(PCS_EINVAL, "device %d outside [0, %d)", device, ndev);
    PCS_HIP(hipSetDevice(device));
    PCS_TRY(scratch_init(device));
    pcs_ctx* c = new (std::nothrow) pcs_ctx();
    if (!c) return set_error(PCS_ENOMEM, "context");
    c->device = device;
    c->sys = *sys;
    hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device);
    if (c->n_cu <= 0) c->n_cu = 256;
    std::vector<uint32_t> sat;
    uint32_t maxrow = 0;
    build_sat(obs_hist, sys->n_k, sys->n_de, sat, maxrow);
    int r = dev_alloc(&c->d_sat, sat.size());
    if (r == PCS_OK) r = dev_alloc(&c->d_sys, 1);
    if (r == PCS_OK) {
        c->dsys = make_devsys(*sys, maxrow, c->d_sat);
        hipError_t e = hipMemcpy(c->d_sat, sat.data(), sat.size() * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(c->d_sys, &c->dsys, sizeof(DevSys), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) r = hip_error(e, "pcs_open");
    }
    if (r != PCS_OK) {
        pcs_close(c);
        return r;
    }
    *out = c;
    return PCS_OK;
}

namespace {

void ctx_release(pcs_ctx* c)
{
    if (c->refs.fetch_sub(1) != 1) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamDestroy(c->stream);
    if (c->d_sat) hipFree(c->d_sat);
    if (c->d_sys) hipFree(c->d_sys);
    delete c;
}

}  // namespace

int pcs_close(pcs_ctx* c)
{
    if (!c) return PCS_OK;
    ctx_release(c);
    return PCS_OK;
}

int pcs_set_tuning(pcs_ctx* c, const pcs_tuning* t)
{
    if (!c || !t) return set_error(PCS_EINVAL, "NULL argument");
    pcs_tuning v = *t;
    if (!v.particles_per_lane) v.particles_per_lane = 1;
    if (!v.waves_per_group) v.waves_per_group = 4;
    if (!v.groups_per_cu) v.groups_per_cu = 8;
    if (!v.stage_records) v.stage_records = 256;
    PCS_TRY(check_tuning(v));
    c->tune = v;
    return PCS_OK;
}

int pcs_get_tuning(pcs_ctx* c, pcs_tuning* t)
{
    if (!c || !t) return set_error(PCS_EINVAL, "NULL argument");
    *t = c->tune;
    return PCS_OK;
}

// Waits for weights enqueued by pcs_batch_pmc_weights_async (they read the
// trials' parameters and the population) and forgets them: a call that
// rewrites those buffers comes first.
static int drop_weights(pcs_batch* b)
{
    if (b->wstream) PCS_HIP(hipStreamSynchronize(b->wstream));
    b->w_pending = false;
    return PCS_OK;
}

int pcs_batch_create(pcs_ctx* c, pcs_batch** out)
{
    if (!c || !out) return set_error(PCS_EINVAL, "NULL argument");
    *out = nullptr;
    PCS_HIP(hipSetDevice(c->device));
    pcs_batch* b = new (std::nothrow) pcs_batch();
    if (!b) return set_error(PCS_ENOMEM, "batch");
    b->ctx = c;
    c->refs.fetch_add(1);
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    for (int i = 0; i < 4 && e == hipSuccess; ++i) e = hipEventCreate(&b->ev[i]);
    if (e == hipSuccess) e = hipHostMalloc((void**)&b->h_seed, sizeof(uint64_t), hipHostMallocDefault);
    int r = e == hipSuccess ? PCS_OK : hip_error(e, "pcs_batch_create");
    if (r == PCS_OK) r = dev_alloc(&b->d_seed, 1);
    if (r == PCS_OK) r = dev_alloc(&b->d_tile_counter, 1);
    if (r != PCS_OK) {
        pcs_batch_destroy(b);
        return r;
    }
    *out = b;
    return PCS_OK;
}

int pcs_batch_destroy(pcs_batch* b)
{
    if (!b) return PCS_OK;
    hipSetDevice(b->ctx->device);
    if (b->stream) hipStreamSynchronize(b->stream);
    for (hipStream_t ps : b->pstream)
        if (ps) hipStreamSynchronize(ps);
    if (b->wstream) hipStreamSynchronize(b->wstream);
    destroy_graph(b);
    void* ptrs[] = {b->d_props, b->d_vframes, b->d_vparticles, b->d_vprop, b->d_vtile, b->d_rec, b->d_srt,
                    b->d_rec2, b->d_srt2, b->d_count, b->d_tracks, b->d_tile_counter, b->d_dist, b->d_exact, b->d_status,
                    b->d_seed, b->d_in, b->d_theta_pop, b->d_cumw, b->d_theta_new, b->d_totals, b->d_wpop,
                    b->d_xn, b->d_xo, b->d_wsum, b->d_part};
    for (void* p : ptrs)
        if (p) hipFree(p);
    if (b->h_seed) hipHostFree(b->h_seed);
    for (hipEvent_t& e : b->ev)
        if (e) hipEventDestroy(e);
    for (int k = 0; k < PCS_PLAIN_STREAMS; ++k) {
        if (b->pstream[k]) hipStreamDestroy(b->pstream[k]);
        if (b->pjoin[k]) hipEventDestroy(b->pjoin[k]);
    }
    if (b->pfork) hipEventDestroy(b->pfork);
    if (b->wfork) hipEventDestroy(b->wfork);
    if (b->wstream) hipStreamDestroy(b->wstream);
    if (b->stream) hipStreamDestroy(b->stream);
    pcs_ctx* c = b->ctx;
    delete b;
    ctx_release(c);
    return PCS_OK;
}

int pcs_batch_load(pcs_batch* b, const pcs_proposal* props, uint32_t n, const uint32_t* vframes,
                   const uint32_t* vparticles, uint32_t n_videos)
{
    if (!b) return set_error(PCS_EINVAL, "batch is NULL");
    if (n && (!props || !vframes || !vparticles)) return set_error(PCS_EINVAL, "NULL proposal or video arrays");
    pcs_ctx* c = b->ctx;
    PCS_HIP(hipSetDevice(c->device));
    PCS_HIP(hipStreamSynchronize(b->stream));
    PCS_TRY(drop_weights(b));
    PCS_TRY(restore_slices(b));
    destroy_graph(b);
    b->ready = false;
    b->ran = false;
    b->dev_planned = false;
    b->host_vtile = false;
    b->props.assign(props, props + n);
    b->hprops.resize(n);
    b->vframes.assign(vframes, vframes + n_videos);
    b->vparticles.assign(vparticles, vparticles + n_videos);
    b->vprop.assign(n_videos, 0xffffffffu);
    b->n_props = n;
    b->n_videos = n_videos;
    uint64_t arena = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const pcs_proposal& p = props[i];
        if (p.reserved != 0) return set_error(PCS_EINVAL, "proposal %u: reserved field must be 0", i);
        if ((uint64_t)p.first_video + p.n_videos > n_videos)
            return set_error(PCS_EINVAL, "proposal %u: videos [%u, %u) outside %u", i, p.first_video,
                             p.first_video + p.n_videos, n_videos);
        PCS_TRY(validate_proposal(c->sys, p));
        for (uint32_t v = p.first_video; v < p.first_video + p.n_videos; ++v) {
            if (b->vprop[v] != 0xffffffffu)
                return set_error(PCS_EINVAL, "video %u belongs to proposals %u and %u", v, b->vprop[v], i);
            b->vprop[v] = i;
        }
        DevProp& d = b->hprops[i];
        std::memset(&d, 0, sizeof(d));
        d.sim_id = p.sim_id;
        d.first_video = p.first_video;
        d.n_videos = p.n_videos;
        d.sys_id = 0;
        d.rec = nullptr;  // set below, once the arena exists
        d.srt = nullptr;
        d.rec_cap = record_capacity(c->sys, p, vframes + p.first_video, vparticles + p.first_video, c->cap_factor);
        arena += d.rec_cap;
        for (int s = 0; s < PCS_MAX_SPECIES; ++s) {
            d.half_az[s] = p.half_az[s];
            d.cum_prob[s] = p.cum_prob[s];
            d.sigma[s] = p.sigma[s];
        }
    }
    uint64_t tiles = 0;
    for (uint32_t v = 0; v < n_videos; ++v) {
        if (b->vprop[v] == 0xffffffffu) return set_error(PCS_EINVAL, "video %u belongs to no proposal", v);
        if (vparticles[v] > (1u << 31)) return set_error(PCS_EINVAL, "video %u: more than 2^31 particles", v);
        tiles += (vparticles[v] + 63u) / 64u;
    }
    if (tiles > 0xffffffffull) return set_error(PCS_EINVAL, "the batch needs %llu tiles, more than 2^32 - 1",
                                                (unsigned long long)tiles);
    b->arena_records = arena;
    plan_tiles(b, (int)c->tune.particles_per_lane);

    // A capacity is zeroed before its buffers are replaced, so that an
    // allocation failing part-way makes the next call allocate them all.
    if (n > b->cap_props) {
        b->cap_props = 0;
        PCS_TRY(dev_alloc(&b->d_props, n));
        PCS_TRY(dev_alloc(&b->d_count, n));
        PCS_TRY(dev_alloc(&b->d_tracks, n));
        PCS_TRY(dev_alloc(&b->d_dist, n));
        PCS_TRY(dev_alloc(&b->d_exact, n));
        PCS_TRY(dev_alloc(&b->d_status, n));
        b->cap_props = n;
    }
    if (n_videos + 1 > b->cap_videos) {
        b->cap_videos = 0;
        PCS_TRY(dev_alloc(&b->d_vframes, n_videos + 1));
        PCS_TRY(dev_alloc(&b->d_vparticles, n_videos + 1));
        PCS_TRY(dev_alloc(&b->d_vprop, n_videos + 1));
        PCS_TRY(dev_alloc(&b->d_vtile, n_videos + 1));
        b->cap_videos = n_videos + 1;
    }
    if (arena > b->cap_arena) {
        b->cap_arena = 0;
        PCS_TRY(dev_alloc(&b->d_rec, arena));
        PCS_TRY(dev_alloc(&b->d_srt, arena));
        b->cap_arena = arena;
    }
    {
        uint64_t off = 0;
        for (uint32_t i = 0; i < n; ++i) {
            b->hprops[i].rec = b->d_rec + off;
            b->hprops[i].srt = b->d_srt + off;
            off += b->hprops[i].rec_cap;
        }
    }
    if (n) {
        PCS_HIP(hipMemcpyAsync(b->d_props, b->hprops.data(), n * sizeof(DevProp), hipMemcpyHostToDevice, b->stream));
    }
    if (n_videos) {
        PCS_HIP(hipMemcpyAsync(b->d_vframes, vframes, n_videos * 4, hipMemcpyHostToDevice, b->stream));
        PCS_HIP(hipMemcpyAsync(b->d_vparticles, vparticles, n_videos * 4, hipMemcpyHostToDevice, b->stream));
        PCS_HIP(hipMemcpyAsync(b->d_vprop, b->vprop.data(), n_videos * 4, hipMemcpyHostToDevice, b->stream));
    }
    PCS_TRY(upload_plan(b));
    PCS_HIP(hipStreamSynchronize(b->stream));
    b->ready = true;
    return PCS_OK;
}

int pcs_batch_run(pcs_batch* b, uint64_t seed, int mode)
{
    if (!b) return set_error(PCS_EINVAL, "batch is NULL");
    if (mode != PCS_MODE_PLAIN && mode != PCS_MODE_BATCHED && mode != PCS_MODE_BATCHED_DIRECT &&
        mode != PCS_MODE_PLAIN_STREAMS)
        return set_error(PCS_EINVAL, "mode %d unknown", mode);
    if (!b->ready)
        return set_error(PCS_ESTATE, "batch holds no simulations: no pcs_batch_load / pcs_batch_propose, or the last one failed");
    pcs_ctx* c = b->ctx;
    PCS_HIP(hipSetDevice(c->device));
    if (b->planned_ppl != (int)c->tune.particles_per_lane) {
        PCS_HIP(hipStreamSynchronize(b->stream));
        if (b->dev_planned) {
            PCS_TRY(device_plan(b));
        } else {
            plan_tiles(b, (int)c->tune.particles_per_lane);
            PCS_TRY(upload_plan(b));
        }
        destroy_graph(b);
    }
    // The seed lives in device memory so one captured graph serves every seed.
    PCS_HIP(hipStreamSynchronize(b->stream));
    PCS_TRY(restore_slices(b));
    *b->h_seed = seed;
    PCS_HIP(hipMemcpyAsync(b->d_seed, b->h_seed, sizeof(uint64_t), hipMemcpyHostToDevice, b->stream));
    b->reruns = 0;
    if (mode == PCS_MODE_BATCHED) {
        if (!b->gexec || !same_tune(b->graph_tune, c->tune)) PCS_TRY(capture_graph(b));
        if (!b->graph_events) PCS_HIP(hipEventRecord(b->ev[0], b->stream));
        PCS_HIP(hipGraphLaunch(b->gexec, b->stream));
        if (!b->graph_events) PCS_HIP(hipEventRecord(b->ev[3], b->stream));
    } else if (mode == PCS_MODE_BATCHED_DIRECT) {
        PCS_TRY(enqueue_batched(b, true));
    } else {
        PCS_HIP(hipEventRecord(b->ev[0], b->stream));
        PCS_HIP(launch_reset(b->d_count, b->n_props, b->d_tile_counter, b->stream));
        PCS_TRY(enqueue_plain(b, nullptr, mode == PCS_MODE_PLAIN_STREAMS));
        PCS_HIP(hipEventRecord(b->ev[3], b->stream));
    }
    b->last_mode = mode;
    b->ran = true;
    return PCS_OK;
}

int pcs_batch_sync(pcs_batch* b)
{
    if (!b) return set_error(PCS_EINVAL, "batch is NULL");
    PCS_HIP(hipSetDevice(b->ctx->device));
    PCS_HIP(hipStreamSynchronize(b->stream));
    return PCS_OK;
}

int pcs_batch_timing(pcs_batch* b, float out[3])
{
    if (!b || !out) return set_error(PCS_EINVAL, "NULL argument");
    if (!b->ran) return set_error(PCS_ESTATE, "batch has not run");
    PCS_HIP(hipEventSynchronize(b->ev[3]));
    out[0] = out[1] = out[2] = -1.0f;
    PCS_HIP(hipEventElapsedTime(&out[0], b->ev[0], b->ev[3]));
    if (b->last_mode == PCS_MODE_BATCHED_DIRECT || (b->last_mode == PCS_MODE_BATCHED && b->graph_events)) {
        PCS_HIP(hipEventElapsedTime(&out[1], b->ev[1], b->ev[2]));
        PCS_HIP(hipEventElapsedTime(&out[2], b->ev[2], b->ev[3]));
    }
    return PCS_OK;
}

int pcs_batch_arena(pcs_batch* b, uint64_t* capacity_records, uint64_t* bytes)
{
    if (!b) return set_error(PCS_EINVAL, "batch is NULL");
    if (b->dev_planned && b->d_totals) {  // the device plan's record slices
        uint64_t tot[2] = {0, 0};
        PCS_HIP(hipMemcpyAsync(tot, b->d_totals, sizeof(tot), hipMemcpyDeviceToHost, b->stream));
        PCS_HIP(hipStreamSynchronize(b->stream));
        b->arena_records = std::min<uint64_t>(tot[0], b->cap_arena);
    }
    if (capacity_records) *capacity_records = b->arena_records;
    if (bytes) *bytes = 2ull * 4ull * b->cap_arena;
    return PCS_OK;
}

int pcs_batch_overflow_reruns(pcs_batch* b, uint32_t* n)
{
    if (!b || !n) return set_error(PCS_EINVAL, "NULL argument");
    *n = b->reruns;
    return PCS_OK;
}

int pcs_batch_read(pcs_batch* b, double* dist_out, uint32_t* hist_out, uint32_t* tracks_out)
{
    if (!b) return set_error(PCS_EINVAL, "batch is NULL");
    if (!b->ran) return set_error(PCS_ESTATE, "batch has not run");
    pcs_ctx* c = b->ctx;
    PCS_HIP(hipSetDevice(c->device));
    PCS_HIP(hipStreamSynchronize(b->stream));
    PCS_TRY(rerun_overflows(b));
    if (b->n_props == 0) return PCS_OK;
    if (dist_out)
        PCS_HIP(hipMemcpyAsync(dist_out, b->d_dist, b->n_props * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    if (tracks_out)
        PCS_HIP(hipMemcpyAsync(tracks_out, b->d_tracks, b->n_props * 4, hipMemcpyDeviceToHost, b->stream));
    if (hist_out) {
        // Dense histograms are formed from the records in chunks of proposals.
        std::vector<uint32_t> cnt(b->n_props);
        PCS_HIP(hipMemcpyAsync(cnt.data(), b->d_count, b->n_props * 4, hipMemcpyDeviceToHost, b->stream));
        PCS_HIP(hipStreamSynchronize(b->stream));
        const size_t per = (size_t)c->sys.n_k * c->sys.n_de;
        const uint32_t chunk = (uint32_t)std::max<size_t>(1, std::min<size_t>(4096, (256u << 20) / (per * 4)));
        uint32_t* d_h = nullptr;
        PCS_TRY(dev_alloc(&d_h, per * chunk));
        int r = PCS_OK;
        for (uint32_t lo = 0; lo < b->n_props && r == PCS_OK; lo += chunk) {
            const uint32_t m = std::min(chunk, b->n_props - lo);
            uint32_t mx = 0;
            for (uint32_t i = lo; i < lo + m; ++i)  // the kernel caps each count at its slice
                mx = std::max(mx, b->dev_planned ? cnt[i] : std::min(cnt[i], b->hprops[i].rec_cap));
            hipError_t e = hipMemsetAsync(d_h, 0, per * m * 4, b->stream);
            if (e == hipSuccess)
                e = launch_dense_hist(b->d_props, lo, m, b->d_count, mx, c->sys.n_k, c->sys.n_de, d_h, b->stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(hist_out + per * lo, d_h, per * m * 4, hipMemcpyDeviceToHost, b->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
            if (e != hipSuccess) r = hip_error(e, "dense histogram");
        }
        hipFree(d_h);
        PCS_TRY(r);
    }
    PCS_HIP(hipStreamSynchronize(b->stream));
    return PCS_OK;
}

int pcs_batch_propose(pcs_batch* b, const pcs_prior* pr, const double* theta, const double* cum_w,
                      const double* tau, uint32_t n_pop, uint32_t n_trials, const uint32_t* video_frames,
                      uint32_t J, uint32_t first_sim_id, uint64_t seed, double* theta_out)
{
    if (!b || !pr) return set_error(PCS_EINVAL, "NULL argument");
    pcs_ctx* c = b->ctx;
    const uint32_t C = pr->n_components, Z = pr->n_depths;
    if (C != c->sys.n_species) return set_error(PCS_EINVAL, "prior has %u components, the model %u", C, c->sys.n_species);
    if (Z != 1u && Z != C) return set_error(PCS_EINVAL, "n_depths %u is neither 1 nor n_components", Z);
    const uint32_t P = 2u * C + Z;
    const bool resident = !theta && !cum_w;
    if (resident) {
        if (!b->pop_resident) return set_error(PCS_ESTATE, "no resident population (pcs_batch_set_population)");
        if (b->pop_P != P) return set_error(PCS_EINVAL, "resident population has %u parameters, the prior %u", b->pop_P, P);
        if (n_pop != 0 && n_pop != b->pop_n)
            return set_error(PCS_EINVAL, "n_pop %u, the resident population has %u", n_pop, b->pop_n);
        n_pop = b->pop_n;
    }
    if (n_pop == 0 || (!resident && (!theta || !cum_w)) || !tau)
        return set_error(PCS_EINVAL, "population is empty or NULL");
    if (J == 0 || J > 4096u) return set_error(PCS_EINVAL, "videos_per_trial %u outside [1, 4096]", J);
    if (n_trials && !video_frames) return set_error(PCS_EINVAL, "video_frames is NULL");
    if ((uint64_t)n_trials * J >= 0x80000000ull) return set_error(PCS_EINVAL, "n_trials x videos_per_trial too large");
    const double lo[3] = {pr->m_lo, pr->log10c_lo, pr->az_lo}, hi[3] = {pr->m_hi, pr->log10c_hi, pr->az_hi};
    for (int k = 0; k < 3; ++k)
        if (!std::isfinite(lo[k]) || !std::isfinite(hi[k]) || !(lo[k] <= hi[k]))
            return set_error(PCS_EINVAL, "prior interval %d is empty or not finite", k);
    if (!(pr->m_lo >= 0.0) || !(pr->az_lo >= 0.0)) return set_error(PCS_EINVAL, "m and az bounds must be >= 0");
    if (!(pr->deltat > 0.0) || !(pr->volume > 0.0)) return set_error(PCS_EINVAL, "deltat and volume must be positive");
    // the Poisson draws take ~lambda / 64 exponentials per lane: bound lambda
    if ((double)C * std::pow(10.0, pr->log10c_hi) * pr->volume > 16777216.0)
        return set_error(PCS_EINVAL, "prior allows more than 2^24 expected particles per video");
    // bounded inputs keep every reflection and sum finite
    for (uint32_t p = 0; p < P; ++p)
        if (!(tau[p] >= 0.0 && tau[p] <= 1e100)) return set_error(PCS_EINVAL, "tau[%u] outside [0, 1e100]", p);
    if (!resident)
        for (size_t i = 0; i < (size_t)P * n_pop; ++i)
            if (!(std::fabs(theta[i]) <= 1e100)) return set_error(PCS_EINVAL, "theta[%zu] not finite or above 1e100", i);
    for (int k = 0; k < 3; ++k)
        if (!(std::fabs(lo[k]) <= 1e100 && std::fabs(hi[k]) <= 1e100))
            return set_error(PCS_EINVAL, "prior bounds above 1e100");
    if (!resident)
        for (uint32_t i = 0; i < n_pop; ++i)
            if (!std::isfinite(cum_w[i]) || (i && cum_w[i] < cum_w[i - 1]))
                return set_error(PCS_EINVAL, "cum_w must be finite and non-decreasing (member %u)", i);
    PCS_HIP(hipSetDevice(c->device));
    PCS_HIP(hipStreamSynchronize(b->stream));
    PCS_TRY(drop_weights(b));
    PCS_TRY(restore_slices(b, false));
    b->ready = false;
    b->ran = false;
    const uint32_t nv = n_trials * J;
    bool relayout = !b->dev_planned || n_trials != b->n_props || J != b->J;
    if (n_trials > b->cap_props) {
        b->cap_props = 0;
        PCS_TRY(dev_alloc(&b->d_props, n_trials));
        PCS_TRY(dev_alloc(&b->d_count, n_trials));
        PCS_TRY(dev_alloc(&b->d_tracks, n_trials));
        PCS_TRY(dev_alloc(&b->d_dist, n_trials));
        PCS_TRY(dev_alloc(&b->d_exact, n_trials));
        PCS_TRY(dev_alloc(&b->d_status, n_trials));
        b->cap_props = n_trials;
        relayout = true;
    }
    if (n_trials > b->cap_in || !b->d_in) {
        b->cap_in = b->cap_new = 0;
        PCS_TRY(dev_alloc(&b->d_in, n_trials));
        PCS_TRY(dev_alloc(&b->d_theta_new, (size_t)3u * PCS_MAX_SPECIES * n_trials));
        b->cap_in = b->cap_new = n_trials;
    }
    if ((size_t)nv + 1u > b->cap_videos) {
        b->cap_videos = 0;
        PCS_TRY(dev_alloc(&b->d_vframes, (size_t)nv + 1u));
        PCS_TRY(dev_alloc(&b->d_vparticles, (size_t)nv + 1u));
        PCS_TRY(dev_alloc(&b->d_vprop, (size_t)nv + 1u));
        PCS_TRY(dev_alloc(&b->d_vtile, (size_t)nv + 1u));
        b->cap_videos = (size_t)nv + 1u;
        relayout = true;
    }
    if (!resident) b->pop_resident = false;  // replaced by an uploaded population without weights
    if (!resident && (n_pop > b->cap_pop || !b->d_theta_pop)) {
        b->cap_pop = 0;
        PCS_TRY(dev_alloc(&b->d_theta_pop, (size_t)3u * PCS_MAX_SPECIES * n_pop));
        PCS_TRY(dev_alloc(&b->d_cumw, n_pop));
        PCS_TRY(dev_alloc(&b->d_wpop, n_pop));
        b->cap_pop = n_pop;
    }
    b->new_P = P;
    if (!b->d_totals) PCS_TRY(dev_alloc(&b->d_totals, 2));
    b->dev_planned = true;
    b->host_vtile = false;
    b->J = J;
    b->n_props = n_trials;
    b->n_videos = nv;
    b->ran = false;
    // host side of the plan: proposal i owns videos [i J, (i + 1) J)
    b->props.assign(n_trials, pcs_proposal{});
    for (uint32_t i = 0; i < n_trials; ++i) {
        b->props[i].sim_id = first_sim_id + i;
        b->props[i].first_video = i * J;
        b->props[i].n_videos = J;
    }
    b->hprops.clear();
    b->vframes.clear();
    b->vparticles.clear();
    hipStream_t st = b->stream;
    if (!resident) {
        PCS_HIP(hipMemcpyAsync(b->d_theta_pop, theta, (size_t)P * n_pop * 8u, hipMemcpyHostToDevice, st));
        PCS_HIP(hipMemcpyAsync(b->d_cumw, cum_w, (size_t)n_pop * 8u, hipMemcpyHostToDevice, st));
    }
    if (nv) PCS_HIP(hipMemcpyAsync(b->d_vframes, video_frames, (size_t)nv * 4u, hipMemcpyHostToDevice, st));
    ProposeArgs pa{};
    pa.theta = b->d_theta_pop;
    pa.cum_w = b->d_cumw;
    pa.theta_out = b->d_theta_new;
    pa.props = b->d_in;
    pa.vparticles = b->d_vparticles;
    pa.n_pop = n_pop;
    pa.n_trials = n_trials;
    pa.J = J;
    pa.C = C;
    pa.Z = Z;
    pa.first_sim_id = first_sim_id;
    pa.seed = seed;
    for (int k = 0; k < 3; ++k) {
        pa.lo[k] = lo[k];
        pa.hi[k] = hi[k];
    }
    pa.deltat = pr->deltat;
    pa.volume = pr->volume;
    for (uint32_t p = 0; p < P; ++p) pa.tau[p] = tau[p];
    PCS_HIP(launch_propose(pa, st));
    PCS_TRY(device_plan(b));
    // The plan's record slices must fit the arena: grow it and plan again
    // when they do not (the first call, or proposals that outgrow the last).
    uint64_t tot[2] = {0, 0};
    PCS_HIP(hipMemcpyAsync(tot, b->d_totals, sizeof(tot), hipMemcpyDeviceToHost, st));
    PCS_HIP(hipStreamSynchronize(st));
    if (tot[1] > 0xffffffffull)  // tile indices are 32-bit
        return set_error(PCS_EINVAL, "the proposed trials need %llu tiles, more than 2^32 - 1",
                         (unsigned long long)tot[1]);
    if (tot[0] > b->cap_arena || !b->d_rec) {
        const uint64_t cap = tot[0] + tot[0] / 8u + 1024u;
        b->cap_arena = 0;
        PCS_TRY(dev_alloc(&b->d_rec, cap));
        PCS_TRY(dev_alloc(&b->d_srt, cap));
        b->cap_arena = cap;
        relayout = true;
        PCS_TRY(device_plan(b));
    }
    b->last_need = tot[0];
    b->arena_records = tot[0];
    if (theta_out && n_trials) {
        for (uint32_t p = 0; p < P; ++p)
            PCS_HIP(hipMemcpyAsync(theta_out + (size_t)p * n_trials, b->d_theta_new + (size_t)p * n_trials,
                                   (size_t)n_trials * 8u, hipMemcpyDeviceToHost, st));
    }
    PCS_HIP(hipStreamSynchronize(st));
    if (relayout) destroy_graph(b);
    // the point pcs_batch_pmc_weights_async starts from: the trials exist
    if (!b->wstream) {
        PCS_HIP(hipStreamCreateWithFlags(&b->wstream, hipStreamNonBlocking));
        PCS_HIP(hipEventCreateWithFlags(&b->wfork, hipEventDisableTiming));
    }
    PCS_HIP(hipEventRecord(b->wfork, st));
    b->ready = true;
    return PCS_OK;
}

int pcs_batch_set_population(pcs_batch* b, const double* theta, const double* w, uint32_t n_pop, uint32_t n_params)
{
    if (!b || !theta || !w) return set_error(PCS_EINVAL, "NULL argument");
    if (n_pop == 0) return set_error(PCS_EINVAL, "population is empty");
    if (n_params < 1 || n_params > 3u * PCS_MAX_SPECIES) return set_error(PCS_EINVAL, "n_params outside [1, 24]");
    for (size_t i = 0; i < (size_t)n_params * n_pop; ++i)
        if (!(std::fabs(theta[i]) <= 1e100)) return set_error(PCS_EINVAL, "theta[%zu] not finite or above 1e100", i);
    for (uint32_t i = 0; i < n_pop; ++i)
        if (!(w[i] >= 0.0) || !std::isfinite(w[i])) return set_error(PCS_EINVAL, "w[%u] negative or not finite", i);
    PCS_HIP(hipSetDevice(b->ctx->device));
    PCS_HIP(hipStreamSynchronize(b->stream));
    PCS_TRY(drop_weights(b));
    b->pop_resident = false;  // until the new population is in place
    if (n_pop > b->cap_pop || !b->d_theta_pop || !b->d_wpop) {
        b->cap_pop = 0;
        PCS_TRY(dev_alloc(&b->d_theta_pop, (size_t)3u * PCS_MAX_SPECIES * n_pop));
        PCS_TRY(dev_alloc(&b->d_cumw, n_pop));
        PCS_TRY(dev_alloc(&b->d_wpop, n_pop));
        b->cap_pop = n_pop;
    }
    hipStream_t st = b->stream;
    PCS_HIP(hipMemcpyAsync(b->d_theta_pop, theta, (size_t)n_params * n_pop * 8u, hipMemcpyHostToDevice, st));
    PCS_HIP(hipMemcpyAsync(b->d_wpop, w, (size_t)n_pop * 8u, hipMemcpyHostToDevice, st));
    PCS_HIP(launch_cumsum(b->d_wpop, n_pop, b->d_cumw, st));
    PCS_HIP(hipStreamSynchronize(st));
    b->pop_resident = true;
    b->pop_n = n_pop;
    b->pop_P = n_params;
    return PCS_OK;
}

int pcs_batch_pmc_weights_async(pcs_batch* b, const double* tau)
{
    if (!b || !tau) return set_error(PCS_EINVAL, "NULL argument");
    if (!b->pop_resident) return set_error(PCS_ESTATE, "no resident population (pcs_batch_set_population)");
    if (!b->ready || !b->dev_planned || b->new_P != b->pop_P)
        return set_error(PCS_ESTATE, "no trials proposed from the resident population (pcs_batch_propose)");
    const uint32_t P = b->pop_P, n_new = b->n_props, n_old = b->pop_n;
    ScaleRows sr{};
    for (uint32_t p = 0; p < P; ++p) {
        if (!(tau[p] >= 0.0) || !std::isfinite(tau[p])) return set_error(PCS_EINVAL, "tau[%u] invalid", p);
        if (tau[p] > 0.0) {
            sr.rows[sr.n_rows] = p;
            sr.tau[sr.n_rows] = tau[p];
            ++sr.n_rows;
        }
    }
    PCS_HIP(hipSetDevice(b->ctx->device));
    PCS_TRY(drop_weights(b));
    b->w_n = n_new;
    b->w_uniform = sr.n_rows == 0;  // every member identical: equal weights
    if (n_new == 0 || b->w_uniform) {
        b->w_pending = true;
        return PCS_OK;
    }
    auto grow = [](double** ptr, size_t& cap, size_t n) -> int {
        if (n <= cap && *ptr) return PCS_OK;
        cap = 0;
        const int r = dev_alloc(ptr, n);
        cap = r == PCS_OK ? n : 0;
        return r;
    };
    PCS_TRY(grow(&b->d_xn, b->cap_xn, (size_t)sr.n_rows * n_new));
    PCS_TRY(grow(&b->d_xo, b->cap_xo, (size_t)sr.n_rows * n_old));
    PCS_TRY(grow(&b->d_wsum, b->cap_wsum, n_new));
    PCS_TRY(grow(&b->d_part, b->cap_part, (size_t)pmc_chunks(n_old) * n_new));
    // after the proposal step (recorded by pcs_batch_propose), not after
    // whatever the batch stream has taken since: a walk enqueued before this
    // call runs beside the weights
    PCS_HIP(hipStreamWaitEvent(b->wstream, b->wfork, 0));
    hipStream_t st = b->wstream;
    PCS_HIP(launch_scale_rows(b->d_theta_new, n_new, sr, b->d_xn, st));
    PCS_HIP(launch_scale_rows(b->d_theta_pop, n_old, sr, b->d_xo, st));
    PCS_HIP(launch_pmc_sums(b->d_xn, n_new, b->d_xo, b->d_wpop, n_old, sr.n_rows, b->d_part, b->d_wsum, st));
    b->w_pending = true;
    return PCS_OK;
}

int pcs_batch_pmc_weights_wait(pcs_batch* b, double* w_out)
{
    if (!b || !w_out) return set_error(PCS_EINVAL, "NULL argument");
    if (!b->w_pending) return set_error(PCS_ESTATE, "no weights enqueued (pcs_batch_pmc_weights_async)");
    PCS_HIP(hipSetDevice(b->ctx->device));
    b->w_pending = false;
    const uint32_t n_new = b->w_n;
    if (n_new == 0) return PCS_OK;
    if (b->w_uniform) {
        for (uint32_t j = 0; j < n_new; ++j) w_out[j] = 1.0 / (double)n_new;
        return PCS_OK;
    }
    PCS_HIP(hipMemcpyAsync(w_out, b->d_wsum, (size_t)n_new * 8u, hipMemcpyDeviceToHost, b->wstream));
    PCS_HIP(hipStreamSynchronize(b->wstream));
    double sum = 0.0;
    for (uint32_t j = 0; j < n_new; ++j) sum += (w_out[j] = 1.0 / w_out[j]);
    for (uint32_t j = 0; j < n_new; ++j) w_out[j] /= sum;
    return PCS_OK;
}

int pcs_batch_pmc_weights(pcs_batch* b, const double* tau, double* w_out)
{
    if (!b || !tau || !w_out) return set_error(PCS_EINVAL, "NULL argument");
    PCS_TRY(pcs_batch_pmc_weights_async(b, tau));
    return pcs_batch_pmc_weights_wait(b, w_out);
}

int pcs_batch_read_proposals(pcs_batch* b, pcs_proposal* props_out, uint32_t* particles_out)
{
    if (!b) return set_error(PCS_EINVAL, "batch is NULL");
    if (!b->ready) return set_error(PCS_ESTATE, "batch holds no simulations (the last load or propose failed)");
    PCS_HIP(hipSetDevice(b->ctx->device));
    PCS_HIP(hipStreamSynchronize(b->stream));
    if (b->dev_planned) {
        if (props_out && b->n_props)
            PCS_HIP(hipMemcpyAsync(props_out, b->d_in, (size_t)b->n_props * sizeof(pcs_proposal), hipMemcpyDeviceToHost,
                                   b->stream));
        if (particles_out && b->n_videos)
            PCS_HIP(hipMemcpyAsync(particles_out, b->d_vparticles, (size_t)b->n_videos * 4u, hipMemcpyDeviceToHost,
                                   b->stream));
        PCS_HIP(hipStreamSynchronize(b->stream));
        return PCS_OK;
    }
    if (props_out) std::copy(b->props.begin(), b->props.end(), props_out);
    if (particles_out) std::copy(b->vparticles.begin(), b->vparticles.end(), particles_out);
    return PCS_OK;
}

int pcs_batch_read_exact(pcs_batch* b, uint64_t* exact_out)
{
    if (!b || !exact_out) return set_error(PCS_EINVAL, "NULL argument");
    if (!b->ran) return set_error(PCS_ESTATE, "batch has not run");
    PCS_HIP(hipSetDevice(b->ctx->device));
    PCS_HIP(hipStreamSynchronize(b->stream));
    if (b->n_props)
        PCS_HIP(hipMemcpyAsync(exact_out, b->d_exact, b->n_props * sizeof(uint64_t), hipMemcpyDeviceToHost, b->stream));
    PCS_HIP(hipStreamSynchronize(b->stream));
    return PCS_OK;
}

int pcs_set_record_capacity_factor(pcs_ctx* c, double factor)
{
    if (!c || !(factor > 0.0) || !std::isfinite(factor)) return set_error(PCS_EINVAL, "factor must be positive");
    c->cap_factor = factor;
    return PCS_OK;
}

int pcs_selftest_math(int device, const uint32_t* w, float* radius, float* s, float* c, uint64_t n)
{
    if (n && (!w || !radius || !s || !c)) return set_error(PCS_EINVAL, "NULL argument");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(PCS_ENODEV, "no GPU visible");
    if (device < 0 || device >= ndev) return set_error(PCS_EINVAL, "device %d outside [0, %d)", device, ndev);
    PCS_HIP(hipSetDevice(device));
    if (n == 0) return PCS_OK;
    uint32_t* dw = nullptr;
    float* dout = nullptr;
    PCS_TRY(dev_alloc(&dw, n));
    if (dev_alloc(&dout, 3 * n) != PCS_OK) {
        hipFree(dw);
        return PCS_ENOMEM;
    }
    hipError_t e = hipMemcpy(dw, w, n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_selftest_math(dw, dout, dout + n, dout + 2 * n, n, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(radius, dout, n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(s, dout + n, n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(c, dout + 2 * n, n * 4, hipMemcpyDeviceToHost);
    hipFree(dw);
    hipFree(dout);
    if (e != hipSuccess) return hip_error(e, "pcs_selftest_math");
    return PCS_OK;
}

int pcs_selftest_sqrt(int device, uint32_t lo_bits, uint32_t hi_bits, uint64_t* n_wrong)
{
    if (!n_wrong) return set_error(PCS_EINVAL, "NULL argument");
    if (lo_bits < 0x00800000u || hi_bits > 0x7f800000u || lo_bits > hi_bits)
        return set_error(PCS_EINVAL, "range [%#x, %#x) is not within the positive normal floats", lo_bits, hi_bits);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(PCS_ENODEV, "no GPU visible");
    if (device < 0 || device >= ndev) return set_error(PCS_EINVAL, "device %d outside [0, %d)", device, ndev);
    PCS_HIP(hipSetDevice(device));
    *n_wrong = 0;
    if (lo_bits == hi_bits) return PCS_OK;
    unsigned long long* dbad = nullptr;
    PCS_TRY(dev_alloc(&dbad, 1));
    unsigned long long h = 0;
    hipError_t e = hipMemset(dbad, 0, sizeof(h));
    if (e == hipSuccess) e = launch_selftest_sqrt(lo_bits, hi_bits - lo_bits, dbad, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(&h, dbad, sizeof(h), hipMemcpyDeviceToHost);
    hipFree(dbad);
    if (e != hipSuccess) return hip_error(e, "pcs_selftest_sqrt");
    *n_wrong = h;
    return PCS_OK;
}

int pcs_selftest_copy(int device, uint64_t bytes, float* ms)
{
    if (bytes == 0 || bytes % 16u) return set_error(PCS_EINVAL, "bytes must be a positive multiple of 16");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(PCS_ENODEV, "no GPU visible");
    if (device < 0 || device >= ndev) return set_error(PCS_EINVAL, "device %d outside [0, %d)", device, ndev);
    PCS_HIP(hipSetDevice(device));
    uint8_t *src = nullptr, *dst = nullptr;
    PCS_TRY(dev_alloc(&src, bytes));
    if (dev_alloc(&dst, bytes) != PCS_OK) {
        hipFree(src);
        return PCS_ENOMEM;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMemset(src, 1, bytes);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
    if (e == hipSuccess) e = launch_selftest_copy(src, dst, bytes, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (ms) *ms = t;
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    hipFree(src);
    hipFree(dst);
    if (e != hipSuccess) return hip_error(e, "pcs_selftest_copy");
    return PCS_OK;
}

int pcs_autotune(pcs_ctx* c, const pcs_proposal* props, uint32_t n, const uint32_t* vframes,
                 const uint32_t* vparticles, uint32_t n_videos, pcs_tuning* best_out, float* ms_out)
{
    if (!c) return set_error(PCS_EINVAL, "context is NULL");
    pcs_batch* b = nullptr;
    PCS_TRY(pcs_batch_create(c, &b));
    const pcs_tuning saved = c->tune;
    pcs_tuning best = saved;
    float best_ms = -1.0f;
    int r = pcs_batch_load(b, props, n, vframes, vparticles, n_videos);
    for (uint32_t ppl : {1u, 2u}) {
        bool warm = false;  // one untimed run per kernel (particles per lane)
        for (uint32_t wpg : {2u, 4u}) {
            for (uint32_t gpc : {2u, 4u, 6u, 8u, 12u, 16u}) {
                if (r != PCS_OK) break;
                if (wpg * gpc > 32u) continue;  // more waves than a CU holds (8 per SIMD)
                pcs_tuning t{ppl, wpg, gpc, saved.stage_records};
                c->tune = t;
                if (!warm) {
                    r = pcs_batch_run(b, 0x5eedull, PCS_MODE_BATCHED_DIRECT);
                    warm = true;
                }
                float t_ms = -1.0f;  // the faster of two timed runs
                for (int k = 1; k <= 2 && r == PCS_OK; ++k) {
                    r = pcs_batch_run(b, 0x5eedull + (uint64_t)k, PCS_MODE_BATCHED_DIRECT);
                    float tm[3];
                    if (r == PCS_OK) r = pcs_batch_timing(b, tm);
                    if (r == PCS_OK && (t_ms < 0.0f || tm[1] < t_ms)) t_ms = tm[1];
                }
                if (r != PCS_OK) break;
                if (best_ms < 0.0f || t_ms < best_ms) {
                    best_ms = t_ms;
                    best = t;
                }
            }
        }
    }
    pcs_batch_destroy(b);
    c->tune = (r == PCS_OK) ? best : saved;
    if (r != PCS_OK) return r;
    if (best_out) *best_out = best;
    if (ms_out) *ms_out = best_ms;
    return PCS_OK;
}

int pcs_generate_tracks(pcs_ctx* c, const pcs_proposal* prop, const uint32_t* vframes, const uint32_t* vparticles,
                        uint32_t n_videos, uint64_t seed, uint32_t* K, float* DE, uint64_t cap, uint64_t* n_out)
{
    if (!c || !prop || !n_out || (n_videos && (!vframes || !vparticles)) || (cap && (!K || !DE)))
        return set_error(PCS_EINVAL, "NULL argument");
    if (prop->reserved != 0) return set_error(PCS_EINVAL, "reserved field must be 0");
    PCS_TRY(validate_proposal(c->sys, *prop));
    PCS_HIP(hipSetDevice(c->device));
    // Exact bound: a track needs kmin detected frames and one undetected
    // frame (or the end), so a video holds at most (F+1)/(kmin+1) per particle.
    uint64_t words = 0;
    std::vector<uint32_t> vtile(n_videos + 1), vprop(n_videos, 0u);
    uint32_t tiles = 0;
    for (uint32_t v = 0; v < n_videos; ++v) {
        words += 2ull * vparticles[v] * ((vframes[v] + 1u) / (c->sys.kmin + 1u));
        vtile[v] = tiles;
        tiles += (vparticles[v] + 63u) / 64u;
    }
    vtile[n_videos] = tiles;
    if (words > 0xFFFFFFFFull) return set_error(PCS_EINVAL, "simulation too large for one track list");
    DevProp dp{};
    dp.sim_id = prop->sim_id;
    dp.first_video = 0;
    dp.n_videos = n_videos;
    dp.rec_cap = (uint32_t)words;
    for (int s = 0; s < PCS_MAX_SPECIES; ++s) {
        dp.half_az[s] = prop->half_az[s];
        dp.cum_prob[s] = prop->cum_prob[s];
        dp.sigma[s] = prop->sigma[s];
    }
    uint32_t *d_rec = nullptr, *d_v = nullptr, *d_count = nullptr;
    DevProp* d_prop = nullptr;
    uint64_t* d_seed = nullptr;
    int r = dev_alloc(&d_rec, words);
    if (r == PCS_OK) r = dev_alloc(&d_v, 4ull * n_videos + 1);
    if (r == PCS_OK) r = dev_alloc(&d_count, 1);
    if (r == PCS_OK) r = dev_alloc(&d_prop, 1);
    if (r == PCS_OK) r = dev_alloc(&d_seed, 1);
    uint32_t count = 0;
    if (r == PCS_OK) {
        dp.rec = d_rec;
        dp.srt = nullptr;
        hipStream_t st = c->stream;
        hipError_t e = hipMemcpyAsync(d_prop, &dp, sizeof(dp), hipMemcpyHostToDevice, st);
        if (e == hipSuccess && n_videos) e = hipMemcpyAsync(d_v, vframes, n_videos * 4, hipMemcpyHostToDevice, st);
        if (e == hipSuccess && n_videos)
            e = hipMemcpyAsync(d_v + n_videos, vparticles, n_videos * 4, hipMemcpyHostToDevice, st);
        if (e == hipSuccess && n_videos)
            e = hipMemcpyAsync(d_v + 2 * n_videos, vprop.data(), n_videos * 4, hipMemcpyHostToDevice, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_v + 3 * n_videos, vtile.data(), (n_videos + 1) * 4, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(d_seed, &seed, 8, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = launch_reset(d_count, 1, nullptr, st);
        if (e == hipSuccess) {
            WalkArgs a{};
            a.sys = c->d_sys;
            a.props = d_prop;
            a.vframes = d_v;
            a.vparticles = d_v + n_videos;
            a.vprop = d_v + 2 * n_videos;
            a.vtile = d_v + 3 * n_videos;
            a.n_videos = n_videos;
            a.tile_lo = 0;
            a.tile_hi = tiles;
            a.stage_cap = 256;
            a.rec_count = d_count;
            a.seed = d_seed;
            e = launch_tracks(a, c->sys.detect_model == PCS_DETECT_GAUSS_Z, c->sys.z_boundary == PCS_Z_REFLECT, st);
        }
        if (e == hipSuccess) e = hipMemcpyAsync(&count, d_count, 4, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        std::vector<uint32_t> words_h;
        if (e == hipSuccess && count) {
            words_h.resize(count);
            e = hipMemcpy(words_h.data(), d_rec, (size_t)count * 4, hipMemcpyDeviceToHost);
        }
        if (e != hipSuccess) r = hip_error(e, "pcs_generate_tracks");
        const uint64_t n = count / 2u;
        for (uint64_t i = 0; i < n && i < cap && r == PCS_OK; ++i) {
            K[i] = words_h[2 * i];
            std::memcpy(&DE[i], &words_h[2 * i + 1], 4);
        }
        *n_out = n;
    }
    void* ptrs[] = {d_rec, d_v, d_count, d_prop, d_seed};
    for (void* p : ptrs)
        if (p) hipFree(p);
    return r;
}

int pcs_pmc_weights(int device, const double* theta_new, uint32_t n_new, const double* theta_old, const double* w_old,
                    uint32_t n_old, const double* tau, uint32_t n_params, double* w_out)
{
    if (device < 0 || device >= kMaxDevices) return set_error(PCS_EINVAL, "device %d outside [0, %d)", device, kMaxDevices);
    if (n_params < 1 || n_params > 3 * PCS_MAX_SPECIES) return set_error(PCS_EINVAL, "n_params outside [1, 24]");
    if ((n_new && (!theta_new || !w_out)) || (n_old && (!theta_old || !w_old)) || !tau)
        return set_error(PCS_EINVAL, "NULL argument");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(PCS_ENODEV, "no GPU visible");
    if (device < 0 || device >= ndev) return set_error(PCS_EINVAL, "device %d outside [0, %d)", device, ndev);
    PCS_HIP(hipSetDevice(device));
    if (n_new == 0) return PCS_OK;
    // scale by 1/tau on the host so the kernel works in standard units
    std::vector<double> xn((size_t)n_params * n_new), xo((size_t)n_params * n_old);
    for (uint32_t p = 0; p < n_params; ++p) {
        for (uint32_t j = 0; j < n_new; ++j) xn[(size_t)p * n_new + j] = theta_new[(size_t)p * n_new + j] / tau[p];
        for (uint32_t i = 0; i < n_old; ++i) xo[(size_t)p * n_old + i] = theta_old[(size_t)p * n_old + i] / tau[p];
    }
    // Device workspace of this GPU, kept between calls (a hipFree would
    // synchronise the whole device every generation); calls on one GPU take
    // it in turn.
    PmcWork& wk = g_pmc[device];
    std::lock_guard<std::mutex> lk(wk.mu);
    int r = PCS_OK;
    auto grow = [&r](double** p, size_t& cap, size_t n) {
        if (r == PCS_OK && n > cap) {
            r = dev_alloc(p, n);
            cap = r == PCS_OK ? n : 0;
        }
    };
    grow(&wk.xn, wk.cap[0], xn.size());
    grow(&wk.xo, wk.cap[1], xo.size() + 1);
    grow(&wk.w, wk.cap[2], (size_t)n_old + 1);
    grow(&wk.s, wk.cap[3], n_new);
    grow(&wk.part, wk.cap[4], (size_t)pmc_chunks(n_old) * n_new);
    if (r == PCS_OK && !wk.stream) {
        const hipError_t e = hipStreamCreateWithFlags(&wk.stream, hipStreamNonBlocking);
        if (e != hipSuccess) r = hip_error(e, "pcs_pmc_weights stream");
    }
    if (r == PCS_OK) {
        hipStream_t st = wk.stream;
        hipError_t e = hipMemcpyAsync(wk.xn, xn.data(), xn.size() * 8, hipMemcpyHostToDevice, st);
        if (e == hipSuccess && n_old) e = hipMemcpyAsync(wk.xo, xo.data(), xo.size() * 8, hipMemcpyHostToDevice, st);
        if (e == hipSuccess && n_old) e = hipMemcpyAsync(wk.w, w_old, (size_t)n_old * 8, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = launch_pmc_sums(wk.xn, n_new, wk.xo, wk.w, n_old, n_params, wk.part, wk.s, st);
        if (e == hipSuccess) e = hipMemcpyAsync(w_out, wk.s, (size_t)n_new * 8, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) r = hip_error(e, "pcs_pmc_weights");
    }
    if (r != PCS_OK) return r;
    double sum = 0.0;
    for (uint32_t j = 0; j < n_new; ++j) sum += (w_out[j] = 1.0 / w_out[j]);
    for (uint32_t j = 0; j < n_new; ++j) w_out[j] /= sum;
    return PCS_OK;
}

int pcs_simulate_static(pcs_ctx* c, const pcs_proposal* props, uint32_t n, const uint32_t* vframes,
                        const uint32_t* vparticles, uint32_t n_videos, uint64_t seed, double* dist_out,
                        uint32_t* tracks_out, uint64_t* exact_out)
{
    if (!c) return set_error(PCS_EINVAL, "context is NULL");
    if (n && (!props || !vframes || !vparticles || !dist_out)) return set_error(PCS_EINVAL, "NULL argument");
    if (n > PCS_STATIC_MAX_SIMS)
        return set_error(PCS_EINVAL, "%u simulations, the static scratch takes at most %u", n, PCS_STATIC_MAX_SIMS);
    if (n_videos > PCS_STATIC_MAX_VIDEOS)
        return set_error(PCS_EINVAL, "%u videos, the static scratch takes at most %u", n_videos, PCS_STATIC_MAX_VIDEOS);
    const pcs_tuning t = c->tune;
    const uint32_t tp = 64u * t.particles_per_lane;
    uint64_t next = 0, tiles = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const pcs_proposal& p = props[i];
        if (p.reserved != 0) return set_error(PCS_EINVAL, "proposal %u: reserved field must be 0", i);
        if (p.first_video != next)
            return set_error(PCS_EINVAL, "proposal %u: first_video %u, expected %llu (videos follow the previous proposal's)",
                             i, p.first_video, (unsigned long long)next);
        next += p.n_videos;
        if (next > n_videos) return set_error(PCS_EINVAL, "proposal %u: videos beyond %u", i, n_videos);
        PCS_TRY(validate_proposal(c->sys, p));
    }
    if (next != n_videos) return set_error(PCS_EINVAL, "videos [%llu, %u) belong to no proposal", (unsigned long long)next, n_videos);
    for (uint32_t v = 0; v < n_videos; ++v) tiles += (vparticles[v] + tp - 1u) / tp;
    if (tiles > 0xffffffffull) return set_error(PCS_EINVAL, "more than 2^32 tiles");
    if (n == 0) return PCS_OK;
    PCS_HIP(hipSetDevice(c->device));
    unsigned char* base = scratch_ptr(c->device);
    if (!base) return set_error(PCS_ESTATE, "static scratch of device %d not set up", c->device);
    StaticLayout L;
    if (!static_layout(n, n_videos, L)) return set_error(PCS_EINVAL, "tables do not fit the static scratch");
    DevProp* d_props = reinterpret_cast<DevProp*>(base + L.props);
    uint32_t* d_vframes = reinterpret_cast<uint32_t*>(base + L.vframes);
    uint32_t* d_vparticles = reinterpret_cast<uint32_t*>(base + L.vparticles);
    uint32_t* d_vtile = reinterpret_cast<uint32_t*>(base + L.vtile);
    uint32_t* d_count = reinterpret_cast<uint32_t*>(base + L.count);
    uint32_t* d_tracks = reinterpret_cast<uint32_t*>(base + L.tracks);
    double* d_dist = reinterpret_cast<double*>(base + L.dist);
    uint64_t* d_exact = reinterpret_cast<uint64_t*>(base + L.exact);
    uint64_t* d_seed = reinterpret_cast<uint64_t*>(base + L.misc);
    uint32_t* d_tile_counter = reinterpret_cast<uint32_t*>(base + L.misc + 8u);
    uint64_t* d_totals = reinterpret_cast<uint64_t*>(base + L.misc + 16u);
    uint32_t* d_rec = reinterpret_cast<uint32_t*>(base + L.rec);
    uint32_t* d_srt = reinterpret_cast<uint32_t*>(base + L.srt);
    hipStream_t st = c->stream;
    const bool gauss = c->sys.detect_model == PCS_DETECT_GAUSS_Z;
    const bool zref = c->sys.z_boundary == PCS_Z_REFLECT;

    PCS_HIP(hipMemcpyAsync(base + L.in, props, (size_t)n * sizeof(pcs_proposal), hipMemcpyHostToDevice, st));
    PCS_HIP(hipMemcpyAsync(d_vframes, vframes, (size_t)n_videos * 4u, hipMemcpyHostToDevice, st));
    PCS_HIP(hipMemcpyAsync(d_vparticles, vparticles, (size_t)n_videos * 4u, hipMemcpyHostToDevice, st));
    PCS_HIP(hipMemcpyAsync(d_seed, &seed, sizeof(seed), hipMemcpyHostToDevice, st));
    PrepArgs pa{};
    pa.sys = c->d_sys;
    pa.in = reinterpret_cast<const pcs_proposal*>(base + L.in);
    pa.vframes = d_vframes;
    pa.vparticles = d_vparticles;
    pa.n = n;
    pa.n_videos = n_videos;
    pa.tile_particles = tp;
    pa.cap_factor = c->cap_factor;
    pa.props = d_props;
    pa.vprop = reinterpret_cast<uint32_t*>(base + L.vprop);
    pa.vtile = d_vtile;
    pa.rec = d_rec;
    pa.srt = d_srt;
    pa.arena_cap = L.arena_cap;
    pa.rec_count = d_count;
    pa.tile_counter = d_tile_counter;
    pa.totals = d_totals;
    PCS_HIP(launch_prepare(pa, st));
    WalkArgs a{};
    a.sys = c->d_sys;
    a.props = d_props;
    a.vframes = d_vframes;
    a.vparticles = d_vparticles;
    a.vprop = pa.vprop;
    a.vtile = d_vtile;
    a.n_videos = n_videos;
    a.tile_lo = 0;
    a.tile_hi = (uint32_t)tiles;
    a.stage_cap = t.stage_records;
    a.rec_count = d_count;
    a.seed = d_seed;
    a.tile_counter = d_tile_counter;
    PCS_HIP(launch_walk(a, (int)t.particles_per_lane, (int)t.waves_per_group, true, (uint32_t)c->n_cu * t.groups_per_cu,
                        gauss, zref, st));
    DistArgs d{};
    d.sys = c->d_sys;
    d.props = d_props;
    d.prop_lo = 0;
    d.n_props = n;
    d.rec_count = d_count;
    d.dist = d_dist;
    d.exact = d_exact;
    d.tracks = d_tracks;
    d.status = reinterpret_cast<int32_t*>(base + L.status);
    PCS_HIP(launch_distance(d, c->sys.n_k, c->sys.n_de, st));
    uint64_t totals[2] = {0, 0};
    PCS_HIP(hipMemcpyAsync(dist_out, d_dist, (size_t)n * 8u, hipMemcpyDeviceToHost, st));
    if (tracks_out) PCS_HIP(hipMemcpyAsync(tracks_out, d_tracks, (size_t)n * 4u, hipMemcpyDeviceToHost, st));
    if (exact_out) PCS_HIP(hipMemcpyAsync(exact_out, d_exact, (size_t)n * 8u, hipMemcpyDeviceToHost, st));
    PCS_HIP(hipMemcpyAsync(totals, d_totals, sizeof(totals), hipMemcpyDeviceToHost, st));
    PCS_HIP(hipStreamSynchronize(st));

    // Slices that overflowed (distance NaN) are rerun, one launch pair each,
    // with a slice of exactly the size the first pass counted, placed in the
    // arena past the planned slices.
    uint64_t used = std::min<uint64_t>(totals[0], L.arena_cap);
    for (uint32_t i = 0; i < n; ++i) {
        if (!(dist_out[i] != dist_out[i])) continue;
        uint32_t cnt = 0, vt[2] = {0, 0};
        DevProp dp;
        PCS_HIP(hipMemcpyAsync(&cnt, d_count + i, 4u, hipMemcpyDeviceToHost, st));
        PCS_HIP(hipMemcpyAsync(&dp, d_props + i, sizeof(DevProp), hipMemcpyDeviceToHost, st));
        PCS_HIP(hipMemcpyAsync(&vt[0], d_vtile + props[i].first_video, 4u, hipMemcpyDeviceToHost, st));
        PCS_HIP(hipMemcpyAsync(&vt[1], d_vtile + props[i].first_video + props[i].n_videos, 4u, hipMemcpyDeviceToHost, st));
        PCS_HIP(hipStreamSynchronize(st));
        if (used + cnt > L.arena_cap)
            return set_error(PCS_EOVERFLOW, "simulation %u: %u track records do not fit the static scratch", i, cnt);
        dp.rec = d_rec + used;
        dp.srt = d_srt + used;
        dp.rec_cap = cnt;
        used += cnt;
        const uint32_t zero = 0u;
        PCS_HIP(hipMemcpyAsync(d_props + i, &dp, sizeof(DevProp), hipMemcpyHostToDevice, st));
        PCS_HIP(hipMemcpyAsync(d_count + i, &zero, 4u, hipMemcpyHostToDevice, st));
        WalkArgs ar = a;
        ar.tile_lo = vt[0];
        ar.tile_hi = vt[1];
        PCS_HIP(launch_walk(ar, (int)t.particles_per_lane, (int)t.waves_per_group, false, 0, gauss, zref, st));
        DistArgs dr = d;
        dr.prop_lo = i;
        dr.n_props = 1;
        PCS_HIP(launch_distance(dr, c->sys.n_k, c->sys.n_de, st));
        PCS_HIP(hipMemcpyAsync(dist_out + i, d_dist + i, 8u, hipMemcpyDeviceToHost, st));
        if (tracks_out) PCS_HIP(hipMemcpyAsync(tracks_out + i, d_tracks + i, 4u, hipMemcpyDeviceToHost, st));
        if (exact_out) PCS_HIP(hipMemcpyAsync(exact_out + i, d_exact + i, 8u, hipMemcpyDeviceToHost, st));
        PCS_HIP(hipStreamSynchronize(st));
        if (dist_out[i] != dist_out[i])
            return set_error(PCS_EOVERFLOW, "simulation %u: rerun with an exact slice overflowed", i);
    }
    return PCS_OK;
}

int pcs_simulate(pcs_ctx* c, const pcs_proposal* props, uint32_t n, const uint32_t* vframes,
                 const uint32_t* vparticles, uint32_t n_videos, uint64_t seed, int mode, double* dist_out,
                 uint32_t* hist_out, uint32_t* tracks_out)
{
    if (!c) return set_error(PCS_EINVAL, "context is NULL");
    pcs_batch* b = nullptr;
    PCS_TRY(pcs_batch_create(c, &b));
    int r = pcs_batch_load(b, props, n, vframes, vparticles, n_videos);
    if (r == PCS_OK) r = pcs_batch_run(b, seed, mode);
    if (r == PCS_OK) r = pcs_batch_read(b, dist_out, hist_out, tracks_out);
    pcs_batch_destroy(b);
    return r;
}

}  // extern "C"
