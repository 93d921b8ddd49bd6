// pcs_server.hip -- serving form of the hot path: continuous batching of
// simulation requests from many concurrent ABC-PMC chains on one GPU.
//
// Replaces the reference's job queue, per-device worker thread, queue pool
// and stream callback (src/gpu/detail/simulation_impl.inl:631-730, 889-901,
// 1023-1136). Design:
//   - models (systems) and their observed summed-area tables are registered
//     once and stay resident in HBM;
//   - submit() copies a request into a bounded host queue and returns;
//   - one scheduler thread admits queued requests while the record arena (the
//     HBM footprint that grows with a simulation's size) has room and fewer
//     than max_inflight simulations are on the GPU, packs every request it
//     admitted in this pass into one launch group (walk + distance kernels),
//     and issues it on the next of several streams, so groups admitted while
//     others run overlap on the device;
//   - the distance kernel writes each result straight into pinned host
//     memory; the scheduler retires groups whose completion event fired,
//     frees their arena ranges and hands completions to poll().
// With PCS_SERVE_DEVICE_LOOP the launch groups are replaced by one persistent
// kernel (serve_loop_kernel) that takes requests from a ring in pinned host
// memory; the scheduler thread then only admits requests into the ring,
// collects results from the completion ring and relaunches the kernel when it
// has left for lack of work.
// The random stream is counter based, keyed per request, so a request's
// result is the same whatever it shares the GPU with.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "pcs_common.h"

using namespace pcs;

namespace {

constexpr int kStreams = 4;
constexpr uint32_t kMaxSystems = 64;

struct SysEntry {
    pcs_system sys;
    uint32_t* d_sat = nullptr;
};

struct Pending {
    pcs_request req;
    std::vector<uint32_t> frames, particles;
    double submit_us;
    uint32_t cap;
};

// First-fit allocator over the record arena (offsets in records).
class Arena {
public:
    void init(uint64_t n) { free_[0] = n; }
    bool alloc(uint64_t n, uint64_t& off)
    {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                off = it->first;
                const uint64_t rest = it->second - n;
                const uint64_t noff = it->first + n;
                free_.erase(it);
                if (rest) free_[noff] = rest;
                return true;
            }
        }
        return false;
    }
    void release(uint64_t off, uint64_t n)
    {
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

private:
    std::map<uint64_t, uint64_t> free_;
};

// Model classes that need their own walk kernel: bit 1 = Gaussian detection,
// bit 0 = reflecting z.
constexpr int kClasses = 4;

uint32_t model_class(const pcs_system& sy)
{
    return (sy.detect_model == PCS_DETECT_GAUSS_Z ? 2u : 0u) + (sy.z_boundary == PCS_Z_REFLECT ? 1u : 0u);
}

struct Group {
    bool busy = false;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    // device tables of this group; one tile table per model class, so the
    // walks of every class present are issued back to back
    DevProp* d_props = nullptr;
    uint32_t *d_vframes = nullptr, *d_vparticles = nullptr, *d_vprop = nullptr;
    uint32_t* d_vtile[kClasses] = {};
    uint32_t* d_count = nullptr;
    // pinned host staging (uploads) and results (written by the distance kernel)
    DevProp* h_props = nullptr;
    uint32_t *h_vframes = nullptr, *h_vparticles = nullptr, *h_vprop = nullptr;
    uint32_t* h_vtile[kClasses] = {};
    double* h_dist = nullptr;
    uint64_t* h_exact = nullptr;
    uint32_t* h_tracks = nullptr;
    int32_t* h_status = nullptr;
    // bookkeeping: the requests of the group (kept for exact reruns of
    // overflowed record slices) and their arena ranges
    std::vector<Pending> reqs;
    std::vector<uint64_t> arena_off;
    uint32_t n = 0;
};

// Host side of the device-resident loop.
struct LoopHost {
    bool on = false;
    uint32_t R = 0, maxv = 0;
    LoopReq* h_ring = nullptr;      // pinned, coherent: requests (host writes)
    uint32_t* h_ring_vid = nullptr; // pinned, coherent: video arrays per slot
    LoopDone* h_done = nullptr;     // pinned, coherent: results (device writes)
    uint32_t* h_stop = nullptr;     // pinned, coherent: 1 = leave now
    LoopSlot* d_slots = nullptr;
    uint32_t* d_slot_vid = nullptr;
    LoopCtl* d_ctl = nullptr;
    uint32_t* d_cu_token = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    bool running = false, stopping = false;
    uint32_t run_nsys = 0;          // systems known to the running launch
    uint64_t next_r = 0;            // number of the next request to publish
    uint64_t idle_ticks = 0;
    uint64_t launches = 0;
    int n_cu = 0;
    struct Book {
        bool busy = false;
        uint64_t r = 0, off = 0;
#if PCS_TRACE
        double publish_us = 0;
#endif
        Pending req;  // kept for an exact rerun of an overflowed record slice
    };
    std::vector<Book> book;         // per slot
    std::vector<uint32_t> live;     // slots in flight
};

}  // namespace

struct pcs_server {
    int device = 0;
    pcs_server_config cfg{};
    std::vector<SysEntry> systems;
    DevSys* d_sys = nullptr;
    uint64_t* d_zero_seed = nullptr;
    uint32_t *d_rec = nullptr, *d_srt = nullptr;
    uint64_t arena_records = 0;
    Arena arena;
    Group groups[kStreams];
    int next_group = 0;
    LoopHost loop;
    uint32_t inflight = 0;
    uint32_t max_queue = 0;
    double cap_factor = 1.0;  // record slice estimate scale (pcs_server_set_record_capacity_factor)
    uint64_t reruns = 0;      // requests rerun after a record slice overflow
    uint64_t group_walked = 0;  // device loop: completed requests whose tiles were walked a tile per group

    std::mutex mu;
    std::condition_variable cv;
    std::deque<Pending> queue;
    std::deque<pcs_completion> done;
    uint32_t outstanding = 0;  // submitted, not yet polled
    bool stop = false;
    int fatal = PCS_OK;
    std::thread worker;
    std::chrono::steady_clock::time_point t0;

    double now_us() const
    {
        return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

template <typename T>
hipError_t dmalloc(T** p, size_t n) { return hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T)); }
template <typename T>
hipError_t hmalloc(T** p, size_t n) { return hipHostMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T), hipHostMallocDefault); }

hipError_t group_alloc(Group& g, uint32_t max_props, uint32_t max_videos)
{
    hipError_t e = hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&g.done, hipEventDisableTiming);
    if (e == hipSuccess) e = dmalloc(&g.d_props, max_props);
    if (e == hipSuccess) e = dmalloc(&g.d_vframes, max_videos);
    if (e == hipSuccess) e = dmalloc(&g.d_vparticles, max_videos);
    if (e == hipSuccess) e = dmalloc(&g.d_vprop, max_videos);
    for (int k = 0; k < kClasses && e == hipSuccess; ++k) e = dmalloc(&g.d_vtile[k], max_videos + 1);
    if (e == hipSuccess) e = dmalloc(&g.d_count, max_props);
    if (e == hipSuccess) e = hmalloc(&g.h_props, max_props);
    if (e == hipSuccess) e = hmalloc(&g.h_vframes, max_videos);
    if (e == hipSuccess) e = hmalloc(&g.h_vparticles, max_videos);
    if (e == hipSuccess) e = hmalloc(&g.h_vprop, max_videos);
    for (int k = 0; k < kClasses && e == hipSuccess; ++k) e = hmalloc(&g.h_vtile[k], max_videos + 1);
    if (e == hipSuccess) e = hmalloc(&g.h_dist, max_props);
    if (e == hipSuccess) e = hmalloc(&g.h_exact, max_props);
    if (e == hipSuccess) e = hmalloc(&g.h_tracks, max_props);
    if (e == hipSuccess) e = hmalloc(&g.h_status, max_props);
    return e;
}

void group_free(Group& g)
{
    if (g.stream) hipStreamSynchronize(g.stream);
    void* dev[] = {g.d_props, g.d_vframes, g.d_vparticles, g.d_vprop, g.d_count};
    for (void* p : dev)
        if (p) hipFree(p);
    void* host[] = {g.h_props, g.h_vframes, g.h_vparticles, g.h_vprop, g.h_dist, g.h_exact, g.h_tracks, g.h_status};
    for (void* p : host)
        if (p) hipHostFree(p);
    for (int k = 0; k < kClasses; ++k) {
        if (g.d_vtile[k]) hipFree(g.d_vtile[k]);
        if (g.h_vtile[k]) hipHostFree(g.h_vtile[k]);
    }
    if (g.done) hipEventDestroy(g.done);
    if (g.stream) hipStreamDestroy(g.stream);
    g = Group{};
}

// Device record of an admitted request: its videos start at first_video of
// the launch, its records at arena offset off; Philox key = request seed.
DevProp make_prop(const pcs_server* s, const Pending& p, uint32_t first_video, uint64_t off)
{
    DevProp d;
    std::memset(&d, 0, sizeof(d));
    d.sim_id = p.req.prop.sim_id;
    d.first_video = first_video;
    d.n_videos = p.req.n_videos;
    d.sys_id = p.req.sys_id;
    d.rec = s->d_rec + off;
    d.srt = s->d_srt + off;
    d.rec_cap = p.cap;
    d.key0 = (uint32_t)p.req.seed;
    d.key1 = (uint32_t)(p.req.seed >> 32);
    for (int c = 0; c < PCS_MAX_SPECIES; ++c) {
        d.half_az[c] = p.req.prop.half_az[c];
        d.cum_prob[c] = p.req.prop.cum_prob[c];
        d.sigma[c] = p.req.prop.sigma[c];
    }
    return d;
}

// Issue one launch group on g from the admitted requests (moved into g).
int launch_group(pcs_server* s, Group& g, std::vector<Pending>& batch, std::vector<uint64_t>& offs)
{
    const uint32_t n = (uint32_t)batch.size();
    uint32_t nv = 0;
    g.arena_off.assign(offs.begin(), offs.end());
    uint32_t n_de_max = 1, n_k_max = 1;
    uint32_t classes = 0;  // bit model_class() present
    uint32_t tiles[kClasses] = {0, 0, 0, 0};
    for (uint32_t i = 0; i < n; ++i) {
        const Pending& p = batch[i];
        const pcs_system& sy = s->systems[p.req.sys_id].sys;
        const uint32_t cls = model_class(sy);
        g.h_props[i] = make_prop(s, p, nv, offs[i]);
        for (uint32_t v = 0; v < p.req.n_videos; ++v, ++nv) {
            g.h_vframes[nv] = p.frames[v];
            g.h_vparticles[nv] = p.particles[v];
            g.h_vprop[nv] = i;
            // tile tables: the videos of other classes have no tiles in this one
            for (uint32_t k = 0; k < (uint32_t)kClasses; ++k) {
                g.h_vtile[k][nv] = tiles[k];
                if (k == cls) tiles[k] += (p.particles[v] + 63u) / 64u;
            }
        }
        n_de_max = std::max(n_de_max, sy.n_de);
        n_k_max = std::max(n_k_max, sy.n_k);
        classes |= 1u << cls;
    }
    g.reqs.clear();
    for (Pending& p : batch) g.reqs.push_back(std::move(p));
    g.n = n;
    hipStream_t st = g.stream;
    hipError_t e = hipMemcpyAsync(g.d_props, g.h_props, n * sizeof(DevProp), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(g.d_vframes, g.h_vframes, nv * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(g.d_vparticles, g.h_vparticles, nv * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(g.d_vprop, g.h_vprop, nv * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_reset(g.d_count, n, nullptr, st);
    // One walk launch per (detection, boundary) class present, each over its
    // own tile table.
    for (uint32_t cls = 0; cls < (uint32_t)kClasses && e == hipSuccess; ++cls) {
        if (!(classes & (1u << cls))) continue;
        g.h_vtile[cls][nv] = tiles[cls];
        e = hipMemcpyAsync(g.d_vtile[cls], g.h_vtile[cls], (nv + 1) * 4, hipMemcpyHostToDevice, st);
        if (e != hipSuccess) break;
        WalkArgs a{};
        a.sys = s->d_sys;
        a.props = g.d_props;
        a.vframes = g.d_vframes;
        a.vparticles = g.d_vparticles;
        a.vprop = g.d_vprop;
        a.vtile = g.d_vtile[cls];
        a.n_videos = nv;
        a.tile_lo = 0;
        a.tile_hi = tiles[cls];
        a.stage_cap = 256;
        a.rec_count = g.d_count;
        a.seed = s->d_zero_seed;
        a.tile_counter = nullptr;
        e = launch_walk(a, 1, 4, false, 0, (cls & 2u) != 0, (cls & 1u) != 0, st);
    }
    if (e == hipSuccess) {
        DistArgs d{};
        d.sys = s->d_sys;
        d.props = g.d_props;
        d.prop_lo = 0;
        d.n_props = n;
        d.rec_count = g.d_count;
        d.dist = g.h_dist;
        d.exact = g.h_exact;
        d.tracks = g.h_tracks;
        d.status = g.h_status;
        e = launch_distance(d, n_k_max, n_de_max, st);
    }
    if (e == hipSuccess) e = hipEventRecord(g.done, st);
    if (e != hipSuccess) return hip_error(e, "serving launch");
    g.busy = true;
    return PCS_OK;
}

// A request whose record slice overflowed (status PCS_EOVERFLOW, tracks =
// the exact number of records it produced) goes back to the front of the
// queue with a slice of exactly that size: the random stream is keyed by the
// request, so the rerun reproduces the same tracks. Returns false when the
// exact slice cannot fit the arena (the overflow is then reported).
bool requeue_overflow(pcs_server* s, Pending& p, uint32_t count, std::vector<Pending>& again)
{
    if (count == 0u || count > s->arena_records || count <= p.cap) return false;
    p.cap = count;
    again.push_back(std::move(p));
    return true;
}

void push_back_reruns(pcs_server* s, std::vector<Pending>& again)
{
    if (again.empty()) return;
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto it = again.rbegin(); it != again.rend(); ++it) s->queue.push_front(std::move(*it));
    s->reruns += again.size();
    again.clear();
    s->cv.notify_all();
}

void retire_group(pcs_server* s, Group& g)
{
    const double t = s->now_us();
    std::vector<pcs_completion> out;
    std::vector<Pending> again;
    out.reserve(g.n);
    for (uint32_t i = 0; i < g.n; ++i) {
        Pending& p = g.reqs[i];
        s->arena.release(g.arena_off[i], p.cap);
        if (g.h_status[i] == PCS_EOVERFLOW && requeue_overflow(s, p, g.h_tracks[i], again)) continue;
        pcs_completion c{};
        c.tag = p.req.tag;
        c.dist = g.h_dist[i];
        c.exact = g.h_exact[i];
        c.tracks = g.h_tracks[i];
        c.status = g.h_status[i];
        c.submit_us = p.submit_us;
        c.done_us = t;
        out.push_back(c);
    }
    s->inflight -= g.n;
    g.busy = false;
    g.n = 0;
    g.reqs.clear();
    push_back_reruns(s, again);
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto& c : out) s->done.push_back(c);
    s->cv.notify_all();
}

// A scheduler thread keeps polling (yielding its core to other threads) for
// this long after its last work before it sleeps on the condition variable:
// a request arriving meanwhile is published at once instead of after the
// thread's wake-up (~10 us).
constexpr auto kSchedSpin = std::chrono::microseconds(2000);

void scheduler(pcs_server* s)
{
    hipSetDevice(s->device);
    std::vector<Pending> batch;
    std::vector<uint64_t> offs;
    auto last_work = std::chrono::steady_clock::now();
    for (;;) {
        bool idle = true;
        // Retire finished groups.
        for (Group& g : s->groups) {
            if (!g.busy) continue;
            const hipError_t q = hipEventQuery(g.done);
            if (q == hipSuccess) {
                retire_group(s, g);
                idle = false;
            } else if (q != hipErrorNotReady) {
                std::lock_guard<std::mutex> lk(s->mu);
                s->fatal = hip_error(q, "serving group");
                s->stop = true;
                s->cv.notify_all();
                return;
            }
        }
        // Admit and launch into the next free group.
        Group* free_g = nullptr;
        for (int k = 0; k < kStreams; ++k) {
            Group& g = s->groups[(s->next_group + k) % kStreams];
            if (!g.busy) {
                free_g = &g;
                s->next_group = (s->next_group + k + 1) % kStreams;
                break;
            }
        }
        batch.clear();
        offs.clear();
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->stop) return;
            if (free_g) {
                uint32_t nv = 0;
                while (!s->queue.empty() && s->inflight + batch.size() < s->cfg.max_inflight) {
                    Pending& p = s->queue.front();
                    if (nv + p.req.n_videos > s->cfg.max_inflight * s->cfg.max_videos) break;
                    uint64_t off;
                    if (!s->arena.alloc(p.cap, off)) break;  // wait for memory to free up
                    offs.push_back(off);
                    nv += p.req.n_videos;
                    batch.push_back(std::move(p));
                    s->queue.pop_front();
                }
            }
        }
        if (!batch.empty()) {
            s->inflight += (uint32_t)batch.size();
            const int r = launch_group(s, *free_g, batch, offs);
            if (r != PCS_OK) {
                std::lock_guard<std::mutex> lk(s->mu);
                s->fatal = r;
                s->stop = true;
                s->cv.notify_all();
                return;
            }
            idle = false;
        }
        if (idle) {
            bool any_busy = false;
            for (Group& g : s->groups) any_busy |= g.busy;
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->stop) return;
            if (!any_busy && s->queue.empty() && std::chrono::steady_clock::now() - last_work > kSchedSpin)
                s->cv.wait_for(lk, std::chrono::milliseconds(2));
            else
                lk.unlock(), std::this_thread::yield();
        } else {
            last_work = std::chrono::steady_clock::now();
        }
    }
}

// ------------------------------------------------------ device-resident loop

uint32_t request_tiles(const Pending& p)
{
    uint64_t t = 0;
    for (uint32_t v = 0; v < p.req.n_videos; ++v) t += (p.particles[v] + 63u) / 64u;
    return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(t, 1), 0xffffffffu);
}

// Write request r into its ring slot and publish it (seq last).
void loop_publish(pcs_server* s, uint32_t slot, uint64_t r, const Pending& p, uint64_t off)
{
    LoopHost& L = s->loop;
    LoopReq& Q = L.h_ring[slot];
    uint32_t* vid = L.h_ring_vid + (uint64_t)slot * 3u * (L.maxv + 1u);
    uint32_t acc = 0;
    for (uint32_t v = 0; v < p.req.n_videos; ++v) {
        vid[3 * v] = p.frames[v];
        vid[3 * v + 1] = p.particles[v];
        vid[3 * v + 2] = acc;
        acc += (p.particles[v] + 63u) / 64u;
    }
    vid[3 * p.req.n_videos] = 0;
    vid[3 * p.req.n_videos + 1] = 0;
    vid[3 * p.req.n_videos + 2] = acc;
    Q.n_videos = p.req.n_videos;
    Q.tiles = std::max(acc, 1u);
    Q.prop = make_prop(s, p, 0, off);
    __atomic_store_n(&Q.seq, r + 1, __ATOMIC_RELEASE);
}

// Move finished requests to the completion queue; returns how many.
uint32_t loop_collect(pcs_server* s, std::vector<pcs_completion>& out)
{
    LoopHost& L = s->loop;
    std::vector<Pending> again;
    uint64_t n_group = 0;
    out.clear();
    for (size_t i = 0; i < L.live.size();) {
        const uint32_t slot = L.live[i];
        LoopHost::Book& b = L.book[slot];
        const LoopDone& d = L.h_done[slot];
        if (__atomic_load_n(&d.seq, __ATOMIC_ACQUIRE) != b.r + 1) {
            ++i;
            continue;
        }
        s->arena.release(b.off, b.req.cap);
        b.busy = false;
        L.live[i] = L.live.back();
        L.live.pop_back();
        if (d.status == PCS_EOVERFLOW && requeue_overflow(s, b.req, d.tracks, again)) continue;
        pcs_completion c{};
        c.tag = b.req.req.tag;
        c.dist = d.dist;
        c.exact = d.exact;
        c.tracks = d.tracks;
        c.status = d.status;
        c.submit_us = b.req.submit_us;
        c.done_us = s->now_us();
#if PCS_TRACE
        // host clock (us): submit, publish to the ring, result seen; device
        // wall clock (ticks of 10 ns): admitted, first claim, distance, its
        // row sweep, done; then whether the tiles were group-walked, and the
        // records
        std::fprintf(stderr, "pcs-trace %llu %.3f %.3f %.3f %llu %llu %llu %llu %llu %u %u\n",
                     (unsigned long long)c.tag, c.submit_us, b.publish_us, c.done_us, (unsigned long long)d.t_admit,
                     (unsigned long long)d.t_claim, (unsigned long long)d.t_dist, (unsigned long long)d.t_sweep,
                     (unsigned long long)d.t_done, d.group_walked, d.tracks);
#endif
        out.push_back(c);
        n_group += d.group_walked != 0u ? 1u : 0u;
    }
    const uint32_t n_again = (uint32_t)again.size();
    push_back_reruns(s, again);
    if (!out.empty()) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->group_walked += n_group;
        for (auto& c : out) s->done.push_back(c);
        s->cv.notify_all();
    }
    return (uint32_t)out.size() + n_again;
}

// Launch the loop kernel for the first nsys systems (LDS and distance width
// sized for the largest of them).
int loop_launch(pcs_server* s, uint32_t nsys)
{
    LoopHost& L = s->loop;
    uint32_t nk = 1, nde = 1;
    for (uint32_t i = 0; i < nsys; ++i) {
        nk = std::max(nk, s->systems[i].sys.n_k);
        nde = std::max(nde, s->systems[i].sys.n_de);
    }
    LoopArgs a{};
    a.sys = s->d_sys;
    a.slots = L.d_slots;
    a.slot_vid = L.d_slot_vid;
    a.ring = L.h_ring;
    a.ring_vid = L.h_ring_vid;
    a.done = L.h_done;
    a.stop = L.h_stop;
    a.ctl = L.d_ctl;
    a.cu_token = L.d_cu_token;
    a.R = L.R;
    a.maxv = L.maxv;
    a.stage_cap = 256;
    a.nk_lds = nk;
    a.nde_lds = nde;
    a.idle_ticks = L.idle_ticks;
    int gpc = 0;
    hipError_t e = serve_loop_fit(a, nde, &gpc);
    if (e != hipSuccess) return hip_error(e, "serving loop occupancy");
    if (gpc < 1) return set_error(PCS_EINVAL, "serving loop does not fit a CU (n_k %u, n_de %u)", nk, nde);
    // group 0 admits requests, the others walk them
    const uint32_t grid = std::max(2u, (uint32_t)L.n_cu * (uint32_t)std::min(gpc, 8));
    // Requests of up to one tile per CU, admitted while at most 12 earlier
    // requests are unfinished, are walked a tile per group: a quiet GPU's
    // polling groups (one per CU) claim them all at once. The exchange buffers
    // need their LDS room.
    a.coop_max_tiles = (!(s->cfg.flags & PCS_SERVE_LOOP_ONE_WAVE_TILES) && serve_loop_coop_fits(a))
                           ? std::min((uint32_t)L.n_cu, (grid - 1u) / 4u) : 0u;
    a.coop_max_open = 12u;
    e = hipMemsetAsync(&L.d_ctl->quit, 0, sizeof(uint32_t), L.stream);
    if (e == hipSuccess) e = launch_serve_loop(a, grid, nde, L.stream);
    if (e == hipSuccess) e = hipEventRecord(L.ev, L.stream);
    if (e != hipSuccess) return hip_error(e, "serving loop launch");
    L.running = true;
    L.run_nsys = nsys;
    ++L.launches;
    return PCS_OK;
}

void fail(pcs_server* s, int r)
{
    std::lock_guard<std::mutex> lk(s->mu);
    s->fatal = r;
    s->stop = true;
    s->cv.notify_all();
}

void scheduler_loop(pcs_server* s)
{
    hipSetDevice(s->device);
    LoopHost& L = s->loop;
    std::vector<pcs_completion> out;
    auto last_work = std::chrono::steady_clock::now();
    auto last_query = last_work;
    for (;;) {
        bool progress = loop_collect(s, out) > 0;
        // Has the kernel left (idle limit, or the stop flag of a restart)?
        // Asked every 100 us at most: the results above are read from host
        // memory, and a HIP call per pass would delay noticing them.
        const auto now = std::chrono::steady_clock::now();
        if (L.running && (L.stopping || now - last_query >= std::chrono::microseconds(100))) {
            last_query = now;
            const hipError_t q = hipEventQuery(L.ev);
            if (q == hipSuccess) {
                L.running = false;
                uint32_t err = 0;
                const hipError_t e = hipMemcpy(&err, &L.d_ctl->error, sizeof(err), hipMemcpyDeviceToHost);
                if (e != hipSuccess || err != 0) {
                    fail(s, e != hipSuccess ? hip_error(e, "serving loop")
                                            : set_error(PCS_ESTATE, "serving loop: a tile claim could not be resolved"));
                    return;
                }
            } else if (q != hipErrorNotReady) {
                fail(s, hip_error(q, "serving loop"));
                return;
            }
        }
        if (!L.running && L.stopping) {
            __atomic_store_n(L.h_stop, 0u, __ATOMIC_RELEASE);
            L.stopping = false;
        }
        // Admit queued requests into free ring slots while the arena has room.
        bool restart = false;
        uint32_t nsys = 0;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (s->stop) return;
            nsys = (uint32_t)s->systems.size();
            while (!s->queue.empty() && L.live.size() < L.R) {
                Pending& p = s->queue.front();
                if (L.running && p.req.sys_id >= L.run_nsys) {
                    restart = true;  // the running launch does not know this model
                    break;
                }
                const uint32_t slot = (uint32_t)(L.next_r % L.R);
                if (L.book[slot].busy) break;  // request next_r - R still running
                uint64_t off;
                if (!s->arena.alloc(p.cap, off)) break;  // wait for memory to free up
                LoopHost::Book& b = L.book[slot];
                b.busy = true;
                b.r = L.next_r;
                b.off = off;
                loop_publish(s, slot, L.next_r, p, off);
#if PCS_TRACE
                b.publish_us = s->now_us();
#endif
                b.req = std::move(p);
                L.live.push_back(slot);
                ++L.next_r;
                s->queue.pop_front();
                progress = true;
            }
        }
        if (restart && L.running && !L.stopping) {
            __atomic_store_n(L.h_stop, 1u, __ATOMIC_RELEASE);
            L.stopping = true;
        }
        if (!L.running && !L.stopping && !L.live.empty()) {
            const int r = loop_launch(s, nsys);
            if (r != PCS_OK) {
                fail(s, r);
                return;
            }
            progress = true;
        }
        if (!progress) {
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->stop) return;
            if (L.live.empty() && s->queue.empty() && std::chrono::steady_clock::now() - last_work > kSchedSpin)
                s->cv.wait_for(lk, std::chrono::milliseconds(2));
            else
                lk.unlock(), std::this_thread::yield();
        } else {
            last_work = std::chrono::steady_clock::now();
        }
    }
}

hipError_t loop_alloc(pcs_server* s)
{
    LoopHost& L = s->loop;
    L.on = true;
    L.R = s->cfg.max_inflight;
    L.maxv = s->cfg.max_videos;
    L.book.assign(L.R, LoopHost::Book{});
    L.live.reserve(L.R);
    const size_t vid = (size_t)L.R * 3u * (L.maxv + 1u);
    hipError_t e = hipHostMalloc((void**)&L.h_ring, L.R * sizeof(LoopReq), hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostMalloc((void**)&L.h_ring_vid, vid * 4, hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostMalloc((void**)&L.h_done, L.R * sizeof(LoopDone), hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostMalloc((void**)&L.h_stop, 64, hipHostMallocCoherent);
    if (e == hipSuccess) e = dmalloc(&L.d_slots, L.R);
    if (e == hipSuccess) e = dmalloc(&L.d_slot_vid, vid);
    if (e == hipSuccess) e = dmalloc(&L.d_ctl, 1);
    if (e == hipSuccess) e = dmalloc(&L.d_cu_token, kLoopCuSlots);
    if (e == hipSuccess) e = hipMemset(L.d_cu_token, 0, kLoopCuSlots * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(L.d_slots, 0, L.R * sizeof(LoopSlot));
    if (e == hipSuccess) e = hipMemset(L.d_ctl, 0, sizeof(LoopCtl));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&L.ev, hipEventDisableTiming);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&L.n_cu, hipDeviceAttributeMultiprocessorCount, s->device);
    int khz = 0;
    if (e == hipSuccess) e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, s->device);
    if (e != hipSuccess) return e;
    std::memset(L.h_ring, 0, L.R * sizeof(LoopReq));
    std::memset(L.h_ring_vid, 0, vid * 4);
    std::memset(L.h_done, 0, L.R * sizeof(LoopDone));
    std::memset(L.h_stop, 0, 64);
    const uint64_t us = s->cfg.loop_idle_us ? s->cfg.loop_idle_us : 20000u;
    L.idle_ticks = us * (uint64_t)(khz > 0 ? khz : 100000) / 1000u;
    return hipSuccess;
}

// Stop the kernel (it leaves after its current tiles) and free the loop.
void loop_free(pcs_server* s)
{
    LoopHost& L = s->loop;
    if (L.h_stop) __atomic_store_n(L.h_stop, 1u, __ATOMIC_RELEASE);
    if (L.stream) hipStreamSynchronize(L.stream);
    void* dev[] = {L.d_slots, L.d_slot_vid, L.d_ctl, L.d_cu_token};
    for (void* p : dev)
        if (p) hipFree(p);
    void* host[] = {L.h_ring, L.h_ring_vid, L.h_done, L.h_stop};
    for (void* p : host)
        if (p) hipHostFree(p);
    if (L.ev) hipEventDestroy(L.ev);
    if (L.stream) hipStreamDestroy(L.stream);
    L = LoopHost{};
}

}  // namespace

extern "C" {

int pcs_server_create(int device, const pcs_server_config* cfg, pcs_server** out)
{
    if (!out) return set_error(PCS_EINVAL, "out is NULL");
    *out = nullptr;
    if (cfg && (cfg->flags & ~(PCS_SERVE_DEVICE_LOOP | PCS_SERVE_LOOP_ONE_WAVE_TILES)))
        return set_error(PCS_EINVAL, "unknown server flags 0x%x", cfg->flags);
    if (cfg && cfg->max_videos > 4096) return set_error(PCS_EINVAL, "max_videos %u above 4096", cfg->max_videos);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(PCS_ENODEV, "no GPU visible");
    if (device < 0 || device >= ndev) return set_error(PCS_EINVAL, "device %d outside [0, %d)", device, ndev);
    pcs_server* s = new (std::nothrow) pcs_server();
    if (!s) return set_error(PCS_ENOMEM, "server");
    s->device = device;
    if (cfg) s->cfg = *cfg;
    if (!s->cfg.arena_bytes) s->cfg.arena_bytes = 1ull << 30;
    if (!s->cfg.max_inflight) s->cfg.max_inflight = 4096;
    if (!s->cfg.max_videos) s->cfg.max_videos = 64;
    s->max_queue = 4 * s->cfg.max_inflight;
    s->systems.reserve(kMaxSystems);  // the scheduler reads entries without the lock
    s->arena_records = s->cfg.arena_bytes / 8;  // records + sort scratch
    s->t0 = std::chrono::steady_clock::now();
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = dmalloc(&s->d_sys, kMaxSystems);
    if (e == hipSuccess) e = dmalloc(&s->d_zero_seed, 1);
    if (e == hipSuccess) e = hipMemset(s->d_zero_seed, 0, 8);
    if (e == hipSuccess) e = dmalloc(&s->d_rec, s->arena_records);
    if (e == hipSuccess) e = dmalloc(&s->d_srt, s->arena_records);
    const bool device_loop = (s->cfg.flags & PCS_SERVE_DEVICE_LOOP) != 0;
    if (device_loop) {
        if (e == hipSuccess) e = loop_alloc(s);
    } else {
        for (int k = 0; k < kStreams && e == hipSuccess; ++k)
            e = group_alloc(s->groups[k], s->cfg.max_inflight, s->cfg.max_inflight * s->cfg.max_videos);
    }
    if (e != hipSuccess) {
        const int r = hip_error(e, "pcs_server_create");
        pcs_server_destroy(s);
        return r;
    }
    s->arena.init(s->arena_records);
    s->worker = std::thread(device_loop ? scheduler_loop : scheduler, s);
    *out = s;
    return PCS_OK;
}

int pcs_server_add_system(pcs_server* s, const pcs_system* sys, const uint32_t* obs_hist, uint32_t* sys_id)
{
    if (!s || !sys_id) return set_error(PCS_EINVAL, "NULL argument");
    const int r = validate_system(sys);
    if (r != PCS_OK) return r;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->systems.size() >= kMaxSystems) return set_error(PCS_EINVAL, "more than %u systems", kMaxSystems);
    std::vector<uint32_t> sat;
    uint32_t maxrow = 0;
    build_sat(obs_hist, sys->n_k, sys->n_de, sat, maxrow);
    SysEntry ent;
    ent.sys = *sys;
    hipError_t e = hipSetDevice(s->device);
    if (e == hipSuccess) e = dmalloc(&ent.d_sat, sat.size());
    if (e == hipSuccess) e = hipMemcpy(ent.d_sat, sat.data(), sat.size() * 4, hipMemcpyHostToDevice);
    const DevSys d = make_devsys(*sys, maxrow, ent.d_sat);
    const uint32_t id = (uint32_t)s->systems.size();
    if (e == hipSuccess) e = hipMemcpy(s->d_sys + id, &d, sizeof(DevSys), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (ent.d_sat) hipFree(ent.d_sat);
        return hip_error(e, "pcs_server_add_system");
    }
    s->systems.push_back(ent);
    *sys_id = id;
    return PCS_OK;
}

int pcs_server_submit(pcs_server* s, const pcs_request* req)
{
    if (!s || !req) return set_error(PCS_EINVAL, "NULL argument");
    if (req->n_videos == 0 || req->n_videos > s->cfg.max_videos)
        return set_error(PCS_EINVAL, "n_videos %u outside [1, %u]", req->n_videos, s->cfg.max_videos);
    if (!req->frames || !req->particles) return set_error(PCS_EINVAL, "NULL video arrays");
    for (uint32_t v = 0; v < req->n_videos; ++v)
        if (req->particles[v] > (1u << 31)) return set_error(PCS_EINVAL, "video %u: more than 2^31 particles", v);
    Pending p;
    p.req = *req;
    p.frames.assign(req->frames, req->frames + req->n_videos);
    p.particles.assign(req->particles, req->particles + req->n_videos);
    p.req.frames = nullptr;
    p.req.particles = nullptr;
    p.req.prop.first_video = 0;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->fatal != PCS_OK) return set_error(s->fatal, "server stopped after an error");
    if (req->sys_id >= s->systems.size()) return set_error(PCS_EINVAL, "unknown sys_id %u", req->sys_id);
    const pcs_system& sy = s->systems[req->sys_id].sys;
    if (req->prop.reserved != 0) return set_error(PCS_EINVAL, "reserved field must be 0");
    int r = validate_proposal(sy, req->prop);
    if (r != PCS_OK) return r;
    p.cap = record_capacity(sy, req->prop, p.frames.data(), p.particles.data(), s->cap_factor);
    if (s->loop.on && request_tiles(p) > kLoopMaxTiles)
        return set_error(PCS_EINVAL, "request needs %u tiles of 64 particles, the loop takes %u", request_tiles(p),
                         kLoopMaxTiles);
    if (p.cap > s->arena_records)
        return set_error(PCS_EINVAL, "request needs %u records, arena holds %llu", p.cap,
                         (unsigned long long)s->arena_records);
    if (s->queue.size() >= s->max_queue) return PCS_EAGAIN;
    p.submit_us = s->now_us();
    s->queue.push_back(std::move(p));
    ++s->outstanding;
    s->cv.notify_all();
    return PCS_OK;
}

int pcs_server_poll(pcs_server* s, pcs_completion* out, uint32_t max, uint32_t* n_out, uint32_t wait_us)
{
    if (!s || !n_out || (max && !out)) return set_error(PCS_EINVAL, "NULL argument");
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->done.empty() && wait_us && s->fatal == PCS_OK)
        s->cv.wait_for(lk, std::chrono::microseconds(wait_us), [s] { return !s->done.empty() || s->stop; });
    uint32_t n = 0;
    while (n < max && !s->done.empty()) {
        out[n++] = s->done.front();
        s->done.pop_front();
    }
    s->outstanding -= n;
    *n_out = n;
    if (n == 0 && s->fatal != PCS_OK) return set_error(s->fatal, "server stopped after an error");
    return PCS_OK;
}

int pcs_server_pending(pcs_server* s, uint32_t* n)
{
    if (!s || !n) return set_error(PCS_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *n = s->outstanding;
    return PCS_OK;
}

int pcs_server_set_record_capacity_factor(pcs_server* s, double factor)
{
    if (!s || !(factor > 0.0) || !std::isfinite(factor)) return set_error(PCS_EINVAL, "factor must be positive");
    std::lock_guard<std::mutex> lk(s->mu);
    s->cap_factor = factor;
    return PCS_OK;
}

int pcs_server_overflow_reruns(pcs_server* s, uint64_t* n)
{
    if (!s || !n) return set_error(PCS_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *n = s->reruns;
    return PCS_OK;
}

int pcs_server_group_walked(pcs_server* s, uint64_t* n)
{
    if (!s || !n) return set_error(PCS_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *n = s->group_walked;
    return PCS_OK;
}

int pcs_server_now_us(pcs_server* s, double* us)
{
    if (!s || !us) return set_error(PCS_EINVAL, "NULL argument");
    *us = s->now_us();
    return PCS_OK;
}

int pcs_server_destroy(pcs_server* s)
{
    if (!s) return PCS_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    hipSetDevice(s->device);
    loop_free(s);
    for (Group& g : s->groups) group_free(g);
    for (SysEntry& e : s->systems)
        if (e.d_sat) hipFree(e.d_sat);
    void* ptrs[] = {s->d_sys, s->d_zero_seed, s->d_rec, s->d_srt};
    for (void* p : ptrs)
        if (p) hipFree(p);
    delete s;
    return PCS_OK;
}

}  // extern "C"
