// Serving engine: random-access block decode from an archive resident in HBM.
//
// The reference decodes a container's blocks with one rayon task per block
// (crates/fqxv/src/container/decompress.rs:45-52); a serving process instead
// sees a stream of requests, each a set of block ids of one resident archive.
// This engine admits requests in arrival order against a byte budget for
// decoded output, packs every request that fits into one batched
// (persistent-kernel) decode, and keeps up to `slots` such launches in flight
// at once on separate streams, so requests that arrive while others run are
// admitted as soon as a slot and output memory free up (continuous batching).
//
// A second mode (fqxv_rans_server_create_persistent) removes launch
// boundaries altogether: one persistent grid (k_rans_serve) stays resident and
// takes blocks from a queue the host appends to, so a request's blocks start
// as soon as a wave is free, whatever else is running; requests are admitted
// against the same output budget and complete when their last block does.
//
// Output of a request: its blocks in request order, each at a 256-byte aligned
// offset inside one region of the caller's output pool (block i at out_off + sum over earlier blocks
// of round_up(n, 256)).  Bytes are those of a plain decode of each block.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <cstdint>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <unordered_map>
#include <vector>

#include "../../include/fqxv_rans_hip.h"
#include "rans_serve.h"

hipError_t fqxv_launch_serve(const uint8_t *, uint8_t *, const fqxv_rans_job *, fqxv_serve_res *, uint32_t,
                             const fqxv_serve_ctl *, fqxv_serve_dev *, uint8_t *, uint64_t, uint32_t, hipStream_t);
hipError_t fqxv_decode_occupancy(int *);
size_t fqxv_ws_per_wave();

typedef uint8_t u8;
typedef uint32_t u32;
typedef uint64_t u64;
typedef int32_t i32;

namespace {

constexpr u64 ALIGN = 256;

u64 round_up(u64 v) { return (v + ALIGN - 1) & ~(ALIGN - 1); }

// First-fit allocator over [0, size) with coalescing frees.
class Pool {
  public:
    explicit Pool(u64 size) { free_[0] = size; }
    bool alloc(u64 n, u64 *off)
    {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            *off = it->first;
            const u64 rest = it->second - n, at = it->first + n;
            free_.erase(it);
            if (rest) free_[at] = rest;
            return true;
        }
        return false;
    }
    void release(u64 off, u64 n)
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
    std::map<u64, u64> free_;
};

struct Req {
    u64 ticket = 0;
    std::vector<u32> ids;
    u64 bytes = 0;
    u64 off = 0;
    double t_arrival = 0, t_admit = 0;
    u32 first = 0; // index of its first job in the slot's batch
};

struct Slot {
    fqxv_rans_ctx *ctx = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    fqxv_rans_job *h_jobs = nullptr;
    u32 *h_order = nullptr;
    i32 *h_status = nullptr;
    fqxv_rans_job *d_jobs = nullptr;
    u32 *d_order = nullptr;
    i32 *d_status = nullptr;
    u64 *d_out_len = nullptr;
    std::vector<Req> reqs;
    bool busy = false;
};

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard()
    {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

} // namespace

// Persistent mode: a published job's owner.
struct JobMeta {
    u64 ticket;
    u32 pos;  // block position within the request
};

// Persistent mode: an admitted request waiting for its blocks.
struct Active {
    Req req;
    u32 remaining = 0;
    std::vector<i32> status;
};

struct fqxv_rans_server {
    int device = 0;
    bool persistent = false;
    const u8 *d_archive = nullptr;
    std::vector<fqxv_rans_job> blocks;
    u8 *d_pool = nullptr;
    u64 budget = 0;
    Pool *pool = nullptr;
    u32 max_batch = 0;
    std::deque<Req> queue;
    std::vector<Slot> slots;
    std::deque<fqxv_rans_done> done;
    std::unordered_map<u64, std::pair<u64, u64>> held; // completed, not yet released
    u64 next_ticket = 1;
    // Persistent mode (queue layout in rans_serve.h).
    u32 ring = 0, waves = 0;
    u64 idle_ticks = 0;
    fqxv_serve_ctl *h_ctl = nullptr;       // pinned, coherent
    fqxv_rans_job *h_jobs = nullptr;       // pinned, coherent, `ring` records
    fqxv_serve_res *h_res = nullptr;       // pinned, coherent, `ring` records
    fqxv_serve_dev *h_dev = nullptr;       // pinned staging for the device block
    fqxv_rans_job *dp_jobs = nullptr;      // device views of the three above
    fqxv_serve_res *dp_res = nullptr;
    fqxv_serve_ctl *dp_ctl = nullptr;
    fqxv_serve_dev *d_dev = nullptr;       // device memory
    u8 *d_ws = nullptr;
    hipStream_t pstream = nullptr;
    hipEvent_t pdone = nullptr;
    bool running = false;
    u64 pub = 0, oldest = 0, head = 0;     // published, oldest unreaped, claimed at last exit
    std::vector<JobMeta> meta;             // per ring slot
    std::vector<char> reaped;              // per ring slot
    std::unordered_map<u64, Active> active;
    std::chrono::steady_clock::time_point t0;
    std::mutex mu;

    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

void free_slot(int device, Slot &s)
{
    DeviceGuard dg(device);
    if (s.stream) (void)hipStreamSynchronize(s.stream);
    if (s.done) (void)hipEventDestroy(s.done);
    if (s.h_jobs) (void)hipHostFree(s.h_jobs);
    if (s.h_order) (void)hipHostFree(s.h_order);
    if (s.h_status) (void)hipHostFree(s.h_status);
    if (s.d_jobs) (void)hipFree(s.d_jobs);
    if (s.d_order) (void)hipFree(s.d_order);
    if (s.d_status) (void)hipFree(s.d_status);
    if (s.d_out_len) (void)hipFree(s.d_out_len);
    if (s.stream) (void)hipStreamDestroy(s.stream);
    if (s.ctx) fqxv_rans_ctx_destroy(s.ctx);
    s = Slot();
}

// Completes finished launches into `done`.  Returns a FQXV_RANS_* code.
int reap(fqxv_rans_server *s)
{
    for (Slot &sl : s->slots) {
        if (!sl.busy) continue;
        const hipError_t q = hipEventQuery(sl.done);
        if (q == hipErrorNotReady) continue;
        if (q != hipSuccess) return FQXV_RANS_ERR_DEVICE;
        const double t = s->now();
        for (const Req &r : sl.reqs) {
            i32 st = FQXV_RANS_OK;
            for (u32 k = 0; k < r.ids.size() && st == FQXV_RANS_OK; ++k) st = sl.h_status[r.first + k];
            fqxv_rans_done d;
            d.ticket = r.ticket;
            d.out_off = r.off;
            d.out_bytes = r.bytes;
            d.t_arrival = r.t_arrival;
            d.t_admit = r.t_admit;
            d.t_done = t;
            d.status = st;
            d.nblocks = (u32)r.ids.size();
            s->held[r.ticket] = {r.off, r.bytes};
            s->done.push_back(d);
        }
        sl.reqs.clear();
        sl.busy = false;
    }
    return FQXV_RANS_OK;
}

// Packs queued requests (arrival order, head-of-line) into free slots.
int admit(fqxv_rans_server *s)
{
    for (Slot &sl : s->slots) {
        if (sl.busy) continue;
        if (s->queue.empty()) break;
        u32 nj = 0;
        const double t = s->now();
        while (!s->queue.empty()) {
            Req &r = s->queue.front();
            if (nj + r.ids.size() > s->max_batch) break;
            u64 off = 0;
            if (!s->pool->alloc(r.bytes, &off)) break;
            r.off = off;
            r.t_admit = t;
            r.first = nj;
            u64 o = off;
            for (u32 id : r.ids) {
                fqxv_rans_job j = s->blocks[id];
                j.dst_off = o;
                sl.h_jobs[nj++] = j;
                o += round_up(j.dst_cap);
            }
            sl.reqs.push_back(std::move(r));
            s->queue.pop_front();
        }
        if (nj == 0) break; // head request waits for output memory
        for (u32 i = 0; i < nj; ++i) sl.h_order[i] = i;
        std::stable_sort(sl.h_order, sl.h_order + nj,
                         [&](u32 a, u32 b) { return sl.h_jobs[a].dst_cap > sl.h_jobs[b].dst_cap; });
        const hipError_t e1 =
            hipMemcpyAsync(sl.d_jobs, sl.h_jobs, nj * sizeof(fqxv_rans_job), hipMemcpyHostToDevice, sl.stream);
        const hipError_t e2 = hipMemcpyAsync(sl.d_order, sl.h_order, nj * sizeof(u32), hipMemcpyHostToDevice, sl.stream);
        if (e1 != hipSuccess || e2 != hipSuccess) return FQXV_RANS_ERR_DEVICE;
        const int rc = fqxv_rans_decode_batch(sl.ctx, s->d_archive, s->d_pool, sl.d_jobs, nj, sl.d_order, sl.d_status,
                                              sl.d_out_len, sl.stream);
        if (rc) return rc;
        if (hipMemcpyAsync(sl.h_status, sl.d_status, nj * sizeof(i32), hipMemcpyDeviceToHost, sl.stream) !=
                hipSuccess ||
            hipEventRecord(sl.done, sl.stream) != hipSuccess)
            return FQXV_RANS_ERR_DEVICE;
        sl.busy = true;
    }
    return FQXV_RANS_OK;
}

// Persistent mode: completes finished jobs (tag = index + 1) and requests
// whose last block finished.
void reap_p(fqxv_rans_server *s)
{
    const double t = s->now();
    for (u64 q = s->oldest; q < s->pub; q++) {
        const u32 k = (u32)(q % s->ring);
        if (s->reaped[k]) continue;
        const fqxv_serve_res &r = s->h_res[k];
        if (__atomic_load_n(&r.tag, __ATOMIC_ACQUIRE) != q + 1) continue;
        s->reaped[k] = 1;
        auto it = s->active.find(s->meta[k].ticket);
        Active &a = it->second;
        a.status[s->meta[k].pos] = __atomic_load_n(&r.status, __ATOMIC_RELAXED);
        if (--a.remaining == 0) {
            i32 st = FQXV_RANS_OK;
            for (i32 v : a.status)
                if (v != FQXV_RANS_OK) {
                    st = v;
                    break;
                }
            fqxv_rans_done d;
            d.ticket = a.req.ticket;
            d.out_off = a.req.off;
            d.out_bytes = a.req.bytes;
            d.t_arrival = a.req.t_arrival;
            d.t_admit = a.req.t_admit;
            d.t_done = t;
            d.status = st;
            d.nblocks = (u32)a.req.ids.size();
            s->held[a.req.ticket] = {a.req.off, a.req.bytes};
            s->done.push_back(d);
            s->active.erase(it);
        }
    }
    while (s->oldest < s->pub && s->reaped[s->oldest % s->ring]) {
        s->reaped[s->oldest % s->ring] = 0;
        s->oldest++;
    }
}

// Persistent mode: publishes queued requests (arrival order, head-of-line)
// while their output fits the budget and their blocks fit the ring; each
// request's blocks go out largest first.  Then makes sure a grid is running.
int admit_p(fqxv_rans_server *s)
{
    const u64 pub0 = s->pub;
    const double t = s->now();
    while (!s->queue.empty()) {
        Req &r = s->queue.front();
        const u32 k = (u32)r.ids.size();
        if (s->pub + k - s->oldest > s->ring) break;
        u64 off = 0;
        if (!s->pool->alloc(r.bytes, &off)) break;
        r.off = off;
        r.t_admit = t;
        std::vector<u64> at(k);
        u64 o = off;
        for (u32 i = 0; i < k; i++) {
            at[i] = o;
            o += round_up(s->blocks[r.ids[i]].dst_cap);
        }
        std::vector<u32> ord(k);
        for (u32 i = 0; i < k; i++) ord[i] = i;
        std::stable_sort(ord.begin(), ord.end(), [&](u32 a, u32 b) {
            return s->blocks[r.ids[a]].dst_cap > s->blocks[r.ids[b]].dst_cap;
        });
        for (u32 i : ord) {
            const u32 slot = (u32)(s->pub % s->ring);
            fqxv_rans_job j = s->blocks[r.ids[i]];
            j.dst_off = at[i];
            s->h_jobs[slot] = j;
            s->meta[slot] = JobMeta{r.ticket, i};
            s->pub++;
        }
        Active a;
        a.remaining = k;
        a.status.assign(k, FQXV_RANS_OK);
        a.req = std::move(r);
        s->active.emplace(a.req.ticket, std::move(a));
        s->queue.pop_front();
    }
    // Job records before the count that publishes them.
    if (s->pub != pub0) __atomic_store_n(&s->h_ctl->tail, s->pub, __ATOMIC_RELEASE);
    if (s->running) {
        const hipError_t q = hipEventQuery(s->pdone);
        if (q == hipErrorNotReady) return FQXV_RANS_OK;
        if (q != hipSuccess) return FQXV_RANS_ERR_DEVICE;
        s->running = false;
        // Every wave has left: the claimed count is final.
        if (hipMemcpy(&s->head, &s->d_dev->head, sizeof(u64), hipMemcpyDeviceToHost) != hipSuccess)
            return FQXV_RANS_ERR_DEVICE;
    }
    if (s->head >= s->pub) return FQXV_RANS_OK; // nothing unclaimed
    // (Re)launch: claims continue at `head`; the tail and stop are mirrored
    // by the dispatcher, seeded here.
    s->h_dev->head = s->head;
    s->h_dev->dtail = s->pub;
    s->h_dev->dstop = 0;
    s->h_dev->alive = s->waves;
    s->h_dev->disp_cu = 0;
    if (hipMemcpyAsync(s->d_dev, s->h_dev, sizeof(fqxv_serve_dev), hipMemcpyHostToDevice, s->pstream) != hipSuccess ||
        fqxv_launch_serve(s->d_archive, s->d_pool, s->dp_jobs, s->dp_res, s->ring, s->dp_ctl, s->d_dev, s->d_ws,
                          s->idle_ticks, s->waves, s->pstream) != hipSuccess ||
        hipEventRecord(s->pdone, s->pstream) != hipSuccess)
        return FQXV_RANS_ERR_DEVICE;
    s->running = true;
    return FQXV_RANS_OK;
}

void free_persistent(fqxv_rans_server *s)
{
    DeviceGuard dg(s->device);
    if (s->h_ctl) {
        // The dispatcher sees stop, decode waves finish their jobs and leave.
        __atomic_store_n(&s->h_ctl->stop, 1u, __ATOMIC_RELEASE);
        if (s->pstream) (void)hipStreamSynchronize(s->pstream);
    }
    if (s->pdone) (void)hipEventDestroy(s->pdone);
    if (s->pstream) (void)hipStreamDestroy(s->pstream);
    if (s->d_ws) (void)hipFree(s->d_ws);
    if (s->d_dev) (void)hipFree(s->d_dev);
    if (s->h_dev) (void)hipHostFree(s->h_dev);
    if (s->h_res) (void)hipHostFree(s->h_res);
    if (s->h_jobs) (void)hipHostFree(s->h_jobs);
    if (s->h_ctl) (void)hipHostFree(s->h_ctl);
}

} // namespace

extern "C" {

void fqxv_rans_server_destroy(fqxv_rans_server *s)
{
    if (!s) return;
    if (s->persistent) free_persistent(s);
    for (Slot &sl : s->slots) free_slot(s->device, sl);
    delete s->pool;
    delete s;
}

int fqxv_rans_server_create(int device, const uint8_t *d_archive, const fqxv_rans_job *blocks, uint32_t nblocks,
                            uint8_t *d_pool, uint64_t budget, uint32_t slots, uint32_t max_batch,
                            fqxv_rans_server **out)
{
    if (!out || !d_archive || !blocks || nblocks == 0 || !d_pool || budget < ALIGN || slots == 0 || slots > 16 ||
        max_batch == 0)
        return FQXV_RANS_ERR_INVALID_ARG;
    *out = nullptr;
    auto *s = new (std::nothrow) fqxv_rans_server();
    if (!s) return FQXV_RANS_ERR_DEVICE;
    s->device = device;
    s->d_archive = d_archive;
    s->blocks.assign(blocks, blocks + nblocks);
    for (const fqxv_rans_job &j : s->blocks)
        if (j.dst_cap > budget) {
            delete s;
            return FQXV_RANS_ERR_INVALID_ARG;
        }
    s->budget = budget & ~(ALIGN - 1);
    s->max_batch = max_batch;
    s->d_pool = d_pool;
    s->pool = new (std::nothrow) Pool(s->budget);
    s->t0 = std::chrono::steady_clock::now();
    DeviceGuard dg(device);
    bool ok = s->pool != nullptr;
    s->slots.resize(slots);
    for (Slot &sl : s->slots) {
        if (!ok) break;
        ok = fqxv_rans_ctx_create(device, &sl.ctx) == FQXV_RANS_OK && hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) == hipSuccess &&
             hipHostMalloc((void **)&sl.h_jobs, max_batch * sizeof(fqxv_rans_job), 0) == hipSuccess &&
             hipHostMalloc((void **)&sl.h_order, max_batch * sizeof(u32), 0) == hipSuccess &&
             hipHostMalloc((void **)&sl.h_status, max_batch * sizeof(i32), 0) == hipSuccess &&
             hipMalloc((void **)&sl.d_jobs, max_batch * sizeof(fqxv_rans_job)) == hipSuccess &&
             hipMalloc((void **)&sl.d_order, max_batch * sizeof(u32)) == hipSuccess &&
             hipMalloc((void **)&sl.d_status, max_batch * sizeof(i32)) == hipSuccess &&
             hipMalloc((void **)&sl.d_out_len, max_batch * sizeof(u64)) == hipSuccess;
    }
    if (!ok) {
        fqxv_rans_server_destroy(s);
        return FQXV_RANS_ERR_DEVICE;
    }
    *out = s;
    return FQXV_RANS_OK;
}

int fqxv_rans_server_create_persistent(int device, const uint8_t *d_archive, const fqxv_rans_job *blocks,
                                       uint32_t nblocks, uint8_t *d_pool, uint64_t budget, uint32_t ring,
                                       uint32_t waves, uint32_t idle_ms, fqxv_rans_server **out)
{
    if (!out || !d_archive || !blocks || nblocks == 0 || !d_pool || budget < ALIGN || ring == 0 || idle_ms == 0 ||
        idle_ms > 60000)
        return FQXV_RANS_ERR_INVALID_ARG;
    *out = nullptr;
    auto *s = new (std::nothrow) fqxv_rans_server();
    if (!s) return FQXV_RANS_ERR_DEVICE;
    s->device = device;
    s->persistent = true;
    s->d_archive = d_archive;
    s->blocks.assign(blocks, blocks + nblocks);
    for (const fqxv_rans_job &j : s->blocks)
        if (j.dst_cap > budget) {
            delete s;
            return FQXV_RANS_ERR_INVALID_ARG;
        }
    s->budget = budget & ~(ALIGN - 1);
    s->max_batch = ring; // a request's blocks must fit the ring at once
    s->ring = ring;
    s->idle_ticks = (u64)idle_ms * 100000; // 100 MHz real-time clock
    s->d_pool = d_pool;
    s->pool = new (std::nothrow) Pool(s->budget);
    s->t0 = std::chrono::steady_clock::now();
    s->meta.resize(ring);
    s->reaped.assign(ring, 0);
    DeviceGuard dg(device);
    hipDeviceProp_t prop;
    int occ = 0;
    bool ok = s->pool != nullptr && hipGetDeviceProperties(&prop, device) == hipSuccess &&
              fqxv_decode_occupancy(&occ) == hipSuccess;
    if (ok) {
        // A resident grid that filled every wave slot would starve the
        // kernels that consume its output (they could not be scheduled until
        // the server went idle): by default it takes half of the chip.
        const u32 slots_all = (u32)std::max(1, occ) * (u32)prop.multiProcessorCount;
        const u32 maxw = slots_all - 1; // one slot for the dispatcher
        s->waves = waves ? std::min(waves, maxw) : std::max(1u, slots_all / 2 - 1);
    }
    const unsigned pin = hipHostMallocMapped | hipHostMallocCoherent;
    ok = ok && hipHostMalloc((void **)&s->h_ctl, sizeof(fqxv_serve_ctl), pin) == hipSuccess &&
         hipHostMalloc((void **)&s->h_jobs, (size_t)ring * sizeof(fqxv_rans_job), pin) == hipSuccess &&
         hipHostMalloc((void **)&s->h_res, (size_t)ring * sizeof(fqxv_serve_res), pin) == hipSuccess &&
         hipHostMalloc((void **)&s->h_dev, sizeof(fqxv_serve_dev), 0) == hipSuccess &&
         hipHostGetDevicePointer((void **)&s->dp_ctl, s->h_ctl, 0) == hipSuccess &&
         hipHostGetDevicePointer((void **)&s->dp_jobs, s->h_jobs, 0) == hipSuccess &&
         hipHostGetDevicePointer((void **)&s->dp_res, s->h_res, 0) == hipSuccess &&
         hipMalloc((void **)&s->d_dev, sizeof(fqxv_serve_dev)) == hipSuccess &&
         hipMalloc((void **)&s->d_ws, (size_t)s->waves * fqxv_ws_per_wave()) == hipSuccess &&
         hipStreamCreateWithFlags(&s->pstream, hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&s->pdone, hipEventDisableTiming) == hipSuccess;
    if (ok) {
        memset(s->h_ctl, 0, sizeof(fqxv_serve_ctl));
        memset(s->h_res, 0, (size_t)ring * sizeof(fqxv_serve_res));
    }
    if (!ok) {
        fqxv_rans_server_destroy(s);
        return FQXV_RANS_ERR_DEVICE;
    }
    *out = s;
    return FQXV_RANS_OK;
}

int fqxv_rans_server_now(fqxv_rans_server *s, double *t)
{
    if (!s || !t) return FQXV_RANS_ERR_INVALID_ARG;
    *t = s->now();
    return FQXV_RANS_OK;
}

int fqxv_rans_server_submit(fqxv_rans_server *s, const uint32_t *ids, uint32_t k, double t_arrival,
                            uint64_t *ticket)
{
    if (!s || !ids || k == 0 || !ticket || k > s->max_batch) return FQXV_RANS_ERR_INVALID_ARG;
    Req r;
    r.ids.assign(ids, ids + k);
    for (u32 id : r.ids) {
        if (id >= s->blocks.size()) return FQXV_RANS_ERR_INVALID_ARG;
        r.bytes += round_up(s->blocks[id].dst_cap);
    }
    r.bytes = std::max(r.bytes, ALIGN); // empty blocks still get a region of their own
    if (r.bytes > s->budget) return FQXV_RANS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    r.ticket = s->next_ticket++;
    r.t_arrival = t_arrival >= 0 ? t_arrival : s->now();
    *ticket = r.ticket;
    s->queue.push_back(std::move(r));
    return FQXV_RANS_OK;
}

int fqxv_rans_server_poll(fqxv_rans_server *s, uint32_t *queued, uint32_t *in_flight)
{
    if (!s) return FQXV_RANS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    DeviceGuard dg(s->device);
    int rc;
    if (s->persistent) {
        reap_p(s);
        rc = admit_p(s);
    } else {
        rc = reap(s);
        if (!rc) rc = admit(s);
    }
    if (queued) *queued = (u32)s->queue.size();
    if (in_flight) {
        u32 n = (u32)s->active.size();
        for (const Slot &sl : s->slots) n += (u32)sl.reqs.size();
        *in_flight = n;
    }
    return rc;
}

int fqxv_rans_server_take(fqxv_rans_server *s, fqxv_rans_done *out, uint32_t max, uint32_t *got)
{
    if (!s || !got || (max && !out)) return FQXV_RANS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    u32 n = 0;
    while (n < max && !s->done.empty()) {
        out[n++] = s->done.front();
        s->done.pop_front();
    }
    *got = n;
    return FQXV_RANS_OK;
}

int fqxv_rans_server_release(fqxv_rans_server *s, uint64_t ticket)
{
    if (!s) return FQXV_RANS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->held.find(ticket);
    if (it == s->held.end()) return FQXV_RANS_ERR_INVALID_ARG;
    s->pool->release(it->second.first, it->second.second);
    s->held.erase(it);
    return FQXV_RANS_OK;
}

int fqxv_rans_server_waves(fqxv_rans_server *s, uint32_t *waves)
{
    if (!s || !waves) return FQXV_RANS_ERR_INVALID_ARG;
    if (s->persistent) {
        *waves = s->waves;
        return FQXV_RANS_OK;
    }
    DeviceGuard dg(s->device);
    hipDeviceProp_t prop;
    int occ = 0;
    if (hipGetDeviceProperties(&prop, s->device) != hipSuccess || fqxv_decode_occupancy(&occ) != hipSuccess)
        return FQXV_RANS_ERR_DEVICE;
    *waves = (uint32_t)std::max(1, occ) * (uint32_t)prop.multiProcessorCount;
    return FQXV_RANS_OK;
}

} // extern "C"
