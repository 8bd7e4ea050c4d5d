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
// Output of a request: its blocks in request order, each at a 256-byte aligned
// offset inside one region of the caller's output pool (block i at out_off + sum over earlier blocks
// of round_up(n, 256)).  Bytes are those of a plain decode of each block.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <unordered_map>
#include <vector>

#include "../../include/fqxv_rans_hip.h"

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

struct fqxv_rans_server {
    int device = 0;
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

} // namespace

extern "C" {

void fqxv_rans_server_destroy(fqxv_rans_server *s)
{
    if (!s) return;
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
    int rc = reap(s);
    if (!rc) rc = admit(s);
    if (queued) *queued = (u32)s->queue.size();
    if (in_flight) {
        u32 n = 0;
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

} // extern "C"
