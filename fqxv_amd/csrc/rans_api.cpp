// C ABI of the gfx950 rANS Nx16 coder (declared in include/fqxv_rans_hip.h).
//
// Mirrors crates/fqxv-rans/src/lib.rs: encode (lib.rs:111-127), decode
// (lib.rs:186-200), peek_output_len / decode_bounded (lib.rs:207-231).  The
// host-buffer entry points stage through device memory owned by the context;
// the batched entry points work on caller-owned device arenas.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

#include "../../include/fqxv_rans_hip.h"
#include "rans_internal.h"

typedef uint8_t u8;
typedef uint32_t u32;
typedef uint64_t u64;
typedef int32_t i32;

// Device-side launchers (rans_decode.hip / rans_encode.hip).
hipError_t fqxv_launch_decode_batch(const u8 *, u8 *, const fqxv_rans_job *, u32, const u32 *, u32 *, i32 *,
                                    u64 *, u8 *, u32, hipStream_t);
hipError_t fqxv_launch_decode_one(const u8 *, u8 *, const fqxv_rans_job &, i32 *, u64 *, u8 *, hipStream_t);
hipError_t fqxv_decode_occupancy(int *);
size_t fqxv_ws_per_wave();
hipError_t fqxv_launch_encode_batch(const u8 *, u8 *, const fqxv_rans_job *, u32, const u32 *, u32 *, i32 *,
                                    u64 *, u8 *, u32, hipStream_t, int, void *);
size_t fqxv_enc_pre_bytes();
hipError_t fqxv_launch_encode_one(const u8 *, u8 *, const fqxv_rans_job &, i32 *, u64 *, u8 *, hipStream_t);
hipError_t fqxv_encode_occupancy(int *);
hipError_t fqxv_launch_perm_split(const u32 *, u64, u8 *, hipStream_t);
size_t fqxv_crc_const_bytes();
void fqxv_crc_const_fill(void *);
hipError_t fqxv_launch_crc(const u8 *, const u64 *, const u64 *, const u32 *, const u32 *, u32, const u32 *, u32,
                           const void *, u32 *, u32 *, u32, const u64 *, const u32 *, u32, const u32 *, u32 *,
                           hipStream_t);
uint32_t fqxv_crc_seg_chunks();
// Waves of the CRC kernels (grid-stride over chunk items and ranges):
// measured against 2048 and 5120.
constexpr u32 CRC_GRID = 16384;
hipError_t fqxv_launch_perm_join(const u8 *, u64, u32 *, hipStream_t);
hipError_t fqxv_launch_pack_jobs(const u8 *, const fqxv_rans_job *, const u64 *, const i32 *, const u64 *, u32, u8 *,
                                 hipStream_t);
hipError_t fqxv_launch_xxh3(const u8 *, const u64 *, const u64 *, const u64 *, u32, u64, u64 *, u64 *,
                            hipStream_t);
hipError_t fqxv_launch_gather_items(const u8 *, const u64 *, const u64 *, const u32 *, u32, u8 *, hipStream_t);

namespace {

constexpr u32 N_COUNTERS = 64;

u64 hdr_bound(u64 n, int order)
{
    if (n == 0) return 9;
    const u64 k = n < 256 ? n : 256;
    return order == 0 ? 9 + 512 + 128 : 9 + 32 + 512 * k + 128;
}

} // namespace

struct fqxv_rans_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    u32 dec_waves = 0, enc_waves = 0;       // grid used by batched calls
    u32 dec_max = 0, enc_max = 0;           // occupancy maxima
    u8 *ws = nullptr;
    u64 ws_waves = 0;
    u32 *counters = nullptr;
    u32 next_counter = 0;
    u8 *d_in = nullptr;
    u64 d_in_cap = 0;
    u8 *d_out = nullptr;
    u64 d_out_cap = 0;
    i32 *d_status = nullptr;
    u64 *d_out_len = nullptr;
    void *crc_const = nullptr;              // slicing tables + zero-byte operators
    u8 *crc_buf = nullptr;                  // per-call range/item arrays
    u64 crc_cap = 0;
    u8 *job_buf = nullptr;                  // host-staged batches: jobs, order, status, out_len
    u64 job_cap = 0;
    u8 *xxh_buf = nullptr;                  // per-call xxh3 range/item arrays and block sums
    u64 xxh_cap = 0;
    u8 *cat_buf = nullptr;                  // concatenated-digest staging and copy items
    u64 cat_cap = 0;
    void *pre = nullptr;                    // encode histogram pre-pass (fixed size, never moves)
};

namespace {

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

int ensure_ws(fqxv_rans_ctx *c, u64 waves)
{
    if (c->ws_waves >= waves) return FQXV_RANS_OK;
    if (c->ws) {
        if (hipDeviceSynchronize() != hipSuccess) return FQXV_RANS_ERR_DEVICE;
        (void)hipFree(c->ws);
        c->ws = nullptr;
        c->ws_waves = 0;
    }
    if (hipMalloc((void **)&c->ws, waves * fqxv_ws_per_wave()) != hipSuccess) {
        c->ws = nullptr;
        return FQXV_RANS_ERR_DEVICE;
    }
    c->ws_waves = waves;
    return FQXV_RANS_OK;
}

// The encode pre-pass buffer: allocated once per context at its fixed size,
// so a captured graph's address of it stays valid.
int ensure_pre(fqxv_rans_ctx *c)
{
    if (c->pre) return FQXV_RANS_OK;
    if (hipMalloc(&c->pre, fqxv_enc_pre_bytes()) != hipSuccess) {
        c->pre = nullptr;
        return FQXV_RANS_ERR_DEVICE;
    }
    return FQXV_RANS_OK;
}

int ensure_buf(u8 **p, u64 *cap, u64 need)
{
    need = std::max<u64>(need, 64);
    if (*cap >= need) return FQXV_RANS_OK;
    if (*p) {
        if (hipDeviceSynchronize() != hipSuccess) return FQXV_RANS_ERR_DEVICE;
        (void)hipFree(*p);
        *p = nullptr;
        *cap = 0;
    }
    // Round up to limit reallocations; the slack also keeps aligned reads
    // near the end of a stream inside the allocation.
    u64 sz = 1;
    while (sz < need + 64) sz <<= 1;
    if (hipMalloc((void **)p, sz) != hipSuccess) {
        *p = nullptr;
        return FQXV_RANS_ERR_DEVICE;
    }
    *cap = sz;
    return FQXV_RANS_OK;
}

// The caller's stream; NULL is the HIP null stream (not the context's
// stream, which only the host-buffer calls use).
hipStream_t pick_stream(fqxv_rans_ctx *, void *stream) { return (hipStream_t)stream; }

u32 *next_counter(fqxv_rans_ctx *c)
{
    u32 *p = c->counters + (c->next_counter % N_COUNTERS);
    c->next_counter++;
    return p;
}

} // namespace

extern "C" {

const char *fqxv_rans_strerror(int code)
{
    switch (code) {
    case FQXV_RANS_OK: return "ok";
    case FQXV_RANS_ERR_TRUNCATED: return "truncated stream";
    case FQXV_RANS_ERR_UNKNOWN_ORDER: return "unknown order tag";
    case FQXV_RANS_ERR_FREQ_EXCEEDS: return "rANS freq table exceeds TOTFREQ";
    case FQXV_RANS_ERR_FREQ_SUM: return "rANS freq table does not sum to TOTFREQ";
    case FQXV_RANS_ERR_OUTPUT_ALLOC: return "rANS output length too large to allocate";
    case FQXV_RANS_ERR_ABSENT_CONTEXT: return "symbol references absent context";
    case FQXV_RANS_ERR_BOUND: return "rANS output length exceeds caller bound";
    case FQXV_RANS_ERR_DST_TOO_SMALL: return "destination buffer too small";
    case FQXV_RANS_ERR_INVALID_ARG: return "invalid argument";
    case FQXV_RANS_ERR_DEVICE: return "HIP device error";
    case FQXV_RANS_ERR_SCRATCH_LIMIT: return "request exceeds the fixed scratch of the no-allocation call";
    case FQXV_NAMES_ERR_TRUNCATED: return "truncated";
    case FQXV_NAMES_ERR_VARINT: return "varint too long";
    case FQXV_NAMES_ERR_VERSION: return "unsupported version";
    case FQXV_NAMES_ERR_RECORD_COUNT: return "record count exceeds token stream";
    case FQXV_NAMES_ERR_OP_SYM_UNDERRUN: return "op sym underrun";
    case FQXV_NAMES_ERR_OP_RUN_TOO_LONG: return "op run too long";
    case FQXV_NAMES_ERR_OP_BOUND: return "op columns exceed the stream bound";
    case FQXV_NAMES_ERR_PLANE_LEN: return "delta plane length mismatch";
    case FQXV_NAMES_ERR_MATCH: return "MATCH without prior token";
    case FQXV_NAMES_ERR_STR_UNDERRUN: return "string data underrun";
    case FQXV_NAMES_ERR_PAD_UNDERRUN: return "pad underrun";
    case FQXV_NAMES_ERR_DELTA_UNDERRUN: return "delta underrun";
    case FQXV_NAMES_ERR_DELTA_PRIOR: return "DELTA without numeric prior";
    case FQXV_NAMES_ERR_UNKNOWN_OP: return "unknown op";
    case FQXV_NAMES_ERR_ALLOC: return "op run too large to allocate";
    default: return "unknown error";
    }
}

size_t fqxv_rans_encode_bound(size_t n, int order)
{
    if (n == 0) return 9;
    return (size_t)(hdr_bound(n, order) + 2 * (u64)n + 16);
}

int fqxv_rans_peek_len(const uint8_t *src, size_t len, uint64_t *n)
{
    if (!src || len < 9) return FQXV_RANS_ERR_TRUNCATED;
    u64 v = 0;
    for (int i = 0; i < 8; i++) v |= (u64)src[1 + i] << (8 * i);
    if (n) *n = v;
    return FQXV_RANS_OK;
}

int fqxv_rans_ctx_create(int device, fqxv_rans_ctx **out)
{
    if (!out) return FQXV_RANS_ERR_INVALID_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return FQXV_RANS_ERR_DEVICE;
    DeviceGuard g(device);
    auto *c = new (std::nothrow) fqxv_rans_ctx();
    if (!c) return FQXV_RANS_ERR_DEVICE;
    c->device = device;
    hipDeviceProp_t prop;
    int dec_occ = 0, enc_occ = 0;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess || fqxv_decode_occupancy(&dec_occ) != hipSuccess ||
        fqxv_encode_occupancy(&enc_occ) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
        hipMalloc((void **)&c->counters, N_COUNTERS * sizeof(u32)) != hipSuccess ||
        hipMalloc((void **)&c->d_status, sizeof(i32)) != hipSuccess ||
        hipMalloc((void **)&c->d_out_len, sizeof(u64)) != hipSuccess) {
        fqxv_rans_ctx_destroy(c);
        return FQXV_RANS_ERR_DEVICE;
    }
    c->dec_waves = c->dec_max = (u32)std::max(1, dec_occ) * (u32)prop.multiProcessorCount;
    c->enc_waves = c->enc_max = (u32)std::max(1, enc_occ) * (u32)prop.multiProcessorCount;
    // FQXV_RANS_GRID caps the persistent grid (tuning and tests).
    if (const char *g = getenv("FQXV_RANS_GRID")) {
        const long v = strtol(g, nullptr, 10);
        if (v > 0) {
            c->dec_waves = std::min<u32>(c->dec_waves, (u32)v);
            c->enc_waves = std::min<u32>(c->enc_waves, (u32)v);
        }
    }
    *out = c;
    return FQXV_RANS_OK;
}

void fqxv_rans_ctx_destroy(fqxv_rans_ctx *c)
{
    if (!c) return;
    {
        DeviceGuard g(c->device);
        if (c->stream) (void)hipStreamSynchronize(c->stream);
        (void)hipDeviceSynchronize();
        if (c->ws) (void)hipFree(c->ws);
        if (c->pre) (void)hipFree(c->pre);
        if (c->counters) (void)hipFree(c->counters);
        if (c->d_in) (void)hipFree(c->d_in);
        if (c->d_out) (void)hipFree(c->d_out);
        if (c->d_status) (void)hipFree(c->d_status);
        if (c->d_out_len) (void)hipFree(c->d_out_len);
        if (c->crc_const) (void)hipFree(c->crc_const);
        if (c->crc_buf) (void)hipFree(c->crc_buf);
        if (c->job_buf) (void)hipFree(c->job_buf);
        if (c->xxh_buf) (void)hipFree(c->xxh_buf);
        if (c->cat_buf) (void)hipFree(c->cat_buf);
        if (c->stream) (void)hipStreamDestroy(c->stream);
    }
    delete c;
}

int fqxv_rans_max_waves(fqxv_rans_ctx *c, uint32_t *waves)
{
    if (!c || !waves) return FQXV_RANS_ERR_INVALID_ARG;
    *waves = std::max(c->dec_max, c->enc_max);
    return FQXV_RANS_OK;
}

int fqxv_rans_ctx_set_grid(fqxv_rans_ctx *c, uint32_t waves)
{
    if (!c) return FQXV_RANS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    c->dec_waves = waves ? std::min(waves, c->dec_max) : c->dec_max;
    c->enc_waves = waves ? std::min(waves, c->enc_max) : c->enc_max;
    return FQXV_RANS_OK;
}

void fqxv_rans_lpt_order(const uint64_t *sizes, uint32_t count, uint32_t *order_out)
{
    std::iota(order_out, order_out + count, 0u);
    std::stable_sort(order_out, order_out + count, [&](u32 a, u32 b) { return sizes[a] > sizes[b]; });
}

// Host-buffer decode through either device form: the single-block launch or
// the persistent batched kernel given a one-job queue (lib.rs:241-255 forces a
// backend the same way; the output is identical).
static int decode_host(fqxv_rans_ctx *c, const uint8_t *src, size_t len, size_t max_len, uint8_t *dst, size_t cap,
                       size_t *out_len, bool batched)
{
    if (out_len) *out_len = 0;
    if (!c || (!src && len) || !out_len) return FQXV_RANS_ERR_INVALID_ARG;
    u64 n = 0;
    const bool has_n = fqxv_rans_peek_len(src, len, &n) == FQXV_RANS_OK;
    if (has_n && n > (u64)max_len) return FQXV_RANS_ERR_BOUND;
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard g(c->device);
    const u64 dcap = has_n ? std::min<u64>(n, (u64)cap) : 0;
    if (has_n && n <= (u64)cap && n > 0 && !dst) return FQXV_RANS_ERR_INVALID_ARG;
    int rc;
    if ((rc = ensure_buf(&c->d_in, &c->d_in_cap, len)) || (rc = ensure_buf(&c->d_out, &c->d_out_cap, dcap)) ||
        (rc = ensure_ws(c, 1)) || (batched && (rc = ensure_buf(&c->job_buf, &c->job_cap, sizeof(fqxv_rans_job)))))
        return rc;
    if (len && hipMemcpyAsync(c->d_in, src, len, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        return FQXV_RANS_ERR_DEVICE;
    fqxv_rans_job jb{0, (u64)len, 0, dcap, UINT64_MAX, 0, 0};
    hipError_t e;
    if (batched) {
        fqxv_rans_job *d_job = (fqxv_rans_job *)c->job_buf;
        e = hipMemcpyAsync(d_job, &jb, sizeof jb, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess)
            e = fqxv_launch_decode_batch(c->d_in, c->d_out, d_job, 1, nullptr, next_counter(c), c->d_status,
                                         c->d_out_len, c->ws, 1, c->stream);
    } else {
        e = fqxv_launch_decode_one(c->d_in, c->d_out, jb, c->d_status, c->d_out_len, c->ws, c->stream);
    }
    if (e != hipSuccess) return FQXV_RANS_ERR_DEVICE;
    i32 st = 0;
    u64 ol = 0;
    if (hipMemcpyAsync(&st, c->d_status, sizeof st, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipMemcpyAsync(&ol, c->d_out_len, sizeof ol, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return FQXV_RANS_ERR_DEVICE;
    if (st != FQXV_RANS_OK) return st;
    if (ol && hipMemcpy(dst, c->d_out, ol, hipMemcpyDeviceToHost) != hipSuccess) return FQXV_RANS_ERR_DEVICE;
    *out_len = (size_t)ol;
    return FQXV_RANS_OK;
}

int fqxv_rans_decode_bounded(fqxv_rans_ctx *c, const uint8_t *src, size_t len, size_t max_len, uint8_t *dst,
                             size_t cap, size_t *out_len)
{
    return decode_host(c, src, len, max_len, dst, cap, out_len, false);
}

int fqxv_rans_decode_with(fqxv_rans_ctx *c, int backend, const uint8_t *src, size_t len, uint8_t *dst, size_t cap,
                          size_t *out_len)
{
    if (out_len) *out_len = 0;
    if (backend != FQXV_RANS_BACKEND_GFX950 && backend != FQXV_RANS_BACKEND_GFX950_BATCHED)
        return FQXV_RANS_ERR_INVALID_ARG;
    return decode_host(c, src, len, SIZE_MAX, dst, cap, out_len, backend == FQXV_RANS_BACKEND_GFX950_BATCHED);
}

int fqxv_rans_backend_detect(void) { return FQXV_RANS_BACKEND_GFX950; }

int fqxv_rans_decode(fqxv_rans_ctx *c, const uint8_t *src, size_t len, uint8_t *dst, size_t cap, size_t *out_len)
{
    return fqxv_rans_decode_bounded(c, src, len, SIZE_MAX, dst, cap, out_len);
}

int fqxv_rans_encode(fqxv_rans_ctx *c, const uint8_t *src, size_t n, int order, uint8_t *dst, size_t cap,
                     size_t *out_len)
{
    if (out_len) *out_len = 0;
    if (!c || (!src && n) || !out_len || !dst) return FQXV_RANS_ERR_INVALID_ARG;
    if (order != 0 && order != 1) return FQXV_RANS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard g(c->device);
    const u64 bound = fqxv_rans_encode_bound(n, order);
    int rc;
    if ((rc = ensure_buf(&c->d_in, &c->d_in_cap, n)) || (rc = ensure_buf(&c->d_out, &c->d_out_cap, bound)) ||
        (rc = ensure_ws(c, 1)))
        return rc;
    if (n && hipMemcpyAsync(c->d_in, src, n, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        return FQXV_RANS_ERR_DEVICE;
    fqxv_rans_job jb{0, (u64)n, 0, bound, 0, order, 0};
    if (fqxv_launch_encode_one(c->d_in, c->d_out, jb, c->d_status, c->d_out_len, c->ws, c->stream) != hipSuccess)
        return FQXV_RANS_ERR_DEVICE;
    i32 st = 0;
    u64 ol = 0;
    if (hipMemcpyAsync(&st, c->d_status, sizeof st, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipMemcpyAsync(&ol, c->d_out_len, sizeof ol, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return FQXV_RANS_ERR_DEVICE;
    if (st != FQXV_RANS_OK) return st;
    if (ol > cap) return FQXV_RANS_ERR_DST_TOO_SMALL;
    if (hipMemcpy(dst, c->d_out, ol, hipMemcpyDeviceToHost) != hipSuccess) return FQXV_RANS_ERR_DEVICE;
    *out_len = (size_t)ol;
    return FQXV_RANS_OK;
}

// mode: 0 decode, 1 encode, 2 encode-best.
static int batch_common(fqxv_rans_ctx *c, const uint8_t *d_src, uint8_t *d_dst, const fqxv_rans_job *d_jobs,
                        uint32_t count, const uint32_t *d_order, int32_t *d_status, uint64_t *d_out_len,
                        void *stream, int mode)
{
    const bool enc = mode != 0;
    if (!c) return FQXV_RANS_ERR_INVALID_ARG;
    if (count == 0) return FQXV_RANS_OK;
    if (!d_src || !d_dst || !d_jobs || !d_status || !d_out_len) return FQXV_RANS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard g(c->device);
    u32 grid = std::min<u32>(count, enc ? c->enc_waves : c->dec_waves);
    int rc = ensure_ws(c, grid);
    if (!rc && enc) rc = ensure_pre(c);
    if (rc) return rc;
    hipStream_t st = pick_stream(c, stream);
    u32 *ctr = next_counter(c);
    const hipError_t e = enc ? fqxv_launch_encode_batch(d_src, d_dst, d_jobs, count, d_order, ctr, d_status,
                                                        d_out_len, c->ws, grid, st, mode == 2, c->pre)
                             : fqxv_launch_decode_batch(d_src, d_dst, d_jobs, count, d_order, ctr, d_status,
                                                        d_out_len, c->ws, grid, st);
    return e == hipSuccess ? FQXV_RANS_OK : FQXV_RANS_ERR_DEVICE;
}

// Host-staged batch: the source span and the jobs go up, one persistent
// launch runs (largest job first), status, lengths and the whole destination
// arena come back.  mode 0 decodes, 2 encodes at both orders keeping the
// smaller stream.  Serves the names decoder and encoder.
int fqxv_run_jobs_host(fqxv_rans_ctx *c, int mode, const uint8_t *h_src, uint64_t src_len,
                       const fqxv_rans_job *h_jobs, uint32_t count, uint64_t dst_total, uint8_t *h_dst,
                       int32_t *h_status, uint64_t *h_out_len)
{
    if (!c || (mode != 0 && mode != 2) || (src_len && !h_src) || (count && (!h_jobs || !h_status || !h_out_len)) ||
        (dst_total && !h_dst))
        return FQXV_RANS_ERR_INVALID_ARG;
    if (count == 0) return FQXV_RANS_OK;
    std::vector<u64> cost(count);
    for (u32 i = 0; i < count; i++) {
        const fqxv_rans_job &j = h_jobs[i];
        if (j.src_off + j.src_len > src_len || j.dst_off + j.dst_cap > dst_total) return FQXV_RANS_ERR_INVALID_ARG;
        cost[i] = j.dst_cap + j.src_len;
    }
    std::vector<u32> order(count);
    fqxv_rans_lpt_order(cost.data(), count, order.data());
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard g(c->device);
    const u32 grid = std::min<u32>(count, mode ? c->enc_waves : c->dec_waves);
    const u64 jb = sizeof(fqxv_rans_job) * (u64)count, ob = 4ull * count, sb = 4ull * count, lb = 8ull * count;
    int rc;
    if ((rc = ensure_buf(&c->d_in, &c->d_in_cap, src_len)) || (rc = ensure_buf(&c->d_out, &c->d_out_cap, dst_total)) ||
        (rc = ensure_buf(&c->job_buf, &c->job_cap, lb + jb + ob + sb)) || (rc = ensure_ws(c, grid)) ||
        (mode && (rc = ensure_pre(c))))
        return rc;
    u64 *d_len = (u64 *)c->job_buf;
    fqxv_rans_job *d_jobs = (fqxv_rans_job *)(d_len + count);
    u32 *d_order = (u32 *)(d_jobs + count);
    i32 *d_st = (i32 *)(d_order + count);
    hipStream_t st = c->stream;
    if ((src_len && hipMemcpyAsync(c->d_in, h_src, src_len, hipMemcpyHostToDevice, st) != hipSuccess) ||
        hipMemcpyAsync(d_jobs, h_jobs, jb, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(d_order, order.data(), ob, hipMemcpyHostToDevice, st) != hipSuccess)
        return FQXV_RANS_ERR_DEVICE;
    const hipError_t e =
        mode ? fqxv_launch_encode_batch(c->d_in, c->d_out, d_jobs, count, d_order, next_counter(c), d_st, d_len,
                                        c->ws, grid, st, 1, c->pre)
             : fqxv_launch_decode_batch(c->d_in, c->d_out, d_jobs, count, d_order, next_counter(c), d_st, d_len,
                                        c->ws, grid, st);
    if (e != hipSuccess) return FQXV_RANS_ERR_DEVICE;
    if (hipMemcpyAsync(h_status, d_st, sb, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(h_out_len, d_len, lb, hipMemcpyDeviceToHost, st) != hipSuccess ||
        (dst_total && hipMemcpyAsync(h_dst, c->d_out, dst_total, hipMemcpyDeviceToHost, st) != hipSuccess) ||
        hipStreamSynchronize(st) != hipSuccess)
        return FQXV_RANS_ERR_DEVICE;
    return FQXV_RANS_OK;
}

int fqxv_jobs_host_packed(fqxv_rans_ctx *c, int mode, const uint8_t *h_src, uint64_t src_len,
                          const fqxv_rans_job *h_jobs, uint32_t count, uint64_t dst_total, uint8_t **h_packed,
                          uint64_t *h_pack_off, int32_t *h_status, uint64_t *h_out_len)
{
    if (!c || !h_packed || (mode != 0 && mode != 2) || (src_len && !h_src) ||
        (count && (!h_jobs || !h_pack_off || !h_status || !h_out_len)))
        return FQXV_RANS_ERR_INVALID_ARG;
    *h_packed = nullptr;
    std::vector<u64> cost(count);
    for (u32 i = 0; i < count; i++) {
        const fqxv_rans_job &j = h_jobs[i];
        if (j.src_off + j.src_len > src_len || j.dst_off + j.dst_cap > dst_total) return FQXV_RANS_ERR_INVALID_ARG;
        cost[i] = j.dst_cap + j.src_len;
    }
    std::vector<u32> order(count);
    if (count) fqxv_rans_lpt_order(cost.data(), count, order.data());
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard g(c->device);
    const u32 grid = std::min<u32>(std::max<u32>(count, 1), mode ? c->enc_waves : c->dec_waves);
    const u64 jb = sizeof(fqxv_rans_job) * (u64)count, ob = 4ull * count, sb = 4ull * count, lb = 8ull * count;
    int rc;
    if ((rc = ensure_buf(&c->d_in, &c->d_in_cap, src_len)) || (rc = ensure_buf(&c->d_out, &c->d_out_cap, dst_total)) ||
        (rc = ensure_buf(&c->job_buf, &c->job_cap, 2 * lb + jb + ob + sb)) || (rc = ensure_ws(c, grid)) ||
        (mode && (rc = ensure_pre(c))))
        return rc;
    auto launch = [&](u32 *ctr, i32 *d_st, u64 *d_len, fqxv_rans_job *d_jobs, u32 *d_order, hipStream_t st) {
        return mode ? fqxv_launch_encode_batch(c->d_in, c->d_out, d_jobs, count, d_order, ctr, d_st, d_len, c->ws,
                                               grid, st, 1, c->pre)
                    : fqxv_launch_decode_batch(c->d_in, c->d_out, d_jobs, count, d_order, ctr, d_st, d_len, c->ws,
                                               grid, st);
    };
    u64 *d_len = (u64 *)c->job_buf;
    u64 *d_poff = d_len + count;
    fqxv_rans_job *d_jobs = (fqxv_rans_job *)(d_poff + count);
    u32 *d_order = (u32 *)(d_jobs + count);
    i32 *d_st = (i32 *)(d_order + count);
    hipStream_t st = c->stream;
    if (count) {
        if ((src_len && hipMemcpyAsync(c->d_in, h_src, src_len, hipMemcpyHostToDevice, st) != hipSuccess) ||
            hipMemcpyAsync(d_jobs, h_jobs, jb, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(d_order, order.data(), ob, hipMemcpyHostToDevice, st) != hipSuccess ||
            launch(next_counter(c), d_st, d_len, d_jobs, d_order, st) != hipSuccess ||
            hipMemcpyAsync(h_status, d_st, sb, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(h_out_len, d_len, lb, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return FQXV_RANS_ERR_DEVICE;
    }
    // Packed offsets of the successful jobs, in job order.
    u64 total = 0;
    for (u32 i = 0; i < count; i++) {
        h_pack_off[i] = total;
        if (h_status[i] == 0) {
            if (h_out_len[i] > h_jobs[i].dst_cap) return FQXV_RANS_ERR_DEVICE;
            total += h_out_len[i];
        }
    }
    u8 *h = (u8 *)malloc(std::max<u64>(total, 1));
    if (!h) return FQXV_RANS_ERR_DEVICE;
    // The sources are consumed: the packed outputs are staged in d_in.
    if (total) {
        if ((rc = ensure_buf(&c->d_in, &c->d_in_cap, total))) {
            free(h);
            return rc;
        }
        if (hipMemcpyAsync(d_poff, h_pack_off, lb, hipMemcpyHostToDevice, st) != hipSuccess ||
            fqxv_launch_pack_jobs(c->d_out, d_jobs, d_len, d_st, d_poff, count, c->d_in, st) != hipSuccess ||
            hipMemcpyAsync(h, c->d_in, total, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            free(h);
            return FQXV_RANS_ERR_DEVICE;
        }
    }
    *h_packed = h;
    return FQXV_RANS_OK;
}

int fqxv_rans_decode_batch(fqxv_rans_ctx *c, const uint8_t *d_src, uint8_t *d_dst, const fqxv_rans_job *d_jobs,
                           uint32_t count, const uint32_t *d_order, int32_t *d_status, uint64_t *d_out_len,
                           void *stream)
{
    return batch_common(c, d_src, d_dst, d_jobs, count, d_order, d_status, d_out_len, stream, 0);
}

int fqxv_rans_encode_batch(fqxv_rans_ctx *c, const uint8_t *d_src, uint8_t *d_dst, const fqxv_rans_job *d_jobs,
                           uint32_t count, const uint32_t *d_order, int32_t *d_status, uint64_t *d_out_len,
                           void *stream)
{
    return batch_common(c, d_src, d_dst, d_jobs, count, d_order, d_status, d_out_len, stream, 1);
}

int fqxv_rans_encode_best_batch(fqxv_rans_ctx *c, const uint8_t *d_src, uint8_t *d_dst, const fqxv_rans_job *d_jobs,
                                uint32_t count, const uint32_t *d_order, int32_t *d_status, uint64_t *d_out_len,
                                void *stream)
{
    return batch_common(c, d_src, d_dst, d_jobs, count, d_order, d_status, d_out_len, stream, 2);
}

size_t fqxv_rans_encode_best_bound(size_t n) { return fqxv_rans_encode_bound(n, 0) + fqxv_rans_encode_bound(n, 1); }

static int one_common(fqxv_rans_ctx *c, const uint8_t *d_src, uint8_t *d_dst, const fqxv_rans_job *job,
                      int32_t *d_status, uint64_t *d_out_len, void *stream, bool enc)
{
    if (!c || !job || !d_src || !d_dst || !d_status || !d_out_len) return FQXV_RANS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard g(c->device);
    int rc = ensure_ws(c, 1);
    if (rc) return rc;
    hipStream_t st = pick_stream(c, stream);
    const hipError_t e = enc ? fqxv_launch_encode_one(d_src, d_dst, *job, d_status, d_out_len, c->ws, st)
                             : fqxv_launch_decode_one(d_src, d_dst, *job, d_status, d_out_len, c->ws, st);
    return e == hipSuccess ? FQXV_RANS_OK : FQXV_RANS_ERR_DEVICE;
}

struct fqxv_rans_graph {
    fqxv_rans_ctx *ctx = nullptr;
    u32 *counter = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

void fqxv_rans_graph_destroy(fqxv_rans_graph *g)
{
    if (!g) return;
    if (g->ctx) {
        DeviceGuard dg(g->ctx->device);
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        if (g->counter) (void)hipFree(g->counter);
    }
    delete g;
}

int fqxv_rans_graph_create(fqxv_rans_ctx *c, int direction, const uint8_t *d_src, uint8_t *d_dst,
                           const fqxv_rans_job *d_jobs, uint32_t count, const uint32_t *d_order,
                           int32_t *d_status, uint64_t *d_out_len, fqxv_rans_graph **out)
{
    if (!c || !out || direction < 0 || direction > 2 || count == 0 || !d_src || !d_dst || !d_jobs ||
        !d_status || !d_out_len)
        return FQXV_RANS_ERR_INVALID_ARG;
    *out = nullptr;
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard dg(c->device);
    const bool enc = direction != 0;
    const u32 grid = std::min<u32>(count, enc ? c->enc_waves : c->dec_waves);
    // The graph bakes in the workspace address: size it for the largest grid
    // any later call on this context can ask for, so it is never reallocated.
    int rc = ensure_ws(c, std::max(c->dec_max, c->enc_max));
    if (!rc && enc) rc = ensure_pre(c);
    if (rc) return rc;
    auto *g = new (std::nothrow) fqxv_rans_graph();
    if (!g) return FQXV_RANS_ERR_DEVICE;
    g->ctx = c;
    hipStream_t cs = nullptr;
    bool ok = hipMalloc((void **)&g->counter, sizeof(u32)) == hipSuccess && hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) == hipSuccess &&
              hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (ok) {
        const hipError_t e = enc ? fqxv_launch_encode_batch(d_src, d_dst, d_jobs, count, d_order, g->counter,
                                                            d_status, d_out_len, c->ws, grid, cs, direction == 2,
                                                            c->pre)
                                 : fqxv_launch_decode_batch(d_src, d_dst, d_jobs, count, d_order, g->counter,
                                                            d_status, d_out_len, c->ws, grid, cs);
        const hipError_t e2 = hipStreamEndCapture(cs, &g->graph);
        ok = e == hipSuccess && e2 == hipSuccess &&
             hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) == hipSuccess;
    }
    if (cs) (void)hipStreamDestroy(cs);
    if (!ok) {
        fqxv_rans_graph_destroy(g);
        return FQXV_RANS_ERR_DEVICE;
    }
    *out = g;
    return FQXV_RANS_OK;
}

int fqxv_rans_graph_launch(fqxv_rans_graph *g, void *stream)
{
    if (!g || !g->exec) return FQXV_RANS_ERR_INVALID_ARG;
    DeviceGuard dg(g->ctx->device);
    return hipGraphLaunch(g->exec, pick_stream(g->ctx, stream)) == hipSuccess ? FQXV_RANS_OK
                                                                             : FQXV_RANS_ERR_DEVICE;
}

int fqxv_rans_perm_split(const uint32_t *d_perm, uint64_t n, uint8_t *d_planes, void *stream)
{
    if (n && (!d_perm || !d_planes)) return FQXV_RANS_ERR_INVALID_ARG;
    return fqxv_launch_perm_split(d_perm, n, d_planes, (hipStream_t)stream) == hipSuccess ? FQXV_RANS_OK
                                                                                        : FQXV_RANS_ERR_DEVICE;
}

int fqxv_rans_perm_join(const uint8_t *d_planes, uint64_t n, uint32_t *d_perm, void *stream)
{
    if (n && (!d_perm || !d_planes)) return FQXV_RANS_ERR_INVALID_ARG;
    return fqxv_launch_perm_join(d_planes, n, d_perm, (hipStream_t)stream) == hipSuccess ? FQXV_RANS_OK
                                                                                       : FQXV_RANS_ERR_DEVICE;
}

int fqxv_rans_crc32c_batch(fqxv_rans_ctx *c, const uint8_t *d_src, const uint64_t *off, const uint64_t *len,
                           uint32_t count, uint32_t *d_crc, void *stream)
{
    if (!c || (count && (!d_src || !off || !len || !d_crc))) return FQXV_RANS_ERR_INVALID_ARG;
    if (count == 0) return FQXV_RANS_OK;
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard g(c->device);
    hipStream_t st = pick_stream(c, stream);
    if (!c->crc_const) {
        std::vector<u8> h(fqxv_crc_const_bytes());
        fqxv_crc_const_fill(h.data());
        if (hipMalloc(&c->crc_const, h.size()) != hipSuccess) {
            c->crc_const = nullptr;
            return FQXV_RANS_ERR_DEVICE;
        }
        if (hipMemcpy(c->crc_const, h.data(), h.size(), hipMemcpyHostToDevice) != hipSuccess)
            return FQXV_RANS_ERR_DEVICE;
    }
    // Chunk items (range, chunk) in range order; first item of each range.
    std::vector<u32> item_range, item_chunk, first(count);
    for (u32 r = 0; r < count; r++) {
        if (len[r] >> 47) return FQXV_RANS_ERR_INVALID_ARG;
        first[r] = (u32)item_range.size();
        const u64 nch = (len[r] + 65535) >> 16;
        if (item_range.size() + nch > 0xFFFFFFFFull) return FQXV_RANS_ERR_INVALID_ARG;
        for (u64 k = 0; k < nch; k++) {
            item_range.push_back(r);
            item_chunk.push_back((u32)k);
        }
    }
    const u64 ni = item_range.size();
    // A range of more than one segment (fqxv_crc_seg_chunks() chunks) is
    // folded in two levels: its segments first (in parallel), then the
    // segment values; otherwise one fold per range over its chunk values.
    const u64 seg = fqxv_crc_seg_chunks();
    std::vector<u64> seg_len;
    std::vector<u32> seg_first, range_seg;
    bool two_level = false;
    for (u32 r = 0; r < count; r++) two_level |= ((len[r] + 65535) >> 16) > seg;
    if (two_level) {
        range_seg.resize(count);
        for (u32 r = 0; r < count; r++) {
            range_seg[r] = (u32)seg_len.size();
            const u64 nch = (len[r] + 65535) >> 16;
            for (u64 k = 0; k < nch; k += seg) {
                const u64 b0 = k << 16, b1 = std::min<u64>((k + seg) << 16, len[r]);
                seg_len.push_back(b1 - b0);
                seg_first.push_back(first[r] + (u32)k);
            }
        }
    }
    const u64 ns = seg_len.size();
    // Device layout: off[count] len[count] seg_len[ns] (u64), first[count]
    // item_range[ni] item_chunk[ni] raw[ni] seg_first[ns] range_seg[count]
    // seg_raw[ns] (u32).
    const u64 need = 16ull * count + 8ull * ns + 4ull * count + 12ull * ni + 8ull * ns + 4ull * count + 64;
    int rc = ensure_buf(&c->crc_buf, &c->crc_cap, need);
    if (rc) return rc;
    u64 *d_off = (u64 *)c->crc_buf, *d_len = d_off + count, *d_seg_len = d_len + count;
    u32 *d_first = (u32 *)(d_seg_len + ns), *d_ir = d_first + count, *d_ic = d_ir + ni, *d_raw = d_ic + ni;
    u32 *d_seg_first = d_raw + ni, *d_range_seg = d_seg_first + ns, *d_seg_raw = d_range_seg + count;
    if (hipMemcpyAsync(d_off, off, 8ull * count, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(d_len, len, 8ull * count, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(d_first, first.data(), 4ull * count, hipMemcpyHostToDevice, st) != hipSuccess ||
        (ni && hipMemcpyAsync(d_ir, item_range.data(), 4 * ni, hipMemcpyHostToDevice, st) != hipSuccess) ||
        (ni && hipMemcpyAsync(d_ic, item_chunk.data(), 4 * ni, hipMemcpyHostToDevice, st) != hipSuccess) ||
        (ns && (hipMemcpyAsync(d_seg_len, seg_len.data(), 8 * ns, hipMemcpyHostToDevice, st) != hipSuccess ||
                hipMemcpyAsync(d_seg_first, seg_first.data(), 4 * ns, hipMemcpyHostToDevice, st) != hipSuccess ||
                hipMemcpyAsync(d_range_seg, range_seg.data(), 4ull * count, hipMemcpyHostToDevice, st) !=
                    hipSuccess)))
        return FQXV_RANS_ERR_DEVICE;
    if (fqxv_launch_crc(d_src, d_off, d_len, d_ir, d_ic, (u32)ni, d_first, count, c->crc_const, d_raw, d_crc,
                        CRC_GRID, d_seg_len, d_seg_first, (u32)ns, d_range_seg, d_seg_raw,
                        st) != hipSuccess)
        return FQXV_RANS_ERR_DEVICE;
    // The staging vectors above are host locals: finish before returning.
    return hipStreamSynchronize(st) == hipSuccess ? FQXV_RANS_OK : FQXV_RANS_ERR_DEVICE;
}

// xxh3 of `count` ranges of a device arena; caller holds the context lock.
static int xxh3_locked(fqxv_rans_ctx *c, const uint8_t *d_src, const uint64_t *off, const uint64_t *len,
                       uint32_t count, uint64_t *d_out, hipStream_t st)
{
    // 1 KiB blocks of the long inputs: range r owns [first[r], first[r + 1]).
    std::vector<u64> first(count + 1);
    u64 ni = 0;
    for (u32 r = 0; r < count; r++) {
        first[r] = ni;
        if (len[r] > 240) ni += (len[r] - 1) >> 10;
    }
    first[count] = ni;
    // Device layout: sums[ni][8] off[count] len[count] first[count + 1].
    const u64 need = 64ull * ni + 24ull * count + 8 + 64;
    int rc = ensure_buf(&c->xxh_buf, &c->xxh_cap, need);
    if (rc) return rc;
    u64 *d_sums = (u64 *)c->xxh_buf, *d_off = d_sums + 8 * ni, *d_len = d_off + count, *d_first = d_len + count;
    if (hipMemcpyAsync(d_off, off, 8ull * count, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(d_len, len, 8ull * count, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(d_first, first.data(), 8ull * (count + 1), hipMemcpyHostToDevice, st) != hipSuccess)
        return FQXV_RANS_ERR_DEVICE;
    if (fqxv_launch_xxh3(d_src, d_off, d_len, d_first, count, ni, d_sums, d_out, st) != hipSuccess)
        return FQXV_RANS_ERR_DEVICE;
    // The staging vectors above are host locals: finish before returning.
    return hipStreamSynchronize(st) == hipSuccess ? FQXV_RANS_OK : FQXV_RANS_ERR_DEVICE;
}

int fqxv_rans_xxh3_batch(fqxv_rans_ctx *c, const uint8_t *d_src, const uint64_t *off, const uint64_t *len,
                         uint32_t count, uint64_t *d_out, void *stream)
{
    if (!c || (count && (!d_src || !off || !len || !d_out))) return FQXV_RANS_ERR_INVALID_ARG;
    if (count == 0) return FQXV_RANS_OK;
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard g(c->device);
    return xxh3_locked(c, d_src, off, len, count, d_out, pick_stream(c, stream));
}

int fqxv_rans_xxh3_concat(fqxv_rans_ctx *c, const uint8_t *d_src, const uint64_t *seg_off, const uint64_t *seg_len,
                          const uint32_t *seg_first, uint32_t count, uint64_t *d_out, void *stream)
{
    if (!c || (count && (!d_src || !seg_first || !d_out))) return FQXV_RANS_ERR_INVALID_ARG;
    if (count == 0) return FQXV_RANS_OK;
    const u32 nseg = seg_first[count];
    if (nseg && (!seg_off || !seg_len)) return FQXV_RANS_ERR_INVALID_ARG;
    // Staging layout: each concatenation at a 256-byte aligned start, then
    // the copy items (pieces of at most 64 KiB).
    std::vector<u64> off(count), len(count), it_src, it_dst;
    std::vector<u32> it_len;
    u64 total = 0;
    for (u32 r = 0; r < count; r++) {
        if (seg_first[r] > seg_first[r + 1] || seg_first[r + 1] > nseg) return FQXV_RANS_ERR_INVALID_ARG;
        off[r] = total;
        u64 l = 0;
        for (u32 s = seg_first[r]; s < seg_first[r + 1]; s++) {
            for (u64 p = 0; p < seg_len[s]; p += 65536) {
                it_src.push_back(seg_off[s] + p);
                it_dst.push_back(total + l + p);
                it_len.push_back((u32)std::min<u64>(65536, seg_len[s] - p));
            }
            l += seg_len[s];
        }
        len[r] = l;
        total += (l + 255) & ~255ull;
    }
    const u64 ni = it_len.size();
    if (ni > 0xFFFFFFFFull) return FQXV_RANS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard g(c->device);
    const u64 items_at = total;
    int rc = ensure_buf(&c->cat_buf, &c->cat_cap, items_at + 20 * ni + 64);
    if (rc) return rc;
    hipStream_t st = pick_stream(c, stream);
    u64 *d_isrc = (u64 *)(c->cat_buf + items_at), *d_idst = d_isrc + ni;
    u32 *d_ilen = (u32 *)(d_idst + ni);
    if (ni && (hipMemcpyAsync(d_isrc, it_src.data(), 8 * ni, hipMemcpyHostToDevice, st) != hipSuccess ||
               hipMemcpyAsync(d_idst, it_dst.data(), 8 * ni, hipMemcpyHostToDevice, st) != hipSuccess ||
               hipMemcpyAsync(d_ilen, it_len.data(), 4 * ni, hipMemcpyHostToDevice, st) != hipSuccess ||
               fqxv_launch_gather_items(d_src, d_isrc, d_idst, d_ilen, (u32)ni, c->cat_buf, st) != hipSuccess))
        return FQXV_RANS_ERR_DEVICE;
    return xxh3_locked(c, c->cat_buf, off.data(), len.data(), count, d_out, st);
}

// Host ranges: packed into the context's input buffer (256-byte aligned
// starts), hashed, digests copied back.
int fqxv_xxh3_host_ranges(fqxv_rans_ctx *c, const uint8_t *const *ptrs, const uint64_t *lens, uint32_t count,
                          uint64_t *out)
{
    if (!c || (count && (!ptrs || !lens || !out))) return FQXV_RANS_ERR_INVALID_ARG;
    if (count == 0) return FQXV_RANS_OK;
    std::vector<u64> off(count);
    u64 total = 0;
    for (u32 i = 0; i < count; i++) {
        off[i] = total;
        total += (lens[i] + 255) & ~255ull;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard g(c->device);
    int rc;
    if ((rc = ensure_buf(&c->d_in, &c->d_in_cap, total)) || (rc = ensure_buf(&c->d_out, &c->d_out_cap, 8ull * count)))
        return rc;
    hipStream_t st = c->stream;
    for (u32 i = 0; i < count; i++)
        if (lens[i] && hipMemcpyAsync(c->d_in + off[i], ptrs[i], lens[i], hipMemcpyHostToDevice, st) != hipSuccess)
            return FQXV_RANS_ERR_DEVICE;
    if ((rc = xxh3_locked(c, c->d_in, off.data(), lens, count, (u64 *)c->d_out, st))) return rc;
    if (hipMemcpy(out, c->d_out, 8ull * count, hipMemcpyDeviceToHost) != hipSuccess) return FQXV_RANS_ERR_DEVICE;
    return FQXV_RANS_OK;
}

int fqxv_rans_decode_one_device(fqxv_rans_ctx *c, const uint8_t *d_src, uint8_t *d_dst, const fqxv_rans_job *job,
                                int32_t *d_status, uint64_t *d_out_len, void *stream)
{
    return one_common(c, d_src, d_dst, job, d_status, d_out_len, stream, false);
}

int fqxv_rans_encode_one_device(fqxv_rans_ctx *c, const uint8_t *d_src, uint8_t *d_dst, const fqxv_rans_job *job,
                                int32_t *d_status, uint64_t *d_out_len, void *stream)
{
    return one_common(c, d_src, d_dst, job, d_status, d_out_len, stream, true);
}

} // extern "C"
