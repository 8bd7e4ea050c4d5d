/*
 * fqxv_rans_hip.h -- C ABI of the MI355X (gfx950) rANS Nx16 coder.
 *
 * Drop-in boundary for the fqxv-rans crate (crates/fqxv-rans/src/lib.rs):
 * the same wire format (scalar.rs:8-20), the same order-0/order-1 models,
 * byte-identical streams.  Plain pointers and sizes only; no HIP or torch
 * types appear in the signatures (a `void *stream` is a hipStream_t).
 *
 * Two families of entry points:
 *   - host-buffer calls that mirror the Rust API one to one
 *     (fqxv_rans_encode / _decode / _decode_bounded);
 *   - device-resident batched calls over an arena of blocks
 *     (fqxv_rans_encode_batch / _decode_batch), the form the container's
 *     per-block dispatch (crates/fqxv/src/container/compress.rs:866-883,
 *     decompress.rs:45-52) hands to one GPU.
 *
 * Return codes: FQXV_RANS_OK or one of the FQXV_RANS_ERR_* values below.
 * Codes 1..7 are the reference's `Error::Malformed(&'static str)` messages
 * (fqxv_rans_strerror gives the exact string); the Rust side maps every one
 * of them to `Error::Malformed`.
 *
 * Thread safety: a context serialises its own host-buffer calls with an
 * internal lock.  Batched and *_one_device calls on one context must not
 * overlap on the device (they share the context's work-queue counters and
 * workspace): use one context per concurrently used stream.
 */
#ifndef FQXV_RANS_HIP_H
#define FQXV_RANS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FQXV_RANS_OK 0
#define FQXV_RANS_ERR_TRUNCATED 1        /* "truncated stream"                        scalar.rs:256-263 */
#define FQXV_RANS_ERR_UNKNOWN_ORDER 2    /* "unknown order tag"                       scalar.rs:153     */
#define FQXV_RANS_ERR_FREQ_EXCEEDS 3     /* "rANS freq table exceeds TOTFREQ"         model.rs:165      */
#define FQXV_RANS_ERR_FREQ_SUM 4         /* "rANS freq table does not sum to TOTFREQ" model.rs:169      */
#define FQXV_RANS_ERR_OUTPUT_ALLOC 5     /* "rANS output length too large to allocate" model.rs:202     */
#define FQXV_RANS_ERR_ABSENT_CONTEXT 6   /* "symbol references absent context"        scalar.rs:206     */
#define FQXV_RANS_ERR_BOUND 7            /* "rANS output length exceeds caller bound" lib.rs:228        */
#define FQXV_RANS_ERR_DST_TOO_SMALL 8    /* encode: destination smaller than fqxv_rans_encode_bound   */
#define FQXV_RANS_ERR_INVALID_ARG 9      /* bad order / NULL pointer / bad job                         */
#define FQXV_RANS_ERR_DEVICE 10          /* HIP runtime failure (allocation, launch, copy)             */
#define FQXV_RANS_ERR_SCRATCH_LIMIT 11   /* request exceeds the fixed scratch of the no-allocation call */

/* Names-stream errors: the reference's Error::Malformed strings of
 * crates/fqxv-tokenizer/src/lib.rs (a sub-stream's rANS failure keeps its
 * FQXV_RANS_ERR_* code). */
#define FQXV_NAMES_ERR_TRUNCATED 32      /* "truncated"                          lib.rs:705        */
#define FQXV_NAMES_ERR_VARINT 33         /* "varint too long"                    lib.rs:708        */
#define FQXV_NAMES_ERR_VERSION 34        /* "unsupported version"                lib.rs:546        */
#define FQXV_NAMES_ERR_RECORD_COUNT 35   /* "record count exceeds token stream"  lib.rs:566        */
#define FQXV_NAMES_ERR_OP_SYM_UNDERRUN 36 /* "op sym underrun"                   lib.rs:226        */
#define FQXV_NAMES_ERR_OP_RUN_TOO_LONG 37 /* "op run too long"                   lib.rs:237        */
#define FQXV_NAMES_ERR_OP_BOUND 38       /* "op columns exceed the stream bound" lib.rs:242        */
#define FQXV_NAMES_ERR_PLANE_LEN 39      /* "delta plane length mismatch"        lib.rs:160        */
#define FQXV_NAMES_ERR_MATCH 40          /* "MATCH without prior token"          lib.rs:606        */
#define FQXV_NAMES_ERR_STR_UNDERRUN 41   /* "string data underrun"               lib.rs:614        */
#define FQXV_NAMES_ERR_PAD_UNDERRUN 42   /* "pad underrun"                       lib.rs:588        */
#define FQXV_NAMES_ERR_DELTA_UNDERRUN 43 /* "delta underrun"                     lib.rs:631        */
#define FQXV_NAMES_ERR_DELTA_PRIOR 44    /* "DELTA without numeric prior"        lib.rs:637        */
#define FQXV_NAMES_ERR_UNKNOWN_OP 45     /* "unknown op"                         lib.rs:642        */
#define FQXV_NAMES_ERR_ALLOC 46          /* "op run too large to allocate"       lib.rs:245    */

/* Order tags (lib.rs:58-64; the stream's first byte, scalar.rs:25-26). */
#define FQXV_RANS_ORDER0 0
#define FQXV_RANS_ORDER1 1

typedef struct fqxv_rans_ctx fqxv_rans_ctx;

/* One block of a batched call.  Offsets index the caller's device arenas.
 *   decode: src = one rANS stream, dst = its output region; `max_len` is the
 *           decode_bounded bound (lib.rs:226-231; UINT64_MAX for plain
 *           decode), `dst_cap` plays the fallible allocation of
 *           model.rs:199-205; `order` is ignored (the stream carries it).
 *   encode: src = raw bytes, dst = output region of at least
 *           fqxv_rans_encode_bound(src_len, order) bytes; `max_len` ignored.
 */
typedef struct fqxv_rans_job {
    uint64_t src_off;
    uint64_t src_len;
    uint64_t dst_off;
    uint64_t dst_cap;
    uint64_t max_len;
    int32_t order;
    uint32_t flags; /* FQXV_RANS_JOB_* bits (decode); other bits must be 0 */
} fqxv_rans_job;

/* Decode: keep an order-1 block's distinct code rows in the wave's device
 * workspace (read through L1/L2) even when they fit in LDS, where they go
 * otherwise.  A staging choice for tuning (fqxv_amd.tune); the output is
 * the same either way.  No counterpart in the reference (a GPU placement). */
#define FQXV_RANS_JOB_O1_ROWS_WORKSPACE 1u

/* Exact message of a return code (the reference's Malformed strings). */
const char *fqxv_rans_strerror(int code);

/* Output capacity that always suffices for encode: header + 2n + 16 bytes of
 * slack (n = 0 -> 9).  Replaces the Vec growth inside scalar::encode_order0/1
 * (scalar.rs:29-111). */
size_t fqxv_rans_encode_bound(size_t n, int order);

/* lib.rs:207-210 `peek_output_len`: the declared n, or FQXV_RANS_ERR_TRUNCATED
 * when fewer than 9 bytes are given. */
int fqxv_rans_peek_len(const uint8_t *src, size_t len, uint64_t *n);

/* Context on HIP device `device` (the library owns a blocking stream for
 * the host-buffer calls and a lazily grown device workspace).  lib.rs has no
 * context: the Rust wrapper keeps one per thread. */
int fqxv_rans_ctx_create(int device, fqxv_rans_ctx **out);
void fqxv_rans_ctx_destroy(fqxv_rans_ctx *ctx);

/* lib.rs:111-127 `encode(src, order)`.  Host buffers.  Writes at most `cap`
 * bytes; *out_len = stream length.  cap >= fqxv_rans_encode_bound() always
 * suffices; a smaller cap fails with FQXV_RANS_ERR_DST_TOO_SMALL when the
 * stream does not fit. */
int fqxv_rans_encode(fqxv_rans_ctx *ctx, const uint8_t *src, size_t n, int order, uint8_t *dst,
                     size_t cap, size_t *out_len);

/* lib.rs:186-200 `decode(src)`.  Host buffers.  A declared n above `cap`
 * fails with FQXV_RANS_ERR_OUTPUT_ALLOC (the reference's fallible
 * allocation).  On any error *out_len = 0 and dst holds no partial output
 * guarantee. */
int fqxv_rans_decode(fqxv_rans_ctx *ctx, const uint8_t *src, size_t len, uint8_t *dst, size_t cap,
                     size_t *out_len);

/* lib.rs:226-231 `decode_bounded(src, max_len)`: FQXV_RANS_ERR_BOUND before
 * any device work when the declared n exceeds max_len. */
int fqxv_rans_decode_bounded(fqxv_rans_ctx *ctx, const uint8_t *src, size_t len, size_t max_len,
                             uint8_t *dst, size_t cap, size_t *out_len);

/* Device forms a decode can be pinned to: the single-block launch, or the
 * persistent batched kernel fed a one-job queue.  Both produce identical
 * output.  They stand where lib.rs:67-97 `Backend` (Scalar/Sse42/Avx2/
 * Avx512) stands. */
#define FQXV_RANS_BACKEND_GFX950 0
#define FQXV_RANS_BACKEND_GFX950_BATCHED 1

/* lib.rs:241-255 `decode_with(src, backend)`: decode through the named
 * device form.  Same arguments and errors as fqxv_rans_decode; an unknown
 * backend is FQXV_RANS_ERR_INVALID_ARG. */
int fqxv_rans_decode_with(fqxv_rans_ctx *ctx, int backend, const uint8_t *src, size_t len, uint8_t *dst,
                          size_t cap, size_t *out_len);

/* lib.rs:80-96 `Backend::detect()`: the form fqxv_rans_decode uses
 * (FQXV_RANS_BACKEND_GFX950). */
int fqxv_rans_backend_detect(void);

/* Batched device-resident encode/decode: one persistent-kernel launch codes
 * all `count` jobs.  All pointers are device pointers (d_jobs included).
 * `d_order` (optional, may be NULL) is a permutation of 0..count-1 giving the
 * order in which waves take jobs (largest first is best; see
 * fqxv_rans_lpt_order).  Per job: d_status[i] = return code,
 * d_out_len[i] = bytes written.  Asynchronous on `stream` (a hipStream_t;
 * NULL is the HIP null stream); the call is capturable into a hipGraph. */
int fqxv_rans_decode_batch(fqxv_rans_ctx *ctx, const uint8_t *d_src, uint8_t *d_dst,
                           const fqxv_rans_job *d_jobs, uint32_t count, const uint32_t *d_order,
                           int32_t *d_status, uint64_t *d_out_len, void *stream);
int fqxv_rans_encode_batch(fqxv_rans_ctx *ctx, const uint8_t *d_src, uint8_t *d_dst,
                           const fqxv_rans_job *d_jobs, uint32_t count, const uint32_t *d_order,
                           int32_t *d_status, uint64_t *d_out_len, void *stream);

/* Batched encode at both orders, keeping the smaller stream per job: the
 * callers' encode_best (crates/fqxv-tokenizer/src/lib.rs:91-95,
 * crates/fqxv/src/container/reorder.rs:597-607 keep order-0 on a tie,
 * crates/fqxv-lroverlap/src/codec.rs:292-303 keeps order-1), which today
 * encodes every stream twice on the CPU.  Per job, `order` is the order kept
 * on a tie and the destination region holds fqxv_rans_encode_best_bound(n)
 * bytes; the chosen stream starts at dst_off (its first byte is its order). */
int fqxv_rans_encode_best_batch(fqxv_rans_ctx *ctx, const uint8_t *d_src, uint8_t *d_dst,
                                const fqxv_rans_job *d_jobs, uint32_t count, const uint32_t *d_order,
                                int32_t *d_status, uint64_t *d_out_len, void *stream);
size_t fqxv_rans_encode_best_bound(size_t n);

/* The plain form: one kernel launch per block (the reference's per-call
 * granularity, lib.rs:111-200), device-resident, asynchronous on `stream`.
 * `job` is a host struct; status and out_len are device pointers. */
int fqxv_rans_decode_one_device(fqxv_rans_ctx *ctx, const uint8_t *d_src, uint8_t *d_dst,
                                const fqxv_rans_job *job, int32_t *d_status, uint64_t *d_out_len,
                                void *stream);
int fqxv_rans_encode_one_device(fqxv_rans_ctx *ctx, const uint8_t *d_src, uint8_t *d_dst,
                                const fqxv_rans_job *job, int32_t *d_status, uint64_t *d_out_len,
                                void *stream);

/* Batched call captured once into a hipGraph and replayed: the work-queue
 * reset kernel and the persistent kernel, arguments fixed at creation (the
 * graph owns its own queue counter).  `direction` is 0 for decode, 1 for
 * encode, 2 for encode-best.  No reference counterpart: the form the
 * container's per-block dispatch (decompress.rs:45-52) takes when the same
 * layout is decoded repeatedly.  A replay must not overlap another call on the same context;
 * the context must outlive its graphs.  The buffers passed at creation are
 * read and written by every replay. */
typedef struct fqxv_rans_graph fqxv_rans_graph;
int fqxv_rans_graph_create(fqxv_rans_ctx *ctx, int direction, const uint8_t *d_src, uint8_t *d_dst,
                           const fqxv_rans_job *d_jobs, uint32_t count, const uint32_t *d_order,
                           int32_t *d_status, uint64_t *d_out_len, fqxv_rans_graph **out);
int fqxv_rans_graph_launch(fqxv_rans_graph *g, void *stream);
void fqxv_rans_graph_destroy(fqxv_rans_graph *g);

/* Persistent-grid size for later batched calls and graphs on `ctx`
 * (workgroups of one wave; 0 restores the occupancy maximum).  Used by the
 * autotuner. */
int fqxv_rans_ctx_set_grid(fqxv_rans_ctx *ctx, uint32_t waves);

/* ---- Serving: random-access decode from an archive resident in HBM ------
 *
 * No reference counterpart (the reference decodes whole containers,
 * decompress.rs:45-52); this is the request loop of a decode service.
 * `blocks[i]` (host array, copied) locates block i of the resident archive
 * `d_archive`: src_off/src_len of its stream, dst_cap = its decoded length n,
 * max_len the decode_bounded bound.  Decoded output lives in the caller's
 * device buffer `d_pool` of `budget` bytes; a request is admitted (in arrival
 * order) when its output fits in the free pool and a slot is free.  Every
 * admitted request that fits is packed into one batched decode of at most
 * `max_batch` blocks; up to `slots` such launches run at once, each on its own
 * stream.  Output of a request: its blocks in request order, block i at
 * d_pool + out_off + sum over earlier blocks of round_up(n, 256).  The
 * region stays valid until
 * fqxv_rans_server_release(ticket).  Times are seconds since creation. */
typedef struct fqxv_rans_server fqxv_rans_server;
typedef struct fqxv_rans_done {
    uint64_t ticket;
    uint64_t out_off;   /* offset of the request's output in the pool */
    uint64_t out_bytes; /* bytes of pool held by the request */
    double t_arrival;   /* as passed to submit (or the submit time) */
    double t_admit;     /* launch of the batch that carried it */
    double t_done;      /* completion seen by fqxv_rans_server_poll */
    int32_t status;     /* FQXV_RANS_OK or the first failing block's code */
    uint32_t nblocks;
} fqxv_rans_done;

int fqxv_rans_server_create(int device, const uint8_t *d_archive, const fqxv_rans_job *blocks, uint32_t nblocks,
                            uint8_t *d_pool, uint64_t budget, uint32_t slots, uint32_t max_batch,
                            fqxv_rans_server **out);
/* The same service without launch boundaries: one persistent grid of
 * `waves` decode waves (0 = half of the chip's wave slots, which leaves room
 * for the kernels that consume the output; at most all slots but one) stays
 * resident and takes
 * blocks from a ring of `ring` jobs that the host appends to (pinned host
 * memory; a dispatcher wave mirrors the published count into device memory),
 * so an admitted request's blocks start as soon as a wave is free, whatever
 * else is running.  A request may name at most `ring` blocks.  The grid
 * leaves as a whole once nothing has been published for `idle_ms`
 * (1..60000) and every job is claimed, and is relaunched when work arrives;
 * destroy stops it.  Same submit/poll/take/release calls. */
int fqxv_rans_server_create_persistent(int device, const uint8_t *d_archive, const fqxv_rans_job *blocks,
                                       uint32_t nblocks, uint8_t *d_pool, uint64_t budget, uint32_t ring,
                                       uint32_t waves, uint32_t idle_ms, fqxv_rans_server **out);
void fqxv_rans_server_destroy(fqxv_rans_server *srv);
int fqxv_rans_server_now(fqxv_rans_server *srv, double *t);
/* Queue a request for blocks ids[0..k-1]; t_arrival < 0 means now. */
int fqxv_rans_server_submit(fqxv_rans_server *srv, const uint32_t *ids, uint32_t k, double t_arrival,
                            uint64_t *ticket);
/* Progress: completes finished launches, then admits queued requests. */
int fqxv_rans_server_poll(fqxv_rans_server *srv, uint32_t *queued, uint32_t *in_flight);
/* Pops up to `max` completion records. */
int fqxv_rans_server_take(fqxv_rans_server *srv, fqxv_rans_done *out, uint32_t max, uint32_t *got);
/* Returns a completed request's output region to the pool. */
int fqxv_rans_server_release(fqxv_rans_server *srv, uint64_t ticket);
/* Decode waves the service runs: the resident grid's (persistent form) or
 * each batched launch's (occupancy maximum), so a batched peak can be
 * measured on the same number of waves. */
int fqxv_rans_server_waves(fqxv_rans_server *srv, uint32_t *waves);

/* The reorder permutation's byte planes around the coder
 * (crates/fqxv/src/container/reorder.rs:597-607 encode side,
 * reorder.rs:1089-1100 decode side): plane k (bytes [k*n, (k+1)*n)) holds
 * byte k of every u32 index.  Device pointers, asynchronous on `stream`. */
int fqxv_rans_perm_split(const uint32_t *d_perm, uint64_t n, uint8_t *d_planes, void *stream);
int fqxv_rans_perm_join(const uint8_t *d_planes, uint64_t n, uint32_t *d_perm, void *stream);

/* CRC-32C (crates/fqxv/src/crc.rs: reflected Castagnoli 0x82F63B78, init and
 * xor-out 0xFFFFFFFF) of `count` byte ranges [off[i], off[i] + len[i]) of the
 * device arena `d_src` -- the container's frame, per-stream, whole-file and
 * footer checksums (crates/fqxv/src/container/mod.rs:53-97).  `off` and
 * `len` are host arrays (len < 2^47); d_crc[i] receives the checksum.  Work
 * runs on `stream`; the call returns when the checksums are written. */
int fqxv_rans_crc32c_batch(fqxv_rans_ctx *ctx, const uint8_t *d_src, const uint64_t *off, const uint64_t *len,
                           uint32_t count, uint32_t *d_crc, void *stream);

/* ---- Names streams (crates/fqxv-tokenizer/src/lib.rs, format 1) --------
 *
 * lib.rs:543-650 `decode(src)` for `count` names streams at once; stream i
 * is src[off[i], off[i] + len[i]) in host memory (the names section of a
 * block payload, crates/fqxv/src/container/mod.rs:88-95).  The framing is
 * walked on the host, every rANS sub-stream of every stream is decoded in
 * one batched device launch, and the names are rebuilt on the host.  The
 * return code reports argument and device failures only; each stream's own
 * result is fqxv_names_status (FQXV_RANS_OK, an FQXV_NAMES_ERR_* code, or a
 * sub-stream's FQXV_RANS_ERR_* code -- the first failing check in the
 * reference's order).  Synchronous. */
typedef struct fqxv_names_batch fqxv_names_batch;
int fqxv_names_decode_batch(fqxv_rans_ctx *ctx, const uint8_t *src, const uint64_t *off, const uint64_t *len,
                            uint32_t count, fqxv_names_batch **out);
int32_t fqxv_names_status(const fqxv_names_batch *b, uint32_t i);
/* Stream i's names as [u64 n] then n x ([u32 len][len bytes]), little-endian:
 * exactly the bytes the block's names digest covers
 * (crates/fqxv/src/container/block.rs:164-176).  Valid until destroy.
 * Returns the stream's status. */
int fqxv_names_records(const fqxv_names_batch *b, uint32_t i, const uint8_t **data, uint64_t *size,
                       uint64_t *n_names);
void fqxv_names_batch_destroy(fqxv_names_batch *b);
/* lib.rs:455-540 `encode(names)` for `count` blocks at once: block i's
 * names, in the records layout of fqxv_names_records, are
 * src[off[i], off[i] + len[i]) in host memory.  Tokenizing runs on the
 * host (threads across blocks); every sub-stream of every block is encoded
 * at both orders in one batched device launch, the smaller kept (order-0 on
 * a tie, lib.rs:88-95) -- the reference encodes each of them twice on the
 * CPU; the framing is written on the host.  Streams come back through
 * fqxv_names_output (byte-identical to the reference's encode).  A block
 * whose records are malformed gets FQXV_RANS_ERR_INVALID_ARG.
 * Synchronous. */
int fqxv_names_encode_batch(fqxv_rans_ctx *ctx, const uint8_t *src, const uint64_t *off, const uint64_t *len,
                            uint32_t count, fqxv_names_batch **out);
/* Block i's names stream of an encode batch; returns the block's status. */
int fqxv_names_output(const fqxv_names_batch *b, uint32_t i, const uint8_t **data, uint64_t *size);
/* The names content digest of every stream of a decoded batch: xxh3-64 over
 * its records (the value block.rs:164-176 stores and block.rs:1097-1109
 * compares), computed on the device.  out[i] = 0 for a failed stream. */
int fqxv_names_digests(fqxv_rans_ctx *ctx, const fqxv_names_batch *b, uint64_t *out);

/* XXH3-64 (seed 0; the XXH3_64bits of xxHash 0.8 that the reference's
 * xxhash-rust 0.8.18 `Xxh3` computes) of `count` byte ranges
 * [off[i], off[i] + len[i]) of the device arena `d_src`: the container's
 * per-stream content digests (crates/fqxv/src/container/block.rs:164-176,
 * mod.rs:85-87) and the reorder layout's output digest (reorder.rs:1580).
 * `off` and `len` are host arrays; d_out[i] (device) receives the digest.
 * Work runs on `stream`; the call returns when the digests are written. */
int fqxv_rans_xxh3_batch(fqxv_rans_ctx *ctx, const uint8_t *d_src, const uint64_t *off, const uint64_t *len,
                         uint32_t count, uint64_t *d_out, void *stream);

/* XXH3-64 of `count` byte strings, each the concatenation of several ranges
 * of the device arena `d_src`: digest i covers segments
 * seg_first[i] .. seg_first[i + 1] - 1, where segment s is
 * [seg_off[s], seg_off[s] + seg_len[s]), in that order.  The sequence and
 * quality digests of crates/fqxv/src/container/block.rs:164-196 (checked at
 * block.rs:1097-1109) hash `[n_reads u64][lens u32...][bytes]`, whose parts
 * live apart: with the header and the lengths in the arena, a block's digest
 * is three segments.  Host arrays (seg_first holds count + 1 entries);
 * d_out[i] (device) receives digest i.  The segments are staged
 * back to back on the device (the context owns the staging buffer) and
 * hashed as one range each.  Work runs on `stream`; the call returns when
 * the digests are written. */
int fqxv_rans_xxh3_concat(fqxv_rans_ctx *ctx, const uint8_t *d_src, const uint64_t *seg_off,
                          const uint64_t *seg_len, const uint32_t *seg_first, uint32_t count, uint64_t *d_out,
                          void *stream);

/* Host helper: job order, largest `sizes[i]` first (ties by index). */
void fqxv_rans_lpt_order(const uint64_t *sizes, uint32_t count, uint32_t *order_out);

/* Persistent-grid size used by the batched calls (workgroups of one wave). */
int fqxv_rans_max_waves(fqxv_rans_ctx *ctx, uint32_t *waves);

/* fqxv_rans_decode_batch for callers that cannot allocate device memory
 * (code under hipGraph capture, serving loops, harnesses that forbid
 * hipMalloc).  No reference counterpart.  Same jobs, order, status and
 * out_len contract, on the current HIP device, asynchronous on `stream`; no
 * context.  The work-queue counter and the per-wave workspaces come from a
 * fixed device array of the library (about 72 MiB, one copy per GPU, held
 * until the process ends) instead of allocations: the persistent grid is at
 * most fqxv_rans_noalloc_max_waves() waves (`waves` 0 = that maximum; more
 * is FQXV_RANS_ERR_SCRATCH_LIMIT), any number of jobs.  Calls that overlap
 * on one GPU share its array and race: order them (one stream).  The call
 * allocates nothing, on the device or the host. */
int fqxv_rans_decode_batch_noalloc(const uint8_t *d_src, uint8_t *d_dst, const fqxv_rans_job *d_jobs,
                                   uint32_t count, const uint32_t *d_order, int32_t *d_status,
                                   uint64_t *d_out_len, uint32_t waves, void *stream);
uint32_t fqxv_rans_noalloc_max_waves(void);

#ifdef __cplusplus
}
#endif
#endif
