// Library-internal entry points shared between the C ABI translation units
// (not exported through include/fqxv_rans_hip.h).
#pragma once
#include <cstdint>

#include "../../include/fqxv_rans_hip.h"

extern "C" {
// Runs `count` jobs whose sources lie in the host buffer h_src[0, src_len)
// against a destination arena of dst_total bytes, in one batched device
// launch -- mode 0 decodes, mode 2 encodes at both orders keeping the
// smaller stream (job.order = the order kept on a tie) -- and copies the
// arena to h_dst and the per-job status / output length back.  Synchronous.
// Job offsets index h_src and the arena.
int fqxv_run_jobs_host(fqxv_rans_ctx *c, int mode, const uint8_t *h_src, uint64_t src_len,
                       const fqxv_rans_job *h_jobs, uint32_t count, uint64_t dst_total, uint8_t *h_dst,
                       int32_t *h_status, uint64_t *h_out_len);

// fqxv_run_jobs_host (mode 0 decode, mode 2 encode-best) with the outputs
// packed: the job regions stay on the device; only the successful jobs'
// output bytes come back, job i's at (*h_packed)[h_pack_off[i]], so a job
// that fails holds no host memory.  *h_packed is allocated with malloc (the
// caller frees it); nothing is allocated on an error.  Synchronous.
int fqxv_jobs_host_packed(fqxv_rans_ctx *c, int mode, const uint8_t *h_src, uint64_t src_len,
                          const fqxv_rans_job *h_jobs, uint32_t count, uint64_t dst_total, uint8_t **h_packed,
                          uint64_t *h_pack_off, int32_t *h_status, uint64_t *h_out_len);

// xxh3-64 of `count` host byte ranges on the device (uploaded, hashed,
// digests copied to out).  Synchronous.
int fqxv_xxh3_host_ranges(fqxv_rans_ctx *c, const uint8_t *const *ptrs, const uint64_t *lens, uint32_t count,
                          uint64_t *out);
}
