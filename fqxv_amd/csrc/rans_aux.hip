// Byte-plane transposes around the coder for the reorder permutation, and
// the packing of batched outputs for host callers.
//
// crates/fqxv/src/container/reorder.rs:597-607 stores the global permutation
// (n u32 read indices) as four byte planes of n bytes each -- plane k holds
// byte k of every index -- and rANS-codes the 4n bytes; reorder.rs:1089-1100
// decodes them back.  These kernels are the two transposes, device-resident:
// HBM-bound (8n bytes per launch), four indices per thread, plane rows
// written and read as whole dwords when n is a multiple of 4.
#include "rans_device.h"

namespace {

constexpr u32 AUX_THREADS = 256;

// `words`: n is a multiple of 4, perm is 16-byte and planes 4-byte aligned.
__global__ __launch_bounds__(AUX_THREADS) void k_perm_split(const u32 *__restrict__ perm, u64 n,
                                                            u8 *__restrict__ planes, bool words)
{
    const u64 stride = (u64)gridDim.x * AUX_THREADS;
    const u64 groups = n / 4;
    for (u64 g = (u64)blockIdx.x * AUX_THREADS + threadIdx.x; g < (n + 3) / 4; g += stride) {
        const u64 i = 4 * g;
        if (words && g < groups) {
            const uint4 q = ((const uint4 *)perm)[g];
#pragma unroll
            for (u32 k = 0; k < 4; k++) {
                const u32 s = 8 * k;
                const u32 w = ((q.x >> s) & 0xff) | (((q.y >> s) & 0xff) << 8) | (((q.z >> s) & 0xff) << 16) |
                              (((q.w >> s) & 0xff) << 24);
                ((u32 *)(planes + k * n))[g] = w;
            }
        } else {
            for (u64 e = i; e < i + 4 && e < n; e++) {
                const u32 v = perm[e];
#pragma unroll
                for (u32 k = 0; k < 4; k++) planes[k * n + e] = (u8)(v >> (8 * k));
            }
        }
    }
}

__global__ __launch_bounds__(AUX_THREADS) void k_perm_join(const u8 *__restrict__ planes, u64 n,
                                                           u32 *__restrict__ perm, bool words)
{
    const u64 stride = (u64)gridDim.x * AUX_THREADS;
    const u64 groups = n / 4;
    for (u64 g = (u64)blockIdx.x * AUX_THREADS + threadIdx.x; g < (n + 3) / 4; g += stride) {
        const u64 i = 4 * g;
        if (words && g < groups) {
            u32 w[4];
#pragma unroll
            for (u32 k = 0; k < 4; k++) w[k] = ((const u32 *)(planes + k * n))[g];
            uint4 q;
            q.x = (w[0] & 0xff) | ((w[1] & 0xff) << 8) | ((w[2] & 0xff) << 16) | ((w[3] & 0xff) << 24);
            q.y = ((w[0] >> 8) & 0xff) | (((w[1] >> 8) & 0xff) << 8) | (((w[2] >> 8) & 0xff) << 16) |
                  (((w[3] >> 8) & 0xff) << 24);
            q.z = ((w[0] >> 16) & 0xff) | (((w[1] >> 16) & 0xff) << 8) | (((w[2] >> 16) & 0xff) << 16) |
                  (((w[3] >> 16) & 0xff) << 24);
            q.w = (w[0] >> 24) | ((w[1] >> 24) << 8) | ((w[2] >> 24) << 16) | ((w[3] >> 24) << 24);
            ((uint4 *)perm)[g] = q;
        } else {
            for (u64 e = i; e < i + 4 && e < n; e++)
                perm[e] = (u32)planes[e] | ((u32)planes[n + e] << 8) | ((u32)planes[2 * n + e] << 16) |
                          ((u32)planes[3 * n + e] << 24);
        }
    }
}

bool aux_words(const void *perm, const void *planes, u64 n)
{
    return (n & 3) == 0 && ((uintptr_t)perm & 15) == 0 && ((uintptr_t)planes & 3) == 0;
}

u32 aux_grid(u64 n)
{
    const u64 groups = (n + 3) / 4;
    const u64 blocks = (groups + AUX_THREADS - 1) / AUX_THREADS;
    return (u32)(blocks < 65536 ? (blocks ? blocks : 1) : 65536);
}

// One wave per job (grid-stride): job i's output, out_len[i] bytes at
// jobs[i].dst_off of the arena, to packed + pack_off[i]; a failed job copies
// nothing.  The arena and the packed buffer do not overlap.
__global__ __launch_bounds__(64) void k_pack_jobs(const u8 *__restrict__ arena, const fqxv_rans_job *__restrict__ jobs,
                                                  const u64 *__restrict__ out_len, const i32 *__restrict__ status,
                                                  const u64 *__restrict__ pack_off, u32 count, u8 *__restrict__ packed)
{
    for (u32 i = blockIdx.x; i < count; i += gridDim.x) {
        if (uni((u32)status[i]) != 0) continue;
        wave_copy_fwd(packed + uni64(pack_off[i]), arena + uni64(jobs[i].dst_off), uni64(out_len[i]));
    }
}

// One wave per item (grid-stride): item i copies len[i] bytes from
// src + src_off[i] to dst + dst_off[i].  Items are pieces of at most 64 KiB
// of the segments that a concatenated digest covers; the source arena and the
// staging buffer do not overlap.
__global__ __launch_bounds__(64) void k_gather_items(const u8 *__restrict__ src, const u64 *__restrict__ src_off,
                                                     const u64 *__restrict__ dst_off, const u32 *__restrict__ len,
                                                     u32 count, u8 *__restrict__ dst)
{
    for (u32 i = blockIdx.x; i < count; i += gridDim.x)
        wave_copy_fwd(dst + uni64(dst_off[i]), src + uni64(src_off[i]), uni(len[i]));
}

} // namespace

hipError_t fqxv_launch_gather_items(const u8 *src, const u64 *src_off, const u64 *dst_off, const u32 *len, u32 count,
                                    u8 *dst, hipStream_t st)
{
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_items, dim3(count < 16384 ? count : 16384), dim3(64), 0, st, src, src_off, dst_off,
                       len, count, dst);
    return hipGetLastError();
}

hipError_t fqxv_launch_pack_jobs(const u8 *arena, const fqxv_rans_job *jobs, const u64 *out_len, const i32 *status,
                                 const u64 *pack_off, u32 count, u8 *packed, hipStream_t st)
{
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack_jobs, dim3(count < 4096 ? count : 4096), dim3(64), 0, st, arena, jobs, out_len, status,
                       pack_off, count, packed);
    return hipGetLastError();
}

hipError_t fqxv_launch_perm_split(const u32 *perm, u64 n, u8 *planes, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_perm_split, dim3(aux_grid(n)), dim3(AUX_THREADS), 0, st, perm, n, planes,
                       aux_words(perm, planes, n));
    return hipGetLastError();
}

hipError_t fqxv_launch_perm_join(const u8 *planes, u64 n, u32 *perm, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_perm_join, dim3(aux_grid(n)), dim3(AUX_THREADS), 0, st, planes, n, perm,
                       aux_words(perm, planes, n));
    return hipGetLastError();
}
