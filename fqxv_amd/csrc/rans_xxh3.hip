// XXH3-64 (seed 0, default 192-byte secret) of many byte ranges of a device
// arena: the container's per-stream content digests
// (crates/fqxv/src/container/block.rs:164-176, mod.rs:85-87) and the reorder
// layout's output digest (reorder.rs:1580), which the reference computes with
// xxhash-rust 0.8.18's Xxh3 (the published XXH3_64bits of xxHash 0.8).
//
// Inputs above 240 bytes run the long loop: 1 KiB blocks of 16 stripes of 64
// bytes; stripe s adds, to accumulator lane i,
//     lo32(k) * hi32(k)  with k = word_i ^ secret[8s + 8i]
// and adds word_i to lane i ^ 1, then after each block every lane is
// scrambled on its own: a = ((a ^ a >> 47) ^ secret[128 + 8i]) * PRIME32_1.
// Within a block everything is a sum, so
//   k_xxh3_blocks: one thread per (block, lane) forms the block's eight
//                  lane sums -- all blocks of all ranges in parallel, an
//                  HBM-bound pass (1 KiB read, 64 B written per block);
//   k_xxh3_final:  eight lanes per range run the scramble chain over those
//                  sums (lane i reads only lane-i sums), add the last partial
//                  block and the last stripe, and merge; inputs of at most
//                  240 bytes take the short forms on one lane.
#include "rans_device.h"

namespace {

constexpr u32 XXH_BLOCK = 1024; // 16 stripes x 64 bytes with the default secret
constexpr u32 XXH_THREADS = 256;
#ifndef XXH_WG
#define XXH_WG 32768
#endif

__constant__ __attribute__((aligned(16))) u8 kSecret[192] = {
    0xb8, 0xfe, 0x6c, 0x39, 0x23, 0xa4, 0x4b, 0xbe, 0x7c, 0x01, 0x81, 0x2c, 0xf7, 0x21, 0xad, 0x1c,
    0xde, 0xd4, 0x6d, 0xe9, 0x83, 0x90, 0x97, 0xdb, 0x72, 0x40, 0xa4, 0xa4, 0xb7, 0xb3, 0x67, 0x1f,
    0xcb, 0x79, 0xe6, 0x4e, 0xcc, 0xc0, 0xe5, 0x78, 0x82, 0x5a, 0xd0, 0x7d, 0xcc, 0xff, 0x72, 0x21,
    0xb8, 0x08, 0x46, 0x74, 0xf7, 0x43, 0x24, 0x8e, 0xe0, 0x35, 0x90, 0xe6, 0x81, 0x3a, 0x26, 0x4c,
    0x3c, 0x28, 0x52, 0xbb, 0x91, 0xc3, 0x00, 0xcb, 0x88, 0xd0, 0x65, 0x8b, 0x1b, 0x53, 0x2e, 0xa3,
    0x71, 0x64, 0x48, 0x97, 0xa2, 0x0d, 0xf9, 0x4e, 0x38, 0x19, 0xef, 0x46, 0xa9, 0xde, 0xac, 0xd8,
    0xa8, 0xfa, 0x76, 0x3f, 0xe3, 0x9c, 0x34, 0x3f, 0xf9, 0xdc, 0xbb, 0xc7, 0xc7, 0x0b, 0x4f, 0x1d,
    0x8a, 0x51, 0xe0, 0x4b, 0xcd, 0xb4, 0x59, 0x31, 0xc8, 0x9f, 0x7e, 0xc9, 0xd9, 0x78, 0x73, 0x64,
    0xea, 0xc5, 0xac, 0x83, 0x34, 0xd3, 0xeb, 0xc3, 0xc5, 0x81, 0xa0, 0xff, 0xfa, 0x13, 0x63, 0xeb,
    0x17, 0x0d, 0xdd, 0x51, 0xb7, 0xf0, 0xda, 0x49, 0xd3, 0x16, 0x55, 0x26, 0x29, 0xd4, 0x68, 0x9e,
    0x2b, 0x16, 0xbe, 0x58, 0x7d, 0x47, 0xa1, 0xfc, 0x8f, 0xf8, 0xb8, 0xd1, 0x7a, 0xd0, 0x31, 0xce,
    0x45, 0xcb, 0x3a, 0x8f, 0x95, 0x16, 0x04, 0x28, 0xaf, 0xd7, 0xfb, 0xca, 0xbb, 0x4b, 0x40, 0x7e,
};

constexpr u64 P32_1 = 0x9E3779B1ull, P32_2 = 0x85EBCA77ull, P32_3 = 0xC2B2AE3Dull;
constexpr u64 P64_1 = 0x9E3779B185EBCA87ull, P64_2 = 0xC2B2AE3D27D4EB4Full, P64_3 = 0x165667B19E3779F9ull;
constexpr u64 P64_4 = 0x85EBCA77C2B2AE63ull, P64_5 = 0x27D4EB2F165667C5ull;
constexpr u64 MX1 = 0x165667919E3779F9ull, MX2 = 0x9FB21C651E98DF25ull;

// Little-endian loads at any alignment.
__device__ __forceinline__ u64 ld64(const u8 *p)
{
    u64 v = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) v |= (u64)p[k] << (8 * k);
    return v;
}
__device__ __forceinline__ u32 ld32(const u8 *p)
{
    return (u32)p[0] | ((u32)p[1] << 8) | ((u32)p[2] << 16) | ((u32)p[3] << 24);
}
__device__ __forceinline__ u64 sec64(u32 o) { return ld64(kSecret + o); }
// o a multiple of 8 (the secret is 16-byte aligned): one load.
__device__ __forceinline__ u64 sec64a(u32 o) { return *(const u64 *)(kSecret + o); }

__device__ __forceinline__ u64 fold64(u64 a, u64 b) { return (a * b) ^ __umul64hi(a, b); }
__device__ __forceinline__ u64 rotl64(u64 x, u32 r) { return (x << r) | (x >> (64 - r)); }
__device__ __forceinline__ u64 av64(u64 h)
{
    h ^= h >> 33;
    h *= P64_2;
    h ^= h >> 29;
    h *= P64_3;
    return h ^ (h >> 32);
}
__device__ __forceinline__ u64 av3(u64 h)
{
    h ^= h >> 37;
    h *= MX1;
    return h ^ (h >> 32);
}
__device__ __forceinline__ u64 mix16(const u8 *p, u32 so) { return fold64(ld64(p) ^ sec64(so), ld64(p + 8) ^ sec64(so + 8)); }

// XXH3_64bits for n <= 240 (one lane).
__device__ u64 xxh3_short(const u8 *p, u64 n)
{
    if (n == 0) return av64(sec64(56) ^ sec64(64));
    if (n <= 3) {
        const u32 c = ((u32)p[0] << 16) | ((u32)p[n >> 1] << 24) | (u32)p[n - 1] | ((u32)n << 8);
        return av64((u64)c ^ (u64)(ld32(kSecret) ^ ld32(kSecret + 4)));
    }
    if (n <= 8) {
        const u64 in64 = (u64)ld32(p + n - 4) + ((u64)ld32(p) << 32);
        u64 h = in64 ^ (sec64(8) ^ sec64(16));
        h ^= rotl64(h, 49) ^ rotl64(h, 24);
        h *= MX2;
        h ^= (h >> 35) + n;
        h *= MX2;
        return h ^ (h >> 28);
    }
    if (n <= 16) {
        const u64 lo = ld64(p) ^ (sec64(24) ^ sec64(32));
        const u64 hi = ld64(p + n - 8) ^ (sec64(40) ^ sec64(48));
        return av3(n + __builtin_bswap64(lo) + hi + fold64(lo, hi));
    }
    u64 acc = n * P64_1;
    if (n <= 128) {
        if (n > 32) {
            if (n > 64) {
                if (n > 96) acc += mix16(p + 48, 96) + mix16(p + n - 64, 112);
                acc += mix16(p + 32, 64) + mix16(p + n - 48, 80);
            }
            acc += mix16(p + 16, 32) + mix16(p + n - 32, 48);
        }
        acc += mix16(p, 0) + mix16(p + n - 16, 16);
        return av3(acc);
    }
    for (u32 i = 0; i < 8; i++) acc += mix16(p + 16 * i, 16 * i);
    acc = av3(acc);
    for (u32 i = 8; i < (u32)(n / 16); i++) acc += mix16(p + 16 * i, 16 * (i - 8) + 3);
    acc += mix16(p + n - 16, 136 - 17);
    return av3(acc);
}

// Lane i's contribution of one stripe at p (secret offset so): its own
// product term, and the word of lane i ^ 1 (fetched by that lane).
__device__ __forceinline__ u64 stripe_lane(u64 w, u64 w_pair, u64 key)
{
    const u64 k = w ^ key;
    return (u64)(u32)k * (k >> 32) + w_pair;
}

// Two stripe words of one lane: words 2m and 2m + 1 of a stripe are each
// other's pair, so they update accumulator lanes 2m and 2m + 1 without any
// cross-lane exchange.
__device__ __forceinline__ void stripe_pair(u64 w0, u64 w1, u64 k0, u64 k1, u64 &a0, u64 &a1)
{
    a0 += stripe_lane(w0, w1, k0);
    a1 += stripe_lane(w1, w0, k1);
}

// Eight-lane workers: worker w takes the blocks [w * per, (w + 1) * per)
// of the concatenation of all ranges' blocks (range r holds blocks
// [first[r], first[r + 1])), finds its first range by binary search and
// walks forward.  Lane j reads bytes 128t + 16j of the block (t < 8): words
// 2m, 2m + 1 (m = j & 3) of stripe 2t + (j >> 2), and sums them into
// accumulator lanes 2m, 2m + 1; lanes j and j ^ 4 (the two stripe parities)
// are added at the end of the block.  The eight lanes of a worker are
// always active together.
__global__ __launch_bounds__(XXH_THREADS) void k_xxh3_blocks(const u8 *__restrict__ src, const u64 *__restrict__ off,
                                                             const u64 *__restrict__ first, u32 count, u64 nitems,
                                                             u64 per, u64 *__restrict__ sums)
{
    const u32 j = threadIdx.x & 7, m = j & 3;
    const u64 w = ((u64)blockIdx.x * XXH_THREADS + threadIdx.x) >> 3;
    u64 it = w * per;
    const u64 end = it + per < nitems ? it + per : nitems;
    if (it >= end) return;
    u64 key0[8], key1[8];
#pragma unroll
    for (u32 t = 0; t < 8; t++) {
        const u32 so = 8 * (2 * t + (j >> 2)) + 16 * m;
        key0[t] = sec64a(so);
        key1[t] = sec64a(so + 8);
    }
    // Largest r with first[r] <= it (first[count] = nitems > it).
    u32 lo = 0, hi = count;
    while (hi - lo > 1) {
        const u32 mid = (lo + hi) >> 1;
        if (first[mid] <= it) lo = mid; else hi = mid;
    }
    u32 r = lo;
    u64 r_end = first[r + 1];
    const u8 *p = src + off[r] + (it - first[r]) * XXH_BLOCK;
    for (; it < end; it++, p += XXH_BLOCK) {
        if (it >= r_end) {
            do r++; while (it >= first[r + 1]);
            r_end = first[r + 1];
            p = src + off[r] + (it - first[r]) * XXH_BLOCK;
        }
        u64 a0 = 0, a1 = 0;
        if ((((uintptr_t)p) & 15) == 0) {
            const uint4 *q = (const uint4 *)p + j;
            uint4 v[8];
#pragma unroll
            for (u32 t = 0; t < 8; t++) v[t] = q[8 * t];
#pragma unroll
            for (u32 t = 0; t < 8; t++)
                stripe_pair((u64)v[t].x | ((u64)v[t].y << 32), (u64)v[t].z | ((u64)v[t].w << 32), key0[t], key1[t],
                            a0, a1);
        } else {
            // 16 bytes from aligned dwords: four when the block is 4-byte
            // aligned, else five and byte shifts.  Each dword read holds at
            // least one byte of the block.
            const u32 ph = (u32)((uintptr_t)p & 3);
            const u32 *A = (const u32 *)(p - ph) + 4 * j;
            u32 d[8][5];
            if (ph == 0) {
#pragma unroll
                for (u32 t = 0; t < 8; t++)
#pragma unroll
                    for (u32 e = 0; e < 4; e++) d[t][e] = A[32 * t + e];
#pragma unroll
                for (u32 t = 0; t < 8; t++)
                    stripe_pair((u64)d[t][0] | ((u64)d[t][1] << 32), (u64)d[t][2] | ((u64)d[t][3] << 32), key0[t],
                                key1[t], a0, a1);
            } else {
#pragma unroll
                for (u32 t = 0; t < 8; t++)
#pragma unroll
                    for (u32 e = 0; e < 5; e++) d[t][e] = A[32 * t + e];
#pragma unroll
                for (u32 t = 0; t < 8; t++) {
                    u32 x[4];
#pragma unroll
                    for (u32 e = 0; e < 4; e++) x[e] = __builtin_amdgcn_alignbyte(d[t][e + 1], d[t][e], ph);
                    stripe_pair((u64)x[0] | ((u64)x[1] << 32), (u64)x[2] | ((u64)x[3] << 32), key0[t], key1[t], a0, a1);
                }
            }
        }
        a0 += (u64)__shfl_xor((long long)a0, 4, 8);
        a1 += (u64)__shfl_xor((long long)a1, 4, 8);
        if (j < 4) {
            sums[it * 8 + 2 * m] = a0;
            sums[it * 8 + 2 * m + 1] = a1;
        }
    }
}

// Eight lanes per range r = t >> 3.
__global__ __launch_bounds__(XXH_THREADS) void k_xxh3_final(const u8 *__restrict__ src, const u64 *__restrict__ off,
                                                            const u64 *__restrict__ len,
                                                            const u64 *__restrict__ first, u32 count,
                                                            const u64 *__restrict__ sums, u64 *__restrict__ out)
{
    const u32 i = threadIdx.x & 7;
    const u64 r = ((u64)blockIdx.x * XXH_THREADS + threadIdx.x) >> 3;
    if (r >= count) return;
    const u8 *p = src + off[r];
    const u64 n = len[r];
    if (n <= 240) {
        if (i == 0) out[r] = xxh3_short(p, n);
        return;
    }
    // Initial accumulators {P32_3, P64_1, P64_2, P64_3, P64_4, P32_2, P64_5, P32_1}.
    u64 acc = i == 0 ? P32_3 : i == 1 ? P64_1 : i == 2 ? P64_2 : i == 3 ? P64_3 : i == 4 ? P64_4 : i == 5 ? P32_2
            : i == 6 ? P64_5 : P32_1;
    const u64 scr = sec64(128 + 8 * i);
    const u64 nb = (n - 1) / XXH_BLOCK;
    const u64 *s = sums + first[r] * 8 + i;
    u64 k = 0;
    // The chain: acc = scramble(acc + sum_k).  Sums are read in groups of
    // 16 blocks, the next group's loads issued before the current group's
    // 16 dependent steps, so a load's latency hides behind a group of steps.
    constexpr u32 G = 16;
    if (nb >= 2 * G) {
        u64 va[G], vb[G];
#pragma unroll
        for (u32 j = 0; j < G; j++) va[j] = s[(k + j) * 8];
        for (; k + 2 * G <= nb; k += G) {
#pragma unroll
            for (u32 j = 0; j < G; j++) vb[j] = s[(k + G + j) * 8];
#pragma unroll
            for (u32 j = 0; j < G; j++) {
                u64 a = acc + va[j];
                a ^= a >> 47;
                a ^= scr;
                acc = a * P32_1;
            }
#pragma unroll
            for (u32 j = 0; j < G; j++) va[j] = vb[j];
        }
        // va holds blocks k .. k + G - 1 (k + G <= nb).
#pragma unroll
        for (u32 j = 0; j < G; j++) {
            u64 a = acc + va[j];
            a ^= a >> 47;
            a ^= scr;
            acc = a * P32_1;
        }
        k += G;
    }
    for (; k < nb; k++) {
        u64 a = acc + s[k * 8];
        a ^= a >> 47;
        a ^= scr;
        acc = a * P32_1;
    }
    // Stripes of the last, partial block, then the last stripe.
    const u8 *tail = p + nb * XXH_BLOCK;
    const u64 ns = ((n - 1) - nb * XXH_BLOCK) / 64;
    for (u64 st = 0; st < ns; st++)
        acc += stripe_lane(ld64(tail + 64 * st + 8 * i), ld64(tail + 64 * st + 8 * (i ^ 1)), sec64(8 * st + 8 * i));
    const u8 *last = p + n - 64;
    acc += stripe_lane(ld64(last + 8 * i), ld64(last + 8 * (i ^ 1)), sec64(192 - 64 - 7 + 8 * i));
    // Merge: pairs (2j, 2j + 1) folded against secret[11 + 16j].
    const u64 other = (u64)__shfl_xor((long long)acc, 1, 8);
    u64 m = (i & 1) ? 0 : fold64(acc ^ sec64(11 + 8 * i), other ^ sec64(11 + 8 * i + 8));
    m += (u64)__shfl_xor((long long)m, 2, 8);
    m += (u64)__shfl_xor((long long)m, 4, 8);
    if (i == 0) out[r] = av3(n * P64_1 + m);
}

} // namespace

hipError_t fqxv_launch_xxh3(const u8 *src, const u64 *off, const u64 *len, const u64 *first, u32 count, u64 nitems,
                            u64 *sums, u64 *out, hipStream_t st)
{
    if (nitems) {
        // About XXH_WG workgroups of 32 workers, at least one block each.
        const u64 workers_max = (u64)XXH_WG * (XXH_THREADS / 8);
        const u64 per = (nitems + workers_max - 1) / workers_max;
        const u64 workers = (nitems + per - 1) / per;
        const u64 blocks = (workers * 8 + XXH_THREADS - 1) / XXH_THREADS;
        hipLaunchKernelGGL(k_xxh3_blocks, dim3((u32)blocks), dim3(XXH_THREADS), 0, st, src, off, first, count, nitems,
                           per, sums);
    }
    if (count) {
        const u64 blocks = ((u64)count * 8 + XXH_THREADS - 1) / XXH_THREADS;
        hipLaunchKernelGGL(k_xxh3_final, dim3((u32)blocks), dim3(XXH_THREADS), 0, st, src, off, len, first, count,
                           sums, out);
    }
    return hipGetLastError();
}
