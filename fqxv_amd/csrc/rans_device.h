// Device-side helpers shared by the gfx950 rANS Nx16 kernels.
//
// Everything here is written for 64-wide wavefronts: one workgroup is one
// wave, one rANS state lives in each of lanes 0..31, and cross-lane work uses
// ballots, v_mbcnt prefix counts and readlane/writelane.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fqxv_rans_hip.h"

typedef uint8_t u8;
typedef uint16_t u16;
typedef uint32_t u32;
typedef uint64_t u64;
typedef int32_t i32;
typedef int64_t i64;

// crates/fqxv-rans/src/model.rs:10-19
constexpr u32 SCALE_BITS = 12;
constexpr u32 TOTFREQ = 1u << SCALE_BITS;
constexpr u32 RANS_L = 1u << 16;
constexpr u32 NSTATES = 32;
// Stream layout sizes (scalar.rs:8-20).
constexpr u32 HDR0 = 9;                      // [u8 order][u64 n]
constexpr u32 O0_HDR = 9 + 512 + 4 * NSTATES; // order-0 header through the states
constexpr u32 BITMAP = 32;

// Per-wave global workspace (bytes).  Order-1 decode keeps its (slot x
// context) u8 code table in the first 2 MiB (at most 1 MiB, starting at a
// per-wave offset below 1 MiB) and, for more than 64 contexts, the u32
// freq/cum table after it; order-1 encode keeps 256x256 pair counts and the
// freq/cum table when the alphabet is too wide for LDS.
constexpr u64 WS_CODES = 256ull * 4096 * 2;  // 2 MiB
constexpr u64 WS_FC = 256ull * 256 * 4;      // 256 KiB
constexpr u64 WS_PER_WAVE = WS_CODES + WS_FC;

__device__ __forceinline__ u32 lane_id() { return __lane_id(); }

// The builtin on the bool itself: one v_cmp writing the lane mask (going
// through an int predicate first costs a v_cndmask and a second v_cmp).
__device__ __forceinline__ u64 ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// Number of set bits of `m` in lanes below this one (v_mbcnt_lo/hi).
__device__ __forceinline__ u32 rank_below(u64 m)
{
    return __builtin_amdgcn_mbcnt_hi((u32)(m >> 32), __builtin_amdgcn_mbcnt_lo((u32)m, 0u));
}

__device__ __forceinline__ u32 uni(u32 v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ u64 uni64(u64 v) { return (u64)uni((u32)v) | ((u64)uni((u32)(v >> 32)) << 32); }

// A job record read by the whole wave, made wave-uniform: every value derived
// from it (lengths, offsets, loop bounds) then lives in scalar registers and
// loops over it stay scalar loops.
__device__ __forceinline__ fqxv_rans_job uni_job(const fqxv_rans_job &j)
{
    fqxv_rans_job u;
    u.src_off = uni64(j.src_off);
    u.src_len = uni64(j.src_len);
    u.dst_off = uni64(j.dst_off);
    u.dst_cap = uni64(j.dst_cap);
    u.max_len = uni64(j.max_len);
    u.order = (i32)uni((u32)j.order);
    u.flags = uni(j.flags);
    return u;
}

// v_writelane_b32 (the LLVM intrinsic; clang exposes no builtin for it).
extern "C" __device__ int fqxv_llvm_writelane(int, int, int) __asm("llvm.amdgcn.writelane.i32");
__device__ __forceinline__ u32 writelane(u32 val, u32 lane, u32 old)
{
    return (u32)fqxv_llvm_writelane((int)uni(val), (int)lane, (int)old);
}

__device__ __forceinline__ u32 wave_sum(u32 v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ u64 wave_sum64(u64 v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ u32 wave_or(u32 v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o);
    return v;
}

// Inclusive prefix sum across the wave.
__device__ __forceinline__ u32 wave_scan_incl(u32 v)
{
    const u32 l = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        u32 t = __shfl_up(v, o);
        if (l >= (u32)o) v += t;
    }
    return v;
}

// Wave-wide argmax with ties broken by the smallest key: returns the key
// (an index) whose value is greatest, preferring the lowest index.
__device__ __forceinline__ void wave_argmax(u32 &val, u32 &key)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        u32 ov = __shfl_xor(val, o);
        u32 ok = __shfl_xor(key, o);
        if (ov > val || (ov == val && ok < key)) {
            val = ov;
            key = ok;
        }
    }
}

// The four bytes at p[0..3] (little endian), bytes at or past `end` read as
// zero and never touched.  Interior reads are aligned dword loads, realigned
// with v_alignbyte; the aligned dwords never start before the 4-byte
// boundary at or below p, which lies inside the same (>= 4-byte aligned)
// allocation.
__device__ __forceinline__ u32 load4(const u8 *p, const u8 *end)
{
    // Pointer arithmetic (not an integer round trip) keeps the global
    // address space, so these stay global_load rather than flat_load.
    const u32 ph = (u32)((uintptr_t)p & 3);
    const u32 *base = (const u32 *)(p - ph);
    if (ph == 0) {
        if (p + 4 <= end) return *base;
    } else if ((const u8 *)(base + 2) <= end) {
        return __builtin_amdgcn_alignbyte(base[1], base[0], ph); // shift in bytes
    }
    u32 v = 0;
    for (u32 i = 0; i < 4; i++)
        if (p + i < end) v |= (u32)p[i] << (8 * i);
    return v;
}

__device__ __forceinline__ void store_bytes(u8 *p, u32 v, u32 nbytes)
{
    for (u32 i = 0; i < nbytes; i++) p[i] = (u8)(v >> (8 * i));
}

// Wave-cooperative forward copy of `len` bytes with arbitrary alignment of
// both ends.  Safe for overlapping ranges when dst <= src: each batch reads
// its whole source window (8 dwords per lane, plus the unaligned neighbour)
// before any lane stores, and a batch's stores end below the next batch's
// source (dst + 2048 (b + 1) <= src + 2048 (b + 1)).
__device__ inline void wave_copy_fwd(u8 *dst, const u8 *src, u64 len)
{
    constexpr u32 CB = 8;
    const u32 l = lane_id();
    const u8 *send = src + len;
    // Destination head up to the next 4-byte boundary.
    u32 head = (u32)((4 - ((uintptr_t)dst & 3)) & 3);
    if (head > len) head = (u32)len;
    if (l < head) dst[l] = src[l];
    dst += head;
    src += head;
    len -= head;
    u32 *d32 = (u32 *)dst;
    const u64 nd = len / 4;
    // Bulk: source dword k as two aligned loads and a byte shift, both
    // inside [src, send) while 4k + 8 <= len.
    const u64 nbulk = len >= 8 ? (len - 8) / 4 + 1 : 0;
    const u32 ph = (u32)((uintptr_t)src & 3);
    const u32 *A = (const u32 *)(src - ph);
    u64 base = 0;
    for (; base + 64 * CB <= nbulk; base += 64 * CB) {
        u32 v[CB];
#pragma unroll
        for (u32 u = 0; u < CB; u++) {
            const u64 k = base + 64 * u + l;
            v[u] = __builtin_amdgcn_alignbyte(A[k + 1], A[k], ph);
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (u32 u = 0; u < CB; u++) d32[base + 64 * u + l] = v[u];
    }
    for (; base < nd; base += 64) {
        const u64 k = base + l;
        u32 v = 0;
        if (k < nd) v = load4(src + 4 * k, send);
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        if (k < nd) d32[k] = v;
    }
    const u32 tail = (u32)(len & 3);
    if (l < tail) dst[4 * nd + l] = src[4 * nd + l];
}
