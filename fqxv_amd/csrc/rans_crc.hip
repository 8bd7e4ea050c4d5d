// CRC-32C of many byte ranges of a device arena, for the container's
// integrity checks (crates/fqxv/src/crc.rs: reflected Castagnoli polynomial
// 0x82F63B78, init and xor-out 0xFFFFFFFF; frame, per-stream, whole-file and
// footer CRCs of crates/fqxv/src/container/mod.rs:53-97).
//
// The CRC register update is linear over GF(2): with raw(m) the register
// after m from a zero register and Z_n the operator "append n zero bytes",
//   register(a ++ b) = Z_|b|(register(a)) ^ raw(b),
//   crc32c(m)        = ~(Z_|m|(0xFFFFFFFF) ^ raw(m)).
// So a range is cut into 64 KiB chunks, one wave per chunk; each lane takes
// 1 KiB of the chunk and computes its raw value with slicing-by-8 tables in
// LDS; the 64 lane values are folded in a 6-level tree, the right half of a
// level-k pair being 2^(10+k) bytes long (one 32x32 operator, index uniform
// across the wave); a second kernel folds each range's chunk values in order
// and applies the init/xor-out terms (one wave per range, lanes taking every
// 64th chunk).  Operators for 2^i zero bytes
// (i < 48) and the eight slicing tables are computed on the host.
#include "rans_device.h"

namespace {

constexpr u32 CRC_CHUNK_LOG = 16;                 // 64 KiB per wave
constexpr u32 CRC_SEG = 1024;                     // bytes per lane
constexpr u32 CRC_OPS = 48;                       // operators for 2^0 .. 2^47 bytes
constexpr u32 SEG_LOG = 12;                       // chunks per segment (2^12 x 64 KiB = 256 MiB)
#ifndef CRC_GROUP
#define CRC_GROUP 4                               // 16-byte loads per group
#endif

struct CrcConst {
    u32 table[8][256];                            // slicing-by-8
    u32 op[CRC_OPS][32];                          // Z_{2^i}: column j = image of bit j
};

// Z(x) for one operator held in global memory (rows read per lane).
__device__ __forceinline__ u32 op_apply(const u32 *__restrict__ op, u32 x)
{
    u32 s = 0;
#pragma unroll 8
    for (u32 j = 0; j < 32; j++)
        if ((x >> j) & 1) s ^= op[j];
    return s;
}

// Z_n(x) for any n < 2^48 (product of the power-of-two operators).
__device__ u32 zeros_apply(const CrcConst *__restrict__ cc, u32 x, u64 n)
{
    for (u32 i = 0; n && i < CRC_OPS; i++, n >>= 1)
        if (n & 1) x = op_apply(cc->op[i], x);
    return x;
}

__device__ __forceinline__ u32 slice8(const u32 (*t)[256], u32 crc, u32 w0, u32 w1)
{
    crc ^= w0;
    return t[7][crc & 0xff] ^ t[6][(crc >> 8) & 0xff] ^ t[5][(crc >> 16) & 0xff] ^ t[4][crc >> 24] ^
           t[3][w1 & 0xff] ^ t[2][(w1 >> 8) & 0xff] ^ t[1][(w1 >> 16) & 0xff] ^ t[0][w1 >> 24];
}

// One wave per chunk item t (range r = item_range[t], chunk k = item_chunk[t]).
__global__ __launch_bounds__(64) void k_crc_chunks(const u8 *__restrict__ src, const u64 *__restrict__ off,
                                                   const u64 *__restrict__ len, const u32 *__restrict__ item_range,
                                                   const u32 *__restrict__ item_chunk, u32 nitems,
                                                   const CrcConst *__restrict__ cc, u32 *__restrict__ raw_out)
{
    __shared__ u32 t[8][256];
    const u32 l = lane_id();
    for (u32 i = l; i < 8 * 256; i += 64) (&t[0][0])[i] = (&cc->table[0][0])[i];
    __syncthreads();
    for (u32 it = blockIdx.x; it < nitems; it += gridDim.x) {
        const u32 r = item_range[it];
        const u64 cstart = (u64)item_chunk[it] << CRC_CHUNK_LOG;
        const u64 rem = len[r] - cstart;
        const u64 clen = rem < (1ull << CRC_CHUNK_LOG) ? rem : (1ull << CRC_CHUNK_LOG);
        const u64 a = (u64)l * CRC_SEG;
        const u32 n = a >= clen ? 0u : (u32)((clen - a) < CRC_SEG ? (clen - a) : CRC_SEG);
        const u8 *p = src + off[r] + cstart + a;
        u32 crc = 0;
        u32 i = 0;
        // Head bytes up to 16-byte alignment; groups of CRC_GROUP 16-byte
        // loads issued together (each lane walks its own segment, so a load
        // instruction touches 64 lines: taking a line's bytes in few loads
        // keeps it from being fetched again after an L1 eviction); then
        // 16-byte, 8-byte and single-byte tails.
        while (i < n && (((uintptr_t)(p + i)) & 15)) crc = t[0][(crc ^ p[i++]) & 0xff] ^ (crc >> 8);
        for (; i + 16 * CRC_GROUP <= n; i += 16 * CRC_GROUP) {
            uint4 w[CRC_GROUP];
#pragma unroll
            for (u32 g = 0; g < CRC_GROUP; g++) w[g] = *(const uint4 *)(p + i + 16 * g);
#pragma unroll
            for (u32 g = 0; g < CRC_GROUP; g++) {
                crc = slice8(t, crc, w[g].x, w[g].y);
                crc = slice8(t, crc, w[g].z, w[g].w);
            }
        }
        for (; i + 16 <= n; i += 16) {
            const uint4 w = *(const uint4 *)(p + i);
            crc = slice8(t, crc, w.x, w.y);
            crc = slice8(t, crc, w.z, w.w);
        }
        for (; i + 8 <= n; i += 8) {
            const uint2 w = *(const uint2 *)(p + i);
            crc = slice8(t, crc, w.x, w.y);
        }
        while (i < n) crc = t[0][(crc ^ p[i++]) & 0xff] ^ (crc >> 8);
        // Fold lanes: level k joins lane l (l % 2^(k+1) == 0) with lane
        // l + 2^k; the right group holds rlen bytes.
        u32 glen = n;
#pragma unroll
        for (u32 k = 0; k < 6; k++) {
            const u32 h = 1u << k;
            const u32 rcrc = (u32)__shfl_down((int)crc, h);
            const u32 rlen = (u32)__shfl_down((int)glen, h);
            if ((l & (2 * h - 1)) == 0 && rlen) {
                crc = rlen == (CRC_SEG << k) ? op_apply(cc->op[10 + k], crc) : (u32)zeros_apply(cc, crc, rlen);
                crc ^= rcrc;
                glen += rlen;
            }
        }
        if (l == 0) raw_out[it] = crc;
    }
}

// One wave per range of `len[r]` bytes made of pieces of 2^piece_log bytes
// (the last one shorter) whose raw values start at raw[first_item[r]].  Lane
// l folds pieces l, l+64, l+128, ... in order (Horner with the operator for
// 64 pieces, held in registers), shifts its value by the bytes that follow
// its last piece, and the wave XORs the lanes.  `finalize`: apply the
// init/xor-out terms and write the CRC; else write the range's raw value.
// Pieces are the 64 KiB chunks, or (for very long ranges, in a second
// level) segments of 2^SEG_LOG chunks folded by a first call.
__global__ __launch_bounds__(64) void k_crc_ranges(const u64 *__restrict__ len, const u32 *__restrict__ first_item,
                                                   u32 count, const CrcConst *__restrict__ cc,
                                                   const u32 *__restrict__ raw, u32 *__restrict__ crc_out,
                                                   u32 piece_log, int finalize)
{
    const u32 l = lane_id();
    u32 op22[32];
#pragma unroll
    for (u32 j = 0; j < 32; j++) op22[j] = cc->op[piece_log + 6][j];
    for (u32 r = blockIdx.x; r < count; r += gridDim.x) {
        const u64 n = len[r];
        const u64 nch = (n + (1ull << piece_log) - 1) >> piece_log;
        u32 acc = 0;
        u64 last = 0; // end of this lane's last chunk
        for (u64 k = l; k < nch; k += 64) {
            const u64 end = (k + 1) << piece_log;
            const u64 cend = end < n ? end : n;
            if (last) {
                const u64 d = cend - last;
                if (d == (64ull << piece_log)) {
                    u32 s2 = 0;
#pragma unroll
                    for (u32 j = 0; j < 32; j++)
                        if ((acc >> j) & 1) s2 ^= op22[j];
                    acc = s2;
                } else {
                    acc = zeros_apply(cc, acc, d);
                }
            }
            acc ^= raw[first_item[r] + k];
            last = cend;
        }
        if (last && last < n) acc = zeros_apply(cc, acc, n - last);
#pragma unroll
        for (u32 h = 32; h >= 1; h >>= 1) acc ^= (u32)__shfl_xor((int)acc, h);
        if (l == 0) crc_out[r] = finalize ? ~(zeros_apply(cc, 0xFFFFFFFFu, n) ^ acc) : acc;
    }
}

} // namespace

size_t fqxv_crc_const_bytes() { return sizeof(CrcConst); }
uint32_t fqxv_crc_seg_chunks() { return 1u << SEG_LOG; }

// Host: the slicing tables and the zero-byte operators (column j of Z_n is
// the register after n zero bytes from the single bit j).
void fqxv_crc_const_fill(void *dst)
{
    CrcConst *c = (CrcConst *)dst;
    for (u32 i = 0; i < 256; i++) {
        u32 v = i;
        for (int b = 0; b < 8; b++) v = (v & 1) ? (v >> 1) ^ 0x82F63B78u : v >> 1;
        c->table[0][i] = v;
    }
    for (u32 k = 1; k < 8; k++)
        for (u32 i = 0; i < 256; i++) c->table[k][i] = (c->table[k - 1][i] >> 8) ^ c->table[0][c->table[k - 1][i] & 0xff];
    // Z_1 (one zero byte): column j = update of the register holding bit j.
    for (u32 j = 0; j < 32; j++) {
        const u32 s = 1u << j;
        c->op[0][j] = c->table[0][s & 0xff] ^ (s >> 8);
    }
    // Z_{2^i} = Z_{2^(i-1)} o Z_{2^(i-1)}.
    for (u32 i = 1; i < CRC_OPS; i++)
        for (u32 j = 0; j < 32; j++) {
            u32 x = c->op[i - 1][j], s = 0;
            for (u32 b = 0; b < 32; b++)
                if ((x >> b) & 1) s ^= c->op[i - 1][b];
            c->op[i][j] = s;
        }
}

// `nseg` == 0: one fold per range over its chunk values.  Else the ranges
// are first cut into `nseg` segments of at most 2^SEG_LOG chunks
// (seg_len / seg_first: bytes and first chunk item of each; range r's
// segments start at range_seg[r]); a first fold gives each segment's raw
// value, a second folds each range's segments.
hipError_t fqxv_launch_crc(const u8 *src, const u64 *off, const u64 *len, const u32 *item_range,
                           const u32 *item_chunk, u32 nitems, const u32 *first_item, u32 count, const void *cc,
                           u32 *raw, u32 *crc_out, u32 grid, const u64 *seg_len, const u32 *seg_first, u32 nseg,
                           const u32 *range_seg, u32 *seg_raw, hipStream_t st)
{
    if (nitems)
        hipLaunchKernelGGL(k_crc_chunks, dim3(grid < nitems ? grid : nitems), dim3(64), 0, st, src, off, len,
                           item_range, item_chunk, nitems, (const CrcConst *)cc, raw);
    if (count && nseg == 0)
        hipLaunchKernelGGL(k_crc_ranges, dim3(grid < count ? grid : count), dim3(64), 0, st, len, first_item, count,
                           (const CrcConst *)cc, raw, crc_out, CRC_CHUNK_LOG, 1);
    if (count && nseg) {
        hipLaunchKernelGGL(k_crc_ranges, dim3(grid < nseg ? grid : nseg), dim3(64), 0, st, seg_len, seg_first, nseg,
                           (const CrcConst *)cc, raw, seg_raw, CRC_CHUNK_LOG, 0);
        hipLaunchKernelGGL(k_crc_ranges, dim3(grid < count ? grid : count), dim3(64), 0, st, len, range_seg, count,
                           (const CrcConst *)cc, seg_raw, crc_out, CRC_CHUNK_LOG + SEG_LOG, 1);
    }
    return hipGetLastError();
}
