// gfx950 rANS Nx16 encoder (order-0 and order-1).
//
// Byte-identical to crates/fqxv-rans/src/scalar.rs:29-140 (encode_order0,
// encode_order1, encode_symbol, finish) with normalize / EncSym from
// model.rs:70-264.
//
// One workgroup = one wave64 encodes one block at a time (persistent grid,
// device work queue).  Per block:
//   1. histogram with LDS atomics (order-1: pair counts remapped to an
//      (alphabet + 1) x alphabet table in LDS when that fits TAB_CAP entries,
//      i.e. up to 42 symbols, global atomics otherwise);
//   2. normalize (model.rs:213-264) as wave reductions -- the first index with
//      the strictly greatest frequency absorbs the residual, exactly as the
//      reference's sequential scan does;
//   3. per (context, symbol) constants: f << 12 | cum and the reciprocal
//      magic = floor((2^32-1)/f), which gives q0 in {q-1, q} for every
//      32-bit state, so one compare picks the exact step (any exact
//      division gives the reference's bytes, model.rs:313-365);
//   4. the reverse walk: lanes 0..31 hold the states; rounds run from the
//      last to the first; a ballot + v_mbcnt rank places each lane's
//      renormalisation word (at most one per step, scalar.rs:120-127).  The
//      reference collects words in descending position order and reverses
//      them (scalar.rs:130-140); here each word is written straight to its
//      final offset counted back from the end of the output region, then the
//      word block is slid down behind the states.
// The input of the walk is staged through LDS 2 KiB at a time, double
// buffered, so the state chain never waits on HBM.
#include "rans_device.h"

namespace {

// Dense order-1 table: rows = the alphabet's symbols plus one row for
// context 0 when symbol 0 is absent, columns = the alphabet.  Its capacity
// keeps the whole shared block within 20 KiB, so two waves fit per SIMD.
constexpr u32 TAB_CAP = 1808;
// Table entry of lanes with no position in a full walk round (lanes 32..63
// of a one-order walk): f = 4096, which no state reaches, so the step leaves
// x unchanged and emits nothing.  Above every table the walks index (dense
// order-1 up to 43 x 42 = 1806 entries, both orders up to 42 x 41 + 41).
constexpr u32 TAB_IDLE = TAB_CAP - 1;
static_assert(43 * 42 <= TAB_IDLE && 42 * 41 + 41 <= TAB_IDLE, "idle entry above every table");
constexpr u32 CH = 2048;                 // walk staging chunk (positions)
constexpr u32 CH_DW = CH / 4 + 1;        // + 4 leading bytes (context of the first)
#ifndef WALK_BATCH
#define WALK_BATCH 16                     // rounds whose table entries are read together
#endif
#ifndef HB
#define HB 16                             // histogram: dwords per lane loaded together
#endif

// `tab` holds the counts of a block first (u32 per entry, from the front:
// sub-histograms, dense pair counts), then its coding table: entry e is the
// pair {f << 12 | cum, reciprocal magic} at words 2e, 2e + 1, so a walk step
// reads its entry with one 64-bit load.  The table is written over the
// counts row by row, from the last row to the first (entry e overwrites
// counts 2e and 2e + 1, which belong to rows already done).
struct EncShared {
    alignas(16) u32 tab[2 * TAB_CAP];
    u32 hist[256];
    u8 colmap[256];         // symbol -> dense column (order-1 dense mode)
    u8 rowmap[256];         // context -> dense row
    u8 ctx_of[256];         // k-th present context (ascending)
    // + slack: lanes past the chunk read inside it; rows of whole 16-byte
    // units (stage_put writes 16 bytes per lane).
    alignas(16) u32 stage[2][(CH_DW + 15) & ~3u];
};
static_assert(sizeof(EncShared) <= 20480, "two one-wave workgroups per SIMD need <= 20 KiB of LDS each");
__device__ __forceinline__ u32 *counts(EncShared &sh) { return sh.tab; }
// a * b + c on 24-bit a, b as one full-rate multiply-add (the compiler
// otherwise picks the 64-bit multiply-add for small known operands).
__device__ __forceinline__ u32 mad24(u32 a, u32 b, u32 c)
{
    u32 d;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
}
typedef __attribute__((address_space(3))) const u64 lds_u64;
// The LDS address of a shared-memory object, and a 64-bit load at one.
__device__ __forceinline__ u32 lds_addr(const u32 *p)
{
    return (u32)(uintptr_t)(__attribute__((address_space(3))) const u32 *)p;
}
__device__ __forceinline__ uint2 lds_load2(u32 a)
{
    const u64 v = *(lds_u64 *)(uintptr_t)a;
    return make_uint2((u32)v, (u32)(v >> 32));
}
__device__ __forceinline__ uint2 *entries(EncShared &sh) { return (uint2 *)sh.tab; }
__device__ __forceinline__ const uint2 *entries(const EncShared &sh) { return (const uint2 *)sh.tab; }

// Workspace of one wave (WS_PER_WAVE bytes), for alphabets past the dense
// table: order-1 counts (256 KiB), then the entries {f << 12 | cum, magic}
// of all 256 x 256 pairs (512 KiB).
constexpr u64 WS_ENT = 256ull * 256 * 4;
static_assert(WS_ENT + 256ull * 256 * 8 <= WS_PER_WAVE, "encode workspace");

// Alphabet sizes whose dense table fits: (a + 1) * a entries.
__device__ __forceinline__ bool dense_fits(u32 a) { return (a + 1) * a <= TAB_CAP; }

// Histogram pre-pass of a batch's long jobs.  A long job's byte and pair
// counts on its own wave take several ms while the whole grid is busy (the
// LDS atomics of every wave compete), at the start of the batch's critical
// path.  The pre-pass cuts the long jobs (the queue's first PRE_MAX
// positions of at least PRE_MIN bytes: with the largest-first order, the
// batch's longest jobs) into PRE_SLICE slices counted by every wave of the
// grid, then the encode kernel takes the counts instead of counting.
#ifndef PRE_MIN_MIB
#define PRE_MIN_MIB 2
#endif
constexpr u64 PRE_MIN = (u64)PRE_MIN_MIB << 20;
constexpr u32 PRE_MAX = 2048;
constexpr u32 PRE_PER_LANE = PRE_MAX / 64; // plan: queue positions per lane
constexpr u64 PRE_SLICE = 256ull << 10;
constexpr u32 PRE_SLOT = 256 + TAB_CAP; // per long job: byte counts, dense pair counts
struct EncPre {
    u32 nbig;                       // long jobs: queue positions [0, nbig)
    u32 nslices;
    u32 ctr[2];                     // slice queues of the two counting passes
    u32 slice_start[PRE_MAX + 4];   // first slice of each long job; [nbig..] = nslices
    u32 slot[PRE_MAX * PRE_SLOT];
};

// Bytes of p[0..3] that fall in [begin, end), others zero.
__device__ __forceinline__ u32 load4b(const u8 *p, const u8 *begin, const u8 *end)
{
    if (p >= begin) return load4(p, end);
    u32 v = 0;
    for (u32 i = 0; i < 4; i++)
        if (p + i >= begin && p + i < end) v |= (u32)p[i] << (8 * i);
    return v;
}

// floor(c * 4096 / total), exact, via double with one correction each way.
__device__ __forceinline__ u32 scaled(u32 c, u64 total)
{
    const u64 num = (u64)c << SCALE_BITS;
    u64 q = (u64)((double)num / (double)total);
    if (q * total > num) q--;
    if ((q + 1) * total <= num) q++;
    return (u32)q;
}

__device__ __forceinline__ u32 pick4(const u32 v[4], u32 i)
{
    return i == 0 ? v[0] : i == 1 ? v[1] : i == 2 ? v[2] : v[3];
}

// model.rs:213-264 over 256 counts, lane l owning symbols 4l..4l+3.
__device__ __forceinline__ void normalize_wave(const u32 c[4], u32 f[4])
{
    const u32 l = lane_id();
    const u64 total = wave_sum64((u64)c[0] + c[1] + c[2] + c[3]);
    for (int i = 0; i < 4; i++) f[i] = 0;
    if (total == 0) return;
    u32 bv = 0, bk = 0xffffffffu;
    for (int i = 0; i < 4; i++) {
        if (c[i] == 0) continue;
        u32 v = scaled(c[i], total);
        if (v == 0) v = 1;
        f[i] = v;
        if (v > bv) {
            bv = v;
            bk = 4 * l + i;
        }
    }
    const u32 sum = wave_sum(f[0] + f[1] + f[2] + f[3]);
    wave_argmax(bv, bk);
    u32 max_s = uni(bk);
    int diff = (int)TOTFREQ - (int)sum;
    while (diff != 0) {
        const u32 owner = max_s >> 2, idx = max_s & 3;
        const u32 cur = uni(__shfl(pick4(f, idx), owner));
        if (diff > 0) {
            if (l == owner) {
                for (int i = 0; i < 4; i++)
                    if ((u32)i == idx) f[i] = cur + (u32)diff;
            }
            diff = 0;
        } else {
            int take = (int)cur - 1;
            if (take < 0) take = 0;
            if (take > -diff) take = -diff;
            if (l == owner) {
                for (int i = 0; i < 4; i++)
                    if ((u32)i == idx) f[i] = cur - (u32)take;
            }
            diff += take;
            if (diff != 0) {
                // First index with the strictly greatest frequency above 1.
                u32 v2 = 1, k2 = 0xffffffffu;
                for (int i = 0; i < 4; i++)
                    if (f[i] > v2) {
                        v2 = f[i];
                        k2 = 4 * l + i;
                    }
                wave_argmax(v2, k2);
                max_s = uni(k2);
                if (max_s == 0xffffffffu) return; // unreachable for sum > TOTFREQ
            }
        }
    }
}

__device__ __forceinline__ u32 magic_of(u32 f) { return f ? 0xffffffffu / f : 0u; }

// Write 256 freqs (lane-distributed) as little-endian u16 at p.
__device__ __forceinline__ void put_freqs(u8 *p, const u32 f[4])
{
    const u32 l = lane_id();
    u8 *q = p + 8 * l;
    for (int i = 0; i < 4; i++) {
        q[2 * i] = (u8)f[i];
        q[2 * i + 1] = (u8)(f[i] >> 8);
    }
}

// Stage positions [p0 - 4, p0 + CH) of `in` (bytes outside [in, in+n) = 0):
// CH_DW dwords, lane l holding stage dwords 8l .. 8l + 8 (the ninth only on
// lane 63, the last dword).  The loads of a chunk are issued a chunk ahead
// of their use and must not be waited for where they are issued.  A chunk
// wholly inside the input (every one but the block's first and, rarely, the
// one below a top chunk of under 4 positions) loads ten aligned dwords per
// lane with three wide loads and no per-lane branch, kept raw until
// stage_bytes realigns them; the others take the byte-exact path (bounds per
// byte, waited at once) into the same registers with no shift.
struct StageLoad {
    u32 a[10];
    u32 ph; // byte shift of the stage's start in its aligned dword
};
typedef u32 v4u32 __attribute__((ext_vector_type(4)));
typedef u32 v2u32 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) const v4u32 gu128;
typedef __attribute__((address_space(1))) const v2u32 gu64;
__device__ __forceinline__ void stage_fetch(const u8 *in, u64 n, u64 p0, StageLoad &sl)
{
    const u32 l = lane_id();
    const u8 *g = in + p0 - 4;
    if (p0 >= 4 && p0 + CH + 4 <= n) {
        // Aligned dwords [8l, 8l + 10) from g rounded down: bytes up to
        // g + 4 * (CH_DW + 1) <= in + p0 + CH + 4, inside [in, in + n).
        sl.ph = (u32)((uintptr_t)g & 3);
        const u32 *A = (const u32 *)(g - sl.ph) + 8 * l;
        const v4u32 v0 = *(gu128 *)(uintptr_t)A, v1 = *(gu128 *)(uintptr_t)(A + 4);
        const v2u32 v2 = *(gu64 *)(uintptr_t)(A + 8);
        sl.a[0] = v0.x, sl.a[1] = v0.y, sl.a[2] = v0.z, sl.a[3] = v0.w;
        sl.a[4] = v1.x, sl.a[5] = v1.y, sl.a[6] = v1.z, sl.a[7] = v1.w;
        sl.a[8] = v2.x, sl.a[9] = v2.y;
    } else {
        sl.ph = 0;
#pragma unroll
        for (u32 j = 0; j < 9; j++) {
            const u32 t = 8 * l + j;
            sl.a[j] = t < CH_DW ? load4b(g + 4 * t, in, in + n) : 0u;
        }
        sl.a[9] = 0;
    }
}
__device__ __forceinline__ void stage_bytes(const StageLoad &sl, u32 (&r)[9])
{
#pragma unroll
    for (u32 j = 0; j < 9; j++) r[j] = __builtin_amdgcn_alignbyte(sl.a[j + 1], sl.a[j], sl.ph);
}

// Dense order-1: every staged byte becomes its dense id (rowmap).  A present
// symbol has the same row and column id, and the zero fill before position 0
// becomes context 0's row, so a staged pair (context, symbol) indexes the
// (alphabet + 1) x alphabet table directly.
__device__ __forceinline__ void stage_remap(const EncShared &sh, u32 (&r)[9])
{
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const u32 w = r[i];
        r[i] = (u32)sh.rowmap[w & 0xff] | (u32)sh.rowmap[(w >> 8) & 0xff] << 8 |
               (u32)sh.rowmap[(w >> 16) & 0xff] << 16 | (u32)sh.rowmap[w >> 24] << 24;
    }
}

__device__ __forceinline__ void stage_put(u32 *buf, const u32 (&r)[9])
{
    const u32 l = lane_id();
    uint4 *d = (uint4 *)(buf + 8 * l);
    d[0] = make_uint4(r[0], r[1], r[2], r[3]);
    d[1] = make_uint4(r[4], r[5], r[6], r[7]);
    if (8 * l + 8 < CH_DW) buf[8 * l + 8] = r[8];
}

// Encode-walk table access: the walk's MODE is 0 (order-0, entry = symbol),
// 1 (dense order-1 in LDS, entry = row * cols + column) or 2 (order-1 in the
// workspace, entry = context * 256 + symbol).
struct TabRef {
    const uint2 *gent;  // mode 2: the workspace entries
    u32 cols;           // mode 1: the alphabet size (row stride)
};


// All 16 values of a batch live before any is used: the table reads of a
// batch then share one wait (the scheduler otherwise waits on each read in
// turn in the larger kernels).
static_assert(WALK_BATCH == 16, "pin16 covers one batch");
__device__ __forceinline__ void pin16(u32 (&a)[16])
{
    asm volatile("" : "+v"(a[0]), "+v"(a[1]), "+v"(a[2]), "+v"(a[3]), "+v"(a[4]), "+v"(a[5]), "+v"(a[6]),
                 "+v"(a[7]), "+v"(a[8]), "+v"(a[9]), "+v"(a[10]), "+v"(a[11]), "+v"(a[12]), "+v"(a[13]),
                 "+v"(a[14]), "+v"(a[15]));
}

// One round's table entry for lane l at position i (staged chunk sb, chunk
// start p0).  Inactive lanes get f = 4096 (never renormalise, x unchanged).
// Every read is unconditional: lanes past the chunk read the stage slack,
// and an inactive lane's table index is forced to 0.  Modes 0 and 1 read
// the LDS tables directly, mode 2 the workspace tables.
template <u32 MODE>
__device__ __forceinline__ void round_entry(const EncShared &sh, const TabRef &tr, const u8 *sb, u64 p0, u64 r,
                                            u64 n, u32 &fcm, u32 &mg)
{
    const u32 l = lane_id();
    const u64 i = r * NSTATES + l;
    const bool act = l < NSTATES && i < n;
    // Lanes 32..63 address as lanes 0..31: every read stays within the
    // chunk's stage (at most 31 positions past its end, inside the slack).
    const u32 t = (u32)(r * NSTATES + (l & (NSTATES - 1)) - p0);
    // Position 0's context byte is the stage's zero fill before the input:
    // context 0 (scalar.rs:64, 85).  Mode 1 stages dense ids (stage_remap),
    // so the entry index is row * cols + column directly.
    const u32 sym = sb[4 + t];
    const u32 ctx = sb[3 + t];
    u32 e;
    if (MODE == 0)
        e = sym;
    else if (MODE == 1)
        e = ctx * tr.cols + sym;
    else
        e = ctx * 256 + sym;
    e = act ? e : 0u;
    const uint2 en = MODE == 2 ? tr.gent[e] : entries(sh)[e];
    fcm = act ? en.x : (TOTFREQ << 12);
    mg = act ? en.y : 1u;
}

// Renormalisation words of one staging chunk.  A chunk of CH positions
// emits at most CH words (one per step), so its words fall in the 2 CH bytes
// below the word pointer at the chunk's start.  The stores are raw buffer
// stores on a descriptor for exactly those bytes (base = word pointer - 2 CH,
// 2 CH records): a lane that emits nothing in a step stores at offset
// 2^31, outside the records, and the buffer range check drops it.  So a step
// stores with no branch and no change of the exec mask (a masked store
// needs the mask written before the store and restored after it, and every
// vector instruction behind it in the step waits for that), and the memory
// path sees only the real words.
constexpr u32 SINK_BYTES = 2 * CH;
constexpr u32 SINK_DROP = 0x80000000u;
struct WordSink {
    __amdgpu_buffer_rsrc_t rs;
    u32 off; // word offset of the lowest word so far from the descriptor's base (wave-uniform)
};
__device__ __forceinline__ WordSink sink_below(u64 wp)
{
    return WordSink{__builtin_amdgcn_make_buffer_rsrc((void *)(uintptr_t)(wp - SINK_BYTES), 0, (int)SINK_BYTES,
                                                      0x00020000),
                    SINK_BYTES / 2};
}
// The word pointer after the chunk's words.
__device__ __forceinline__ u64 sink_end(u64 wp, const WordSink &sk) { return wp - SINK_BYTES + 2ull * sk.off; }
// Store x's low half as word `w` of the sink (w in words), or drop it.
__device__ __forceinline__ void sink_store(const WordSink &sk, bool emit, u32 w, u32 x)
{
    __builtin_amdgcn_raw_buffer_store_b16((u16)x, sk.rs, (int)(emit ? w << 1 : SINK_DROP), 0, 0);
}

// x -> (v / f) << 12 | v % f, + cum, given v (x, or x >> 16 after a word)
// and the magic of f.  With q = v / f and g = 4096 - f, the result is
// v + cum + q g.  The magic's quotient q0 is q or q - 1, and q0 = q - 1
// exactly when v >= (q0 + 1) f, so v + q0 g is formed beside that test,
// which selects g or 0 to add with cum in one three-way add: five dependent
// operations after v (the high multiply, a 24-bit multiply-add, the
// compare, the select, the add).  v < 2^20 f keeps q < 2^20, so the
// products are 24-bit multiplies and nothing wraps, except (q0 + 1) f =
// 2^32 for f = 4096, where g = 0 and the select adds nothing either way.
__device__ __forceinline__ u32 step_map(u32 v, u32 f, u32 g, u32 cum, u32 mg)
{
    const u32 q0 = __umulhi(v, mg);
    const u32 a = mad24(q0, g, v);
    const u32 t = mad24(q0, f, f);
    return a + (v >= t ? g : 0u) + cum;
}
// One encode step of every lane (scalar.rs:120-127 per lane; lanes emit in
// descending position order within the round).  The round's k words go
// just below the lowest word so far, the emitting lane of rank r (emitting
// lanes below it) at offset r among them, so the highest lane's word comes
// first in the reference's reversed order.
__device__ __forceinline__ void round_step(u32 &x, WordSink &sk, u32 fcm, u32 mg)
{
    const u32 f = fcm >> 12, cum = fcm & 0xfff, g = TOTFREQ - f;
    // x >= x_max = f << 20  <=>  (x >> 20) >= f  (model.rs:73); inactive
    // lanes (and lanes 32..63) carry f = 4096, which no 32-bit state reaches.
    const bool need = (x >> 20) >= f;
    const u32 v = need ? x >> 16 : x;
    const u64 m = ballot(need);
    sk.off -= (u32)__popcll(m);
    sink_store(sk, need, sk.off + __builtin_amdgcn_mbcnt_lo((u32)m, 0u), x);
    x = step_map(v, f, g, cum, mg);
}

// The state update of round_step alone; the round's store is left to
// store_rounds, which takes the state before the step (xo) and v (vo).  A
// lane emitted a word exactly when v != x: x >= 2^16 always, so x >> 16 < x.
__device__ __forceinline__ void step_only(u32 &x, u32 fcm, u32 mg, u32 &xo, u32 &vo)
{
    const u32 f = fcm >> 12, cum = fcm & 0xfff, g = TOTFREQ - f;
    const bool need = (x >> 20) >= f;
    const u32 v = need ? x >> 16 : x;
    xo = x;
    vo = v;
    x = step_map(v, f, g, cum, mg);
}

// The words of a batch of rounds, in round order, as round_step places them.
// Rounds with xo = vo (a batch not yet run: all zero) store nothing.
__device__ __forceinline__ void store_rounds(WordSink &sk, const u32 (&xo)[WALK_BATCH], const u32 (&vo)[WALK_BATCH])
{
#pragma unroll
    for (u32 k = 0; k < WALK_BATCH; k++) {
        const bool need = vo[k] != xo[k];
        const u64 m = ballot(need);
        sk.off -= (u32)__popcll(m);
        sink_store(sk, need, sk.off + __builtin_amdgcn_mbcnt_lo((u32)m, 0u), xo[k]);
    }
}

// The reverse walk over all n positions; returns the number of words, which
// end at `wend` (2-byte aligned).  The table entries of WALK_BATCH rounds are
// read before their state updates, so the LDS latency is paid once per batch
// on the state chain.
template <u32 MODE>
__device__ __forceinline__ u64 walk(EncShared &sh, const u8 *in, u64 n, const TabRef &tr, u8 *wend, u32 &x_out)
{
    // Wave-uniform (scalar registers): each chunk's word sink is based on it.
    const u64 wtop = uni64((u64)(uintptr_t)wend);
    u64 wp = wtop; // byte address of the lowest word written so far
    u32 x = RANS_L; // a register for the whole walk, written back at the end
    const u64 nch = (n + CH - 1) / CH;
    const u32 l = lane_id();
    const bool lo = l < NSTATES;
    if (l == 0) entries(sh)[TAB_IDLE] = make_uint2(TOTFREQ << 12, 1u);
    // LDS address of entry e: e * estride + ebase, which is the idle entry
    // for every e on lanes 32..63 (one multiply-add, no select).
    const u32 estride = lo ? (u32)sizeof(uint2) : 0u;
    const u32 ebase = lds_addr(sh.tab) + (lo ? 0u : TAB_IDLE * (u32)sizeof(uint2));
    StageLoad pf;
    u32 sr[9];
    stage_fetch(in, n, (nch - 1) * CH, pf);
    stage_bytes(pf, sr);
    if (MODE == 1) stage_remap(sh, sr);
    stage_put(sh.stage[(nch - 1) & 1], sr);
    __syncthreads();
    for (u64 ch = nch; ch-- > 0;) {
        const u64 p0 = ch * CH;
        if (ch > 0) stage_fetch(in, n, p0 - CH, pf);
        const u8 *sb = (const u8 *)sh.stage[ch & 1];
        const u64 pend = p0 + CH < n ? p0 + CH : n;
        const u64 r_hi = (pend - 1) / NSTATES, r_lo = p0 / NSTATES;
        u64 r = r_hi + 1; // rounds r-1 down to r_lo remain
        WordSink sk = sink_below(wp);
        // The block's top round may be partial: it takes the exact step
        // first, so every batch below is full (32 positions in [0, n)).
        if (pend == n && (n % NSTATES) != 0) {
            u32 fcm, mg;
            round_entry<MODE>(sh, tr, sb, p0, r - 1, n, fcm, mg);
            round_step(x, sk, fcm, mg);
            r--;
        }
        // The steps of a batch record their stores, which are made while the
        // next batch's reads are in flight (and after the last batch).
        u32 xo[WALK_BATCH], vo[WALK_BATCH];
#pragma unroll
        for (u32 k = 0; k < WALK_BATCH; k++) xo[k] = vo[k] = 0;
        while (r >= r_lo + WALK_BATCH) {
            // Three phases, each's reads in flight together: the staged
            // bytes, then the table entries they index, then the steps.
            u32 sy[WALK_BATCH], cx[WALK_BATCH], fv[WALK_BATCH], mv[WALK_BATCH];
            const u32 tl = (u32)(r * NSTATES - p0) + (l & (NSTATES - 1));
#pragma unroll
            for (u32 k = 0; k < WALK_BATCH; k++) {
                const u32 t = tl - NSTATES * (k + 1);
                sy[k] = sb[4 + t];
                cx[k] = MODE == 0 ? 0u : sb[3 + t];
            }
            pin16(sy);
            if (MODE != 0) pin16(cx);
            store_rounds(sk, xo, vo);
#pragma unroll
            for (u32 k = 0; k < WALK_BATCH; k++) {
                u32 e = MODE == 0 ? sy[k] : (MODE == 1 ? mad24(cx[k], tr.cols, sy[k]) : cx[k] * 256 + sy[k]);
                uint2 en;
                if (MODE == 2) {
                    e = lo ? e : 0u;
                    en = tr.gent[e];
                } else {
                    en = lds_load2(__umul24(e, estride) + ebase);
                }
                fv[k] = en.x;
                mv[k] = en.y;
            }
            pin16(fv);
            pin16(mv);
#pragma unroll
            for (u32 k = 0; k < WALK_BATCH; k++) {
                if (MODE == 2)
                    step_only(x, lo ? fv[k] : (TOTFREQ << 12), lo ? mv[k] : 1u, xo[k], vo[k]);
                else
                    step_only(x, fv[k], mv[k], xo[k], vo[k]);
            }
            r -= WALK_BATCH;
        }
        store_rounds(sk, xo, vo);
        while (r > r_lo) {
            u32 fcm, mg;
            round_entry<MODE>(sh, tr, sb, p0, r - 1, n, fcm, mg);
            round_step(x, sk, fcm, mg);
            r--;
        }
        wp = sink_end(wp, sk);
        if (ch > 0) {
            stage_bytes(pf, sr);
            if (MODE == 1) stage_remap(sh, sr);
            __syncthreads();
            stage_put(sh.stage[(ch - 1) & 1], sr);
            __syncthreads();
        }
    }
    x_out = x;
    return (wtop - wp) / 2;
}

// Tail of every encode: states, then the words slid down behind them.
__device__ void finish(u8 *dst, u64 hdr, u8 *wend, u64 W, u32 x)
{
    const u32 l = lane_id();
    if (l < NSTATES) store_bytes(dst + hdr - 4 * NSTATES + 4 * l, x, 4);
    // Words were stored by other lanes; order them before the copy reads.
    __threadfence();
    wave_copy_fwd(dst + hdr, wend - 2 * W, 2 * W);
}

// Both orders in one reverse walk (encode-best when neither order's size
// bounds win outright): lanes 0..31 carry the order-1 states, lanes 32..63
// the order-0 states of the same positions.  Dense order-1 only: the staged
// bytes are dense ids, the order-1 entry is ctx * cols + sym and the order-0
// entry base0 + sym, both in the LDS table.  Each half has its own word
// pointer; a lane's rank counts the emitting lanes below it in its half.
__device__ __forceinline__ void round_entry_both(const EncShared &sh, const u8 *sb, u64 p0, u64 r, u64 n,
                                                 u32 cols, u32 base0, u32 &fcm, u32 &mg)
{
    const u32 l = lane_id();
    const u64 i = r * NSTATES + (l & (NSTATES - 1));
    const bool act = i < n;
    const u32 t = (u32)(i - p0);
    const u32 sym = sb[4 + t];
    const u32 ctx = sb[3 + t];
    u32 e = l < NSTATES ? ctx * cols + sym : base0 + sym;
    e = act ? e : 0u;
    const uint2 en = entries(sh)[e];
    fcm = act ? en.x : (TOTFREQ << 12);
    mg = act ? en.y : 1u;
}

// Each half stores its words through its own sink; the other half's lanes
// are dropped by the range check.
__device__ __forceinline__ void round_step_both(u32 &x, WordSink &sk1, WordSink &sk0, u32 fcm, u32 mg)
{
    const u32 f = fcm >> 12, cum = fcm & 0xfff, g = TOTFREQ - f;
    const bool need = (x >> 20) >= f;
    const u32 v = need ? x >> 16 : x;
    const u64 m = ballot(need);
    const u64 m1 = m & 0xffffffffull, m0 = m & ~0xffffffffull;
    sk1.off -= (u32)__popcll(m1);
    sk0.off -= (u32)__popcll(m0);
    // Rank among the emitting lanes of the lane's own half.
    const bool lo = lane_id() < NSTATES;
    const u32 rank = __builtin_amdgcn_mbcnt_hi((u32)(m >> 32), 0u) + __builtin_amdgcn_mbcnt_lo((u32)m1, 0u) * (u32)lo;
    const u32 o = (lo ? sk1.off : sk0.off) + rank;
    sink_store(sk1, need && lo, o, x);
    sink_store(sk0, need && !lo, o, x);
    x = step_map(v, f, g, cum, mg);
}

// store_rounds for the two-order walk: each half through its own sink.
__device__ __forceinline__ void store_rounds_both(WordSink &sk1, WordSink &sk0, const u32 (&xo)[WALK_BATCH],
                                                  const u32 (&vo)[WALK_BATCH])
{
    const bool lo = lane_id() < NSTATES;
#pragma unroll
    for (u32 k = 0; k < WALK_BATCH; k++) {
        const bool need = vo[k] != xo[k];
        const u64 m = ballot(need);
        const u64 m1 = m & 0xffffffffull;
        sk1.off -= (u32)__popcll(m1);
        sk0.off -= (u32)__popcll(m & ~0xffffffffull);
        const u32 rank =
            __builtin_amdgcn_mbcnt_hi((u32)(m >> 32), 0u) + __builtin_amdgcn_mbcnt_lo((u32)m1, 0u) * (u32)lo;
        const u32 o = (lo ? sk1.off : sk0.off) + rank;
        sink_store(sk1, need && lo, o, xo[k]);
        sink_store(sk0, need && !lo, o, xo[k]);
    }
}

// Returns the order-1 and order-0 word counts; the words end at wend1 and
// wend0 (2-byte aligned); x_out: order-1 states in lanes 0..31, order-0 in
// lanes 32..63.
__device__ __forceinline__ void walk_both(EncShared &sh, const u8 *in, u64 n, u32 cols, u32 base0, u8 *wend1,
                                          u8 *wend0, u32 &x_out, u64 &W1, u64 &W0)
{
    const u64 top1 = uni64((u64)(uintptr_t)wend1), top0 = uni64((u64)(uintptr_t)wend0);
    u64 wp1 = top1, wp0 = top0; // byte addresses of each half's lowest word
    u32 x = RANS_L;
    const u64 nch = (n + CH - 1) / CH;
    const u32 l = lane_id();
    // Entry of a staged pair: order-1 half ctx * cols + sym, order-0 half
    // base0 + sym, as (ctx * ecols + sym) * 8 + ebase per lane.
    const bool lo = l < NSTATES;
    const u32 ecols = lo ? cols : 0u;
    const u32 ebase = lds_addr(sh.tab) + (lo ? 0u : base0 * (u32)sizeof(uint2));
    StageLoad pf;
    u32 sr[9];
    stage_fetch(in, n, (nch - 1) * CH, pf);
    stage_bytes(pf, sr);
    stage_remap(sh, sr);
    stage_put(sh.stage[(nch - 1) & 1], sr);
    __syncthreads();
    for (u64 ch = nch; ch-- > 0;) {
        const u64 p0 = ch * CH;
        if (ch > 0) stage_fetch(in, n, p0 - CH, pf);
        const u8 *sb = (const u8 *)sh.stage[ch & 1];
        const u64 pend = p0 + CH < n ? p0 + CH : n;
        const u64 r_hi = (pend - 1) / NSTATES, r_lo = p0 / NSTATES;
        u64 r = r_hi + 1;
        WordSink sk1 = sink_below(wp1), sk0 = sink_below(wp0);
        // A partial top round first (exact step), then full batches only:
        // every lane of a full round is active.
        if (pend == n && (n % NSTATES) != 0) {
            u32 fcm, mg;
            round_entry_both(sh, sb, p0, r - 1, n, cols, base0, fcm, mg);
            round_step_both(x, sk1, sk0, fcm, mg);
            r--;
        }
        // Stores deferred to the next batch as in walk.
        u32 xo[WALK_BATCH], vo[WALK_BATCH];
#pragma unroll
        for (u32 k = 0; k < WALK_BATCH; k++) xo[k] = vo[k] = 0;
        while (r >= r_lo + WALK_BATCH) {
            u32 sy[WALK_BATCH], cx[WALK_BATCH], fv[WALK_BATCH], mv[WALK_BATCH];
            const u32 tl = (u32)(r * NSTATES - p0) + (l & (NSTATES - 1));
#pragma unroll
            for (u32 k = 0; k < WALK_BATCH; k++) {
                const u32 t = tl - NSTATES * (k + 1);
                sy[k] = sb[4 + t];
                cx[k] = sb[3 + t];
            }
            pin16(sy);
            pin16(cx);
            store_rounds_both(sk1, sk0, xo, vo);
#pragma unroll
            for (u32 k = 0; k < WALK_BATCH; k++) {
                const uint2 en = lds_load2(mad24(mad24(cx[k], ecols, sy[k]), (u32)sizeof(uint2), ebase));
                fv[k] = en.x;
                mv[k] = en.y;
            }
            pin16(fv);
            pin16(mv);
#pragma unroll
            for (u32 k = 0; k < WALK_BATCH; k++) step_only(x, fv[k], mv[k], xo[k], vo[k]);
            r -= WALK_BATCH;
        }
        store_rounds_both(sk1, sk0, xo, vo);
        while (r > r_lo) {
            u32 fcm, mg;
            round_entry_both(sh, sb, p0, r - 1, n, cols, base0, fcm, mg);
            round_step_both(x, sk1, sk0, fcm, mg);
            r--;
        }
        wp1 = sink_end(wp1, sk1);
        wp0 = sink_end(wp0, sk0);
        if (ch > 0) {
            stage_bytes(pf, sr);
            stage_remap(sh, sr);
            __syncthreads();
            stage_put(sh.stage[(ch - 1) & 1], sr);
            __syncthreads();
        }
    }
    x_out = x;
    W1 = (top1 - wp1) / 2;
    W0 = (top0 - wp0) / 2;
}

// Input dword t (bytes [4t, 4t+4)) as two aligned loads and a byte shift.
// Both loads stay within [in, in + n) for t < bulk_dwords(n); the dword
// A[t - 1] read for the byte before dword t (t >= 1) starts at or after the
// aligned dword holding in[0].
__device__ __forceinline__ u64 bulk_dwords(u64 n) { return n >= 8 ? (n - 8) / 4 + 1 : 0; }

// Histogram of in[0..n) into sh.hist.  Lanes count into HSUB sub-histograms
// (lane % HSUB, rows of 257 so a byte value sits in a different bank in
// each; as many as the table's counts hold, which hold nothing yet): with
// few distinct bytes, fewer lanes then meet on one counter.  The rows are
// summed at the end.
#ifndef HSUB
#define HSUB 14
#endif
static_assert(HSUB * 257 <= 2 * TAB_CAP, "sub-histograms live in the table's counts");
__device__ __forceinline__ void hist0(EncShared &sh, const u8 *in, u64 n)
{
    const u32 l = lane_id();
    u32 *cnt = counts(sh);
    u32 *sub = cnt + (l % HSUB) * 257;
    for (u32 s = l; s < HSUB * 257; s += 64) cnt[s] = 0;
    __syncthreads();
    const u64 nd = (n + 3) / 4, nbulk = bulk_dwords(n);
    const u32 ph = (u32)((uintptr_t)in & 3);
    const u32 *A = (const u32 *)(in - ph);
    // Bulk: HB dwords per lane are loaded before any of them is counted, so
    // the loads of a batch are in flight together.
    u64 t0 = 0;
    for (; t0 + 64 * HB <= nbulk; t0 += 64 * HB) {
        u32 w[HB];
#pragma unroll
        for (u32 u = 0; u < HB; u++) {
            const u64 t = t0 + 64 * u + l;
            w[u] = __builtin_amdgcn_alignbyte(A[t + 1], A[t], ph);
        }
#pragma unroll
        for (u32 u = 0; u < HB; u++)
#pragma unroll
            for (u32 b = 0; b < 4; b++) atomicAdd(&sub[(w[u] >> (8 * b)) & 0xff], 1u);
    }
    // The rest (fewer than 64 * HB + 2 dwords), one dword per lane at a time.
    for (u64 t = t0 + l; t < nd; t += 64) {
        const u32 w = load4(in + 4 * t, in + n);
        const u32 nb = (4 * t + 4 <= n) ? 4 : (u32)(n - 4 * t);
        for (u32 b = 0; b < nb; b++) atomicAdd(&sub[(w >> (8 * b)) & 0xff], 1u);
    }
    __syncthreads();
    for (u32 s = l; s < 256; s += 64) {
        u32 c = 0;
#pragma unroll
        for (u32 k = 0; k < HSUB; k++) c += cnt[k * 257 + s];
        sh.hist[s] = c;
    }
    __syncthreads();
}

// Pair counts: row(ctx) x col(sym) for every position, with the loads
// batched as in hist0.  ctx0 is the context of position 0: 0 for a whole
// input (scalar.rs:64, 85), the byte before it for a slice of one.
//
// Dense: the (cols + 1) x cols table of counts is kept in as many copies as
// the table's space holds (up to PSUB; lane l counts into copy l % copies),
// which are summed into the first copy at the end: frequent pairs then put
// fewer lanes on one counter.  The counts start at zero here.
// Otherwise: global atomics on the 256 x 256 gcounts (zeroed by the caller).
#ifndef PSUB
#define PSUB 16
#endif
template <bool DENSE>
__device__ __forceinline__ void count_pair(EncShared &sh, u32 *cnt, u32 cols, u32 prev, u32 s)
{
    if (DENSE)
        atomicAdd(&cnt[(u32)sh.rowmap[prev] * cols + sh.colmap[s]], 1u);
    else
        atomicAdd(&cnt[prev * 256 + s], 1u);
}

template <bool DENSE>
__device__ __forceinline__ void hist1(EncShared &sh, const u8 *in, u64 n, u32 *gcounts, u32 cols, u32 ctx0 = 0)
{
    const u32 l = lane_id();
    const u32 ne = DENSE ? (cols + 1) * cols : 0u;
    const u32 stride = ne + 1;  // copies start in different banks
    u32 copies = DENSE ? 2 * TAB_CAP / stride : 1u;
    copies = copies < PSUB ? copies : PSUB;
    u32 *cnt = DENSE ? counts(sh) + (l % copies) * stride : gcounts;
    if (DENSE) {
        for (u32 e = l; e < copies * stride; e += 64) counts(sh)[e] = 0;
        __syncthreads();
    }
    const u64 nd = (n + 3) / 4, nbulk = bulk_dwords(n);
    const u32 ph = (u32)((uintptr_t)in & 3);
    const u32 *A = (const u32 *)(in - ph);
    u64 t0 = 0;
    for (; t0 + 64 * HB <= nbulk; t0 += 64 * HB) {
        u32 w[HB], pv[HB];
#pragma unroll
        for (u32 u = 0; u < HB; u++) {
            const u64 t = t0 + 64 * u + l;
            const u32 a0 = A[t];
            w[u] = __builtin_amdgcn_alignbyte(A[t + 1], a0, ph);
            // Byte 4t - 1: the top byte of input bytes [4t - 4, 4t).
            pv[u] = t > 0 ? __builtin_amdgcn_alignbyte(a0, A[t - 1], ph) >> 24 : ctx0;
        }
#pragma unroll
        for (u32 u = 0; u < HB; u++) {
            if (DENSE) {
                // A present symbol has the same row and column id, so each
                // byte is mapped once; the byte before the dword through
                // rowmap (context 0 of position 0 may be absent).
                u32 id[5];
                id[0] = sh.rowmap[pv[u]];
#pragma unroll
                for (u32 b = 0; b < 4; b++) id[b + 1] = sh.colmap[(w[u] >> (8 * b)) & 0xff];
#pragma unroll
                for (u32 b = 0; b < 4; b++) atomicAdd(&cnt[id[b] * cols + id[b + 1]], 1u);
            } else {
                u32 prev = pv[u];
#pragma unroll
                for (u32 b = 0; b < 4; b++) {
                    const u32 s = (w[u] >> (8 * b)) & 0xff;
                    count_pair<DENSE>(sh, cnt, cols, prev, s);
                    prev = s;
                }
            }
        }
    }
    for (u64 t = t0 + l; t < nd; t += 64) {
        const u32 w = load4(in + 4 * t, in + n);
        u32 prev = t > 0 ? in[4 * t - 1] : ctx0;
        const u32 nb = (4 * t + 4 <= n) ? 4 : (u32)(n - 4 * t);
        for (u32 b = 0; b < nb; b++) {
            const u32 s = (w >> (8 * b)) & 0xff;
            count_pair<DENSE>(sh, cnt, cols, prev, s);
            prev = s;
        }
    }
    __syncthreads();
    if (DENSE && copies > 1) {
        // Copy k sits above the first copy's ne counts, so each lane sums
        // its own entries in place.
        for (u32 e = l; e < ne; e += 64) {
            u32 c = counts(sh)[e];
            for (u32 k = 1; k < copies; k++) c += counts(sh)[k * stride + e];
            counts(sh)[e] = c;
        }
        __syncthreads();
    }
}

// Size bounds of a stream before its walk, for encode_best_job's early
// choice.  A symbol step maps x to floor(x/f)*4096 + x%f + cum with x >= 16f
// (x >= 2^16 unrenormalised, x >= 2^20 f >> 16 after one word), so it
// raises log2(x) by log2(4096/f) + e, log2(15/16) < e <= log2(17/16); a word
// lowers log2(x) by 16 to 16 + log2(16/15) bits (x >= 2^20 before the
// shift); each of the 32 states starts at 2^16 and ends in [2^16, 2^32).
// With S = sum over symbols of log2(4096/f), the W words satisfy
//   (S - 0.093110 n - 512) / 16.093110 < W <= (S + 0.087463 n) / 16.
// S is summed in 2^-20 fixed point, each term's log rounded outwards by one
// unit, so every lane gets the same exact integers.
struct SizeBounds {
    u64 lo, hi; // bytes
};
constexpr u32 LOG_FRAC = 20;
__device__ __forceinline__ void acc_log_bits(u32 c, u32 f, u64 &dn, u64 &up)
{
    if (c == 0) return;
    const double L = (12.0 - log2((double)f)) * (double)(1u << LOG_FRAC); // f >= 1 for a counted symbol
    const u64 lf = (u64)floor(L), lc = (u64)ceil(L);
    dn += (u64)c * (lf > 0 ? lf - 1 : 0);
    up += (u64)c * (lc + 1);
}
__device__ __forceinline__ SizeBounds size_bounds(u64 dn, u64 up, u64 n, u64 hdr)
{
    dn = wave_sum64(dn);
    up = wave_sum64(up);
    constexpr u64 ONE = 1ull << LOG_FRAC;
    const u64 w_max = (up + n * 91751) / (16 * ONE);        // 0.087463 * 2^20 = 91712.4
    const u64 sub = n * 97728 + 512 * ONE;                   // 0.093110 * 2^20 = 97632.7
    const double w_lo = dn > sub ? floor((double)(dn - sub) / (16.0932 * (double)ONE)) - 1.0 : 0.0;
    const u64 w_min = w_lo > 0.0 ? (u64)w_lo : 0;
    return SizeBounds{hdr + 2 * w_min, hdr + 2 * w_max};
}

// encode_best_job's early choice: the order-0 bounds in, and after the
// order-1 tables are built, 1 (order-1 is surely smaller: walk it, skip
// order-0), 0 (order-0 surely smaller: skip the order-1 walk) or -1 (walk
// both and compare exactly).
struct BestGate {
    SizeBounds b0;
    int decided; // 2 after encode_job walked both orders (l0 set)
    u8 *o0;      // the order-0 stream's region (j0) and its capacity
    u64 cap0;
    u64 l0;
};

// Counts of a long job taken by the pre-pass (k_enc_pre_count): a slot of
// PRE_SLOT words, the byte histogram then the dense pair counts.
__device__ __forceinline__ void load_hist(EncShared &sh, const u32 *pre)
{
    for (u32 s = lane_id(); s < 256; s += 64) sh.hist[s] = pre[s];
    __syncthreads();
}

// `have_hist`: sh.hist already holds this input's byte histogram (computed
// by encode_best_job before either order).  `gate`: see BestGate; only the
// GATED instance (encode-best's order-1 call) has the early choice and the
// two-order walk compiled in, so the plain encode kernel carries neither.
// `pre`: this job's pre-pass slot (long jobs), or null.
template <bool GATED>
__device__ int encode_job(EncShared &sh, const u8 *src, u8 *dst, const fqxv_rans_job &jb_in, u64 &out_len,
                          u8 *ws, bool have_hist = false, BestGate *gate = nullptr, const u32 *pre = nullptr)
{
    const fqxv_rans_job jb = uni_job(jb_in);
    const u32 l = lane_id();
    const u8 *in = src + jb.src_off;
    u8 *o = dst + jb.dst_off;
    const u64 n = jb.src_len;
    const int order = jb.order;
    if (order != 0 && order != 1) return FQXV_RANS_ERR_INVALID_ARG;
    if (n >= (1ull << 32)) return FQXV_RANS_ERR_INVALID_ARG;
    const u64 k_max = n < 256 ? n : 256;
    const u64 bound = n == 0 ? 9 : (order == 0 ? O0_HDR : HDR0 + BITMAP + 512 * k_max + 4 * NSTATES) + 2 * n + 16;
    if (jb.dst_cap < bound) return FQXV_RANS_ERR_DST_TOO_SMALL;
    if (l < HDR0) o[l] = l == 0 ? (u8)order : (u8)(n >> (8 * (l - 1)));
    if (n == 0) {
        out_len = 9;
        return 0;
    }
    // Words are written downwards from the end of the region, 2-byte
    // aligned (the bound leaves 16 spare bytes), then slid behind the header.
    u8 *wend = (u8 *)((uintptr_t)(o + jb.dst_cap) & ~(uintptr_t)1);
    if (!have_hist) {
        if (pre)
            load_hist(sh, pre);
        else
            hist0(sh, in, n);
    }

    if (order == 0) {
        u32 c[4], f[4];
        for (int i = 0; i < 4; i++) c[i] = sh.hist[4 * l + i];
        normalize_wave(c, f);
        put_freqs(o + HDR0, f);
        const u32 tot = f[0] + f[1] + f[2] + f[3];
        u32 cum = wave_scan_incl(tot) - tot;
        for (int i = 0; i < 4; i++) {
            entries(sh)[4 * l + i] = make_uint2((f[i] << 12) | cum, magic_of(f[i]));
            cum += f[i];
        }
        __syncthreads();
        const TabRef tr{nullptr, 0};
        u32 x;
        const u64 W = walk<0>(sh, in, n, tr, wend, x);
        finish(o, O0_HDR, wend, W, x);
        out_len = O0_HDR + 2 * W;
        return 0;
    }

    // Order-1.  Alphabet A = symbols present; contexts are {0} u A.
    u32 pres = 0;
    for (int i = 0; i < 4; i++) pres += sh.hist[4 * l + i] != 0;
    const u32 asz = uni(wave_sum(pres));
    const bool dense = dense_fits(asz);
    u32 *gcounts = (u32 *)ws;
    uint2 *gent = (uint2 *)(ws + WS_ENT);
    if (dense) {
        const u32 excl = wave_scan_incl(pres) - pres;
        u32 d = excl;
        for (int i = 0; i < 4; i++) {
            const u32 s = 4 * l + i;
            const bool p = sh.hist[s] != 0;
            sh.colmap[s] = p ? (u8)d : (u8)0;
            sh.rowmap[s] = p ? (u8)d : (u8)asz;
            d += p;
        }
        if (pre) {
            for (u32 e = l; e < (asz + 1) * asz; e += 64) counts(sh)[e] = pre[256 + e];
            __syncthreads();
        } else {
            hist1<true>(sh, in, n, nullptr, asz);
        }
    } else {
        for (u64 e = l; e < 256 * 256; e += 64) gcounts[e] = 0;
        __threadfence();
        hist1<false>(sh, in, n, gcounts, 0);
        __threadfence();
    }
    // Present contexts (row sum > 0), ascending.
    u32 pbits = 0;
    for (int i = 0; i < 4; i++) {
        const u32 ctx = 4 * l + i;
        u32 any = 0;
        if (dense) {
            const u32 row = sh.rowmap[ctx];
            const bool valid = sh.hist[ctx] != 0 || ctx == 0;
            if (valid)
                for (u32 cI = 0; cI < asz; cI++) any |= counts(sh)[row * asz + cI];
        } else {
            for (u32 sI = 0; sI < 256; sI++) any |= gcounts[ctx * 256 + sI];
        }
        pbits |= (any != 0) << i;
    }
    const u32 pc = __popc(pbits);
    const u32 pex = wave_scan_incl(pc) - pc;
    const u32 K = uni(wave_sum(pc));
    {
        u32 d = pex;
        for (int i = 0; i < 4; i++)
            if ((pbits >> i) & 1) sh.ctx_of[d++] = (u8)(4 * l + i);
    }
    // Bitmap bytes: lane l holds the nibble for contexts 4l..4l+3.
    {
        const u32 other = __shfl_xor(pbits, 1);
        if ((l & 1) == 0) o[HDR0 + l / 2] = (u8)(pbits | (other << 4));
    }
    __syncthreads();
    const u64 hdr = HDR0 + BITMAP + 512ull * K + 4 * NSTATES;
    u64 bits_dn = 0, bits_up = 0;
    // Dense rows go from the last to the first, each entry written over
    // counts of rows already done: context 0's own row (the last) first when
    // symbol 0 is absent, then the present symbols' rows in descending order.
    const bool row0_last = dense && sh.hist[0] == 0;
    for (u32 it = 0; it < K; it++) {
        const u32 d = row0_last ? (it == 0 ? 0u : K - it) : K - 1 - it;
        const u32 ctx = sh.ctx_of[d];
        u32 c[4], f[4];
        for (int i = 0; i < 4; i++) {
            const u32 s = 4 * l + i;
            if (dense)
                c[i] = sh.hist[s] ? counts(sh)[(u32)sh.rowmap[ctx] * asz + sh.colmap[s]] : 0u;
            else
                c[i] = gcounts[ctx * 256 + s];
        }
        normalize_wave(c, f);
        if (GATED && gate)
            for (int i = 0; i < 4; i++) acc_log_bits(c[i], f[i], bits_dn, bits_up);
        put_freqs(o + HDR0 + BITMAP + 512ull * d, f);
        const u32 tot = f[0] + f[1] + f[2] + f[3];
        u32 cum = wave_scan_incl(tot) - tot;
        __syncthreads(); // all lanes read this row's counts before it is overwritten
        for (int i = 0; i < 4; i++) {
            const u32 s = 4 * l + i;
            const uint2 en = make_uint2((f[i] << 12) | cum, magic_of(f[i]));
            if (dense) {
                if (sh.hist[s]) entries(sh)[(u32)sh.rowmap[ctx] * asz + sh.colmap[s]] = en;
            } else {
                gent[ctx * 256 + s] = en;
            }
            cum += f[i];
        }
        __syncthreads();
    }
    if (GATED && gate) {
        const SizeBounds b1 = size_bounds(bits_dn, bits_up, n, hdr);
        gate->decided = b1.hi < gate->b0.lo ? 1 : (gate->b0.hi < b1.lo ? 0 : -1);
        if (gate->decided == 0) {
            out_len = 0;
            return 0;
        }
    }
    if (GATED && gate && gate->decided == -1 && dense && (asz + 1) * asz + asz <= TAB_CAP) {
        // Neither order can be ruled out: walk both at once.  The order-0
        // table (from the byte histogram, indexed by dense id) goes behind the
        // order-1 table; the order-0 stream's header is written to its own
        // region as encode_job would for order 0.
        const u32 base0 = (asz + 1) * asz;
        u8 *o0 = gate->o0;
        u32 c[4], f[4];
        for (int i = 0; i < 4; i++) c[i] = sh.hist[4 * l + i];
        normalize_wave(c, f);
        if (l < HDR0) o0[l] = l == 0 ? (u8)0 : (u8)(n >> (8 * (l - 1)));
        put_freqs(o0 + HDR0, f);
        const u32 tot = f[0] + f[1] + f[2] + f[3];
        u32 cum = wave_scan_incl(tot) - tot;
        for (int i = 0; i < 4; i++) {
            const u32 s = 4 * l + i;
            if (f[i]) entries(sh)[base0 + sh.colmap[s]] = make_uint2((f[i] << 12) | cum, magic_of(f[i]));
            cum += f[i];
        }
        __syncthreads();
        u8 *wend0 = (u8 *)((uintptr_t)(o0 + gate->cap0) & ~(uintptr_t)1);
        u32 x;
        u64 W1, W0;
        walk_both(sh, in, n, asz, base0, wend, wend0, x, W1, W0);
        if (l < NSTATES)
            store_bytes(o + hdr - 4 * NSTATES + 4 * l, x, 4);
        else
            store_bytes(o0 + O0_HDR - 4 * NSTATES + 4 * (l - NSTATES), x, 4);
        __threadfence();
        wave_copy_fwd(o + hdr, wend - 2 * W1, 2 * W1);
        wave_copy_fwd(o0 + O0_HDR, wend0 - 2 * W0, 2 * W0);
        out_len = hdr + 2 * W1;
        gate->l0 = O0_HDR + 2 * W0;
        gate->decided = 2;
        return 0;
    }
    if (!dense) __threadfence();
    const TabRef tr = dense ? TabRef{nullptr, asz} : TabRef{gent, 0};
    u32 x;
    const u64 W = dense ? walk<1>(sh, in, n, tr, wend, x) : walk<2>(sh, in, n, tr, wend, x);
    finish(o, hdr, wend, W, x);
    out_len = hdr + 2 * W;
    return 0;
}

// Blocks from this size on take encode_best_job's early choice.
constexpr u64 BEST_GATE_MIN = 1ull << 14;

__device__ __forceinline__ u64 encode_bound_dev(u64 n, int order)
{
    const u64 k_max = n < 256 ? n : 256;
    return n == 0 ? 9 : (order == 0 ? O0_HDR : HDR0 + BITMAP + 512 * k_max + 4 * NSTATES) + 2 * n + 16;
}

// Both orders, the smaller stream kept (the callers' encode_best:
// crates/fqxv-tokenizer/src/lib.rs:91-95 and container/reorder.rs:597-607 keep
// order-0 on a tie, crates/fqxv-lroverlap/src/codec.rs:292-303 keeps order-1).
// `jb.order` names the order kept on a tie.  Order-0 is written at the front
// of the region, order-1 behind it; an order-1 winner slides down to the
// front.  The stream's first byte records the choice.
__device__ int encode_best_job(EncShared &sh, const u8 *src, u8 *dst, const fqxv_rans_job &jb_in, u64 &out_len,
                               u8 *ws, const u32 *pre = nullptr)
{
    const fqxv_rans_job jb = uni_job(jb_in);
    const int tie = jb.order;
    if (tie != 0 && tie != 1) return FQXV_RANS_ERR_INVALID_ARG;
    if (jb.src_len >= (1ull << 32)) return FQXV_RANS_ERR_INVALID_ARG;
    const u64 b0 = encode_bound_dev(jb.src_len, 0);
    if (jb.dst_cap < b0 + encode_bound_dev(jb.src_len, 1)) return FQXV_RANS_ERR_DST_TOO_SMALL;
    fqxv_rans_job j0 = jb, j1 = jb;
    j0.order = 0;
    j0.dst_cap = b0;
    j1.order = 1;
    j1.dst_off = jb.dst_off + b0;
    j1.dst_cap = jb.dst_cap - b0;
    u64 l0 = 0, l1 = 0;
    int err;
    int decided = -1;
    if (jb.src_len >= BEST_GATE_MIN) {
        // Byte histogram once; the order-0 size bounds from its normalised
        // table; order-1 tables next, and a walk only for an order that can
        // still be the smaller one.
        if (pre)
            load_hist(sh, pre);
        else
            hist0(sh, src + jb.src_off, jb.src_len);
        const u32 l = lane_id();
        u32 c[4], f[4];
        for (int i = 0; i < 4; i++) c[i] = sh.hist[4 * l + i];
        normalize_wave(c, f);
        u64 dn = 0, up = 0;
        for (int i = 0; i < 4; i++) acc_log_bits(c[i], f[i], dn, up);
        BestGate g{size_bounds(dn, up, jb.src_len, O0_HDR), -1, dst + j0.dst_off, j0.dst_cap, 0};
        err = encode_job<true>(sh, src, dst, j1, l1, ws, true, &g, pre);
        if (err) return err;
        decided = g.decided;
        if (decided == 2) {
            l0 = g.l0; // both walked in one pass: the exact compare decides
            decided = -1;
        } else if (decided != 1) {
            __syncthreads();
            err = encode_job<false>(sh, src, dst, j0, l0, ws, true);
            if (err) return err;
        }
    } else {
        err = encode_job<false>(sh, src, dst, j0, l0, ws);
        if (err) return err;
        __syncthreads();
        err = encode_job<false>(sh, src, dst, j1, l1, ws, true);
        if (err) return err;
    }
    const bool one = decided >= 0 ? decided == 1 : (tie == 1 ? l1 <= l0 : l1 < l0);
    if (one) {
        __threadfence();
        wave_copy_fwd(dst + jb.dst_off, dst + j1.dst_off, l1);
    }
    out_len = one ? l1 : l0;
    return 0;
}

} // namespace

// Pre-pass plan: workgroup b clears long-job slot b; wave 0 of workgroup 0
// finds the long jobs (queue positions 0.. while src_len >= PRE_MIN) and
// numbers their slices.  Lane l looks at positions PRE_PER_LANE * l + i.
__global__ __launch_bounds__(256) void k_enc_pre_plan(const fqxv_rans_job *__restrict__ jobs,
                                                      const u32 *__restrict__ order, u32 count, EncPre *pre)
{
    u32 *slot = pre->slot + (u64)blockIdx.x * PRE_SLOT;
    for (u32 i = threadIdx.x; i < PRE_SLOT; i += 256) slot[i] = 0;
    if (blockIdx.x != 0 || threadIdx.x >= 64) return;
    const u32 l = lane_id();
    const u32 qmax = count < PRE_MAX ? count : PRE_MAX;
    u32 ns[PRE_PER_LANE];
    u32 first_short = ~0u;
#pragma unroll
    for (u32 i = 0; i < PRE_PER_LANE; i++) {
        const u32 q = PRE_PER_LANE * l + i;
        u64 n = 0;
        if (q < qmax) n = jobs[order ? order[q] : q].src_len;
        const bool big = q < qmax && n >= PRE_MIN && n < (1ull << 32);
        if (!big && first_short == ~0u) first_short = q;
        ns[i] = big ? (u32)((n + PRE_SLICE - 1) / PRE_SLICE) : 0u;
    }
    const u64 m = ballot(first_short != ~0u);
    const u32 nbig = m ? __builtin_amdgcn_readlane(first_short, (u32)__builtin_ctzll(m)) : qmax;
    u32 tot = 0;
#pragma unroll
    for (u32 i = 0; i < PRE_PER_LANE; i++) {
        if (PRE_PER_LANE * l + i >= nbig) ns[i] = 0;
        tot += ns[i];
    }
    u32 at = wave_scan_incl(tot) - tot;
#pragma unroll
    for (u32 i = 0; i < PRE_PER_LANE; i++) {
        pre->slice_start[PRE_PER_LANE * l + i] = at;
        at += ns[i];
    }
    const u32 total = uni(__shfl(at, 63));
    if (l < 4) pre->slice_start[PRE_MAX + l] = total;
    if (l == 0) {
        pre->nbig = nbig;
        pre->nslices = total;
        pre->ctr[0] = 0;
        pre->ctr[1] = 0;
    }
}

// One counting pass over the long jobs' slices, grid-strided through a
// queue: PASS 0 the byte counts, PASS 1 the dense pair counts (order-1 jobs,
// or every job for encode-best, whose alphabet fits the dense table; the
// pass-0 counts give the dense ids exactly as encode_job forms them).  Each
// slice is counted in LDS with encode_job's own histogram code and added to
// its job's slot with global atomics (nonzero counts only).
template <u32 PASS>
__global__ __launch_bounds__(64) void k_enc_pre_count(const u8 *__restrict__ src,
                                                      const fqxv_rans_job *__restrict__ jobs,
                                                      const u32 *__restrict__ order, EncPre *pre, int best)
{
    __shared__ EncShared sh;
    const u32 l = lane_id();
    const u32 nbig = uni(pre->nbig), nslices = uni(pre->nslices);
    for (;;) {
        u32 sidx = 0;
        if (l == 0) sidx = atomicAdd(&pre->ctr[PASS], 1u);
        sidx = uni(__shfl(sidx, 0));
        if (sidx >= nslices) break;
        // The long job holding slice sidx: last b with slice_start[b] <= sidx.
        u32 lo = 0, hi = nbig - 1;
        while (lo < hi) {
            const u32 mid = (lo + hi + 1) / 2;
            if (uni(pre->slice_start[mid]) <= sidx)
                lo = mid;
            else
                hi = mid - 1;
        }
        const u32 b = lo;
        const fqxv_rans_job jb = uni_job(jobs[order ? uni(order[b]) : b]);
        const u8 *in = src + jb.src_off;
        const u64 k = sidx - uni(pre->slice_start[b]);
        const u64 p0 = k * PRE_SLICE;
        const u64 len = jb.src_len - p0 < PRE_SLICE ? jb.src_len - p0 : PRE_SLICE;
        u32 *slot = pre->slot + (u64)b * PRE_SLOT;
        if (PASS == 0) {
            hist0(sh, in + p0, len);
            for (u32 s = l; s < 256; s += 64)
                if (sh.hist[s]) atomicAdd(&slot[s], sh.hist[s]);
        } else {
            if (!best && jb.order != 1) continue;
            u32 pres = 0;
            for (u32 i = 0; i < 4; i++) pres += slot[4 * l + i] != 0;
            const u32 asz = uni(wave_sum(pres));
            if (!dense_fits(asz)) continue;
            const u32 excl = wave_scan_incl(pres) - pres;
            u32 d = excl;
            for (u32 i = 0; i < 4; i++) {
                const u32 s = 4 * l + i;
                const bool p = slot[s] != 0;
                sh.colmap[s] = p ? (u8)d : (u8)0;
                sh.rowmap[s] = p ? (u8)d : (u8)asz;
                d += p;
            }
            const u32 ne = (asz + 1) * asz;
            hist1<true>(sh, in + p0, len, nullptr, asz, p0 ? (u32)in[p0 - 1] : 0u);
            for (u32 e = l; e < ne; e += 64)
                if (counts(sh)[e]) atomicAdd(&slot[256 + e], counts(sh)[e]);
        }
        __syncthreads();
    }
}

// BEST: encode-best (both orders, the smaller kept) instead of the job's order.
// `pre` (may be null): the pre-pass counts of the queue's first nbig jobs.
template <bool BEST>
__global__ __launch_bounds__(64) void k_rans_encode(const u8 *__restrict__ src, u8 *__restrict__ dst,
                                                    const fqxv_rans_job *__restrict__ jobs, u32 count,
                                                    const u32 *__restrict__ order, u32 *counter,
                                                    i32 *status, u64 *out_len, u8 *ws, const EncPre *pre)
{
    __shared__ EncShared sh;
    u8 *myws = ws + (u64)blockIdx.x * WS_PER_WAVE;
    const u32 nbig = pre ? uni(pre->nbig) : 0u;
    for (;;) {
        u32 w = 0;
        if (lane_id() == 0) w = atomicAdd(counter, 1u);
        w = uni(__shfl(w, 0));
        if (w >= count) break;
        const u32 j = order ? uni(order[w]) : w;
        const fqxv_rans_job jb = uni_job(jobs[j]);
        const u32 *ps = w < nbig ? pre->slot + (u64)w * PRE_SLOT : nullptr;
        u64 ol = 0;
        const int err = BEST ? encode_best_job(sh, src, dst, jb, ol, myws, ps)
                             : encode_job<false>(sh, src, dst, jb, ol, myws, false, nullptr, ps);
        if (lane_id() == 0) {
            status[j] = err;
            out_len[j] = err ? 0 : ol;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(64) void k_rans_encode_one(const u8 *__restrict__ src, u8 *__restrict__ dst,
                                                        fqxv_rans_job jb, i32 *status, u64 *out_len, u8 *ws)
{
    __shared__ EncShared sh;
    u64 ol = 0;
    const int err = encode_job<false>(sh, src, dst, jb, ol, ws);
    if (lane_id() == 0) {
        *status = err;
        *out_len = err ? 0 : ol;
    }
}

size_t fqxv_enc_shared_bytes() { return sizeof(EncShared); }

// ------------------------------------------------------------ host launchers
__global__ void k_reset_counter_enc(u32 *counter) { *counter = 0; }

size_t fqxv_enc_pre_bytes() { return sizeof(EncPre); }

// `pre` (may be null: no pre-pass): a device buffer of fqxv_enc_pre_bytes().
hipError_t fqxv_launch_encode_batch(const u8 *src, u8 *dst, const fqxv_rans_job *jobs, u32 count,
                                    const u32 *order, u32 *counter, i32 *status, u64 *out_len, u8 *ws,
                                    u32 grid, hipStream_t st, int best, void *pre)
{
    hipLaunchKernelGGL(k_reset_counter_enc, dim3(1), dim3(1), 0, st, counter);
    EncPre *p = (EncPre *)pre;
    if (p) {
        hipLaunchKernelGGL(k_enc_pre_plan, dim3(PRE_MAX), dim3(256), 0, st, jobs, order, count, p);
        hipLaunchKernelGGL(k_enc_pre_count<0>, dim3(grid), dim3(64), 0, st, src, jobs, order, p, best);
        hipLaunchKernelGGL(k_enc_pre_count<1>, dim3(grid), dim3(64), 0, st, src, jobs, order, p, best);
    }
    if (best)
        hipLaunchKernelGGL(k_rans_encode<true>, dim3(grid), dim3(64), 0, st, src, dst, jobs, count, order, counter,
                           status, out_len, ws, (const EncPre *)p);
    else
        hipLaunchKernelGGL(k_rans_encode<false>, dim3(grid), dim3(64), 0, st, src, dst, jobs, count, order,
                           counter, status, out_len, ws, (const EncPre *)p);
    return hipGetLastError();
}

hipError_t fqxv_launch_encode_one(const u8 *src, u8 *dst, const fqxv_rans_job &jb, i32 *status,
                                  u64 *out_len, u8 *ws, hipStream_t st)
{
    hipLaunchKernelGGL(k_rans_encode_one, dim3(1), dim3(64), 0, st, src, dst, jb, status, out_len, ws);
    return hipGetLastError();
}

hipError_t fqxv_encode_occupancy(int *blocks_per_cu)
{
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_rans_encode<false>, 64, 0);
}
