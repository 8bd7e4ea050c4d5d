// gfx950 rANS Nx16 decoder (order-0 and order-1).
//
// Stream format and semantics: crates/fqxv-rans/src/scalar.rs:143-238 (decode,
// decode_order0, decode_order1, read_states, step_state) with the validation
// of model.rs:160-174 and the output-length checks of lib.rs:207-231.
//
// Mapping to the hardware.  One workgroup = one wave64 decodes one block at a
// time; a persistent grid pulls blocks from a device work queue.  Lanes 0..31
// hold the 32 interleaved states.  Per round of 32 symbols:
//   * order-0: slot -> packed LDS entry ((f-1)<<20 | (slot-cum)<<8 | sym),
//     x' = f*(x>>12) + (slot-cum) (one v_mad_u32_u24; never overflows);
//   * renormalisation words are consumed in lane order: a ballot marks the
//     lanes with x' < 2^16, v_mbcnt gives each its rank, and lane l takes
//     word[pos + rank] from an LDS ring refilled one 512-byte chunk at a
//     time, two chunks ahead of the read position.  The refill loads raw
//     aligned dwords (no branches) and realigns them with v_alignbyte; the
//     ring's first 32 words are mirrored past its end so a round's reads
//     never wrap;
//   * decoded bytes are staged in LDS and written 1 KiB (32 rounds) at a
//     time with one 16-byte store per lane;
//   * order-1: symbol i's context is symbol i-1 (scalar.rs:203), a serial
//     chain through the round.  Each present context (K <= 64) owns a lane.
//     Each lane copies the code row of its own state's slot (u8 codes of the
//     symbol every context decodes there; the rows are stored once per
//     distinct row, found through a slot -> row index in LDS) into an LDS
//     matrix, one row per position; each lane then follows four positions
//     from its own context with three dependent LDS reads, so the chain
//     across the 32 positions is 8 v_readlane steps (v_readlane issues
//     about once per 25 clocks on gfx950, so the round avoids it everywhere
//     else).  The state update is lane-parallel from a (context x symbol)
//     table in LDS.  More than 64 contexts take a column gather and a
//     32-step chain with four contexts packed per lane.
#include "rans_device.h"
#include "rans_serve.h"

#include <mutex>

namespace {

constexpr u32 RING_CHUNK_WORDS = 256;                // 512 bytes per chunk
constexpr u32 RING_CHUNKS = 4;
constexpr u32 RING_WORDS = RING_CHUNK_WORDS * RING_CHUNKS;
constexpr u32 RING_MIRROR = 64;                      // words mirrored past the end
constexpr u32 OUT_ROUNDS = 32;                       // rounds per output flush
constexpr u32 OUT_BYTES = OUT_ROUNDS * NSTATES;      // 1 KiB
constexpr u32 UMAX = 64;                             // order-1 fast path lane space
constexpr u32 O1_BLOCK = 8;                          // order-1 rounds per branch-free block
// Order-1 code matrix rows: 64 codes plus 16 bytes of padding, so the
// 16-byte row writes of consecutive lanes start on different LDS banks.
constexpr u32 MROW = 80;

// Order-1 region past the order-0 slot table: the whole workgroup stays within
// 40 KiB, so four one-wave workgroups still share a CU's 160 KiB of LDS.
constexpr u32 O1_EXTRA_WORDS = 1824;
constexpr u32 TAB_WORDS = TOTFREQ + O1_EXTRA_WORDS;

struct DecShared {
    u32 tab[TAB_WORDS];                               // o0 slot table; o1: fc2[K][nu], then the code rows
    u32 ring[(RING_WORDS + RING_MIRROR) / 2];         // renormalisation word ring
    union {
        u32 outst[OUT_BYTES / 4];                     // output staging (decode loop)
        u32 fc[256];                                  // f << 16 | cum (table builds)
    };
    u16 dense[256];                                   // order-1: context byte -> dense id
    u8 ulane[256];                                    // order-1: symbol -> lane in U
    u8 mat[64 * MROW];                                // order-1: the round's code rows, one per position
                                                      // (set-up: slot boundary bitmap, 128 words)
    alignas(16) u16 rowidx[TOTFREQ];                  // order-1: slot -> distinct code row
};
static_assert(sizeof(DecShared) <= 40960, "four decode workgroups per CU");

// Words region reader (see the file comment).  Chunk c covers words
// [256c, 256c + 256).  While the read position is in chunk c, chunks c, c+1
// and c+2 are resident.  Synchronous form: entering chunk c loads chunk c+2
// and stores it into the ring at once.  Asynchronous form: entering chunk c
// stores chunk c+2 (whose loads were issued on entering chunk c-1) and
// issues the loads of chunk c+3, so a chunk's HBM latency is spread over a
// whole chunk of rounds (a synchronous refill cost ~0.5 us every 27 rounds of
// 4.7-bit data).  Loops that wait on other global loads (order-1 rows in the
// workspace) use the synchronous form: a later vmcnt wait would also wait
// for the refill's loads.
struct WordRing {
    const u32 *A;  // 4-byte-aligned address at or below the words region
    u32 ph;        // byte offset of the region within A[0]
    i64 last;      // index of the last dword of A holding a stream byte (>= -1)
    u64 chunk;
    u32 r0, r1, r2, r3;

    __device__ void init(const u8 *wbase, const u8 *end)
    {
        ph = (u32)((uintptr_t)wbase & 3);
        A = (const u32 *)(wbase - ph);
        last = ((i64)(end - wbase) + (i64)ph - 1) >> 2;
    }
    // Raw loads for chunk c: dwords 128c + l (+1) and 128c + 64 + l (+1),
    // clamped to the last dword that holds a stream byte.  Words past the end
    // of the stream are never consumed (the caller checks the count).
    __device__ __forceinline__ void fetch(u64 c)
    {
        const i64 q = (i64)(128 * c) + (i64)lane_id();
        r0 = A[q < last ? q : last];
        r1 = A[q + 1 < last ? q + 1 : last];
        r2 = A[q + 64 < last ? q + 64 : last];
        r3 = A[q + 65 < last ? q + 65 : last];
    }
    __device__ __forceinline__ void put(u32 *ring, u64 c) const
    {
        const u32 l = lane_id();
        const u32 v0 = __builtin_amdgcn_alignbyte(r1, r0, ph);
        const u32 v1 = __builtin_amdgcn_alignbyte(r3, r2, ph);
        const u32 slot = (u32)(c % RING_CHUNKS) * (RING_CHUNK_WORDS / 2);
        ring[slot + l] = v0;
        ring[slot + 64 + l] = v1;
        if (slot == 0 && l < RING_MIRROR / 2) ring[RING_WORDS / 2 + l] = v0;
    }
    template <bool ASYNC>
    __device__ void reset(u32 *ring, u64 c0)
    {
        chunk = c0;
        for (u64 c = c0; c < c0 + 3; c++) {
            fetch(c);
            put(ring, c);
        }
        if (ASYNC) fetch(c0 + 3);
        __syncthreads();
    }
    // Called after every step of at most RING_CHUNK_WORDS words.
    template <bool ASYNC>
    __device__ __forceinline__ void advance(u32 *ring, u64 pos)
    {
        if ((pos / RING_CHUNK_WORDS) != chunk) {
            chunk++;
            if (ASYNC) {
                put(ring, chunk + 2);
                fetch(chunk + 3);
            } else {
                fetch(chunk + 2);
                put(ring, chunk + 2);
            }
        }
    }
    // Word pos + rank, rank < 32.
    __device__ __forceinline__ u32 word(const u32 *ring, u64 pos, u32 rank) const
    {
        return ((const u16 *)ring)[(u32)(pos % RING_WORDS) + rank];
    }
    __device__ u32 word_direct(u64 idx) const
    {
        const u8 *p = (const u8 *)A + ph + 2 * idx;
        return (u32)p[0] | ((u32)p[1] << 8);
    }
};

// Output staging: lane l < 32 writes its byte of round `row` into the stage;
// every 32 rounds the stage is written out, 16 bytes per lane.
__device__ __forceinline__ void stage_byte(DecShared &sh, u32 row, u32 sym)
{
    ((u8 *)sh.outst)[row * NSTATES + lane_id()] = (u8)sym;
}

__device__ __forceinline__ void flush(const DecShared &sh, u8 *dst, u32 nbytes)
{
    const u32 l = lane_id();
    if (nbytes == OUT_BYTES && ((uintptr_t)dst & 15) == 0) {
        const uint4 v = ((const uint4 *)sh.outst)[l];
        ((uint4 *)dst)[l] = v;
    } else {
        const u8 *st = (const u8 *)sh.outst;
        for (u32 i = 0; i < 16; i++) {
            const u32 b = 16 * l + i;
            if (b < nbytes) dst[b] = st[b];
        }
    }
}

// Read one frequency table (256 little-endian u16 at p), validate it
// (model.rs:160-174), and leave f << 16 | cum in sh.fc.  Returns 0 or the
// error code.  Lane l owns symbols 4l..4l+3.
__device__ int load_freq_table(DecShared &sh, const u8 *p, const u8 *end)
{
    const u32 l = lane_id();
    const u32 w0 = load4(p + 8 * l, end), w1 = load4(p + 8 * l + 4, end);
    u32 f[4] = {w0 & 0xffff, w0 >> 16, w1 & 0xffff, w1 >> 16};
    const u32 mine = f[0] + f[1] + f[2] + f[3];
    const u32 incl = wave_scan_incl(mine);
    const u32 total = uni(__shfl(incl, 63));
    // Frequencies are non-negative, so some running sum exceeds TOTFREQ
    // exactly when the total does.
    if (total > TOTFREQ) return FQXV_RANS_ERR_FREQ_EXCEEDS;
    if (total != TOTFREQ) return FQXV_RANS_ERR_FREQ_SUM;
    u32 c = incl - mine;
    for (int i = 0; i < 4; i++) {
        sh.fc[4 * l + i] = (f[i] << 16) | c;
        c += f[i];
    }
    __syncthreads();
    return 0;
}

// Owner symbol of every slot from sh.fc: lane l visits slots [64l, 64l+64).
// emit(slot, sym, f, cum) is called once per slot.
template <typename Emit>
__device__ __forceinline__ void for_each_slot(const DecShared &sh, Emit emit)
{
    const u32 l = lane_id();
    const u32 k0 = 64 * l;
    // Largest s with cum[s] <= k0 is the owner of slot k0 (it has f > 0).
    u32 lo = 0, hi = 255;
    while (lo < hi) {
        u32 mid = (lo + hi + 1) / 2;
        if ((sh.fc[mid] & 0xffff) <= k0)
            lo = mid;
        else
            hi = mid - 1;
    }
    u32 s = lo;
    u32 fcs = sh.fc[s];
    for (u32 k = k0; k < k0 + 64; k++) {
        while ((fcs & 0xffff) + (fcs >> 16) <= k) fcs = sh.fc[++s];
        emit(k, s, fcs >> 16, fcs & 0xffff);
    }
}

__device__ __forceinline__ u32 read_state(const u8 *p, const u8 *end)
{
    return lane_id() < NSTATES ? load4(p + 4 * lane_id(), end) : 0u;
}

// Exact scalar step for a round in which some state would need more than
// one word (only a malformed stream gets there: a valid state never reaches
// v == 0, see step_state, scalar.rs:229-238).  Lanes go in ascending order,
// each pulling words until v >= RANS_L.
__device__ int slow_round(u32 &v, u32 act_mask_lo, u64 &pos, u64 avail, const WordRing &wr)
{
    const u32 l = lane_id();
    int err = 0;
    for (u32 j = 0; j < NSTATES; j++) {
        if (!((act_mask_lo >> j) & 1)) continue;
        u32 vj = __builtin_amdgcn_readlane(v, j);
        while (vj < RANS_L) {
            if (pos >= avail) {
                err = FQXV_RANS_ERR_TRUNCATED;
                break;
            }
            vj = (vj << 16) | wr.word_direct(pos);
            pos++;
        }
        if (err) break;
        if (l == j) v = vj;
    }
    return err;
}

// Renormalise the active lanes' v and commit them to x.  Returns 0 or the
// truncation error.  Words are taken in lane order (scalar.rs:171-176).
template <bool ASYNC>
__device__ __forceinline__ int renorm(DecShared &sh, WordRing &wr, u32 &x, u32 v, bool act, u64 &pos,
                                      u64 avail)
{
    if (ballot(act && v == 0)) {
        int err = slow_round(v, (u32)ballot(act), pos, avail, wr);
        if (err) return err;
        if (act) x = v;
        wr.reset<ASYNC>(sh.ring, pos / RING_CHUNK_WORDS);
        return 0;
    }
    const bool need = act && v < RANS_L;
    const u64 m = ballot(need);
    const u32 k = (u32)__popcll(m);
    if (pos + k > avail) return FQXV_RANS_ERR_TRUNCATED;
    const u32 w = wr.word(sh.ring, pos, rank_below(m));
    if (act) x = need ? ((v << 16) | w) : v;
    pos += k;
    wr.advance<ASYNC>(sh.ring, pos);
    return 0;
}

// ---------------------------------------------------------------- order-0
// Rounds per branch-free order-0 block: one block consumes at most 256 words
// (one ring chunk), so the ring's refill keeps up with a block at a time.
constexpr u32 O0_BLOCK = 8;

__device__ __attribute__((always_inline)) int decode_o0(DecShared &sh, const u8 *s, const u8 *end, u8 *out, u64 n, u64 cap)
{
    const u32 l = lane_id();
    const u64 len = (u64)(end - s);
    if (len < HDR0 + 512) return FQXV_RANS_ERR_TRUNCATED;
    int err = load_freq_table(sh, s + HDR0, end);
    if (err) return err;
    if (len < O0_HDR) return FQXV_RANS_ERR_TRUNCATED;
    u32 x = read_state(s + HDR0 + 512, end);
    if (n > cap) return FQXV_RANS_ERR_OUTPUT_ALLOC;

    for_each_slot(sh, [&](u32 k, u32 sym, u32 f, u32 cum) {
        sh.tab[k] = ((f - 1) << 20) | ((k - cum) << 8) | sym;
    });

    WordRing wr;
    wr.init(s + O0_HDR, end);
    const u64 avail = (len - O0_HDR) / 2;
    wr.reset<true>(sh.ring, 0);

    const bool lo = l < NSTATES;
    const u64 full = n / NSTATES;
    const u32 rem = (u32)(n % NSTATES);
    u64 pos = 0;
    u32 row = 0;
    u8 *o = out;
    u64 r = 0;
    const u16 *ring16 = (const u16 *)sh.ring;
    while (r < full) {
        if ((row % O0_BLOCK) == 0) {
            // Branch-free blocks of O0_BLOCK rounds, as many as the rounds
            // left and the words left allow (counted once, in 32 bits: a
            // block consumes at most 32 * O0_BLOCK words).  No truncation is
            // possible inside, the ring holds every word a block can reach
            // (at most 256 consumed and 64 candidates past the last round's
            // start; the ring keeps 512 ahead of the chunk start), and a
            // state that would need a second word (only v == 0 does, and
            // only then is the renormalised state below 2^16) sends the
            // block back to the exact per-round path below.
            //
            // Pipeline: p is the slot entry of this lane's state; lane i
            // holds the candidate word cw = word[cb + i] and its slot entry
            // cp (a state renormalised with word w lands on slot w & 0xfff).
            // A round's renormalising lanes take their word and next entry
            // from the candidates with ds_bpermute at (pos - cb) + rank, so
            // after the ballot one LDS-crossbar trip stands between a round
            // and the next; the word read and its table lookup happen a
            // round earlier, from the position the round starts at (a round
            // consumes at most 32 words, so the next round's words lie in
            // [pos, pos + 64)).  Lanes that do not renormalise look up
            // v & 0xfff while the ballot resolves.  Lanes 32..63 run the
            // same steps on a meaningless state: table indices stay in
            // range, their bits are masked out of every count and their
            // stores go to `dense` (unused by order-0), so the round needs
            // no lane masks.
            const u64 bw = (avail - pos) / (NSTATES * O0_BLOCK), br = (full - r) / O0_BLOCK;
            u32 nblk = (u32)((bw < br ? bw : br) < 0x7fffffffull ? (bw < br ? bw : br) : 0x7fffffffull);
            if (nblk) {
                u32 p = sh.tab[x & (TOTFREQ - 1)];
                u64 cb = pos;
                u32 cw = ring16[(u32)(pos % RING_WORDS) + l];
                u32 cp = sh.tab[cw & (TOTFREQ - 1)];
                for (; nblk; nblk--) {
                    const u32 x0 = x;
                    const u64 pos0 = pos;
                    u32 bad = 0;
#pragma unroll
                    for (u32 t = 0; t < O0_BLOCK; t++) {
                        // The next round's candidate words are read first;
                        // their table entries are looked up after this
                        // round's own reads, when they have arrived.
                        u32 nw = ring16[(u32)(pos % RING_WORDS) + l];
                        // Pin the schedule (the compiler otherwise waits on
                        // the word read at once: one LDS round trip per
                        // round).
                        asm volatile("" : "+v"(x), "+v"(p)::"memory");
                        const u32 v = ((p >> 20) + 1) * (x >> SCALE_BITS) + ((p >> 8) & 0xfff);
                        const bool need = v < RANS_L;
                        const u32 m = (u32)ballot(need);
                        const u32 ia = ((u32)(pos - cb) + __builtin_amdgcn_mbcnt_lo(m, 0u)) << 2;
                        u32 wn = (u32)__builtin_amdgcn_ds_bpermute((int)ia, (int)cw);
                        u32 pn = (u32)__builtin_amdgcn_ds_bpermute((int)ia, (int)cp);
                        u32 pv = sh.tab[v & (TOTFREQ - 1)];
                        asm volatile("" : "+v"(nw)::"memory");
                        const u32 np = sh.tab[nw & (TOTFREQ - 1)];
                        asm volatile("" : "+v"(wn), "+v"(pn), "+v"(pv)::"memory");
                        *(lo ? (u8 *)sh.outst + (row + t) * NSTATES + l : (u8 *)sh.dense + (l - NSTATES)) = (u8)p;
                        x = need ? ((v << 16) | wn) : v;
                        p = need ? pn : pv;
                        cb = pos;
                        cw = nw;
                        cp = np;
                        pos += (u32)__popc(m);
                        bad |= (u32)ballot(x < RANS_L);
                    }
                    if (bad) {
                        x = x0;
                        pos = pos0;
                        break;
                    }
                    r += O0_BLOCK;
                    row += O0_BLOCK;
                    wr.advance<true>(sh.ring, pos);
                    if (row == OUT_ROUNDS) {
                        flush(sh, o, OUT_BYTES);
                        o += OUT_BYTES;
                        row = 0;
                    }
                }
                if (r >= full) break;
            }
        }
        // Exact per-round path: truncation, multi-word states, a block that
        // met a second word, rounds before a block boundary.
        const u32 pe = sh.tab[x & (TOTFREQ - 1)];
        const u32 v = ((pe >> 20) + 1) * (x >> SCALE_BITS) + ((pe >> 8) & 0xfff);
        if ((err = renorm<true>(sh, wr, x, v, lo, pos, avail))) return err;
        if (lo) stage_byte(sh, row, pe);
        r++;
        if (++row == OUT_ROUNDS) {
            flush(sh, o, OUT_BYTES);
            o += OUT_BYTES;
            row = 0;
        }
    }
    if (rem) {
        const bool act = l < rem;
        const u32 pe = sh.tab[x & (TOTFREQ - 1)];
        const u32 v = ((pe >> 20) + 1) * (x >> SCALE_BITS) + ((pe >> 8) & 0xfff);
        if ((err = renorm<true>(sh, wr, x, v, act, pos, avail))) return err;
        if (act) stage_byte(sh, row, pe);
    }
    flush(sh, o, row * NSTATES + rem);
    return 0;
}

// ---------------------------------------------------------------- order-1
// Block setup shared by both order-1 paths: presence bitmap, then every
// table read and validated in ascending context order (scalar.rs:181-194).
// Leaves K, sh.dense, and the symbol presence of lane l's symbols in symp.
// Also marks in the first 128 words of sh.mat every slot > 0 at which some
// context's symbol starts (its cum): the code row of a slot, the symbol every
// context decodes there, changes only at those slots.
__device__ int o1_validate(DecShared &sh, const u8 *s, const u8 *end, u32 &K, u32 &symp)
{
    const u32 l = lane_id();
    const u64 len = (u64)(end - s);
    u32 *bnd = (u32 *)sh.mat;
    bnd[2 * l] = 0;
    bnd[2 * l + 1] = 0;
    if (len < HDR0 + BITMAP) return FQXV_RANS_ERR_TRUNCATED;
    const u32 bm = s[HDR0 + l / 2];
    const u32 bits = (bm >> (4 * (l & 1))) & 0xf;
    const u32 cnt = __popc(bits);
    const u32 excl = wave_scan_incl(cnt) - cnt;
    K = uni(wave_sum(cnt));
    u32 d = excl;
    for (int i = 0; i < 4; i++) {
        const bool pr = (bits >> i) & 1;
        sh.dense[4 * l + i] = pr ? (u16)d : (u16)0xffff;
        d += pr;
    }
    __syncthreads();
    symp = 0;
    const u8 *t = s + HDR0 + BITMAP;
    for (u32 c = 0; c < K; c++, t += 512) {
        if ((u64)(t + 512 - s) > len) return FQXV_RANS_ERR_TRUNCATED;
        int err = load_freq_table(sh, t, end);
        if (err) return err;
        for (int i = 0; i < 4; i++) {
            const u32 fcv = sh.fc[4 * l + i], cum = fcv & 0xffff;
            symp |= (u32)((fcv >> 16) != 0) << i;
            if ((fcv >> 16) != 0 && cum != 0) atomicOr(&bnd[cum >> 5], 1u << (cum & 31));
        }
    }
    __syncthreads();
    return 0;
}

// Where a wave's order-1 code table starts in its workspace: a per-workgroup
// offset (multiplicative hash of the workgroup id, 128-byte steps, below
// 1 MiB).  Workspaces are WS_PER_WAVE = 2.25 MiB apart, so tables at offset 0
// of each would all start on the same few L2 set indices and the tables of
// the resident waves would compete for the same sets: measured 2.5% slower
// on a batch of 10,000 mixed blocks, and a 4 MiB workspace stride 9% slower.
constexpr u64 TABLE_SKEW_MAX = 8191ull << 7;
static_assert(TABLE_SKEW_MAX + (u64)TOTFREQ * 256 <= WS_CODES, "skewed order-1 tables fit the code region");
__device__ __forceinline__ u8 *table_base(u8 *ws)
{
    return ws + ((((u64)blockIdx.x * 2654435761u) & 0xffffffffu) >> 19 << 7);
}

// Distinct code rows.  The row of slot k (the code every context decodes at
// k) changes only at the boundary slots o1_validate marked in sh.mat, so the
// rows are numbered in slot order: rowidx[k] = number of boundaries in
// [1, k].  There are at most 1 + the sum over contexts of their symbol counts
// (tens to a few hundred), against 4096 slots.  Returns the row count.
__device__ u32 build_rowidx(DecShared &sh)
{
    const u32 l = lane_id();
    const u32 *bm = (const u32 *)sh.mat;
    const u32 w0 = bm[2 * l], w1 = bm[2 * l + 1];
    const u32 cnt = __popc(w0) + __popc(w1);
    u32 base = wave_scan_incl(cnt) - cnt;
#pragma unroll
    for (u32 h = 0; h < 2; h++) {
        const u32 w = h ? w1 : w0;
#pragma unroll
        for (u32 q = 0; q < 4; q++) {
            u32 v[4];
#pragma unroll
            for (u32 i = 0; i < 4; i++) {
                const u32 b = 8 * q + 2 * i; // slots 64l + 32h + b, b + 1
                const u32 lo16 = base + __popc(w & ((2u << b) - 1));
                const u32 hi16 = base + __popc(w & ((2u << (b + 1)) - 1));
                v[i] = lo16 | (hi16 << 16);
            }
            *(uint4 *)&sh.rowidx[64 * l + 32 * h + 8 * q] = make_uint4(v[0], v[1], v[2], v[3]);
        }
        base += __popc(w);
    }
    __syncthreads();
    return uni(sh.rowidx[TOTFREQ - 1]) + 1;
}

// Fast path: K <= 64 contexts and |contexts u symbols| = nu <= 64.
// t8[row * S + d] (u8) = lane(sym) for each distinct code row, S = K rounded
// up to 16, the bytes past K zero (context 0, a valid lane): in LDS behind
// fc2 when it fits, else in the wave's workspace.  LDS: rowidx[slot] = the
// slot's row, fc2[d * nu + lane(sym)] = (f-1) << 20 | cum << 8 | sym, and the
// round's matrix mat[j][c] (rows of MROW bytes) = the code context c decodes
// to at position j.
//
// Per round: each lane looks up its slot's row and loads 64 bytes there
// (issued at the end of the round before) and stores them as its matrix
// row; lane c then reads mat[4g][c], mat[4g+1][.], mat[4g+2][.],
// mat[4g+3][.] along its own chain (8 groups at once), packs the four codes,
// and the chain takes one v_readlane per group.  Measured against a column
// gather (one v_readlane for each slot's row offset, then 32 dependent
// v_readlane steps): 555 vs 811 ns per round on one 4 MiB block, and
// composing the candidate maps with ds_bpermute instead of LDS rows was
// slower than either.  Rounds run in branch-free blocks of four with one
// check per block.  A table of one row per slot (4096 * S bytes, 196 KiB at
// K = 41) was read the same way; with four blocks per CU decoding at once
// those tables overflowed L2 (hit rate 40% against 97% at one per CU) and a
// round slowed from 555 to 810 ns.  The distinct rows number 21 (DNA), about
// 100 (names) and about 300 (K = 41 qualities) on the test corpora.
template <bool LDS_ROWS, u32 CW>
__device__ __forceinline__ int decode_o1_fast(DecShared &sh, const u8 *s, const u8 *end, u8 *out, u64 n, u32 K,
                                              u32 nu, u32 x, u64 hdr, u8 *ws, u32 rows_off)
{
    const u32 l = lane_id();
    // The code rows: in LDS behind fc2 when they fit (LDS_ROWS), else in the
    // wave's workspace.
    u8 *t8 = LDS_ROWS ? (u8 *)sh.tab + rows_off : table_base(ws);
    u32 *fc2 = sh.tab;
    // Distinct code rows (see build_rowidx): t8 holds one row per number.
    // The small table stays in L1/L2 while many blocks decode at once.
    const u32 nrows = build_rowidx(sh);
    // Rows are 16-byte aligned: S >= K bytes each.  A round reads 64 bytes
    // from each row start: past S they are the codes of the next rows, and
    // the last row's over-read stays inside the region (the caller leaves
    // 64 bytes for it in LDS; in the workspace the table ends at least 1 MiB
    // before the end).  The last 16 bytes of every row are zeroed first, so
    // the bytes past K read as context 0 rather than a stale value.
    const u32 S = (K + 15) & ~15u;
    for (u32 i = l; i < nrows; i += 64) *(uint4 *)(t8 + i * S + (S - 16)) = make_uint4(0, 0, 0, 0);
    __builtin_amdgcn_s_waitcnt(0);
    {
        const u32 *bm = (const u32 *)sh.mat;
        // Row starts in this lane's slots [64l, 64l + 64): the boundaries,
        // and slot 0 (row 0) on lane 0.
        const u64 starts = ((u64)bm[2 * l + 1] << 32 | bm[2 * l]) | (l == 0 ? 1ull : 0ull);
        const u32 row0 = l == 0 ? 0u : (u32)sh.rowidx[64 * l] - (u32)(bm[2 * l] & 1u) + 1u;
        const u8 *t = s + HDR0 + BITMAP;
        for (u32 d = 0; d < K; d++, t += 512) {
            (void)load_freq_table(sh, t, end); // validated already
#pragma unroll
            for (u32 i = 0; i < 4; i++) {
                const u32 sym = 4 * l + i, fcv = sh.fc[sym], f = fcv >> 16;
                if (f) fc2[d * nu + sh.ulane[sym]] = ((f - 1) << 20) | ((fcv & 0xffff) << 8) | sym;
            }
            if (starts) {
                // Owner of slot 64l: the largest symbol with cum <= 64l.
                const u32 k0 = 64 * l;
                u32 lo = 0, hi = 255;
                while (lo < hi) {
                    const u32 mid = (lo + hi + 1) / 2;
                    if ((sh.fc[mid] & 0xffff) <= k0)
                        lo = mid;
                    else
                        hi = mid - 1;
                }
                u32 sy = lo, fcs = sh.fc[sy], ri = row0;
                for (u64 rs = starts; rs; rs &= rs - 1, ri++) {
                    const u32 k = k0 + (u32)__builtin_ctzll(rs);
                    while ((fcs & 0xffff) + (fcs >> 16) <= k) fcs = sh.fc[++sy];
                    t8[ri * S + d] = sh.ulane[sy];
                }
            }
            __syncthreads();
        }
    }
    if (!LDS_ROWS) __threadfence();

    WordRing wr;
    wr.init(s + hdr, end);
    const u64 avail = ((u64)(end - s) - hdr) / 2;
    wr.reset<LDS_ROWS>(sh.ring, 0);

    u32 prev = sh.dense[0]; // context of symbol 0 is 0 (scalar.rs:203)
    const bool lo = l < NSTATES;
    const u64 rounds = (n + NSTATES - 1) / NSTATES;
    const u64 full = n / NSTATES;
    u64 pos = 0;
    u32 row = 0;
    u8 *o = out;
    u64 r = 0;
    u8 *const mrow = sh.mat + MROW * l;
    // Matrix bytes past the S copied per row stay zero (context 0).
#pragma unroll
    for (u32 i = 0; i < 4; i++) ((uint4 *)mrow)[i] = make_uint4(0, 0, 0, 0);
    // Lane l < 32 copies the code row of its own slot into row `l` of the
    // LDS matrix: mat[j][c] is the code context c decodes to at position j.
    // Wide per-lane loads and stores, no cross-lane broadcast of addresses.
    // A round's rows are loaded as soon as its states are known, at the end
    // of the round before, so their latency overlaps that round's
    // bookkeeping.  Rows of at most CW codes: CW / 16 chunks of 16 bytes.
    constexpr u32 NCH = CW >= 16 ? CW / 16 : 1;
    constexpr u32 REP = 64 / CW;          // lanes per context: groups split REP ways
    constexpr u32 GPL = (NSTATES / 4) / REP; // groups of four positions per lane
    typedef u32 v4u __attribute__((ext_vector_type(4)));
    v4u q[NCH];
    // The row address is formed from the LDS array or the workspace pointer
    // directly (through a generic pointer the loads become flat loads).
    u8 *const tws = table_base(ws);
    auto load_rows = [&](u32 ri) {
        if (lo) {
            const u32 off = __umul24(ri, S);
            if constexpr (LDS_ROWS) {
                const v4u *src_ = (const v4u *)((const u8 *)sh.tab + rows_off + off);
#pragma unroll
                for (u32 i = 0; i < NCH; i++) q[i] = src_[i];
            } else {
                const v4u *src_ = (const v4u *)(tws + off);
#pragma unroll
                for (u32 i = 0; i < NCH; i++) q[i] = src_[i];
            }
        }
    };
    auto store_rows = [&]() {
        if (lo) {
#pragma unroll
            for (u32 i = 0; i < NCH; i++) ((v4u *)mrow)[i] = q[i];
        }
        __builtin_amdgcn_wave_barrier();
        asm volatile("" ::: "memory");
    };
    // Lane L works for code context c = L % CW on the groups g = L / CW,
    // L / CW + REP, ... (REP copies of the context lanes when the code space
    // is small, so a round's matrix walk takes 8 / REP reads per level).
    const u32 lc = REP > 1 ? l % CW : l;
    const u32 mbase = REP > 1 ? MROW * 4 * (l / CW) : 0;
    const u16 *ring16 = (const u16 *)sh.ring;
    u32 ri = sh.rowidx[x & (TOTFREQ - 1)];
    load_rows(ri);
    while (r < rounds) {
        u32 nblk = (u32)(((full - r) / O1_BLOCK) < 0x7fffffffull ? (full - r) / O1_BLOCK : 0x7fffffffull);
        if ((row % O1_BLOCK) == 0 && nblk) {
            // Branch-free blocks of O1_BLOCK rounds.  A round that meets an
            // absent context, a state needing a second word or the end of
            // the word stream marks its lanes bad; the whole block is then
            // undone and redone by the exact path one round at a time.  The
            // ring holds every word a block can reach (at most 256 words
            // consumed and 64 candidates past the last round's start; the
            // ring keeps 512 ahead of the chunk start); words read past the
            // stream end are never kept (the block is undone).
            //
            // Candidates, as in order-0: lane i holds the word at cb + i and
            // the code row of the slot it would give (rowidx[w & 0xfff]); a
            // round's renormalising lanes take both with one ds_bpermute at
            // (pos - cb) + rank, the others look up rowidx[v & 0xfff] while
            // the ballot resolves.  Lanes 32..63 carry meaningless states
            // (indices stay in range, counts use the low 32 ballot bits,
            // their output bytes go to `dense`, unused once the tables are
            // built).
            u64 cb = pos;
            u32 cw = ring16[(u32)(pos % RING_WORDS) + l];
            u32 cri = sh.rowidx[cw & (TOTFREQ - 1)];
            for (; nblk; nblk--) {
                const u32 x0 = x, prev0 = prev;
                const u64 pos0 = pos;
                u32 bad = 0;
#pragma unroll
                for (u32 t = 0; t < O1_BLOCK; t++) {
                    u32 nw = ring16[(u32)(pos % RING_WORDS) + l];
                    asm volatile("" : "+v"(x)::"memory");
                    const u32 slot = x & (TOTFREQ - 1);
                    store_rows();
                    // Lane (c, h) walks four positions from context c at
                    // position 4g for each of its groups g (three dependent
                    // LDS reads per group, all groups at once), so the
                    // chain takes one v_readlane per group: dg holds the
                    // four codes with the last one in byte 0, where the
                    // next step's lane select reads it.
                    u32 dg[GPL], e0[GPL], e1[GPL], e2[GPL], e3[GPL];
#pragma unroll
                    for (u32 i = 0; i < GPL; i++) e0[i] = sh.mat[mbase + MROW * (4 * REP * i) + lc];
#pragma unroll
                    for (u32 i = 0; i < GPL; i++) e1[i] = sh.mat[mbase + MROW * (4 * REP * i + 1) + e0[i]];
#pragma unroll
                    for (u32 i = 0; i < GPL; i++) e2[i] = sh.mat[mbase + MROW * (4 * REP * i + 2) + e1[i]];
#pragma unroll
                    for (u32 i = 0; i < GPL; i++) e3[i] = sh.mat[mbase + MROW * (4 * REP * i + 3) + e2[i]];
                    // All last reads in flight before the first is used.
#pragma unroll
                    for (u32 i = 0; i < GPL; i++) asm volatile("" : "+v"(e3[i]));
                    if constexpr (REP > 1) {
#pragma unroll
                        for (u32 i = 0; i < GPL; i++) dg[i] = e3[i] | (e0[i] << 8) | (e1[i] << 16) | (e2[i] << 24);
                    }
                    u32 c = prev, dv = 0, cv = 0;
#pragma unroll
                    for (u32 g = 0; g < NSTATES / 4; g++) {
                        if constexpr (REP == 1) {
                            // One group per chain step: each group's word is
                            // packed after the step before it (tied to that
                            // step's result), so the packing fills the
                            // chain's readlane hazard slots instead of
                            // delaying its start (K = 41: 495 -> 479 ns per
                            // round; with replicated context lanes the
                            // packing up front measured faster).
                            asm volatile("" : "+v"(e0[g]), "+v"(e1[g]), "+v"(e2[g]), "+v"(e3[g]) : "s"(c));
                            dg[g] = e3[g] | (e0[g] << 8) | (e1[g] << 16) | (e2[g] << 24);
                        }
                        cv = writelane(c, 4 * g, cv);
                        // Byte 0 of c is below CW, so adding the lane group
                        // never carries into the lane select's six bits.
                        c = __builtin_amdgcn_readlane(dg[g / REP], uni(c + CW * (g % REP)));
                        dv = writelane(c, 4 * g, dv);
                    }
                    prev = c & 0xffu;
                    // Every lane of a group of four takes the group's dword;
                    // lane 4g + k reads its symbol's code and its context's
                    // code there.
                    const u32 dq = (u32)__builtin_amdgcn_update_dpp(0, (int)dv, 0x00, 0xf, 0xf, false);
                    const u32 cq = (u32)__builtin_amdgcn_update_dpp(0, (int)cv, 0x00, 0xf, 0xf, false);
                    // Lane 4g takes its context from cv (the group's entry
                    // context), lanes 4g+1..3 from dv: one byte select.
                    const u32 k4 = l & 3;
                    const u32 sc = (dq >> (8 * ((k4 + 1) & 3))) & 0xffu;
                    const u32 cc = (((dq & 0xffffff00u) | (cq & 0xffu)) >> (8 * k4)) & 0xffu;
                    const u32 ent = fc2[__umul24(cc & 63, nu) + (sc & 63)];
                    const u32 v = ((ent >> 20) + 1) * (x >> SCALE_BITS) + slot - ((ent >> 8) & 0xfff);
                    const bool need = v < RANS_L;
                    const u32 m = (u32)ballot(need);
                    const u32 ia = ((u32)(pos - cb) + __builtin_amdgcn_mbcnt_lo(m, 0u)) << 2;
                    u32 qn = (u32)__builtin_amdgcn_ds_bpermute((int)ia, (int)(cw | (cri << 16)));
                    u32 riv = sh.rowidx[v & (TOTFREQ - 1)];
                    asm volatile("" : "+v"(nw)::"memory");
                    const u32 nri = sh.rowidx[nw & (TOTFREQ - 1)];
                    asm volatile("" : "+v"(qn), "+v"(riv)::"memory");
                    *(lo ? (u8 *)sh.outst + (row + t) * NSTATES + l : (u8 *)sh.dense + (l - NSTATES)) = (u8)ent;
                    x = need ? ((v << 16) | (qn & 0xffffu)) : v;
                    ri = need ? (qn >> 16) : riv;
                    cb = pos;
                    cw = nw;
                    cri = nri;
                    pos += (u32)__popc(m);
                    bad |= (u32)ballot(x < RANS_L || cc >= K);
                    load_rows(ri);
                }
                if (bad || pos > avail) {
                    x = x0;
                    prev = prev0;
                    pos = pos0;
                    ri = sh.rowidx[x & (TOTFREQ - 1)];
                    load_rows(ri);
                    break;
                }
                r += O1_BLOCK;
                row += O1_BLOCK;
                wr.advance<LDS_ROWS>(sh.ring, pos);
                if (row == OUT_ROUNDS) {
                    flush(sh, o, OUT_BYTES);
                    o += OUT_BYTES;
                    row = 0;
                }
            }
            if (r >= rounds) break;
        }
        const u32 slot = x & (TOTFREQ - 1);
        store_rows();
        // Exact path: absent contexts, truncation, multi-word states, tail.
        if (prev >= K) return FQXV_RANS_ERR_ABSENT_CONTEXT;
        const u32 cntr = r < full ? NSTATES : (u32)(n - r * NSTATES);
        u32 c = prev;
        u32 cv = 0;
#pragma unroll
        for (u32 j = 0; j < NSTATES; j++) {
            const u32 nxt = sh.mat[MROW * j + (c & 63)];
            cv = writelane(c | (nxt << 8), j, cv);
            c = nxt;
        }
        prev = c;
        // First position whose context is no context: that symbol fails.
        const u64 nb = ballot(lo && l < cntr && (cv & 0xff) >= K);
        const u32 jb = nb ? (u32)__builtin_ctzll(nb) : NSTATES;
        const bool act = lo && l < cntr && l < jb;
        const u32 ent = fc2[(cv & 0x3f) * nu + ((cv >> 8) & 0x3f)];
        const u32 v = ((ent >> 20) + 1) * (x >> SCALE_BITS) + slot - ((ent >> 8) & 0xfff);
        int err = renorm<LDS_ROWS>(sh, wr, x, v, act, pos, avail);
        if (err) return err;
        if (act) stage_byte(sh, row, ent);
        if (jb < cntr) return FQXV_RANS_ERR_ABSENT_CONTEXT;
        ri = sh.rowidx[x & (TOTFREQ - 1)];
        load_rows(ri);
        r++;
        if (++row == OUT_ROUNDS) {
            // The last round may be partial: never write past out + n.
            const u64 left = n - (u64)(o - out);
            flush(sh, o, left < OUT_BYTES ? (u32)left : OUT_BYTES);
            o += OUT_BYTES;
            row = 0;
        }
    }
    if (row) flush(sh, o, (u32)(n - (u64)(o - out)));
    return 0;
}

// Register path: a code space of at most eight lane ids (nu <= 8: DNA-like
// blocks).  The map "context -> code of the symbol decoded" of one slot is
// eight 3-bit codes, held per lane as two byte maps (contexts 0..3 and
// 4..7), so the composition of two maps is two v_perm_b32 (each selector
// byte picks a byte of the 8-byte table).  Per round, lane j takes the map of
// its own state's slot and an inclusive scan over lanes 0..31 (row_shr 1, 2,
// 4, 8 within 16-lane rows, then row_bcast 15 into the second row: five DPP
// steps) leaves lane j the composition of maps 0..j; applied to the round's
// incoming context it is position j's symbol, and a one-lane wave shift
// gives each position its context.  The state entry comes from the row of
// entries of the slot's code row (eight contexts, loaded while the scan runs)
// by the context's three bits, so the chain has no LDS access between the
// round's state and the next one except the renormalisation word.
//
// LDS (sh.tab, words): MT[slot] the packed map of each slot (bits 8i..8i+3
// the code of context i, bits 8i+4..8i+7 that of context 4 + i), ER[row * 8
// + c] the entry (f-1)<<20 | cum<<8 | sym of context c at each distinct code
// row, PRC[row] the packed map of each row.  Codes of contexts >= K are 0
// and their entries 0: a position that meets one is flagged and redone by
// the exact path, which reports the absent context.
constexpr u32 PERM_MT = 0;
constexpr u32 PERM_ER = TOTFREQ;
constexpr u32 PERM_ROWS_MAX = 1 + 8 * 8;  // 1 + sum over <= 8 contexts of <= 8 symbols
constexpr u32 PERM_PRC = PERM_ER + 8 * PERM_ROWS_MAX;
static_assert(PERM_PRC + PERM_ROWS_MAX <= TAB_WORDS, "register-path tables fit the decode table area");

// DPP temporaries of the scan, one pair per step.  A step's lanes without a
// source lane are the same in every round (lanes l % 16 < d for row_shr:d,
// rows 0 and 2 for the row broadcast), and DPP leaves them unwritten, so the
// identity map put there once before the first round stays (no per-round
// re-initialisation of the DPP destinations).
struct ScanTemps {
    u32 l1, h1, l2, h2, l4, h4, l8, h8, lb, hb;
    __device__ void init()
    {
        l1 = l2 = l4 = l8 = lb = 0x03020100u;
        h1 = h2 = h4 = h8 = hb = 0x07060504u;
    }
};

// One scan step: lane j's map becomes (its map) o (the map of the lane CTRL
// names), applied right to left; lanes without a source lane keep theirs
// (their temporaries hold the identity).
template <int CTRL, int RMASK>
__device__ __forceinline__ void compose_step(u32 &lo, u32 &hi, u32 &tl, u32 &th)
{
    tl = (u32)__builtin_amdgcn_update_dpp((int)tl, (int)lo, CTRL, RMASK, 0xf, false);
    th = (u32)__builtin_amdgcn_update_dpp((int)th, (int)hi, CTRL, RMASK, 0xf, false);
    const u32 nl = __builtin_amdgcn_perm(hi, lo, tl);
    const u32 nh = __builtin_amdgcn_perm(hi, lo, th);
    lo = nl;
    hi = nh;
}

// Packed map w of position j (lanes 0..31) -> code of the symbol at j and
// code of its context, given the round's incoming context code cin (< 8).
__device__ __forceinline__ void perm_round_codes(u32 w, u32 cin, ScanTemps &tp, u32 &sc, u32 &cc)
{
    u32 lo = w & 0x0f0f0f0fu, hi = (w >> 4) & 0x0f0f0f0fu;
    compose_step<0x111, 0xf>(lo, hi, tp.l1, tp.h1); // row_shr:1
    compose_step<0x112, 0xf>(lo, hi, tp.l2, tp.h2); // row_shr:2
    compose_step<0x114, 0xf>(lo, hi, tp.l4, tp.h4); // row_shr:4
    compose_step<0x118, 0xf>(lo, hi, tp.l8, tp.h8); // row_shr:8
    compose_step<0x142, 0xa>(lo, hi, tp.lb, tp.hb); // row_bcast:15 into rows 1 and 3
    // Selector byte 0 = cin, bytes 1..3 = 12 (a zero byte).
    sc = __builtin_amdgcn_perm(hi, lo, 0x0c0c0c00u | cin);
    cc = (u32)__builtin_amdgcn_update_dpp((int)cin, (int)sc, 0x138, 0xf, 0xf, false); // wave_shr:1
}

__device__ __forceinline__ u32 perm_entry(uint4 ea, uint4 eb, u32 cc)
{
    const bool b0 = cc & 1;
    const u32 s01 = b0 ? ea.y : ea.x, s23 = b0 ? ea.w : ea.z;
    const u32 s45 = b0 ? eb.y : eb.x, s67 = b0 ? eb.w : eb.z;
    const u32 s03 = (cc & 2) ? s23 : s01, s47 = (cc & 2) ? s67 : s45;
    return (cc & 4) ? s47 : s03;
}

__device__ __forceinline__ int decode_o1_perm(DecShared &sh, const u8 *s, const u8 *end, u8 *out, u64 n, u32 K,
                                              u32 x, u64 hdr)
{
    const u32 l = lane_id();
    u32 *const MT = sh.tab + PERM_MT;
    u32 *const ER = sh.tab + PERM_ER;
    u32 *const PRC = sh.tab + PERM_PRC;
    const u32 nrows = build_rowidx(sh);
    for (u32 i = l; i < 8 * nrows; i += 64) ER[i] = 0;
    for (u32 i = l; i < nrows; i += 64) PRC[i] = 0;
    __syncthreads();
    {
        // Each row starts at one slot (slot 0 or a boundary), owned by one
        // lane, which writes that row's codes and entries context by context.
        const u32 *bm = (const u32 *)sh.mat;
        const u64 starts = ((u64)bm[2 * l + 1] << 32 | bm[2 * l]) | (l == 0 ? 1ull : 0ull);
        const u32 row0 = l == 0 ? 0u : (u32)sh.rowidx[64 * l] - (u32)(bm[2 * l] & 1u) + 1u;
        const u8 *t = s + HDR0 + BITMAP;
        for (u32 d = 0; d < K; d++, t += 512) {
            (void)load_freq_table(sh, t, end); // validated already
            if (starts) {
                // Owner of slot 64l: the largest symbol with cum <= 64l.
                const u32 k0 = 64 * l;
                u32 lo = 0, hi = 255;
                while (lo < hi) {
                    const u32 mid = (lo + hi + 1) / 2;
                    if ((sh.fc[mid] & 0xffff) <= k0)
                        lo = mid;
                    else
                        hi = mid - 1;
                }
                u32 sy = lo, fcs = sh.fc[sy], ri = row0;
                const u32 sh4 = d < 4 ? 8 * d : 8 * (d - 4) + 4;
                for (u64 rs = starts; rs; rs &= rs - 1, ri++) {
                    const u32 k = k0 + (u32)__builtin_ctzll(rs);
                    while ((fcs & 0xffff) + (fcs >> 16) <= k) fcs = sh.fc[++sy];
                    PRC[ri] |= (u32)sh.ulane[sy] << sh4;
                    ER[8 * ri + d] = (((fcs >> 16) - 1) << 20) | ((fcs & 0xffff) << 8) | sy;
                }
            }
            __syncthreads();
        }
        // MT[k] = PRC[rowidx[k]], lane l for slots [64l, 64l + 64).
#pragma unroll 2
        for (u32 k = 64 * l; k < 64 * l + 64; k += 8) {
            const uint4 r4 = *(const uint4 *)&sh.rowidx[k];
            const u32 rr[4] = {r4.x, r4.y, r4.z, r4.w};
            u32 v[8];
#pragma unroll
            for (u32 i = 0; i < 4; i++) {
                v[2 * i] = PRC[rr[i] & 0xffff];
                v[2 * i + 1] = PRC[rr[i] >> 16];
            }
            *(uint4 *)&MT[k] = make_uint4(v[0], v[1], v[2], v[3]);
            *(uint4 *)&MT[k + 4] = make_uint4(v[4], v[5], v[6], v[7]);
        }
        __syncthreads();
    }

    WordRing wr;
    wr.init(s + hdr, end);
    const u64 avail = ((u64)(end - s) - hdr) / 2;
    wr.reset<true>(sh.ring, 0);

    u32 prev = sh.dense[0]; // context of symbol 0 is 0 (scalar.rs:203); >= K if absent
    const bool lo = l < NSTATES;
    const u64 rounds = (n + NSTATES - 1) / NSTATES;
    const u64 full = n / NSTATES;
    u64 pos = 0;
    u32 row = 0;
    u8 *o = out;
    u64 r = 0;
    const u16 *ring16 = (const u16 *)sh.ring;
    // This lane's slot: its packed map and its code row's entries.
    u32 mt = MT[x & (TOTFREQ - 1)];
    u32 ri = sh.rowidx[x & (TOTFREQ - 1)];
    uint4 ea = *(const uint4 *)&ER[8 * ri], eb = *(const uint4 *)&ER[8 * ri + 4];
    ScanTemps tp;
    tp.init();
    while (r < rounds) {
        u32 nblk = (u32)(((full - r) / O1_BLOCK) < 0x7fffffffull ? (full - r) / O1_BLOCK : 0x7fffffffull);
        if ((row % O1_BLOCK) == 0 && nblk && prev < K) {
            // Branch-free blocks of O1_BLOCK rounds, undone and redone by the
            // exact path when a round meets an absent context, a state that
            // needs a second word, or the end of the words (as in
            // decode_o1_fast).  Candidates: lane i holds the word at cb + i,
            // the packed map and the row of the slot it would give; a round's
            // renormalising lanes take them with ds_bpermute at
            // (pos - cb) + rank, the others look up v & 0xfff while the ballot
            // resolves.  Lanes 32..63 carry meaningless states (indices stay
            // in range, counts use the low 32 ballot bits, their output
            // bytes go to `dense`, unused once the tables are built).
            u64 cb = pos;
            u32 cw = ring16[(u32)(pos % RING_WORDS) + l];
            u32 cmt = MT[cw & (TOTFREQ - 1)];
            u32 cri = sh.rowidx[cw & (TOTFREQ - 1)];
            for (; nblk; nblk--) {
                const u32 x0 = x, prev0 = prev;
                const u64 pos0 = pos;
                u32 bad = 0;
#pragma unroll
                for (u32 t = 0; t < O1_BLOCK; t++) {
                    u32 nw = ring16[(u32)(pos % RING_WORDS) + l];
                    asm volatile("" : "+v"(x), "+v"(mt)::"memory");
                    const u32 slot = x & (TOTFREQ - 1);
                    u32 sc, cc;
                    perm_round_codes(mt, prev, tp, sc, cc);
                    const u32 ent = perm_entry(ea, eb, cc);
                    const u32 v = ((ent >> 20) + 1) * (x >> SCALE_BITS) + slot - ((ent >> 8) & 0xfff);
                    const bool need = v < RANS_L;
                    const u32 m = (u32)ballot(need);
                    const u32 ia = ((u32)(pos - cb) + __builtin_amdgcn_mbcnt_lo(m, 0u)) << 2;
                    u32 qn = (u32)__builtin_amdgcn_ds_bpermute((int)ia, (int)(cw | (cri << 16)));
                    u32 qm = (u32)__builtin_amdgcn_ds_bpermute((int)ia, (int)cmt);
                    u32 mv = MT[v & (TOTFREQ - 1)];
                    u32 rv = sh.rowidx[v & (TOTFREQ - 1)];
                    asm volatile("" : "+v"(nw)::"memory");
                    const u32 nmt = MT[nw & (TOTFREQ - 1)];
                    const u32 nri = sh.rowidx[nw & (TOTFREQ - 1)];
                    asm volatile("" : "+v"(qn), "+v"(qm), "+v"(mv), "+v"(rv)::"memory");
                    *(lo ? (u8 *)sh.outst + (row + t) * NSTATES + l : (u8 *)sh.dense + (l - NSTATES)) = (u8)ent;
                    x = need ? ((v << 16) | (qn & 0xffffu)) : v;
                    ri = need ? (qn >> 16) : rv;
                    mt = need ? qm : mv;
                    ea = *(const uint4 *)&ER[8 * ri];
                    eb = *(const uint4 *)&ER[8 * ri + 4];
                    cb = pos;
                    cw = nw;
                    cmt = nmt;
                    cri = nri;
                    pos += (u32)__popc(m);
                    bad |= (u32)ballot(x < RANS_L || cc >= K);
                    prev = __builtin_amdgcn_readlane(sc, NSTATES - 1);
                }
                if (bad || pos > avail) {
                    x = x0;
                    prev = prev0;
                    pos = pos0;
                    mt = MT[x & (TOTFREQ - 1)];
                    ri = sh.rowidx[x & (TOTFREQ - 1)];
                    ea = *(const uint4 *)&ER[8 * ri];
                    eb = *(const uint4 *)&ER[8 * ri + 4];
                    break;
                }
                r += O1_BLOCK;
                row += O1_BLOCK;
                wr.advance<true>(sh.ring, pos);
                if (row == OUT_ROUNDS) {
                    flush(sh, o, OUT_BYTES);
                    o += OUT_BYTES;
                    row = 0;
                }
            }
            if (r >= rounds) break;
        }
        // Exact path: absent contexts, truncation, multi-word states, tail.
        if (prev >= K) return FQXV_RANS_ERR_ABSENT_CONTEXT;
        const u32 slot = x & (TOTFREQ - 1);
        const u32 cntr = r < full ? NSTATES : (u32)(n - r * NSTATES);
        u32 sc, cc;
        perm_round_codes(mt, prev, tp, sc, cc);
        // First position whose context is no context: that symbol fails.
        const u64 nb = ballot(lo && l < cntr && cc >= K);
        const u32 jb = nb ? (u32)__builtin_ctzll(nb) : NSTATES;
        const bool act = lo && l < cntr && l < jb;
        const u32 ent = perm_entry(ea, eb, cc);
        const u32 v = ((ent >> 20) + 1) * (x >> SCALE_BITS) + slot - ((ent >> 8) & 0xfff);
        int err = renorm<true>(sh, wr, x, v, act, pos, avail);
        if (err) return err;
        if (act) stage_byte(sh, row, ent);
        if (jb < cntr) return FQXV_RANS_ERR_ABSENT_CONTEXT;
        prev = __builtin_amdgcn_readlane(sc, NSTATES - 1);
        mt = MT[x & (TOTFREQ - 1)];
        ri = sh.rowidx[x & (TOTFREQ - 1)];
        ea = *(const uint4 *)&ER[8 * ri];
        eb = *(const uint4 *)&ER[8 * ri + 4];
        r++;
        if (++row == OUT_ROUNDS) {
            // The last round may be partial: never write past out + n.
            const u64 left = n - (u64)(o - out);
            flush(sh, o, left < OUT_BYTES ? (u32)left : OUT_BYTES);
            o += OUT_BYTES;
            row = 0;
        }
    }
    if (row) flush(sh, o, (u32)(n - (u64)(o - out)));
    return 0;
}

// Wide path: more than 64 contexts, or 64 contexts and one more symbol
// (|contexts u symbols| <= 256).  The same column gather and v_readlane chain
// as the fast path, with four contexts packed per lane: byte q of lane l of
// column j is the code of the symbol context 64q + l decodes at slot_j, so a
// chain step is one v_readlane at lane (c & 63) (the lane select keeps the low
// six bits) and a scalar extract of byte c >> 6.  Bytes of contexts >= K are
// never chosen: a code >= K leaves the fast round and the exact path stops on
// it.  The (context x symbol) entry table lives in the workspace:
// t32[slot * 64 + (d & 63)] byte (d >> 6) (1 MiB), and
// fcw[d * 256 + code] = (f-1) << 20 | cum << 8 | sym.
__device__ int decode_o1_multi(DecShared &sh, const u8 *s, const u8 *end, u8 *out, u64 n, u32 K, u32 x, u64 hdr,
                               u8 *ws)
{
    const u32 l = lane_id();
    u8 *t8 = table_base(ws);
    u32 *fcw = (u32 *)(ws + WS_CODES);
    const u8 *t = s + HDR0 + BITMAP;
    for (u32 d = 0; d < K; d++, t += 512) {
        (void)load_freq_table(sh, t, end); // validated already
        const u32 col = (d & 63) * 4 + (d >> 6);
        for_each_slot(sh, [&](u32 k, u32 sym, u32 f, u32 cum) {
            const u32 ul = sh.ulane[sym];
            t8[k * 256 + col] = (u8)ul;
            if (k == cum) fcw[d * 256 + ul] = ((f - 1) << 20) | (cum << 8) | sym;
        });
        __syncthreads();
    }
    __threadfence();
    const __amdgpu_buffer_rsrc_t t8rs = __builtin_amdgcn_make_buffer_rsrc(t8, (short)0, (int)(TOTFREQ * 256), 0x00020000);

    WordRing wr;
    wr.init(s + hdr, end);
    const u64 avail = ((u64)(end - s) - hdr) / 2;
    wr.reset<false>(sh.ring, 0);

    u32 prev = sh.dense[0]; // context of symbol 0 is 0 (scalar.rs:203)
    const bool lo = l < NSTATES;
    const u64 rounds = (n + NSTATES - 1) / NSTATES;
    const u64 full = n / NSTATES;
    u64 pos = 0;
    u32 row = 0;
    u8 *o = out;
    u64 r = 0;
    // One chain step: the code context c decodes to, from its packed column.
    auto step = [](u32 cv, u32 c) -> u32 { return (__builtin_amdgcn_readlane(cv, c) >> ((c >> 3) & 24)) & 0xffu; };
    while (r < rounds) {
        const u32 slot = x & (TOTFREQ - 1);
        const u32 rowoff = slot * 256;
        u32 cand[NSTATES];
#pragma unroll
        for (u32 j = 0; j < NSTATES; j++)
            cand[j] = __builtin_amdgcn_raw_buffer_load_b32(t8rs, l * 4, __builtin_amdgcn_readlane(rowoff, j), 0);
        if (r < full && prev < K && pos + NSTATES <= avail) {
            // Fast round (see decode_o1_fast).
            const u32 x0 = x;
            const u64 pos0 = pos;
            u32 c = prev;
            u32 sv = 0;
#pragma unroll
            for (u32 j = 0; j < NSTATES; j++) {
                c = step(cand[j], c);
                sv = writelane(c, j, sv);
            }
            // Position j's context is position j-1's symbol.
            u32 ctxv = (u32)__builtin_amdgcn_update_dpp((int)sv, (int)sv, 0x111, 0xf, 0xf, false);
            ctxv = writelane(prev, 0, ctxv);
            ctxv = writelane(__builtin_amdgcn_readlane(sv, 15), 16, ctxv);
            const u32 ent = lo ? fcw[(ctxv & 0xff) * 256 + (sv & 0xff)] : 0u;
            const u32 v = ((ent >> 20) + 1) * (x >> SCALE_BITS) + slot - ((ent >> 8) & 0xfff);
            const bool need = lo && v < RANS_L;
            const u64 m = ballot(need);
            const u32 w = ((const u16 *)sh.ring)[(u32)(pos % RING_WORDS) + __builtin_amdgcn_mbcnt_lo((u32)m, 0u)];
            if (lo) x = need ? ((v << 16) | w) : v;
            pos += (u32)__popcll(m);
            if (lo) stage_byte(sh, row, ent);
            if (ballot(lo && (x < RANS_L || (ctxv & 0xff) >= K)) == 0) {
                prev = c;
                r++;
                wr.advance<false>(sh.ring, pos);
                if (++row == OUT_ROUNDS) {
                    flush(sh, o, OUT_BYTES);
                    o += OUT_BYTES;
                    row = 0;
                }
                continue;
            }
            x = x0;
            pos = pos0;
        }
        // Exact path: absent contexts, truncation, multi-word states, tail.
        if (prev >= K) return FQXV_RANS_ERR_ABSENT_CONTEXT;
        const u32 cntr = r < full ? NSTATES : (u32)(n - r * NSTATES);
        u32 c = prev;
        u32 cv = 0;
#pragma unroll
        for (u32 j = 0; j < NSTATES; j++) {
            const u32 nxt = c < K ? step(cand[j], c) : 0xffu;
            cv = writelane(c | (nxt << 8), j, cv);
            c = nxt;
        }
        prev = c;
        // First position whose context is no context: that symbol fails.
        const u64 nb = ballot(lo && l < cntr && (cv & 0xff) >= K);
        const u32 jb = nb ? (u32)__builtin_ctzll(nb) : NSTATES;
        const bool act = lo && l < cntr && l < jb;
        const u32 ent = act ? fcw[(cv & 0xff) * 256 + ((cv >> 8) & 0xff)] : 0u;
        const u32 v = ((ent >> 20) + 1) * (x >> SCALE_BITS) + slot - ((ent >> 8) & 0xfff);
        int err = renorm<false>(sh, wr, x, v, act, pos, avail);
        if (err) return err;
        if (act) stage_byte(sh, row, ent);
        if (jb < cntr) return FQXV_RANS_ERR_ABSENT_CONTEXT;
        r++;
        if (++row == OUT_ROUNDS) {
            // The last round may be partial: never write past out + n.
            const u64 left = n - (u64)(o - out);
            flush(sh, o, left < OUT_BYTES ? (u32)left : OUT_BYTES);
            o += OUT_BYTES;
            row = 0;
        }
    }
    if (row) flush(sh, o, (u32)(n - (u64)(o - out)));
    return 0;
}

// rows_lds: the fast path may keep its code rows in LDS (job flags).
__device__ __attribute__((always_inline)) int decode_o1(DecShared &sh, const u8 *s, const u8 *end, u8 *out, u64 n, u64 cap, u8 *ws,
                                                     bool rows_lds)
{
    const u32 l = lane_id();
    u32 K, symp;
    int err = o1_validate(sh, s, end, K, symp);
    if (err) return err;
    const u64 hdr = HDR0 + BITMAP + 512ull * K + 4 * NSTATES;
    if ((u64)(end - s) < hdr) return FQXV_RANS_ERR_TRUNCATED;
    const u32 x = read_state(s + hdr - 4 * NSTATES, end);
    if (n > cap) return FQXV_RANS_ERR_OUTPUT_ALLOC;
    // Lane space U: contexts first (lane = dense id), then the symbols that
    // are no context, ascending.
    u32 extra = 0;
    for (int i = 0; i < 4; i++) extra += ((symp >> i) & 1) && sh.dense[4 * l + i] == 0xffff;
    const u32 ex0 = wave_scan_incl(extra) - extra;
    const u32 nu = K + uni(wave_sum(extra));
    {
        u32 e = K + ex0;
        for (int i = 0; i < 4; i++) {
            const u32 sI = 4 * l + i;
            const u32 dd = sh.dense[sI];
            const bool p = (symp >> i) & 1;
            sh.ulane[sI] = dd != 0xffff ? (u8)(dd < 255 ? dd : 255) : (p ? (u8)(e < 255 ? e : 255) : (u8)0xff);
            e += p && dd == 0xffff;
        }
    }
    __syncthreads();
    if (nu <= 8) return decode_o1_perm(sh, s, end, out, n, K, x, hdr);
    if (K <= UMAX && nu <= UMAX) {
        // Distinct code rows (1 + the boundary slots o1_validate marked) and
        // where they go: LDS behind fc2[K][nu] when they fit with the 64-byte
        // over-read of the last row, else the workspace.
        const u32 *bm = (const u32 *)sh.mat;
        const u32 nrows = 1 + uni(wave_sum(__popc(bm[2 * l]) + __popc(bm[2 * l + 1])));
        const u32 S = (K + 15) & ~15u;
        const u32 rows_off = (K * nu * 4 + 15) & ~15u;
        // The code space (nu lane ids) in 16, 32 or 64 lanes.
        const bool in_lds = rows_lds && rows_off + nrows * S + 64 <= TAB_WORDS * 4;
        if (in_lds) {
            if (nu <= 16) return decode_o1_fast<true, 16>(sh, s, end, out, n, K, nu, x, hdr, ws, rows_off);
            if (nu <= 32) return decode_o1_fast<true, 32>(sh, s, end, out, n, K, nu, x, hdr, ws, rows_off);
            return decode_o1_fast<true, 64>(sh, s, end, out, n, K, nu, x, hdr, ws, rows_off);
        }
        if (nu <= 16) return decode_o1_fast<false, 16>(sh, s, end, out, n, K, nu, x, hdr, ws, 0);
        return decode_o1_fast<false, 64>(sh, s, end, out, n, K, nu, x, hdr, ws, 0);
    }
    return decode_o1_multi(sh, s, end, out, n, K, x, hdr, ws);
}

// One job: header dispatch (scalar.rs:143-155, lib.rs:226-231).  Returns
// the status; *n_out is the declared output length.
__device__ __attribute__((always_inline)) int decode_job_core(DecShared &sh, const u8 *src, u8 *dst, const fqxv_rans_job &jb, u8 *ws, u64 &n_out)
{
    const u8 *s = src + jb.src_off;
    const u8 *end = s + jb.src_len;
    int err = 0;
    u64 n = 0;
    if (jb.flags & ~FQXV_RANS_JOB_O1_ROWS_WORKSPACE) {
        err = FQXV_RANS_ERR_INVALID_ARG;
    } else if (jb.src_len < HDR0) {
        err = FQXV_RANS_ERR_TRUNCATED;
    } else {
        const u32 b = lane_id() < HDR0 ? s[lane_id()] : 0;
        const u32 order = uni(__shfl(b, 0));
        for (int i = 0; i < 8; i++) n |= (u64)uni(__shfl(b, 1 + i)) << (8 * i);
        if (n > jb.max_len) {
            err = FQXV_RANS_ERR_BOUND;
        } else if (n != 0) {
            if (order == 0)
                err = decode_o0(sh, s, end, dst + jb.dst_off, n, jb.dst_cap);
            else if (order == 1) {
                // Long order-1 blocks are the batch's critical path: their
                // wave issues first on a shared SIMD.
                const bool crit = n >= (1u << 20);
                if (crit) __builtin_amdgcn_s_setprio(3);
                err = decode_o1(sh, s, end, dst + jb.dst_off, n, jb.dst_cap, ws,
                                (jb.flags & FQXV_RANS_JOB_O1_ROWS_WORKSPACE) == 0);
                if (crit) __builtin_amdgcn_s_setprio(0);
            } else
                err = FQXV_RANS_ERR_UNKNOWN_ORDER;
        }
    }
    n_out = n;
    return err;
}

__device__ __attribute__((always_inline)) void decode_job(DecShared &sh, const u8 *src, u8 *dst, const fqxv_rans_job &jb, i32 *status,
                           u64 *out_len, u8 *ws)
{
    u64 n = 0;
    const int err = decode_job_core(sh, src, dst, jb, ws, n);
    if (lane_id() == 0) {
        *status = err;
        *out_len = err ? 0 : n;
    }
}

} // namespace

__global__ __launch_bounds__(64) void k_rans_decode(const u8 *__restrict__ src, u8 *__restrict__ dst,
                                                    const fqxv_rans_job *__restrict__ jobs, u32 count,
                                                    const u32 *__restrict__ order, u32 *counter,
                                                    i32 *status, u64 *out_len, u8 *ws)
{
    __shared__ DecShared sh;
    u8 *myws = ws + (u64)blockIdx.x * WS_PER_WAVE;
    for (;;) {
        u32 w = 0;
        if (lane_id() == 0) w = atomicAdd(counter, 1u);
        w = uni(__shfl(w, 0));
        if (w >= count) break;
        const u32 j = order ? order[w] : w;
        const fqxv_rans_job jb = jobs[j];
        decode_job(sh, src, dst, jb, status + j, out_len + j, myws);
        __syncthreads();
    }
}

// Single-job launch (the per-block "plain" form): no queue.
__global__ __launch_bounds__(64) void k_rans_decode_one(const u8 *__restrict__ src, u8 *__restrict__ dst,
                                                        fqxv_rans_job jb, i32 *status, u64 *out_len, u8 *ws)
{
    __shared__ DecShared sh;
    decode_job(sh, src, dst, jb, status, out_len, ws);
}

// Persistent decode server (queue layout in rans_serve.h).  Workgroup 0 is
// the dispatcher: it mirrors the host's published count into device memory,
// and raises the device stop flag when the host stops the server or when the
// whole server has been idle -- nothing published for `idle_ticks` of the
// 100 MHz real-time clock and every published job claimed -- then leaves once
// every decode wave has left.  Decode waves claim a published job
// (compare-and-swap on head, only below the mirrored tail, so no claim ever
// waits for a job), decode it exactly as k_rans_decode does, and write the
// result to the host record with a system-scope release (the decoded bytes
// are visible before the tag).  They leave on the device stop flag, so the
// grid leaves as a whole and is relaunched whole by the host when work
// arrives (a grid that shrank one wave at a time could not grow back while
// its remaining waves stayed busy).  The dispatcher leaves when no decode
// wave is left.  Every wait in here ends: the device stop flag is raised at
// the latest `idle_ticks` after the host stops publishing.  Polls are relaxed: an acquire here
// would invalidate caches on every poll and evict the tables of the waves
// that are decoding.  The host's records live in uncached pinned memory and
// are read after the count that publishes them; the one release is each
// job's completion tag, which must make the decoded bytes visible.
__global__ __launch_bounds__(64) void k_rans_serve(const u8 *__restrict__ src, u8 *__restrict__ dst,
                                                   const fqxv_rans_job *ring_jobs, fqxv_serve_res *ring_res,
                                                   u32 ring, const fqxv_serve_ctl *ctl, fqxv_serve_dev *dev,
                                                   u8 *ws, u64 idle_ticks)
{
    __shared__ DecShared sh;
    const u32 l = lane_id();
    // This wave's compute unit: HW_ID bits 15:8 (CU, SH, SE) and the XCC id.
    const u32 cu_key = 0x10000u | (((u32)__builtin_amdgcn_s_getreg((31 << 11) | 4) >> 8) & 0xffu) |
                       (((u32)__builtin_amdgcn_s_getreg((31 << 11) | 20) & 0xfu) << 8);
    if (blockIdx.x == 0) {
        if (l == 0) __hip_atomic_store(&dev->disp_cu, cu_key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        u64 last = ~0ull, t_change = __builtin_amdgcn_s_memrealtime();
        bool stopped = false;
        for (;;) {
            u32 go = 1;
            if (l == 0) {
                const u64 t = __hip_atomic_load(&ctl->tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                const u32 st = __hip_atomic_load(&ctl->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                const u64 now = __builtin_amdgcn_s_memrealtime();
                if (t != last) {
                    __hip_atomic_store(&dev->dtail, t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    last = t;
                    t_change = now;
                }
                if (!stopped &&
                    (st || (now - t_change > idle_ticks &&
                            __hip_atomic_load(&dev->head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= t))) {
                    __hip_atomic_store(&dev->dstop, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    stopped = true;
                }
                if (__hip_atomic_load(&dev->alive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) go = 0;
            }
            if (uni(__shfl(go, 0)) == 0) return;
            __builtin_amdgcn_s_sleep(2);
        }
    }
    // A decode wave on the dispatcher's compute unit leaves at once: the
    // dispatcher's uncached reads of host memory (about 2 us each, back to
    // back) share that unit's vector memory path, and a block decoded beside
    // them measured 2.6 times slower (192 against 73 ms for a 4 MiB order-1
    // block).  The wait for the dispatcher's id is bounded (it is the first
    // workgroup and is resident before any decode wave runs).
    {
        u32 leave = 0;
        if (l == 0) {
            u32 d = 0;
            for (u32 i = 0; i < 100000; i++) {
                d = __hip_atomic_load(&dev->disp_cu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (d) break;
                __builtin_amdgcn_s_sleep(10);
            }
            if (d == cu_key) {
                __hip_atomic_fetch_add(&dev->alive, ~0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                leave = 1;
            }
        }
        if (uni(__shfl(leave, 0))) return;
    }
    u8 *myws = ws + (u64)(blockIdx.x - 1) * WS_PER_WAVE;
    for (;;) {
        u64 q = 0;
        u32 act = 0; // 1: job q, 2: leave
        if (l == 0) {
            // Idle waves back off (about 3 us doubling to about 100 us between
            // polls) so that hundreds of them do not crowd one L2 channel.
            u32 naps = 0;
            for (;;) {
                u64 hd = __hip_atomic_load(&dev->head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const u64 tl = __hip_atomic_load(&dev->dtail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (hd < tl) {
                    if (__hip_atomic_compare_exchange_strong(&dev->head, &hd, hd + 1, __ATOMIC_RELAXED,
                                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                        q = hd;
                        act = 1;
                        break;
                    }
                    continue;
                }
                if (__hip_atomic_load(&dev->dstop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    act = 2;
                    break;
                }
                for (u32 k = 0; k <= naps; k++) __builtin_amdgcn_s_sleep(127);
                naps = naps < 31 ? 2 * naps + 1 : 31;
            }
            if (act == 2) __hip_atomic_fetch_add(&dev->alive, ~0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (uni(__shfl(act, 0)) == 2) return;
        q = (u64)uni(__shfl((u32)q, 0)) | ((u64)uni(__shfl((u32)(q >> 32), 0)) << 32);
        // The job record, written by the host before it raised the tail.
        const u64 *jw = (const u64 *)(ring_jobs + (q % ring));
        u64 w = 0;
        if (l < 6) w = __hip_atomic_load(jw + l, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        fqxv_rans_job jb;
        u64 *jd = (u64 *)&jb;
#pragma unroll
        for (u32 k = 0; k < 6; k++)
            jd[k] = (u64)uni(__shfl((u32)w, k)) | ((u64)uni(__shfl((u32)(w >> 32), k)) << 32);
        u64 n = 0;
        const int err = decode_job_core(sh, src, dst, jb, myws, n);
        if (l == 0) {
            fqxv_serve_res *r = ring_res + (q % ring);
            __hip_atomic_store(&r->status, (i32)err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(&r->out_len, err ? 0ull : n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(&r->tag, q + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        __syncthreads();
    }
}

__global__ void k_reset_counter(u32 *counter) { *counter = 0; }

size_t fqxv_dec_shared_bytes() { return sizeof(DecShared); }

// ------------------------------------------------------------ host launchers
hipError_t fqxv_launch_decode_batch(const u8 *src, u8 *dst, const fqxv_rans_job *jobs, u32 count,
                                    const u32 *order, u32 *counter, i32 *status, u64 *out_len, u8 *ws,
                                    u32 grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, st, counter);
    hipLaunchKernelGGL(k_rans_decode, dim3(grid), dim3(64), 0, st, src, dst, jobs, count, order, counter,
                       status, out_len, ws);
    return hipGetLastError();
}

hipError_t fqxv_launch_decode_one(const u8 *src, u8 *dst, const fqxv_rans_job &jb, i32 *status,
                                  u64 *out_len, u8 *ws, hipStream_t st)
{
    hipLaunchKernelGGL(k_rans_decode_one, dim3(1), dim3(64), 0, st, src, dst, jb, status, out_len, ws);
    return hipGetLastError();
}

hipError_t fqxv_launch_serve(const u8 *src, u8 *dst, const fqxv_rans_job *ring_jobs, fqxv_serve_res *ring_res,
                             u32 ring, const fqxv_serve_ctl *ctl, fqxv_serve_dev *dev, u8 *ws, u64 idle_ticks,
                             u32 decode_waves, hipStream_t st)
{
    hipLaunchKernelGGL(k_rans_serve, dim3(decode_waves + 1), dim3(64), 0, st, src, dst, ring_jobs, ring_res, ring, ctl,
                       dev, ws, idle_ticks);
    return hipGetLastError();
}

// Bytes of device workspace each persistent decode or encode wave owns.
size_t fqxv_ws_per_wave() { return WS_PER_WAVE; }

hipError_t fqxv_decode_occupancy(int *blocks_per_cu)
{
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_rans_decode, 64, 0);
}

// ------------------------------------------ batched decode without allocation
// For callers that cannot allocate (code under graph capture, tight serving
// loops): the batched decode with its work-queue counter and per-wave
// workspace taken from a fixed device array instead of memory the library
// allocates.  The array holds the counter (first 256 bytes) and
// NOALLOC_WAVES per-wave workspaces: 256 + 32 x 2.25 MiB = 72 MiB, held by
// every GPU the library runs on until the process ends.  All calls on one
// GPU share that GPU's copy, so calls that overlap on one GPU race (they
// must be ordered, e.g. issued on one stream); the counter is reset by a
// kernel at the start of every call, nothing else needs to start at zero.
constexpr u32 NOALLOC_WAVES = 32;
constexpr u64 NOALLOC_SCRATCH = 256 + (u64)NOALLOC_WAVES * WS_PER_WAVE;
static __device__ __align__(256) u8 scratch[NOALLOC_SCRATCH];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex noalloc_mu;
u8 *noalloc_ptr[NOALLOC_MAX_DEVICES];

// Device address of this GPU's copy of `scratch`, looked up once per device
// with that device current (the host-side name is only a stand-in).
hipError_t noalloc_scratch(int dev, u8 **out)
{
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(noalloc_mu);
    if (!noalloc_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        noalloc_ptr[dev] = (u8 *)p;
    }
    *out = noalloc_ptr[dev];
    return hipSuccess;
}
} // namespace

extern "C" uint32_t fqxv_rans_noalloc_max_waves(void) { return NOALLOC_WAVES; }

extern "C" int fqxv_rans_decode_batch_noalloc(const uint8_t *d_src, uint8_t *d_dst, const fqxv_rans_job *d_jobs,
                                              uint32_t count, const uint32_t *d_order, int32_t *d_status,
                                              uint64_t *d_out_len, uint32_t waves, void *stream)
{
    if (waves > NOALLOC_WAVES) return FQXV_RANS_ERR_SCRATCH_LIMIT;
    if (count == 0) return FQXV_RANS_OK;
    if (!d_src || !d_dst || !d_jobs || !d_status || !d_out_len) return FQXV_RANS_ERR_INVALID_ARG;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return FQXV_RANS_ERR_DEVICE;
    u8 *base = nullptr;
    if (noalloc_scratch(dev, &base) != hipSuccess) return FQXV_RANS_ERR_DEVICE;
    const u32 w = waves ? waves : NOALLOC_WAVES;
    const u32 grid = count < w ? count : w;
    const hipError_t e = fqxv_launch_decode_batch(d_src, d_dst, d_jobs, count, d_order, (u32 *)base, d_status,
                                                  d_out_len, base + 256, grid, (hipStream_t)stream);
    return e == hipSuccess ? FQXV_RANS_OK : FQXV_RANS_ERR_DEVICE;
}
