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
//     word[pos + rank] from an LDS ring that is refilled one 512-byte chunk
//     at a time, three chunks ahead of the read position;
//   * order-1: symbol i's context is symbol i-1 (scalar.rs:203), a serial
//     chain through the round.  Every present context (K <= 64, dense id =
//     lane) resolves its candidate symbol for all 32 slots in parallel, then
//     the chain is 32 v_readlane steps through those candidate vectors.
#include "rans_device.h"

namespace {

constexpr u32 RING_CHUNK_WORDS = 256;            // words per ring chunk
constexpr u32 RING_CHUNKS = 4;
constexpr u32 RING_WORDS = RING_CHUNK_WORDS * RING_CHUNKS;
constexpr u32 RING_DW = RING_WORDS / 2;

struct DecShared {
    u32 tab[TOTFREQ];     // order-0 packed slot table
    u32 ring[RING_DW];    // renormalisation word ring
    u32 fc[256];          // f << 16 | cum for the table being built
    u16 dense[256];       // order-1: context byte -> dense id (0xffff absent)
};

// Words region reader: a ring of RING_CHUNKS chunks in LDS.  Chunk c covers
// words [256c, 256c + 256) of the region.  When the read position enters
// chunk c, chunks c, c+1, c+2 are resident; chunk c+3 is in flight in
// registers.
struct WordRing {
    const u8 *base;   // first byte of the words region
    const u8 *end;    // one past the last stream byte
    u64 chunk;        // chunk holding the read position
    u32 pf0, pf1;     // prefetched dwords of chunk `chunk + 3`

    __device__ void fetch(u64 c, u32 &d0, u32 &d1) const
    {
        const u8 *p = base + c * (2 * RING_CHUNK_WORDS) + 4 * lane_id();
        d0 = load4(p, end);
        d1 = load4(p + 256, end);
    }
    __device__ void put(u32 *ring, u64 c, u32 d0, u32 d1) const
    {
        u32 *slot = ring + (u32)(c % RING_CHUNKS) * (RING_CHUNK_WORDS / 2);
        slot[lane_id()] = d0;
        slot[lane_id() + 64] = d1;
    }
    __device__ void reset(u32 *ring, u64 c0)
    {
        chunk = c0;
        for (u64 c = c0; c < c0 + 3; c++) {
            u32 a, b;
            fetch(c, a, b);
            put(ring, c, a, b);
        }
        fetch(c0 + 3, pf0, pf1);
        __syncthreads();
    }
    // Called after the read position advanced; at most one boundary per round.
    __device__ __forceinline__ void advance(u32 *ring, u64 pos)
    {
        if ((pos / RING_CHUNK_WORDS) != chunk) {
            chunk++;
            put(ring, chunk + 2, pf0, pf1);
            fetch(chunk + 3, pf0, pf1);
            __syncthreads();
        }
    }
    __device__ __forceinline__ u32 word(const u32 *ring, u64 idx) const
    {
        return ((const u16 *)ring)[idx % RING_WORDS];
    }
    __device__ u32 word_direct(u64 idx) const
    {
        const u8 *p = base + 2 * idx;
        return (u32)p[0] | ((u32)p[1] << 8);
    }
};

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

// Owner symbol of every slot from sh.fc: lane l fills slots [64l, 64l+64).
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

__device__ __forceinline__ void read_states(const u8 *p, const u8 *end, u32 &x)
{
    x = lane_id() < NSTATES ? load4(p + 4 * lane_id(), end) : 0u;
}

// Exact scalar step for rounds in which some state would need more than one
// word (only a malformed stream can get there: a valid state never reaches
// v == 0, see step_state, scalar.rs:229-238).  Lanes are processed in
// ascending order, each pulling words until v >= RANS_L.
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

// ---------------------------------------------------------------- order-0
__device__ int decode_o0(DecShared &sh, const u8 *s, const u8 *end, u8 *out, u64 n, u64 cap)
{
    const u32 l = lane_id();
    const u64 len = (u64)(end - s);
    if (len < HDR0 + 512) return FQXV_RANS_ERR_TRUNCATED;
    int err = load_freq_table(sh, s + HDR0, end);
    if (err) return err;
    if (len < O0_HDR) return FQXV_RANS_ERR_TRUNCATED;
    u32 x;
    read_states(s + HDR0 + 512, end, x);
    if (n > cap) return FQXV_RANS_ERR_OUTPUT_ALLOC;

    for_each_slot(sh, [&](u32 k, u32 sym, u32 f, u32 cum) {
        sh.tab[k] = ((f - 1) << 20) | ((k - cum) << 8) | sym;
    });

    WordRing wr;
    wr.base = s + O0_HDR;
    wr.end = end;
    const u64 avail = (len - O0_HDR) / 2;
    wr.reset(sh.ring, 0);

    const u64 full = n / NSTATES;
    const u32 rem = (u32)(n % NSTATES);
    const u64 rounds = full + (rem ? 1 : 0);
    u64 pos = 0;
    u8 *o = out + l;
    for (u64 r = 0; r < rounds; r++, o += NSTATES) {
        const bool act = l < NSTATES && (r < full || l < rem);
        const u32 p = sh.tab[x & (TOTFREQ - 1)];
        const u32 f = (p >> 20) + 1;
        u32 v = f * (x >> SCALE_BITS) + ((p >> 8) & 0xfff);
        if (ballot(act && v == 0)) {
            err = slow_round(v, (u32)ballot(act), pos, avail, wr);
            if (err) return err;
            if (act) {
                x = v;
                *o = (u8)p;
            }
            wr.reset(sh.ring, pos / RING_CHUNK_WORDS);
            continue;
        }
        const bool need = act && v < RANS_L;
        const u64 m = ballot(need);
        const u32 k = (u32)__popcll(m);
        if (pos + k > avail) return FQXV_RANS_ERR_TRUNCATED;
        const u32 w = wr.word(sh.ring, pos + rank_below(m));
        if (act) {
            x = need ? ((v << 16) | w) : v;
            *o = (u8)p;
        }
        pos += k;
        wr.advance(sh.ring, pos);
    }
    return 0;
}

// ---------------------------------------------------------------- order-1
// Workspace layout per wave, for dense context ids d:
//   codes[slot * KS + d] (u16) = sym | chain(sym) << 8, slot-major so that
//     the 64 lanes resolving one slot for 64 contexts read one cache line;
//     chain(sym) is sym's dense id when K <= 64, else 0xff (absent) / 0;
//   fcw[d * 256 + sym] (u32) = f << 12 | cum.
__device__ int decode_o1(DecShared &sh, const u8 *s, const u8 *end, u8 *out, u64 n, u64 cap, u8 *ws)
{
    const u32 l = lane_id();
    const u64 len = (u64)(end - s);
    if (len < HDR0 + BITMAP) return FQXV_RANS_ERR_TRUNCATED;
    // Presence bitmap (scalar.rs:181-194): lane l owns contexts 4l..4l+3.
    const u32 bm = s[HDR0 + l / 2];
    const u32 bits = (bm >> (4 * (l & 1))) & 0xf;
    const u32 cnt = __popc(bits);
    const u32 excl = wave_scan_incl(cnt) - cnt;
    const u32 K = uni(wave_sum(cnt));
    {
        u32 d = excl;
        for (int i = 0; i < 4; i++) {
            const bool pr = (bits >> i) & 1;
            sh.dense[4 * l + i] = pr ? (u16)d : (u16)0xffff;
            d += pr;
        }
    }
    __syncthreads();
    u16 *codes = (u16 *)ws;
    u32 *fcw = (u32 *)(ws + WS_CODES);
    const u32 KS = K <= 64 ? 64 : 256;
    const u8 *t = s + HDR0 + BITMAP;
    for (u32 d = 0; d < K; d++, t += 512) {
        if ((u64)(t + 512 - s) > len) return FQXV_RANS_ERR_TRUNCATED;
        int err = load_freq_table(sh, t, end);
        if (err) return err;
        for (int i = 0; i < 4; i++) {
            const u32 e = sh.fc[4 * l + i];
            fcw[d * 256 + 4 * l + i] = ((e >> 16) << 12) | (e & 0xffff);
        }
        u16 *cd = codes + d;
        for_each_slot(sh, [&](u32 k, u32 sym, u32, u32) {
            const u32 ds = sh.dense[sym];
            const u32 chain = K <= 64 ? (ds < 64 ? ds : 0xffu) : 0u;
            cd[(u64)k * KS] = (u16)(sym | (chain << 8));
        });
        __syncthreads();
    }
    const u64 hdr = HDR0 + BITMAP + 512ull * K + 4 * NSTATES;
    if (len < hdr) return FQXV_RANS_ERR_TRUNCATED;
    u32 x;
    read_states(t, end, x);
    if (n > cap) return FQXV_RANS_ERR_OUTPUT_ALLOC;
    // The tables were written through the vector memory path by all lanes;
    // make them visible (and drop any stale L1 lines) before the reads.
    __threadfence();

    WordRing wr;
    wr.base = s + hdr;
    wr.end = end;
    const u64 avail = (len - hdr) / 2;
    wr.reset(sh.ring, 0);

    // Context of symbol 0 is 0 (scalar.rs:203).
    u32 prev = sh.dense[0];
    if (K <= 64 && prev == 0xffff) prev = 0xff;
    const u64 full = n / NSTATES;
    const u32 rem = (u32)(n % NSTATES);
    const u64 rounds = full + (rem ? 1 : 0);
    u64 pos = 0;
    u8 *o = out + l;

    if (K > 64) {
        // Wide alphabet: serial chain through global tables.  Correct but
        // slow; no synthetic or fixture stream reaches it.
        for (u64 r = 0; r < rounds; r++, o += NSTATES) {
            const u32 cntr = r < full ? NSTATES : rem;
            u32 symv = 0, ctxv = 0;
            u32 bad = cntr;
            for (u32 j = 0; j < cntr; j++) {
                if (prev == 0xffff) {
                    bad = j;
                    break;
                }
                const u32 sj = __builtin_amdgcn_readlane(x, j) & (TOTFREQ - 1);
                const u32 sym = uni(codes[(u64)sj * KS + prev]) & 0xff;
                if (l == j) {
                    symv = sym;
                    ctxv = prev;
                }
                prev = sh.dense[sym];
            }
            const bool act = l < bad;
            const u32 fc = act ? fcw[ctxv * 256 + symv] : (1u << 12);
            const u32 f = fc >> 12, cum = fc & 0xfff;
            u32 v = f * (x >> SCALE_BITS) + (x & (TOTFREQ - 1)) - cum;
            if (ballot(act && v == 0)) {
                int err = slow_round(v, (u32)ballot(act), pos, avail, wr);
                if (err) return err;
                wr.reset(sh.ring, pos / RING_CHUNK_WORDS);
            } else {
                const bool need = act && v < RANS_L;
                const u64 m = ballot(need);
                const u32 k = (u32)__popcll(m);
                if (pos + k > avail) return FQXV_RANS_ERR_TRUNCATED;
                const u32 w = wr.word(sh.ring, pos + rank_below(m));
                if (need) v = (v << 16) | w;
                pos += k;
                wr.advance(sh.ring, pos);
            }
            if (act) {
                x = v;
                *o = (u8)symv;
            }
            if (bad < cntr) return FQXV_RANS_ERR_ABSENT_CONTEXT;
        }
        return 0;
    }

    const u16 *mycodes = codes + (l < K ? l : 0);
    for (u64 r = 0; r < rounds; r++, o += NSTATES) {
        const u32 cntr = r < full ? NSTATES : rem;
        const u32 slot = x & (TOTFREQ - 1);
        // Candidate code of every present context for each of the 32 slots.
        u32 cand[NSTATES];
#pragma unroll
        for (u32 j = 0; j < NSTATES; j++) {
            const u32 sj = __builtin_amdgcn_readlane(slot, j);
            cand[j] = mycodes[sj * 64];
        }
        // The chain: symbol j's context is symbol j-1.
        u32 symv = 0, ctxv = 0;
        u32 bad = cntr;
        u32 cur = prev;
#pragma unroll
        for (u32 j = 0; j < NSTATES; j++) {
            if (j < cntr && bad == cntr) {
                if (cur >= 64) {
                    bad = j;
                } else {
                    const u32 code = __builtin_amdgcn_readlane(cand[j], cur);
                    symv = writelane(code & 0xff, j, symv);
                    ctxv = writelane(cur, j, ctxv);
                    cur = code >> 8;
                }
            }
        }
        prev = cur;
        const bool act = l < bad;
        const u32 fc = act ? fcw[ctxv * 256 + symv] : (1u << 12);
        const u32 f = fc >> 12, cum = fc & 0xfff;
        u32 v = f * (x >> SCALE_BITS) + slot - cum;
        if (ballot(act && v == 0)) {
            int err = slow_round(v, (u32)ballot(act), pos, avail, wr);
            if (err) return err;
            wr.reset(sh.ring, pos / RING_CHUNK_WORDS);
        } else {
            const bool need = act && v < RANS_L;
            const u64 m = ballot(need);
            const u32 k = (u32)__popcll(m);
            if (pos + k > avail) return FQXV_RANS_ERR_TRUNCATED;
            const u32 w = wr.word(sh.ring, pos + rank_below(m));
            if (need) v = (v << 16) | w;
            pos += k;
            wr.advance(sh.ring, pos);
        }
        if (act) {
            x = v;
            *o = (u8)symv;
        }
        if (bad < cntr) return FQXV_RANS_ERR_ABSENT_CONTEXT;
    }
    return 0;
}

// One job: header dispatch (scalar.rs:143-155, lib.rs:226-231).
__device__ void decode_job(DecShared &sh, const u8 *src, u8 *dst, const fqxv_rans_job &jb, i32 *status,
                           u64 *out_len, u8 *ws)
{
    const u8 *s = src + jb.src_off;
    const u8 *end = s + jb.src_len;
    int err = 0;
    u64 n = 0;
    if (jb.src_len < HDR0) {
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
            else if (order == 1)
                err = decode_o1(sh, s, end, dst + jb.dst_off, n, jb.dst_cap, ws);
            else
                err = FQXV_RANS_ERR_UNKNOWN_ORDER;
        }
    }
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

hipError_t fqxv_decode_occupancy(int *blocks_per_cu)
{
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_rans_decode, 64, 0);
}
