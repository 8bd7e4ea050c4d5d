// Batched decode of read-name streams (crates/fqxv-tokenizer/src/lib.rs,
// format version 1).
//
// A names stream is a version byte, a varint record count, eight rANS
// sub-streams framed as [varint len][bytes] (token counts, string lengths,
// string bytes, numeric literals, paddings, and the op run-length triple),
// then per token column (64) a varint plane width and that many framed
// delta byte-plane sub-streams (lib.rs:98-169, 515-537).  The framing is
// varints and lengths only -- none of it depends on decoded data -- so:
//   1. the host walks the framing of every stream and records its steps in
//      the reference's read order (sub-stream, column width, or the framing
//      error that ends the walk);
//   2. every rANS sub-stream of every stream is decoded in ONE batched,
//      persistent device launch;
//   3. the host replays each stream's steps with the decoded bytes -- the
//      checks of lib.rs:543-588 in order, op run-length expansion
//      (lib.rs:202-245), byte-plane assembly (lib.rs:137-168) -- and
//      rebuilds the names (lib.rs:590-650), threads across streams.
// Names come out in the layout the container's names digest covers
// (crates/fqxv/src/container/block.rs:164-176): [u64 n] then per name
// [u32 len][bytes], little-endian.
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <thread>
#include <vector>

#include "../../include/fqxv_rans_hip.h"
#include "rans_internal.h"

typedef uint8_t u8;
typedef uint32_t u32;
typedef uint64_t u64;
typedef int32_t i32;
typedef int64_t i64;

namespace {

constexpr u8 FORMAT_VERSION = 1;                       // lib.rs:55
constexpr u8 OP_MATCH = 0, OP_STR = 1, OP_NUM = 2, OP_DELTA = 3; // lib.rs:58-61
constexpr u32 NCOL = 64;                               // MAX_COL + 1 (lib.rs:66)
constexpr u64 MAX_STREAM_EXPANSION = 1ull << 16;       // lib.rs:75
constexpr u64 MIN_STREAM_BOUND = 1ull << 20;           // lib.rs:76
constexpr u32 NFLAT = 8;                               // sub-streams before the planes

// A cursor with the tokenizer's error codes (fqxv-bytes Reader through
// lib.rs:700-712): a varint that runs off the end is "varint too long".
struct Cur {
    const u8 *p;
    u64 n, pos = 0;
    Cur(const u8 *p_, u64 n_) : p(p_), n(n_) {}
    int u8v(u8 &v)
    {
        if (pos >= n) return FQXV_NAMES_ERR_TRUNCATED;
        v = p[pos++];
        return 0;
    }
    int varint(u64 &v)
    {
        v = 0;
        for (u32 shift = 0;; shift += 7) {
            if (pos >= n || shift >= 64) return FQXV_NAMES_ERR_VARINT;
            const u8 b = p[pos++];
            v |= (u64)(b & 0x7f) << shift;
            if (!(b & 0x80)) return 0;
        }
    }
    int take(u64 len, u64 &at)
    {
        if (len > n - pos) return FQXV_NAMES_ERR_TRUNCATED;
        at = pos;
        pos += len;
        return 0;
    }
};

enum Kind : u8 { EV_SUB, EV_WIDTH, EV_ERR };

struct Event {
    Kind kind;
    u8 col;
    int err;       // EV_ERR: the framing error
    u64 width;     // EV_WIDTH
    u64 off, len;  // EV_SUB: sub-stream bytes, offset in the caller's buffer
    u64 declared;  // peeked output length (when len >= 9)
    i64 job;       // device job, or -1 when the host settles it
    int host_st;   // status when job < 0
};

struct Stream {
    std::vector<Event> ev;
    u64 n_records = 0, max_out = 0;
    // results
    int status = 0;
    std::vector<u8> rec;   // [u64 n][u32 len][bytes]...
    u64 n_names = 0;
};

u64 peek_n(const u8 *s) { u64 v; memcpy(&v, s + 1, 8); return v; }

// Phase 1: the framing walk (lib.rs:547-588 read order).
void frame(const u8 *src, u64 base, u64 len, Stream &S)
{
    Cur r(src + base, len);
    auto fail = [&](int e) { S.ev.push_back(Event{EV_ERR, 0, e, 0, 0, 0, 0, -1, 0}); };
    u8 ver;
    int e;
    if ((e = r.u8v(ver))) return fail(e);
    if (ver != FORMAT_VERSION) return fail(FQXV_NAMES_ERR_VERSION);
    if ((e = r.varint(S.n_records))) return fail(e);
    S.max_out = len > UINT64_MAX / MAX_STREAM_EXPANSION ? UINT64_MAX : len * MAX_STREAM_EXPANSION;
    S.max_out = std::max(S.max_out, MIN_STREAM_BOUND);
    auto sub = [&](u8 col) -> int {
        u64 n, at;
        int e2;
        if ((e2 = r.varint(n)) || (e2 = r.take(n, at))) return e2;
        Event ev{EV_SUB, col, 0, 0, base + at, n, 0, -1, 0};
        if (n >= 9) ev.declared = peek_n(src + base + at);
        S.ev.push_back(ev);
        return 0;
    };
    for (u32 k = 0; k < NFLAT; k++)
        if ((e = sub(0))) return fail(e);
    for (u32 c = 0; c < NCOL; c++) {
        u64 w;
        if ((e = r.varint(w))) return fail(e);
        S.ev.push_back(Event{EV_WIDTH, (u8)c, 0, w, 0, 0, 0, -1, 0});
        for (u64 pl = 0; pl < w; pl++)
            if ((e = sub((u8)c))) return fail(e);
    }
}

// Decimal of an i64 (Rust's `to_string`), left-padded with '0' to `width`
// (lib.rs:684-698); natural_digits (lib.rs:253-263).
u32 natural_digits(i64 v)
{
    if (v <= 0) return 1;
    u32 d = 0;
    for (u64 x = (u64)v; x; x /= 10) d++;
    return d;
}

void put_num(std::vector<u8> &out, i64 v, u64 width)
{
    char buf[24];
    int k = 0;
    u64 m = v < 0 ? (u64)0 - (u64)v : (u64)v;
    do {
        buf[k++] = (char)('0' + m % 10);
        m /= 10;
    } while (m);
    if (v < 0) buf[k++] = '-';
    if (width > (u64)k) out.insert(out.end(), width - k, (u8)'0');
    for (int i = k - 1; i >= 0; i--) out.push_back((u8)buf[i]);
}

struct Tok {
    bool is_num;
    u64 off, len; // within the name
    i64 value;
};

// Decoded sub-streams of the device phases, indexed by job.
struct JobOut {
    std::vector<const u8 *> ptr;
    std::vector<i32> st;
    std::vector<u64> len;
};

// A stream's state between its flat sub-streams and its delta planes.
struct Mid {
    const u8 *fp[NFLAT] = {};
    u64 fn[NFLAT] = {};
    std::vector<std::vector<u8>> ops;
    u64 n_delta[NCOL] = {};
    size_t e = 0; // first event of the planes part
};

// Bytes of one decoded sub-stream, or its error, in the reference's order of
// checks (lib.rs:105-110 / fqxv-rans lib.rs:226-231).
int bytes_of(const Event &ev, u64 bound, const JobOut &J, const u8 *&p, u64 &n)
{
    if (ev.len >= 9 && ev.declared > bound) return FQXV_RANS_ERR_BOUND;
    const int st = ev.job >= 0 ? J.st[ev.job] : ev.host_st;
    if (st) return st;
    p = ev.job >= 0 ? J.ptr[ev.job] : nullptr;
    n = ev.job >= 0 ? J.len[ev.job] : 0;
    return 0;
}

// Phase 3a for one stream: the flat sub-streams, the record-count check and
// the op run-length expansion (lib.rs:547-588, 202-245).
int replay_flat(const Stream &S, Mid &M, const JobOut &J)
{
    size_t e = 0;
    const std::vector<Event> &ev = S.ev;
    const u8 **fp = M.fp;
    u64 *fn = M.fn;
    for (u32 k = 0; k < NFLAT; k++, e++) {
        if (e >= ev.size()) return FQXV_NAMES_ERR_TRUNCATED; // unreachable: the walk ends in EV_ERR
        if (ev[e].kind == EV_ERR) return ev[e].err;
        int rc = bytes_of(ev[e], S.max_out, J, fp[k], fn[k]);
        if (rc) return rc;
        if (k == 0 && S.n_records > fn[0]) return FQXV_NAMES_ERR_RECORD_COUNT; // lib.rs:565-567
    }
    M.e = e;
    // Op run-length expansion (lib.rs:202-245).
    std::vector<std::vector<u8>> &ops = M.ops;
    ops.assign(NCOL, {});
    {
        Cur cc(fp[5], fn[5]), cl(fp[7], fn[7]);
        const u8 *syms = fp[6];
        const u64 nsyms = fn[6];
        u64 sym_pos = 0, total = 0;
        for (u32 c = 0; c < NCOL; c++) {
            u64 n_runs;
            int rc = cc.varint(n_runs);
            if (rc) return rc;
            for (u64 i = 0; i < n_runs; i++) {
                if (sym_pos >= nsyms) return FQXV_NAMES_ERR_OP_SYM_UNDERRUN;
                const u8 s = syms[sym_pos++];
                u64 run;
                if ((rc = cl.varint(run))) return rc;
                if (run > S.n_records || ops[c].size() + run > S.n_records) return FQXV_NAMES_ERR_OP_RUN_TOO_LONG;
                if (run > S.max_out - total) return FQXV_NAMES_ERR_OP_BOUND;
                total += run;
                ops[c].insert(ops[c].end(), run, s);
            }
        }
    }
    for (u32 c = 0; c < NCOL; c++) M.n_delta[c] = (u64)std::count(ops[c].begin(), ops[c].end(), OP_DELTA);
    return 0;
}

// Phase 3b for one stream: the delta planes and the rebuild.
int replay_planes(Stream &S, Mid &M, const JobOut &J)
{
    size_t e = M.e;
    const std::vector<Event> &ev = S.ev;
    const u8 *const *fp = M.fp;
    const u64 *fn = M.fn;
    const std::vector<std::vector<u8>> &ops = M.ops;
    const u64 *n_delta = M.n_delta;
    // Delta byte planes (lib.rs:137-168).
    std::vector<std::vector<u64>> deltas(NCOL);
    for (u32 c = 0; c < NCOL; c++) {
        if (e >= ev.size()) return FQXV_NAMES_ERR_TRUNCATED;
        if (ev[e].kind == EV_ERR) return ev[e].err;
        const u64 w = ev[e].width;
        e++;
        if (w == 0) continue;
        deltas[c].assign(n_delta[c], 0);
        for (u64 pl = 0; pl < w; pl++, e++) {
            if (e >= ev.size()) return FQXV_NAMES_ERR_TRUNCATED;
            if (ev[e].kind == EV_ERR) return ev[e].err;
            const u8 *p;
            u64 n;
            int rc = bytes_of(ev[e], n_delta[c], J, p, n);
            if (rc) return rc;
            if (n != n_delta[c]) return FQXV_NAMES_ERR_PLANE_LEN;
            const u32 sh = (u32)((8 * pl) & 63); // release-mode `<<` of lib.rs:163
            for (u64 i = 0; i < n; i++) deltas[c][i] |= (u64)p[i] << sh;
        }
    }
    // Rebuild (lib.rs:590-650).
    Cur c_counts(fp[0], fn[0]), c_sl(fp[1], fn[1]), c_num(fp[3], fn[3]);
    const u8 *str_data = fp[2], *pads = fp[4];
    const u64 str_n = fn[2], pads_n = fn[4];
    u64 str_pos = 0, pad_pos = 0;
    u64 opos[NCOL] = {}, dpos[NCOL] = {};
    std::vector<u8> &out = S.rec;
    out.clear();
    out.reserve(8 + std::min<u64>(S.n_records, 1ull << 20) * 40);
    out.resize(8);
    std::vector<Tok> prev, cur;
    u64 prev_start = 0;
    for (u64 r = 0; r < S.n_records; r++) {
        u64 n_toks;
        int rc = c_counts.varint(n_toks);
        if (rc) return rc;
        const u64 start = out.size() + 4;
        out.resize(start);
        cur.clear();
        for (u64 t = 0; t < n_toks; t++) {
            const u32 col = (u32)std::min<u64>(cur.size(), NCOL - 1);
            if (opos[col] >= ops[col].size()) return FQXV_NAMES_ERR_TRUNCATED;
            const u8 op = ops[col][opos[col]++];
            const u64 here = out.size() - start;
            if (op == OP_MATCH) {
                if (cur.size() >= prev.size()) return FQXV_NAMES_ERR_MATCH;
                Tok p = prev[cur.size()];
                const u64 at = out.size();
                out.resize(at + p.len);
                memmove(out.data() + at, out.data() + prev_start + p.off, p.len);
                p.off = here;
                cur.push_back(p);
            } else if (op == OP_STR) {
                u64 ln;
                if ((rc = c_sl.varint(ln))) return rc;
                if (ln > str_n - str_pos) return FQXV_NAMES_ERR_STR_UNDERRUN;
                out.insert(out.end(), str_data + str_pos, str_data + str_pos + ln);
                str_pos += ln;
                cur.push_back(Tok{false, here, ln, 0});
            } else if (op == OP_NUM) {
                u64 z;
                if ((rc = c_num.varint(z))) return rc;
                const i64 v = (i64)z;
                if (pad_pos >= pads_n) return FQXV_NAMES_ERR_PAD_UNDERRUN;
                const u64 width = natural_digits(v) + pads[pad_pos++];
                put_num(out, v, width);
                cur.push_back(Tok{true, here, out.size() - start - here, v});
            } else if (op == OP_DELTA) {
                if (dpos[col] >= deltas[col].size()) return FQXV_NAMES_ERR_DELTA_UNDERRUN;
                const u64 z = deltas[col][dpos[col]++];
                const i64 d = (i64)((z >> 1) ^ ((u64)0 - (z & 1)));
                if (cur.size() >= prev.size() || !prev[cur.size()].is_num) return FQXV_NAMES_ERR_DELTA_PRIOR;
                const i64 v = (i64)((u64)prev[cur.size()].value + (u64)d);
                if (pad_pos >= pads_n) return FQXV_NAMES_ERR_PAD_UNDERRUN;
                const u64 width = natural_digits(v) + pads[pad_pos++];
                put_num(out, v, width);
                cur.push_back(Tok{true, here, out.size() - start - here, v});
            } else {
                return FQXV_NAMES_ERR_UNKNOWN_OP;
            }
        }
        const u32 nl = (u32)(out.size() - start); // `name.len() as u32`
        memcpy(out.data() + start - 4, &nl, 4);
        prev.swap(cur);
        prev_start = start;
    }
    memcpy(out.data(), &S.n_records, 8);
    S.n_names = S.n_records;
    return 0;
}


// ---- encode (lib.rs:88-135, 175-200, 455-540, 654-681) --------------------

void put_varint(std::vector<u8> &out, u64 v)
{
    while (v >= 0x80) {
        out.push_back((u8)(v | 0x80));
        v >>= 7;
    }
    out.push_back((u8)v);
}

struct TokE {
    bool is_num;
    const u8 *p;
    u64 len;
    i64 value;
};

// Maximal digit / non-digit runs; a digit run of at most 18 bytes is
// numeric (lib.rs:654-681).
void tokenize(const u8 *name, u64 n, std::vector<TokE> &out)
{
    out.clear();
    u64 i = 0;
    while (i < n) {
        const bool dig = name[i] >= '0' && name[i] <= '9';
        const u64 s = i;
        while (i < n && (name[i] >= '0' && name[i] <= '9') == dig) i++;
        const u64 len = i - s;
        const bool is_num = dig && len <= 18;
        i64 v = 0;
        if (is_num)
            for (u64 k = s; k < i; k++) v = v * 10 + (name[k] - '0');
        out.push_back(TokE{is_num, name + s, len, v});
    }
}

// The raw sub-streams of one block: the eight flat streams (counts,
// str_lens, str_data, num_vals, pads, op run counts, op symbols, op run
// lengths) and each column's delta byte planes.
struct Parts {
    u64 n = 0;
    std::vector<u8> flat[NFLAT];
    std::vector<std::vector<u8>> planes[NCOL];
    int status = 0;
    std::vector<u8> out; // the names stream
};

int encode_parts(const u8 *rec, u64 size, Parts &P)
{
    if (size < 8) return FQXV_RANS_ERR_INVALID_ARG;
    memcpy(&P.n, rec, 8);
    std::vector<u8> &counts = P.flat[0], &str_lens = P.flat[1], &str_data = P.flat[2], &num_vals = P.flat[3],
                    &pads = P.flat[4];
    std::vector<u8> op_cols[NCOL];
    std::vector<u64> delta_vals[NCOL];
    std::vector<TokE> prev, cur;
    u64 pos = 8;
    for (u64 r = 0; r < P.n; r++) {
        u32 nl;
        if (size - pos < 4) return FQXV_RANS_ERR_INVALID_ARG;
        memcpy(&nl, rec + pos, 4);
        pos += 4;
        if (size - pos < nl) return FQXV_RANS_ERR_INVALID_ARG;
        tokenize(rec + pos, nl, cur);
        pos += nl;
        for (u64 i = 0; i < cur.size(); i++) {
            const TokE &t = cur[i];
            const u32 col = (u32)std::min<u64>(i, NCOL - 1);
            const TokE *p = i < prev.size() ? &prev[i] : nullptr;
            if (p && p->len == t.len && memcmp(p->p, t.p, t.len) == 0) {
                op_cols[col].push_back(OP_MATCH);
            } else if (t.is_num) {
                const u64 pad = t.len - natural_digits(t.value);
                if (p && p->is_num) {
                    op_cols[col].push_back(OP_DELTA);
                    const i64 d = t.value - p->value;
                    delta_vals[col].push_back(((u64)d << 1) ^ (u64)(d >> 63));
                } else {
                    op_cols[col].push_back(OP_NUM);
                    put_varint(num_vals, (u64)t.value);
                }
                pads.push_back((u8)pad);
            } else {
                op_cols[col].push_back(OP_STR);
                put_varint(str_lens, t.len);
                str_data.insert(str_data.end(), t.p, t.p + t.len);
            }
        }
        put_varint(counts, cur.size());
        prev.swap(cur);
    }
    if (pos != size) return FQXV_RANS_ERR_INVALID_ARG;
    // Op run lengths (lib.rs:175-200).
    for (u32 c = 0; c < NCOL; c++) {
        const std::vector<u8> &col = op_cols[c];
        u64 runs = 0;
        for (u64 i = 0; i < col.size();) {
            u64 run = 1;
            while (i + run < col.size() && col[i + run] == col[i]) run++;
            P.flat[6].push_back(col[i]);
            put_varint(P.flat[7], run);
            runs++;
            i += run;
        }
        put_varint(P.flat[5], runs);
    }
    // Delta byte planes (lib.rs:118-135).
    for (u32 c = 0; c < NCOL; c++) {
        const std::vector<u64> &v = delta_vals[c];
        if (v.empty()) continue;
        const u64 maxv = *std::max_element(v.begin(), v.end());
        u32 bits = 0;
        for (u64 x = maxv; x; x >>= 1) bits++;
        const u32 width = std::max(1u, (bits + 7) / 8);
        P.planes[c].resize(width);
        for (u32 k = 0; k < width; k++) {
            std::vector<u8> &pl = P.planes[c][k];
            pl.resize(v.size());
            for (u64 i = 0; i < v.size(); i++) pl[i] = (u8)(v[i] >> (8 * k));
        }
    }
    return 0;
}

// The status a sub-stream gets before its output would be allocated, in the
// decoder's order of checks (scalar.rs:143-194 with model.rs:160-174; the
// bound was checked already): 0 when the whole header through the 32 states
// is present and valid (the stream needs a device job), -1 for n == 0 (an
// empty output whatever the tag, scalar.rs:147-149), or the error.  Streams
// that fail here never allocate in the reference, so they get no device
// output region either.
int header_status(const u8 *s, u64 len)
{
    auto table = [&](u64 at) -> int {
        u32 total = 0;
        for (u32 i = 0; i < 256; i++) total += (u32)s[at + 2 * i] | ((u32)s[at + 2 * i + 1] << 8);
        return total > 4096 ? FQXV_RANS_ERR_FREQ_EXCEEDS : total != 4096 ? FQXV_RANS_ERR_FREQ_SUM : 0;
    };
    if (len < 9) return FQXV_RANS_ERR_TRUNCATED;
    if (peek_n(s) == 0) return -1;
    const u8 order = s[0];
    int e;
    if (order == 0) {
        if (len < 9 + 512) return FQXV_RANS_ERR_TRUNCATED;
        if ((e = table(9))) return e;
        return len < 9 + 512 + 128 ? FQXV_RANS_ERR_TRUNCATED : 0;
    }
    if (order != 1) return FQXV_RANS_ERR_UNKNOWN_ORDER;
    if (len < 9 + 32) return FQXV_RANS_ERR_TRUNCATED;
    u32 k = 0;
    for (u32 i = 0; i < 32; i++) k += (u32)__builtin_popcount(s[9 + i]);
    for (u32 c = 0; c < k; c++) {
        const u64 at = 41 + 512ull * c;
        if (len < at + 512) return FQXV_RANS_ERR_TRUNCATED;
        if ((e = table(at))) return e;
    }
    return len < 41 + 512ull * k + 128 ? FQXV_RANS_ERR_TRUNCATED : 0;
}

// Output bytes of one device launch of the names decoder.  A job declaring
// more runs in a launch of its own, sized to it; when that region cannot be
// allocated the job fails like the reference's fallible allocation
// (model.rs:199-205), and the rest of the batch is unaffected.
constexpr u64 NAMES_CHUNK = 1ull << 30;

template <class F> void parallel_for(u32 count, F f)
{
    const u32 hw = std::max(1u, std::thread::hardware_concurrency());
    const u32 nthr = std::min<u32>({count, hw, 16u});
    std::atomic<u32> next{0};
    auto work = [&]() {
        for (u32 i; (i = next.fetch_add(1)) < count;) f(i);
    };
    if (nthr <= 1) {
        work();
        return;
    }
    std::vector<std::thread> th;
    for (u32 t = 0; t < nthr; t++) th.emplace_back(work);
    for (auto &t : th) t.join();
}

// Runs jobs[first..) on the device, in launches of at most NAMES_CHUNK
// output bytes (a larger job alone, in a region of its own size; when that
// region cannot be allocated, the job fails like the reference's fallible
// allocation, model.rs:199-205).  Only the successful jobs' bytes come back
// to the host (packed), so a corrupt sub-stream declaring a large output
// holds no host memory.  Fills J for those jobs; J keeps the buffers.
int run_phase(fqxv_rans_ctx *ctx, const u8 *src, const std::vector<fqxv_rans_job> &jobs, size_t first, JobOut &J,
              std::vector<std::unique_ptr<u8, void (*)(void *)>> &keep)
{
    J.ptr.resize(jobs.size(), nullptr);
    J.st.resize(jobs.size(), 0);
    J.len.resize(jobs.size(), 0);
    for (size_t a = first; a < jobs.size();) {
        size_t b = a;
        u64 total = 0, lo = UINT64_MAX, hi = 0;
        while (b < jobs.size()) {
            const u64 cap = (jobs[b].dst_cap + 255) & ~255ull;
            if (b > a && total + cap > NAMES_CHUNK) break;
            total += cap;
            lo = std::min(lo, jobs[b].src_off);
            hi = std::max(hi, jobs[b].src_off + jobs[b].src_len);
            b++;
        }
        std::vector<fqxv_rans_job> cj(jobs.begin() + a, jobs.begin() + b);
        u64 at = 0;
        for (fqxv_rans_job &j : cj) {
            j.src_off -= lo;
            j.dst_off = at;
            at += (j.dst_cap + 255) & ~255ull;
        }
        std::vector<u64> poff(cj.size());
        u8 *packed = nullptr;
        const int rc = fqxv_jobs_host_packed(ctx, 0, src + lo, hi - lo, cj.data(), (u32)cj.size(), total, &packed,
                                             poff.data(), J.st.data() + a, J.len.data() + a);
        if (rc == FQXV_RANS_ERR_DEVICE && b - a == 1 && cj[0].dst_cap > NAMES_CHUNK) {
            J.st[a] = FQXV_RANS_ERR_OUTPUT_ALLOC;
            J.len[a] = 0;
            a = b;
            continue;
        }
        if (rc) return rc;
        keep.emplace_back(packed, free);
        for (size_t j = a; j < b; j++)
            if (J.st[j] == 0) J.ptr[j] = packed + poff[j - a];
        a = b;
    }
    return 0;
}

} // namespace

struct fqxv_names_batch {
    std::vector<Stream> s;   // decode results
    std::vector<Parts> e;    // encode results
};

extern "C" {

static int names_decode_batch(fqxv_rans_ctx *ctx, const uint8_t *src, const uint64_t *off, const uint64_t *len,
                              uint32_t count, fqxv_names_batch **out)
{
    if (!out) return FQXV_RANS_ERR_INVALID_ARG;
    *out = nullptr;
    if (!ctx || (count && (!src || !off || !len))) return FQXV_RANS_ERR_INVALID_ARG;
    std::unique_ptr<fqxv_names_batch> B(new fqxv_names_batch());
    B->s.resize(count);
    // Phase 1.
    for (u32 i = 0; i < count; i++) frame(src, off[i], len[i], B->s[i]);
    // Phase 2a: the flat sub-streams of every stream, each bounded by its
    // stream's max_out.  A sub-stream whose header is incomplete or invalid
    // is settled here with the decoder's own error (it never allocates in the
    // reference either), so only complete headers get a device job.
    std::vector<fqxv_rans_job> jobs;
    JobOut J;
    auto add_job = [&](Event &ev, u64 bound) {
        if (ev.len < 9) {
            ev.host_st = FQXV_RANS_ERR_TRUNCATED; // scalar.rs: tag + u64 n need 9 bytes
            return;
        }
        if (ev.declared > bound) {
            ev.host_st = FQXV_RANS_ERR_BOUND;
            return;
        }
        const int h = header_status(src + ev.off, ev.len);
        if (h) {
            ev.host_st = h < 0 ? 0 : h; // n == 0: empty output, no job
            return;
        }
        ev.job = (i64)jobs.size();
        jobs.push_back(fqxv_rans_job{ev.off, ev.len, 0, ev.declared, UINT64_MAX, 0, 0});
    };
    for (Stream &S : B->s) {
        u32 k = 0;
        for (Event &ev : S.ev) {
            if (ev.kind != EV_SUB || k == NFLAT) break;
            add_job(ev, S.max_out);
            k++;
        }
    }
    std::vector<std::unique_ptr<u8, void (*)(void *)>> keep;
    int rc = run_phase(ctx, src, jobs, 0, J, keep);
    if (rc) return rc;
    // Phase 3a, threads across streams; then the planes of the streams that
    // got through, each bounded by its column's delta count (lib.rs:137-168),
    // so no plane region is larger than the output it must match.
    std::vector<Mid> mid(count);
    parallel_for(count, [&](u32 i) {
        Stream &S = B->s[i];
        try {
            S.status = replay_flat(S, mid[i], J);
        } catch (const std::bad_alloc &) {
            S.status = FQXV_NAMES_ERR_ALLOC; // the reference's fallible reserves (lib.rs:244-245)
        }
    });
    const size_t first_plane = jobs.size();
    for (u32 i = 0; i < count; i++) {
        Stream &S = B->s[i];
        if (S.status) continue;
        const Mid &M = mid[i];
        u32 c = 0;
        u64 left = 0; // planes left in column c - 1
        for (size_t e = M.e; e < S.ev.size(); e++) {
            Event &ev = S.ev[e];
            if (ev.kind == EV_ERR) break;
            if (ev.kind == EV_WIDTH) {
                c++;
                left = ev.width;
                continue;
            }
            if (c == 0 || left == 0) break; // unreachable: planes follow their width
            add_job(ev, M.n_delta[c - 1]);
            left--;
        }
    }
    if ((rc = run_phase(ctx, src, jobs, first_plane, J, keep))) return rc;
    // Phase 3b.
    parallel_for(count, [&](u32 i) {
        Stream &S = B->s[i];
        if (!S.status) {
            try {
                S.status = replay_planes(S, mid[i], J);
            } catch (const std::bad_alloc &) {
                S.status = FQXV_NAMES_ERR_ALLOC; // the reference's fallible reserves (lib.rs:150-152, 244-245)
            }
        }
        if (S.status) {
            S.rec.clear();
            S.n_names = 0;
        }
        mid[i] = Mid();
        S.ev.clear();
        S.ev.shrink_to_fit();
    });
    *out = B.release();
    return FQXV_RANS_OK;
}

int fqxv_names_decode_batch(fqxv_rans_ctx *ctx, const uint8_t *src, const uint64_t *off, const uint64_t *len,
                            uint32_t count, fqxv_names_batch **out)
{
    try {
        return names_decode_batch(ctx, src, off, len, count, out);
    } catch (const std::bad_alloc &) {
        return FQXV_RANS_ERR_DEVICE;
    }
}

static int names_encode_batch(fqxv_rans_ctx *ctx, const uint8_t *src, const uint64_t *off, const uint64_t *len,
                              uint32_t count, fqxv_names_batch **out)
{
    std::unique_ptr<fqxv_names_batch> B(new fqxv_names_batch());
    B->e.resize(count);
    // Tokenize every block (host, threads across blocks).
    parallel_for(count, [&](u32 i) {
        try {
            B->e[i].status = encode_parts(src + off[i], len[i], B->e[i]);
        } catch (const std::bad_alloc &) {
            B->e[i].status = FQXV_RANS_ERR_DEVICE;
        }
    });
    // One encode-best job per sub-stream of every block (order-0 on a tie,
    // lib.rs:88-95).
    std::vector<fqxv_rans_job> jobs;
    u64 src_total = 0, dst_total = 0;
    for (Parts &P : B->e) {
        if (P.status) continue;
        auto add = [&](const std::vector<u8> &st) {
            const u64 n = st.size();
            jobs.push_back(fqxv_rans_job{src_total, n, dst_total, fqxv_rans_encode_best_bound(n), 0, 0, 0});
            src_total += (n + 15) & ~15ull;
            dst_total += (fqxv_rans_encode_best_bound(n) + 255) & ~255ull;
        };
        for (u32 k = 0; k < NFLAT; k++) add(P.flat[k]);
        for (u32 c = 0; c < NCOL; c++)
            for (const auto &pl : P.planes[c]) add(pl);
    }
    // Sub-streams back to back (16-byte aligned, the padding zeroed); the
    // encoded streams come back packed, not in their bound-sized regions.
    std::unique_ptr<u8[]> raw(new u8[std::max<u64>(src_total, 1)]);
    {
        size_t j = 0;
        for (Parts &P : B->e) {
            if (P.status) continue;
            auto put = [&](const std::vector<u8> &st) {
                u8 *d = raw.get() + jobs[j].src_off;
                if (!st.empty()) memcpy(d, st.data(), st.size());
                memset(d + st.size(), 0, ((st.size() + 15) & ~15ull) - st.size());
                j++;
            };
            for (u32 k = 0; k < NFLAT; k++) put(P.flat[k]);
            for (u32 c = 0; c < NCOL; c++)
                for (const auto &pl : P.planes[c]) put(pl);
        }
    }
    std::vector<i32> st(jobs.size());
    std::vector<u64> ol(jobs.size()), poff(jobs.size());
    u8 *packed = nullptr;
    const int rc = fqxv_jobs_host_packed(ctx, 2, raw.get(), src_total, jobs.data(), (u32)jobs.size(), dst_total,
                                                &packed, poff.data(), st.data(), ol.data());
    if (rc) return rc;
    std::unique_ptr<u8, decltype(&free)> enc(packed, &free);
    // Frame every block (lib.rs:515-537).
    std::vector<size_t> first(count);
    {
        size_t j = 0;
        for (u32 i = 0; i < count; i++) {
            first[i] = j;
            Parts &P = B->e[i];
            if (P.status) continue;
            j += NFLAT;
            for (u32 c = 0; c < NCOL; c++) j += P.planes[c].size();
        }
    }
    parallel_for(count, [&](u32 i) {
        Parts &P = B->e[i];
        if (P.status) return;
        size_t j = first[i];
        std::vector<u8> &o = P.out;
        o.push_back(FORMAT_VERSION);
        put_varint(o, P.n);
        auto emit = [&]() {
            if (st[j]) {
                P.status = st[j];
                return;
            }
            put_varint(o, ol[j]);
            o.insert(o.end(), enc.get() + poff[j], enc.get() + poff[j] + ol[j]);
            j++;
        };
        for (u32 k = 0; k < NFLAT && !P.status; k++) emit();
        for (u32 c = 0; c < NCOL && !P.status; c++) {
            put_varint(o, P.planes[c].size());
            for (size_t k = 0; k < P.planes[c].size() && !P.status; k++) emit();
        }
        for (auto &f : P.flat) std::vector<u8>().swap(f);
        for (auto &c : P.planes) std::vector<std::vector<u8>>().swap(c);
        if (P.status) o.clear();
    });
    *out = B.release();
    return FQXV_RANS_OK;
}

int fqxv_names_encode_batch(fqxv_rans_ctx *ctx, const uint8_t *src, const uint64_t *off, const uint64_t *len,
                            uint32_t count, fqxv_names_batch **out)
{
    if (!out) return FQXV_RANS_ERR_INVALID_ARG;
    *out = nullptr;
    if (!ctx || (count && (!src || !off || !len))) return FQXV_RANS_ERR_INVALID_ARG;
    try {
        return names_encode_batch(ctx, src, off, len, count, out);
    } catch (const std::bad_alloc &) {
        return FQXV_RANS_ERR_DEVICE;
    }
}

int fqxv_names_output(const fqxv_names_batch *b, uint32_t i, const uint8_t **data, uint64_t *size)
{
    if (!b || i >= b->e.size() || !data || !size) return FQXV_RANS_ERR_INVALID_ARG;
    const Parts &P = b->e[i];
    *data = P.out.data();
    *size = P.out.size();
    return P.status;
}

int32_t fqxv_names_status(const fqxv_names_batch *b, uint32_t i)
{
    if (!b) return FQXV_RANS_ERR_INVALID_ARG;
    if (i < b->s.size()) return b->s[i].status;
    if (i < b->e.size()) return b->e[i].status;
    return FQXV_RANS_ERR_INVALID_ARG;
}

int fqxv_names_records(const fqxv_names_batch *b, uint32_t i, const uint8_t **data, uint64_t *size,
                       uint64_t *n_names)
{
    if (!b || i >= b->s.size() || !data || !size || !n_names) return FQXV_RANS_ERR_INVALID_ARG;
    const Stream &S = b->s[i];
    *data = S.rec.data();
    *size = S.rec.size();
    *n_names = S.n_names;
    return S.status;
}

int fqxv_names_digests(fqxv_rans_ctx *ctx, const fqxv_names_batch *b, uint64_t *out)
{
    if (!ctx || !b || (!out && !b->s.empty())) return FQXV_RANS_ERR_INVALID_ARG;
    const u32 n = (u32)b->s.size();
    std::vector<const u8 *> ptrs(n);
    std::vector<u64> lens(n);
    for (u32 i = 0; i < n; i++) {
        ptrs[i] = b->s[i].rec.data();
        lens[i] = b->s[i].status ? 0 : b->s[i].rec.size();
    }
    const int rc = fqxv_xxh3_host_ranges(ctx, ptrs.data(), lens.data(), n, out);
    if (rc) return rc;
    for (u32 i = 0; i < n; i++)
        if (b->s[i].status) out[i] = 0;
    return FQXV_RANS_OK;
}

void fqxv_names_batch_destroy(fqxv_names_batch *b) { delete b; }

} // extern "C"
