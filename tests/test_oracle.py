"""Pin the C oracle (oracle/rans_oracle.c) before trusting it as the checker.

Each check is chosen so that a plausible restatement mistake fails it:
  * streams the reference encoder wrote (re-encode identity);
  * closed forms derived from the rANS definition, not from the oracle code;
  * hand-derived normalize results for the residual paths;
  * the reference's own model tests (model.rs:266-393) and API tests
    (lib.rs:257-606).
"""
import itertools
import struct
import zlib

import numpy as np
import pytest

from conftest import rng_bytes

M = 4096


def hist(data: bytes) -> np.ndarray:
    return np.bincount(np.frombuffer(data, dtype=np.uint8), minlength=256).astype(np.uint32)


# ------------------------------------------------------------ golden vectors
def test_fixture_streams_present(fixture_streams):
    # 48 streams: 46 order-0 (names sub-streams, reorder and long-read side
    # streams) and the only 2 order-1 streams (reorder.fqxv, K=2 and K=5).
    assert len(fixture_streams) == 48
    assert sum(e["order"] == 1 for e, _ in fixture_streams) == 2
    assert {e["contexts"] for e, _ in fixture_streams if e["order"] == 1} == {2, 5}


def test_fixture_reencode_identity(oracle, fixture_streams):
    """decode(s) then encode at s's order reproduces s byte for byte."""
    for e, s in fixture_streams:
        out = oracle.decode(s)
        assert len(out) == e["n"]
        assert zlib.crc32(out) & 0xFFFFFFFF == e["decoded_crc32"]
        assert oracle.encode(out, e["order"]) == s, e


def test_fixture_freq_tables_are_normalized_histograms(oracle, fixture_streams):
    """The stored order-0 tables (written by the reference's normalize) equal
    normalize(histogram(decoded)) -- pins model.rs:213-264 on real data."""
    for e, s in fixture_streams:
        if e["order"] != 0:
            continue
        out = oracle.decode(s)
        stored = np.frombuffer(s[9:9 + 512], dtype="<u2")
        assert np.array_equal(oracle.normalize(hist(out)), stored), e


def test_fixture_order1_tables(oracle, fixture_streams):
    for e, s in fixture_streams:
        if e["order"] != 1:
            continue
        out = np.frombuffer(oracle.decode(s), dtype=np.uint8)
        bm = np.unpackbits(np.frombuffer(s[9:41], dtype=np.uint8), bitorder="little")
        ctxs = np.nonzero(bm)[0]
        prev = np.concatenate([[0], out[:-1]])
        for d, c in enumerate(ctxs):
            cnt = np.bincount(out[prev == c], minlength=256).astype(np.uint32)
            stored = np.frombuffer(s[41 + 512 * d:41 + 512 * (d + 1)], dtype="<u2")
            assert np.array_equal(oracle.normalize(cnt), stored)


# ---------------------------------------------------------------- closed forms
def test_empty_is_nine_bytes(oracle):
    for order in (0, 1):
        assert oracle.encode(b"", order) == bytes([order]) + bytes(8)
        assert oracle.decode(bytes([order]) + bytes(8)) == b""
    # n == 0 decodes to empty whatever the order tag (scalar.rs:147-149)
    assert oracle.decode(bytes([7]) + bytes(8)) == b""


@pytest.mark.parametrize("n", [1, 5, 1000])
def test_single_symbol_order0(oracle, n):
    # freq['A'] = 4096: x_max = 2^32 never renormalises and the map is the
    # identity, so every state stays 65536 and there are no words.
    s = oracle.encode(b"A" * n, 0)
    assert len(s) == 649
    assert s[0] == 0 and struct.unpack("<Q", s[1:9])[0] == n
    f = np.frombuffer(s[9:521], dtype="<u2")
    assert f[65] == 4096 and f.sum() == 4096
    assert struct.unpack("<32I", s[521:649]) == (65536,) * 32


def test_single_symbol_order1(oracle):
    assert len(oracle.encode(b"A" * 1000, 1)) == 9 + 32 + 2 * 512 + 128
    assert len(oracle.encode(b"A", 1)) == 9 + 32 + 512 + 128


def _closed_form_states(data: bytes, order: int, stream: bytes):
    """n <= 32: no renormalisation; lane j's state is C(65536) for its
    symbol, with C(x) = floor(x/f)*4096 + x mod f + cum (the rANS map)."""
    n = len(data)
    if order == 0:
        tables = {None: np.frombuffer(stream[9:521], dtype="<u2").astype(np.int64)}
        st_off = 521
    else:
        bm = np.unpackbits(np.frombuffer(stream[9:41], dtype=np.uint8), bitorder="little")
        ctxs = list(np.nonzero(bm)[0])
        tables = {int(c): np.frombuffer(stream[41 + 512 * d:41 + 512 * (d + 1)], dtype="<u2").astype(np.int64)
                  for d, c in enumerate(ctxs)}
        st_off = 41 + 512 * len(ctxs)
    states = struct.unpack("<32I", stream[st_off:st_off + 128])
    exp = []
    for j in range(32):
        if j >= n:
            exp.append(65536)
            continue
        t = tables[None] if order == 0 else tables[0 if j == 0 else data[j - 1]]
        s = data[j]
        f = int(t[s])
        cum = int(t[:s].sum())
        x = 65536
        exp.append((x // f) * 4096 + x % f + cum)
    return states, tuple(exp), len(stream) - st_off - 128


def test_short_inputs_closed_form(oracle):
    r = np.random.default_rng(5)
    for order in (0, 1):
        for n in list(range(1, 33)):
            for alpha in (2, 7, 256):
                data = r.integers(0, alpha, n).astype(np.uint8).tobytes()
                s = oracle.encode(data, order)
                got, exp, nbytes_words = _closed_form_states(data, order, s)
                assert got == exp, (order, n, alpha)
                assert nbytes_words == 0
                assert oracle.decode(s) == data


def test_brute_force_tiny_inputs(oracle):
    """All inputs of length <= 3 over a 3-letter alphabet."""
    for n in range(0, 4):
        for tup in itertools.product(b"ACG", repeat=n):
            data = bytes(tup)
            for order in (0, 1):
                s = oracle.encode(data, order)
                assert oracle.decode(s) == data
                if n:
                    got, exp, nw = _closed_form_states(data, order, s)
                    assert got == exp and nw == 0


# ------------------------------------------------------------------ normalize
def test_normalize_properties(oracle):
    r = np.random.default_rng(11)
    cases = [np.full(256, 7), np.r_[np.full(1, 10_000_000), np.ones(255)], np.zeros(256)]
    cases[2][0], cases[2][255] = 3, 1
    for _ in range(300):
        k = r.integers(1, 257)
        c = np.zeros(256, dtype=np.int64)
        idx = r.choice(256, k, replace=False)
        c[idx] = r.integers(1, [2, 100, 10**6][r.integers(0, 3)] + 1, k)
        cases.append(c)
    for c in cases:
        f = oracle.normalize(c).astype(np.int64)
        assert f.sum() == 4096
        assert np.all((f >= 1) == (np.asarray(c) > 0))


def test_normalize_hand_derived(oracle):
    # single symbol (model.rs:283-289)
    c = np.zeros(256, np.uint32)
    c[65] = 100
    assert oracle.normalize(c)[65] == 4096
    # highly skewed (model.rs:297-302): first pass gives 4095 + 255 ones =
    # 4350, residual -254 taken from symbol 0 -> 3841.
    c = np.ones(256, np.uint32)
    c[0] = 10_000_000
    f = oracle.normalize(c)
    assert f[0] == 3841 and np.all(f[1:] == 1)
    # tie on the negative residual: symbols 5 and 9 both get 2045 in the
    # first pass, 250 singletons get 1; sum 4340, residual -244 comes off the
    # FIRST maximum (index 5).
    c = np.zeros(256, np.uint32)
    c[5] = c[9] = 100000
    single = [i for i in range(256) if i not in (5, 9)][:250]
    c[single] = 1
    f = oracle.normalize(c)
    assert f[5] == 1801 and f[9] == 2045 and np.all(f[single] == 1)
    # spill across buckets: 100 big symbols at 40 and 156 singletons give
    # 4156; symbol 0 gives up 39 (down to 1), the residual 21 spills to the
    # next strictly greatest bucket (symbol 1: 40 -> 19).
    c = np.ones(256, np.uint32)
    c[:100] = 1_000_000
    f = oracle.normalize(c)
    assert f[0] == 1 and f[1] == 19 and np.all(f[2:100] == 40) and np.all(f[100:] == 1)
    # positive residual: 3 equal symbols -> 1365 each, +1 to the first.
    c = np.zeros(256, np.uint32)
    c[[10, 20, 30]] = 1
    f = oracle.normalize(c)
    assert (f[10], f[20], f[30]) == (1366, 1365, 1365)


# -------------------------------------------------------- EncSym and division
def test_encsym_matches_division(oracle):
    """model.rs:313-344: apply == (v/f << 12) + v%f + cum at the interval
    endpoints and a 512-step sweep, for every frequency."""
    for f in range(1, M + 1):
        cum = min(100, M - f)
        e = oracle.encsym(f, cum)
        x_max = e.x_max
        lo = max(x_max >> 16, 1)
        vs = {lo, x_max - 1}
        if x_max > (1 << 31):
            vs |= {1 << 31, (1 << 31) + 1}
        step = max((x_max - lo) // 64, 1)
        vs |= set(range(lo, x_max, step))
        for v in vs:
            assert oracle.encsym_apply(e, v) == ((v // f) << 12) + v % f + cum, (f, v)


def test_magic_division_exact():
    """The device encoder divides with magic = floor((2^32-1)/f): q0 =
    mulhi(v, magic) is q or q-1, so one correction is exact for every
    renormalised state v in [x_max >> 16, x_max) (model.rs:313-365)."""
    for f in range(1, M + 1):
        magic = (2**32 - 1) // f
        x_max = f << 20
        lo = x_max >> 16
        v = np.unique(np.r_[np.linspace(lo, min(x_max, 2**32) - 1, 4096).astype(np.uint64),
                            np.arange(lo, lo + 64, dtype=np.uint64),
                            np.arange(min(x_max, 2**32) - 64, min(x_max, 2**32), dtype=np.uint64)])
        q0 = (v * np.uint64(magic)) >> np.uint64(32)
        q = v // np.uint64(f)
        assert np.all((q0 == q) | (q0 + np.uint64(1) == q)), f


# ----------------------------------------------------------------- round trips
@pytest.mark.parametrize("n", [1, 2, 31, 32, 33, 63, 64, 65, 100])
def test_roundtrip_short_and_odd_lengths(oracle, n):
    data = bytes((i * 7 + 3) & 0xFF for i in range(n))
    for order in (0, 1):
        assert oracle.decode(oracle.encode(data, order)) == data


def test_roundtrip_misc(oracle):
    datas = [b"Q", b"A" * 1000, bytes(range(256)) * 40, rng_bytes(1, 100003),
             rng_bytes(2, 5000, 6), rng_bytes(3, 70000, 40)]
    for data in datas:
        for order in (0, 1):
            assert oracle.decode(oracle.encode(data, order)) == data


def test_compresses_skewed_data(oracle):
    # lib.rs:335-349
    data = bytes(ord("I") - (i % 8) * (i % 3 == 0) for i in range(50000))
    assert len(oracle.encode(data, 0)) < len(data) // 2


def test_order1_beats_order0_on_correlated(oracle):
    # lib.rs:476-487
    x, data = 0, bytearray()
    for i in range(100000):
        x = (x + i % 4) & 0xFF
        data.append(ord("A") + x % 4)
    assert len(oracle.encode(bytes(data), 1)) < len(oracle.encode(bytes(data), 0))


# ---------------------------------------------------------------- malformed
def test_rejects_malformed_freq_table(oracle):
    # lib.rs:489-502
    src = b"\x00" + struct.pack("<Q", 1) + struct.pack("<H", 4196) + bytes(510)
    rc, _, _ = oracle.decode_ex(src)
    assert rc == oracle.ERR_FREQ_EXCEEDS
    src = b"\x00" + struct.pack("<Q", 1) + struct.pack("<H", 4000) + bytes(510)
    assert oracle.decode_ex(src)[0] == oracle.ERR_FREQ_SUM


def test_rejects_huge_output_length(oracle):
    # lib.rs:504-521: a valid table, zero states, n = u64::MAX
    src = b"\x00" + struct.pack("<Q", 2**64 - 1) + struct.pack("<H", 4096) + bytes(510) + bytes(128)
    assert oracle.decode_ex(src, cap=1 << 20)[0] == oracle.ERR_OUTPUT_ALLOC


def test_decode_bounded_bomb(oracle):
    # lib.rs:294-312
    enc = oracle.encode(b"A" * 64, 0)
    bomb = bytearray(enc)
    bomb[1:9] = struct.pack("<Q", 1 << 31)
    with pytest.raises(oracle.OracleError) as ei:
        oracle.decode_bounded(bytes(bomb), 1 << 20)
    assert ei.value.code == oracle.ERR_BOUND
    assert oracle.decode_bounded(enc, 1 << 20) == b"A" * 64
    with pytest.raises(oracle.OracleError):
        oracle.decode_bounded(enc[:1], 1 << 20)


def test_unknown_order_and_absent_context(oracle):
    assert oracle.decode_ex(b"\x05" + struct.pack("<Q", 3))[0] == oracle.ERR_UNKNOWN_ORDER
    # order-1 stream whose bitmap lacks context 0 but n > 0
    s = bytearray(oracle.encode(b"AB" * 10, 1))
    s[9] &= 0xFE
    # dropping the ctx-0 bit shifts the tables; the first table read is then
    # ctx 'A''s, decoding symbol 0 looks up absent context 0.
    rc = oracle.decode_ex(bytes(s[:41]) + bytes(s[41 + 512:]))[0]
    assert rc == oracle.ERR_ABSENT_CONTEXT


def test_garbage_never_crashes(oracle):
    r = np.random.default_rng(9)
    for _ in range(3000):
        b = bytearray(r.integers(0, 256, r.integers(0, 300)).astype(np.uint8).tobytes())
        for tag in (None, 0, 1):
            if tag is not None:
                if b:
                    b[0] = tag
                else:
                    b.append(tag)
            n = oracle.peek_len(bytes(b))
            if n is not None and n > 1 << 16:
                continue
            oracle.decode_ex(bytes(b))


def test_mutations_never_crash(oracle):
    r = np.random.default_rng(10)
    for it in range(1500):
        data = r.integers(0, 40, r.integers(0, 2000)).astype(np.uint8).tobytes()
        enc = bytearray(oracle.encode(data, it & 1))
        for _ in range(r.integers(0, 12)):
            if enc:
                enc[r.integers(0, len(enc))] = r.integers(0, 256)
        if enc and r.random() < 0.5:
            del enc[r.integers(0, len(enc)):]
        n = oracle.peek_len(bytes(enc))
        if n is not None and n > len(data):
            continue
        oracle.decode_ex(bytes(enc), cap=len(data))


def test_encode_best_follows_each_callers_tie_rule(oracle):
    """encode_best keeps the shorter stream; on a tie the tokenizer and the
    permutation planes keep order-0 and the long-read codec order-1
    (crates/fqxv-tokenizer/src/lib.rs:91-95, container/reorder.rs:597-607,
    crates/fqxv-lroverlap/src/codec.rs:292-303).  The empty input is a tie
    (both streams are the 9-byte header), so the tag byte shows the rule."""
    assert oracle.encode_best(b"", 0) == bytes(9)
    assert oracle.encode_best(b"", 1) == b"\x01" + bytes(8)
    from fqxv_amd import synth
    markov = synth.fill(synth.K_MARKOV, 200000, 3).tobytes()
    dna = synth.fill(synth.K_DNA, 20000, 4).tobytes()
    for data in (markov, dna, b"A" * 100, bytes(range(256)) * 8):
        e0, e1 = oracle.encode(data, 0), oracle.encode(data, 1)
        for tie in (0, 1):
            got = oracle.encode_best(data, tie)
            assert len(got) == min(len(e0), len(e1))
            assert got == (e1 if len(e1) < len(e0) else e0)
    assert oracle.encode_best(markov, 0)[0] == 1   # context helps Markov qualities
    assert oracle.encode_best(dna, 0)[0] == 0      # i.i.d. bases: order-1 only adds tables


def test_permutation_planes_closed_form_and_errors(oracle):
    """reorder.rs:597-607 / 1089-1100: plane k holds byte k of each index.
    Hand example, split/join inverse, and the two container errors."""
    p = np.array([0x04030201, 0x08070605], dtype=np.uint32)
    assert oracle.perm_split(p).tolist() == [0x01, 0x05, 0x02, 0x06, 0x03, 0x07, 0x04, 0x08]
    rng = np.random.default_rng(9)
    for n in (0, 1, 7, 1000):
        q = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
        assert np.array_equal(oracle.perm_join(oracle.perm_split(q), n), q)
    perm = rng.permutation(5000).astype(np.uint32)
    s = oracle.encode_permutation(perm)
    assert np.array_equal(oracle.decode_permutation(s, 5000), perm)
    assert len(s) == min(len(oracle.encode(oracle.perm_split(perm).tobytes(), o)) for o in (0, 1))
    with pytest.raises(oracle.PermError, match="^bad permutation$"):
        oracle.decode_permutation(s[:20], 5000)
    with pytest.raises(oracle.PermError, match="^bad permutation$"):
        oracle.decode_permutation(s, 4999)          # declared 20000 > bound 19996
    with pytest.raises(oracle.PermError, match="^permutation length mismatch$"):
        oracle.decode_permutation(s, 5001)          # 20000 bytes, 20004 expected


def test_crc32c_matches_reference_written_checksums(oracle):
    """crates/fqxv/src/crc.rs restated: the standard check value, and the 14
    CRC-32C values stored in two reference-written archives (header, frame,
    per-stream, whole-file and footer checksums; tests/golden/crc_vectors.json
    from make_crc_vectors.py)."""
    import json
    import os
    assert oracle.crc32c(b"123456789") == 0xE3069283
    assert oracle.crc32c(b"") == 0
    g = os.path.join(os.path.dirname(__file__), "golden")
    vecs = json.load(open(os.path.join(g, "crc_vectors.json")))
    assert len(vecs) == 14
    for e in vecs:
        b = open(os.path.join(g, f"crc_{e['archive']}.bin"), "rb").read()
        assert oracle.crc32c(b[e["off"]:e["off"] + e["len"]]) == e["crc"], e
    rng = np.random.default_rng(5)
    x = rng.integers(0, 256, 5000, dtype=np.uint8).tobytes()
    for cut in (0, 1, 999, 5000):
        a, b = x[:cut], x[cut:]
        assert oracle.crc32c_combine(oracle.crc32c(a), oracle.crc32c(b), len(b)) == oracle.crc32c(x)


def _size_bound_words(stream: bytes, data: bytes):
    """(S, n, W) of a stream: S = sum over symbols of log2(4096 / f) with the
    stream's own normalised tables, W = its renormalisation word count."""
    n = len(data)
    a = np.frombuffer(data, np.uint8).astype(np.int64)
    if stream[0] == 0:
        f = np.frombuffer(stream[9:9 + 512], "<u2").astype(np.float64)
        S = float(np.sum(12.0 - np.log2(f[a])))
        hdr = 9 + 512 + 128
    else:
        bm = np.frombuffer(stream[9:41], np.uint8)
        ctxs = [i for i in range(256) if (bm[i // 8] >> (i % 8)) & 1]
        k = len(ctxs)
        tab = np.frombuffer(stream[41:41 + 512 * k], "<u2").reshape(k, 256).astype(np.float64)
        row = np.full(256, -1, np.int64)
        row[ctxs] = np.arange(k)
        prev = np.concatenate([[0], a[:-1]])
        S = float(np.sum(12.0 - np.log2(tab[row[prev], a])))
        hdr = 9 + 32 + 512 * k + 128
    return S, n, (len(stream) - hdr) // 2


def test_stream_size_bounds_hold(oracle):
    """The device's encode-best skips the walk of an order whose stream is
    surely the longer one, from bounds on its word count W known before the
    walk: (S - 0.093110 n - 512) / 16.093110 < W <= (S + 0.087463 n) / 16,
    S = sum of log2(4096 / f) over the symbols.  Checked here on the oracle's
    streams for skewed, flat, tiny-frequency and single-symbol inputs, both
    orders, where W is exact."""
    from fqxv_amd import synth
    rng = np.random.default_rng(77)
    cases = [synth.fill(synth.K_MARKOV, 60000, 1).tobytes(), synth.fill(synth.K_DNA, 60000, 2).tobytes(),
             synth.fill(synth.K_NAMES, 60000, 3).tobytes(), rng.integers(0, 256, 70000).astype(np.uint8).tobytes(),
             b"A" * 50000, bytes([0] * 40000 + [1]), bytes(rng.choice([0, 1, 2, 255], 65536, p=[.97, .01, .01, .01]).astype(np.uint8)),
             bytes(rng.integers(0, 2, 100000).astype(np.uint8) * 200)]
    for j, d in enumerate(cases):
        for order in (0, 1):
            S, n, W = _size_bound_words(oracle.encode(d, order), d)
            assert (S - 0.093110 * n - 512) / 16.093110 < W <= (S + 0.087463 * n) / 16, (j, order, S, n, W)


def _device_size_bounds(stream: bytes, data: bytes):
    """The device's integer form of those bounds (acc_log_bits / size_bounds
    in fqxv_amd/csrc/rans_encode.hip): each symbol's log2(4096 / f) in
    2^-20 fixed point, rounded outwards by one unit, summed per symbol count,
    then the constants 91751 and 97728 and the divisor 16.0932.  Returns
    (lo, hi) in stream bytes."""
    import math
    ONE = 1 << 20
    n = len(data)
    a = np.frombuffer(data, np.uint8).astype(np.int64)
    if stream[0] == 0:
        tabs = np.frombuffer(stream[9:9 + 512], "<u2").astype(np.int64)[None, :]
        rows = np.zeros(n, np.int64)
        hdr = 9 + 512 + 128
    else:
        bm = np.frombuffer(stream[9:41], np.uint8)
        ctxs = [i for i in range(256) if (bm[i // 8] >> (i % 8)) & 1]
        tabs = np.frombuffer(stream[41:41 + 512 * len(ctxs)], "<u2").reshape(len(ctxs), 256).astype(np.int64)
        row = np.full(256, -1, np.int64)
        row[ctxs] = np.arange(len(ctxs))
        rows = row[np.concatenate([[0], a[:-1]])]
        hdr = 9 + 32 + 512 * len(ctxs) + 128
    counts = np.zeros(tabs.shape, np.int64)
    np.add.at(counts, (rows, a), 1)
    dn = up = 0
    for (r, s), c in np.ndenumerate(counts):
        if c == 0:
            continue
        L = (12.0 - math.log2(float(tabs[r, s]))) * ONE
        lf, lc = math.floor(L), math.ceil(L)
        dn += int(c) * (lf - 1 if lf > 0 else 0)
        up += int(c) * (lc + 1)
    w_max = (up + n * 91751) // (16 * ONE)
    sub = n * 97728 + 512 * ONE
    w_lo = math.floor((dn - sub) / (16.0932 * ONE)) - 1.0 if dn > sub else 0.0
    w_min = int(w_lo) if w_lo > 0 else 0
    return hdr + 2 * w_min, hdr + 2 * w_max


def test_device_integer_size_bounds_bracket_true_sizes(oracle):
    """The integer constants the device uses bracket each order's true stream
    size on the cases above (so a later change to them is caught here), and
    they are at least as wide as the exact real-valued bounds."""
    from fqxv_amd import synth
    rng = np.random.default_rng(78)
    cases = [synth.fill(synth.K_MARKOV, 60000, 1).tobytes(), synth.fill(synth.K_DNA, 60000, 2).tobytes(),
             synth.fill(synth.K_NAMES, 60000, 3).tobytes(), rng.integers(0, 256, 70000).astype(np.uint8).tobytes(),
             b"A" * 50000, bytes([0] * 40000 + [1]),
             bytes(rng.choice([0, 1, 2, 255], 65536, p=[.97, .01, .01, .01]).astype(np.uint8)),
             synth.fill(synth.K_MARKOV, 1 << 20, 4).tobytes()]
    for j, d in enumerate(cases):
        for order in (0, 1):
            s = oracle.encode(d, order)
            lo, hi = _device_size_bounds(s, d)
            assert lo <= len(s) <= hi, (j, order, lo, len(s), hi)
            S, n, W = _size_bound_words(s, d)
            hdr = len(s) - 2 * W
            assert lo <= hdr + 2 * max(0.0, (S - 0.093110 * n - 512) / 16.093110) + 2
            assert hi >= hdr + 2 * int((S + 0.087463 * n) / 16)
