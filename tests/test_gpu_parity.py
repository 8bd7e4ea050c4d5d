"""GPU parity: the gfx950 library against the oracle, byte for byte.

Every case runs the HIP kernels through the C ABI (libfqxv_rans_hip.so) and
compares with the oracle on the same inputs: encoded streams must be
identical, decoded bytes must be identical, and malformed streams must fail
with the same error code.
"""
import struct

import numpy as np
import pytest

from conftest import rng_bytes

pytestmark = pytest.mark.gpu


def _cases():
    cases = [b"", b"Q", b"AB", b"A" * 1000, bytes(range(256)) * 40]
    cases += [bytes((i * 7 + 3) & 0xFF for i in range(n)) for n in (1, 2, 31, 32, 33, 63, 64, 65, 100)]
    # lib.rs:439-447: dominant symbol (f > 2048), single symbol (f == 4096),
    # two-symbol skew.
    cases.append(b"A" * 60000 + bytes(ord("B") + i % 20 for i in range(4000)))
    cases.append(b"Q" * 40000)
    cases.append(bytes(ord("!") + (i % 3 == 0) * 7 for i in range(80000)))
    cases.append(bytes((i * 131 % 251) for i in range(100003)))
    for seed, (n, a) in enumerate([(5000, 40), (2000, 6), (100003, 256), (70001, 41), (4096, 2),
                                   (1 << 16, 64), (1 << 16, 65), (30000, 200)]):
        cases.append(rng_bytes(100 + seed, n, a))
    return cases


def test_host_api_matches_oracle(coder, oracle):
    for data in _cases():
        for order in (0, 1):
            ref = oracle.encode(data, order)
            got = coder.encode(data, order)
            assert got == ref, (len(data), order)
            assert coder.decode(ref) == data, (len(data), order)


def test_fixture_streams_decode_and_reencode(coder, oracle, fixture_streams):
    for e, s in fixture_streams:
        out = coder.decode(s)
        assert out == oracle.decode(s)
        assert coder.encode(out, e["order"]) == s, e


def test_normalize_residual_paths_on_device(coder, oracle):
    """Inputs whose histograms take normalize's negative-residual, tie and
    spill paths (hand-derived in test_oracle)."""
    def with_counts(c):
        return bytes(np.repeat(np.arange(256, dtype=np.uint8), c))
    c1 = np.ones(256, np.int64)
    c1[0] = 1_000_000
    c2 = np.zeros(256, np.int64)
    c2[5] = c2[9] = 100000
    c2[[i for i in range(256) if i not in (5, 9)][:250]] = 1
    c3 = np.ones(256, np.int64)
    c3[:100] = 10000
    for c in (c1, c2, c3):
        data = with_counts(c)
        for order in (0, 1):
            assert coder.encode(data, order) == oracle.encode(data, order)


def test_decode_bounded_and_errors(coder, oracle):
    import fqxv_amd
    enc = oracle.encode(b"A" * 64, 0)
    bomb = bytearray(enc)
    bomb[1:9] = struct.pack("<Q", 1 << 31)
    with pytest.raises(fqxv_amd.Malformed) as ei:
        coder.decode_bounded(bytes(bomb), 1 << 20)
    assert str(ei.value) == "rANS output length exceeds caller bound"
    assert coder.decode_bounded(enc, 1 << 20) == b"A" * 64
    with pytest.raises(fqxv_amd.Malformed):
        coder.decode_bounded(enc[:1], 1 << 20)
    huge = b"\x00" + struct.pack("<Q", 2**64 - 1) + struct.pack("<H", 4096) + bytes(510) + bytes(128)
    with pytest.raises(fqxv_amd.Malformed) as ei:
        coder.decode(huge)
    assert str(ei.value) == "rANS output length too large to allocate"


def _err_code(fn):
    import fqxv_amd
    try:
        fn()
        return 0
    except fqxv_amd.Malformed as e:
        return e.code


def test_malformed_streams_same_error_as_oracle(coder, oracle):
    r = np.random.default_rng(77)
    checked = 0
    for it in range(400):
        data = r.integers(0, 40, r.integers(0, 3000)).astype(np.uint8).tobytes()
        enc = bytearray(oracle.encode(data, it & 1))
        for _ in range(r.integers(0, 8)):
            if enc:
                enc[r.integers(0, len(enc))] = r.integers(0, 256)
        if enc and r.random() < 0.5:
            del enc[r.integers(0, len(enc)):]
        b = bytes(enc)
        n = oracle.peek_len(b)
        if n is not None and n > max(len(data), 1) * 4:
            continue
        rc, out, _ = oracle.decode_ex(b)
        got = _err_code(lambda: coder.decode(b))
        assert got == rc, (it, rc, got)
        if rc == 0:
            assert coder.decode(b) == out
        checked += 1
    assert checked > 200


def test_batched_equals_single_and_oracle(coder, oracle):
    import torch
    blocks = [c for c in _cases() if len(c) < 200000] * 2
    orders = [i & 1 for i in range(len(blocks))]
    sizes = np.array([len(b) for b in blocks], dtype=np.uint64)
    offs = np.zeros(len(blocks), dtype=np.uint64)
    offs[1:] = np.cumsum(sizes)[:-1]
    arena = torch.from_numpy(np.frombuffer(b"".join(blocks) + bytes(64), dtype=np.uint8).copy()).cuda()
    r = coder.encode_batch(arena, offs, sizes, orders)
    torch.cuda.synchronize()
    st = r.status.cpu().numpy()
    ol = r.out_len.cpu().numpy()
    host = r.dst.cpu().numpy()
    assert np.all(st == 0)
    for j, b in enumerate(blocks):
        got = host[int(r.dst_off[j]):int(r.dst_off[j]) + int(ol[j])].tobytes()
        assert got == oracle.encode(b, orders[j]), j
    d = coder.decode_batch(r.dst, r.dst_off, ol, sizes)
    torch.cuda.synchronize()
    assert np.all(d.status.cpu().numpy() == 0)
    out = d.dst.cpu().numpy()
    for j, b in enumerate(blocks):
        assert out[int(d.dst_off[j]):int(d.dst_off[j]) + len(b)].tobytes() == b, j


def _corpus_roundtrip(coder, oracle, corpus, check_blocks):
    """Batched encode on device, byte-identity vs oracle on `check_blocks`,
    full device round trip."""
    import torch
    arena = torch.from_numpy(corpus.data).cuda()
    r = coder.encode_batch(arena, corpus.offs, corpus.sizes, corpus.orders)
    torch.cuda.synchronize()
    st = r.status.cpu().numpy()
    assert np.all(st == 0)
    ol = r.out_len.cpu().numpy()
    host = None
    for j in check_blocks:
        if host is None:
            host = r.dst.cpu().numpy()
        got = host[int(r.dst_off[j]):int(r.dst_off[j]) + int(ol[j])].tobytes()
        assert got == oracle.encode(corpus.block(j), int(corpus.orders[j])), j
    d = coder.decode_batch(r.dst, r.dst_off, ol, corpus.sizes)
    torch.cuda.synchronize()
    assert np.all(d.status.cpu().numpy() == 0)
    assert np.array_equal(d.out_len.cpu().numpy(), corpus.sizes.astype(np.int64))
    n = corpus.total
    assert torch.equal(d.dst[:n], arena[:n])
    return r, ol


def test_workload_binned_1mib(coder, oracle):
    from fqxv_amd import synth
    c = synth.workload("binned_1mib")
    _corpus_roundtrip(coder, oracle, c, [0])


def test_workload_normal_64x1mib(coder, oracle):
    from fqxv_amd import synth
    c = synth.workload("normal_64x1mib")
    _corpus_roundtrip(coder, oracle, c, range(0, 64, 4))


def test_workload_markov_64x4mib(coder, oracle):
    from fqxv_amd import synth
    c = synth.workload("markov_64x4mib")
    _corpus_roundtrip(coder, oracle, c, [0, 1, 63])


def test_workload_mixed_prefix(coder, oracle):
    from fqxv_amd import synth
    c = synth.workload("mixed_10k", blocks=400)
    _corpus_roundtrip(coder, oracle, c, range(0, 400, 7))


def test_batched_decode_writes_only_its_region(coder, oracle):
    """Regression: a block whose last, partial round fills the 32nd row of
    the output stage must not write past its region.  Regions are packed
    with 1-byte sentinel gaps at odd offsets; every gap must survive."""
    import torch
    rng = np.random.default_rng(12)
    lens = [993, 1000, 1023, 1024, 1025, 2017, 2048, 31, 33, 5000, 1000, 999]
    blocks = [rng.integers(0, 30, n).astype(np.uint8).tobytes() for n in lens] * 4
    orders = [(i // len(lens)) & 1 for i in range(len(blocks))]
    sizes = np.array([len(b) for b in blocks], dtype=np.uint64)
    offs = np.zeros(len(blocks), dtype=np.uint64)
    offs[1:] = np.cumsum(sizes)[:-1]
    arena = torch.from_numpy(np.frombuffer(b"".join(blocks) + bytes(64), dtype=np.uint8).copy()).cuda()
    r = coder.encode_batch(arena, offs, sizes, orders)
    torch.cuda.synchronize()
    assert np.all(r.status.cpu().numpy() == 0)
    ol = r.out_len.cpu().numpy()
    doff = np.zeros(len(blocks), dtype=np.uint64)
    acc = 3
    for j, n in enumerate(sizes):
        doff[j] = acc
        acc += int(n) + 1
    dst = torch.full((acc + 64,), 0xEE, dtype=torch.uint8, device="cuda")
    d = coder.decode_batch(r.dst, r.dst_off, ol, sizes, dst=dst, dst_off=doff)
    torch.cuda.synchronize()
    assert np.all(d.status.cpu().numpy() == 0)
    out = dst.cpu().numpy()
    for j, b in enumerate(blocks):
        o = int(doff[j])
        assert out[o:o + len(b)].tobytes() == b, j
        assert out[o + len(b)] == 0xEE, f"block {j} wrote past its region"
    assert np.all(out[:3] == 0xEE)


def test_encode_best_batch_matches_oracle(coder, oracle):
    """Both orders per block on the device, the shorter kept, each block's
    tie rule honoured (empty blocks are ties); decoded back on the device."""
    import torch
    from fqxv_amd import synth
    blocks = [b"", synth.fill(synth.K_MARKOV, 30000, 3).tobytes(), synth.fill(synth.K_DNA, 30000, 4).tobytes(),
              b"", b"Q" * 5000, synth.fill(synth.K_NAMES, 12345, 5).tobytes(), bytes(range(256)) * 20,
              synth.fill(synth.K_MARKOV, 1 << 20, 6).tobytes(), b"x"]
    ties = np.array([0, 0, 1, 1, 0, 1, 0, 1, 1], dtype=np.int32)
    sizes = np.array([len(b) for b in blocks], dtype=np.uint64)
    offs = np.zeros(len(blocks), dtype=np.uint64)
    offs[1:] = np.cumsum(sizes)[:-1]
    arena = torch.from_numpy(np.frombuffer(b"".join(blocks) + bytes(64), dtype=np.uint8).copy()).cuda()
    r = coder.encode_best_batch(arena, offs, sizes, ties)
    torch.cuda.synchronize()
    assert np.all(r.status.cpu().numpy() == 0)
    ol = r.out_len.cpu().numpy()
    host = r.dst.cpu().numpy()
    for j, b in enumerate(blocks):
        got = host[int(r.dst_off[j]):int(r.dst_off[j]) + int(ol[j])].tobytes()
        assert got == oracle.encode_best(b, int(ties[j])), j
    d = coder.decode_batch(r.dst, r.dst_off, ol, sizes)
    torch.cuda.synchronize()
    assert np.all(d.status.cpu().numpy() == 0)
    out = d.dst.cpu().numpy()
    for j, b in enumerate(blocks):
        assert out[int(d.dst_off[j]):int(d.dst_off[j]) + len(b)].tobytes() == b, j


@pytest.mark.parametrize("k", [64, 65, 100, 128, 129, 192, 193, 255, 256])
def test_order1_wide_alphabets(coder, oracle, k):
    """Order-1 with more than 64 contexts (four contexts packed per lane of
    each gathered column): device encode == oracle, device decode == input,
    for K on both sides of every 64-context boundary."""
    import torch
    rng = np.random.default_rng(900 + k)
    n = 40000 + k
    data = rng.integers(0, k, n).astype(np.uint8).tobytes()
    if k == 64:
        data = data + bytes([200])  # 64 contexts and one symbol that is no context
    ref = oracle.encode(data, 1)
    assert coder.encode(data, 1) == ref
    assert coder.decode(ref) == data
    # Batched, next to a short one of the same width.
    short = rng.integers(0, k, 777).astype(np.uint8).tobytes()
    blobs = [ref, oracle.encode(short, 1)]
    offs = np.array([0, len(ref)], np.uint64)
    lens = np.array([len(b) for b in blobs], np.uint64)
    arena = torch.from_numpy(np.frombuffer(b"".join(blobs) + bytes(64), np.uint8).copy()).cuda()
    r = coder.decode_batch(arena, offs, lens, [len(data), len(short)])
    torch.cuda.synchronize()
    assert list(r.status.cpu().numpy()) == [0, 0]
    out = r.dst.cpu().numpy()
    assert out[:len(data)].tobytes() == data
    assert out[int(r.dst_off[1]):int(r.dst_off[1]) + len(short)].tobytes() == short


def test_encode_best_early_choice_matches_oracle(coder, oracle):
    """Blocks of 16 KiB and more take the early choice: an order whose size
    bounds say it is surely the longer stream is not walked.  Cases where
    the bounds pick order-1 (large Markov qualities), order-0 (small blocks,
    where the order-1 tables cost more than they save; flat bytes) and
    neither (i.i.d. bases, names: both walked, exact compare), each under
    both tie rules: device == oracle, and decoded back."""
    import torch
    from fqxv_amd import synth
    base = [synth.fill(synth.K_MARKOV, 20000, 21).tobytes(), synth.fill(synth.K_MARKOV, 300000, 22).tobytes(),
            synth.fill(synth.K_DNA, 200000, 23).tobytes(), synth.fill(synth.K_NAMES, 200000, 24).tobytes(),
            np.random.default_rng(25).integers(0, 256, 70000).astype(np.uint8).tobytes(), b"A" * 50000,
            synth.fill(synth.K_DNA, 20000, 26).tobytes(), bytes(range(256)) * 300]
    blocks = base + base
    ties = np.array([0] * len(base) + [1] * len(base), dtype=np.int32)
    sizes = np.array([len(b) for b in blocks], dtype=np.uint64)
    offs = np.zeros(len(blocks), dtype=np.uint64)
    offs[1:] = np.cumsum(sizes)[:-1]
    arena = torch.from_numpy(np.frombuffer(b"".join(blocks) + bytes(64), dtype=np.uint8).copy()).cuda()
    r = coder.encode_best_batch(arena, offs, sizes, ties)
    torch.cuda.synchronize()
    assert np.all(r.status.cpu().numpy() == 0)
    ol = r.out_len.cpu().numpy()
    host = r.dst.cpu().numpy()
    for j, b in enumerate(blocks):
        got = host[int(r.dst_off[j]):int(r.dst_off[j]) + int(ol[j])].tobytes()
        assert got == oracle.encode_best(b, int(ties[j])), j
    d = coder.decode_batch(r.dst, r.dst_off, ol, sizes)
    torch.cuda.synchronize()
    assert np.all(d.status.cpu().numpy() == 0)
    out = d.dst.cpu().numpy()
    for j, b in enumerate(blocks):
        assert out[int(d.dst_off[j]):int(d.dst_off[j]) + len(b)].tobytes() == b, j


def test_encode_best_random_blocks_match_oracle(coder, oracle):
    """Encode-best over 48 seeded blocks of 16 KiB to 512 KiB (log-uniform),
    kinds and tie rules drawn at random -- Markov qualities, i.i.d. bases,
    names, flat bytes, a few-symbol skewed source -- so the early choice
    meets every outcome at many sizes: device == oracle for every block."""
    import torch
    from fqxv_amd import synth
    rng = np.random.default_rng(31337)
    blocks, ties = [], []
    for j in range(48):
        n = int(np.exp(rng.uniform(np.log(16384), np.log(524288))))
        kind = int(rng.integers(0, 5))
        if kind < 3:
            b = synth.fill((synth.K_MARKOV, synth.K_DNA, synth.K_NAMES)[kind], n, 500 + j).tobytes()
        elif kind == 3:
            b = rng.integers(0, 256, n).astype(np.uint8).tobytes()
        else:
            b = rng.choice([7, 8, 9, 200], n, p=[0.9, 0.05, 0.04, 0.01]).astype(np.uint8).tobytes()
        blocks.append(b)
        ties.append(int(rng.integers(0, 2)))
    ties = np.array(ties, dtype=np.int32)
    sizes = np.array([len(b) for b in blocks], dtype=np.uint64)
    offs = np.zeros(len(blocks), dtype=np.uint64)
    offs[1:] = np.cumsum(sizes)[:-1]
    arena = torch.from_numpy(np.frombuffer(b"".join(blocks) + bytes(64), dtype=np.uint8).copy()).cuda()
    r = coder.encode_best_batch(arena, offs, sizes, ties)
    torch.cuda.synchronize()
    assert np.all(r.status.cpu().numpy() == 0)
    ol = r.out_len.cpu().numpy()
    host = r.dst.cpu().numpy()
    for j, b in enumerate(blocks):
        got = host[int(r.dst_off[j]):int(r.dst_off[j]) + int(ol[j])].tobytes()
        assert got == oracle.encode_best(b, int(ties[j])), j


def test_encode_best_both_orders_one_walk_edges(coder, oracle):
    """When neither order's size bounds win, both orders are walked in one
    pass (order-1 states in lanes 0..31, order-0 in 32..63) if both tables
    fit the shared table (alphabets up to 41 symbols), else one after the
    other.  Sources with 41 and 42 symbols whose order-1 gain sweeps from
    none to large (the previous symbol repeated with probability p), at sizes
    past the early-choice threshold with partial last rounds, alternating tie
    rules: device == oracle for every block, and decoded back."""
    import torch
    rng = np.random.default_rng(4141)
    blocks, ties = [], []
    # (p, n): the first three are decided by the bounds (order-0), the next
    # six are not (by the oracle's streams and the device's integer bounds:
    # six undecided blocks per alphabet), the last picks order-1.
    cases = [(0.0, 16385), (0.1, 70001), (0.5, 70001), (0.2, 300007), (0.22, 300007), (0.24, 300007),
             (0.26, 300007), (0.18, 700001), (0.2, 700001), (0.22, 700001)]
    for a in (41, 42):
        for p, n in cases:
            src = rng.integers(0, a, n)
            rep = rng.random(n) < p
            for i in np.flatnonzero(rep):
                if i:
                    src[i] = src[i - 1]
            blocks.append((src + 33).astype(np.uint8).tobytes())
            ties.append(len(blocks) % 2)
    ties = np.array(ties, dtype=np.int32)
    sizes = np.array([len(b) for b in blocks], dtype=np.uint64)
    offs = np.zeros(len(blocks), dtype=np.uint64)
    offs[1:] = np.cumsum(sizes)[:-1]
    arena = torch.from_numpy(np.frombuffer(b"".join(blocks) + bytes(64), dtype=np.uint8).copy()).cuda()
    r = coder.encode_best_batch(arena, offs, sizes, ties)
    torch.cuda.synchronize()
    assert np.all(r.status.cpu().numpy() == 0)
    ol = r.out_len.cpu().numpy()
    host = r.dst.cpu().numpy()
    for j, b in enumerate(blocks):
        got = host[int(r.dst_off[j]):int(r.dst_off[j]) + int(ol[j])].tobytes()
        assert got == oracle.encode_best(b, int(ties[j])), j
    d = coder.decode_batch(r.dst, r.dst_off, ol, sizes)
    torch.cuda.synchronize()
    assert np.all(d.status.cpu().numpy() == 0)
    out = d.dst.cpu().numpy()
    for j, b in enumerate(blocks):
        assert out[int(d.dst_off[j]):int(d.dst_off[j]) + len(b)].tobytes() == b, j


def test_order1_row_staging_flag(coder, oracle):
    """An order-1 block decodes to the same bytes with its code rows in LDS
    (flags 0; the rows fit for these alphabets) and in the workspace
    (FQXV_RANS_JOB_O1_ROWS_WORKSPACE), equal to the oracle; a job with an
    unknown flag bit fails with the invalid-argument code and writes nothing."""
    import torch
    from fqxv_amd import _lib, synth
    blocks = [synth.fill(k, n, 900 + i).tobytes() for i, (k, n) in
              enumerate([(synth.K_MARKOV, 300001), (synth.K_DNA, 70000), (synth.K_NAMES, 100003)])]
    blocks.append(rng_bytes(907, 50000, 64))  # 64 contexts: rows past the LDS region
    streams = [oracle.encode(b, 1) for b in blocks]
    src = np.frombuffer(b"".join(streams) + bytes(64), dtype=np.uint8)
    s_len = np.array([len(x) for x in streams], dtype=np.uint64)
    s_off = np.zeros(len(streams), dtype=np.uint64)
    s_off[1:] = np.cumsum(s_len)[:-1]
    sizes = np.array([len(b) for b in blocks], dtype=np.uint64)
    d_off = np.zeros(len(blocks), dtype=np.uint64)
    d_off[1:] = np.cumsum(sizes)[:-1]
    dsrc = torch.from_numpy(src.copy()).cuda()
    outs = []
    for flags in (0, _lib.JOB_O1_ROWS_WORKSPACE):
        jobs = coder.jobs_array(s_off, s_len, d_off, sizes, flags=flags)
        dst = torch.full((int(sizes.sum()) + 64,), 0xEE, dtype=torch.uint8, device="cuda")
        g = coder.graph("decode", dsrc, dst, jobs)
        g.launch()
        torch.cuda.synchronize()
        assert np.all(g.status.cpu().numpy() == 0), flags
        g.close()
        host = dst.cpu().numpy()
        for j, b in enumerate(blocks):
            assert host[int(d_off[j]):int(d_off[j]) + len(b)].tobytes() == b, (flags, j)
        outs.append(host)
    assert np.array_equal(outs[0], outs[1])
    jobs = coder.jobs_array(s_off[:1], s_len[:1], d_off[:1], sizes[:1], flags=2)
    dst = torch.full((int(sizes[0]) + 64,), 0xEE, dtype=torch.uint8, device="cuda")
    g = coder.graph("decode", dsrc, dst, jobs)
    g.launch()
    torch.cuda.synchronize()
    assert int(g.status.cpu()[0]) == _lib.ERR_INVALID_ARG
    assert bool((dst == 0xEE).all())
    g.close()


def test_order1_widest_tables_on_every_wave(coder, oracle):
    """Each wave's order-1 code table starts at its own offset in the wave's
    workspace.  A batch of more widest-table blocks than there are waves --
    256 contexts (the 1 MiB table of the wide path) and 64 contexts (the
    fast path's widest rows), alternating -- has every wave, so every
    offset, build and gather a widest table: decode == input for each."""
    import torch
    coder.set_grid(0)
    waves = coder.max_waves()
    rng = np.random.default_rng(4242)
    blocks = []
    for i in range(waves + 64):
        k = 256 if i % 2 == 0 else 64
        b = rng.integers(0, k, 1500 + i % 97).astype(np.uint8)
        b[:k] = np.arange(k)  # every symbol, so every context, occurs
        blocks.append(b.tobytes())
    streams = [oracle.encode(b, 1) for b in blocks]
    lens = np.array([len(x) for x in streams], np.uint64)
    offs = np.zeros_like(lens)
    offs[1:] = np.cumsum(lens)[:-1]
    arena = torch.from_numpy(np.frombuffer(b"".join(streams) + bytes(64), np.uint8).copy()).cuda()
    r = coder.decode_batch(arena, offs, lens, [len(b) for b in blocks])
    torch.cuda.synchronize()
    assert (r.status.cpu().numpy() == 0).all()
    out = r.dst.cpu().numpy()
    for j, b in enumerate(blocks):
        o = int(r.dst_off[j])
        assert out[o:o + len(b)].tobytes() == b, j


@pytest.mark.parametrize("k", [1, 2, 40, 41, 42, 43, 44, 63, 64])
@pytest.mark.parametrize("base", [0, 33])
def test_order1_encode_dense_capacity(coder, oracle, k, base):
    """Order-1 encode on both sides of the dense LDS table's capacity (an
    (alphabet + 1) x alphabet table of at most 1808 entries: 42 symbols) and
    with symbol 0 present or absent (absent: context 0 takes the extra row).
    Single and batched device encode == oracle; batched decode == input."""
    import torch
    rng = np.random.default_rng(1300 + 2 * k + base)
    blocks = []
    for n in (5 + k, 3000 + k, 70000 + k):
        b = (base + rng.integers(0, k, n)).astype(np.uint8)
        b[: min(k, n)] = base + np.arange(min(k, n))  # every symbol occurs
        blocks.append(b.tobytes())
    for b in blocks:
        assert coder.encode(b, 1) == oracle.encode(b, 1)
    sizes = np.array([len(b) for b in blocks], dtype=np.uint64)
    offs = np.zeros(len(blocks), dtype=np.uint64)
    offs[1:] = np.cumsum(sizes)[:-1]
    arena = torch.from_numpy(np.frombuffer(b"".join(blocks) + bytes(64), dtype=np.uint8).copy()).cuda()
    r = coder.encode_batch(arena, offs, sizes, [1] * len(blocks))
    torch.cuda.synchronize()
    assert np.all(r.status.cpu().numpy() == 0)
    ol = r.out_len.cpu().numpy()
    host = r.dst.cpu().numpy()
    for j, b in enumerate(blocks):
        assert host[int(r.dst_off[j]):int(r.dst_off[j]) + int(ol[j])].tobytes() == oracle.encode(b, 1), j
    d = coder.decode_batch(r.dst, r.dst_off, ol, sizes)
    torch.cuda.synchronize()
    assert np.all(d.status.cpu().numpy() == 0)
    assert torch.equal(d.dst[: int(sizes.sum())], arena[: int(sizes.sum())])


@pytest.mark.parametrize("shift", [0, 1, 2, 3])
def test_encode_histogram_edges(coder, oracle, shift):
    """Block sizes around the histogram's 2 KiB load batches and its 8-byte
    bulk limit, at every source alignment, both orders: batched device
    encode == oracle."""
    import torch
    rng = np.random.default_rng(77 + shift)
    sizes = [0, 1, 4, 7, 8, 9, 12, 2047, 2048, 2049, 2055, 2056, 2057, 2060, 4103, 4104, 4105, 6000, 20483]
    blocks = []
    for i, n in enumerate(sizes):
        alpha = 3 + (i % 5) * 9  # 3 .. 39 symbols
        blocks.append((33 + rng.integers(0, alpha, n)).astype(np.uint8).tobytes())
    for order in (0, 1):
        lens = np.array([len(b) for b in blocks], dtype=np.uint64)
        offs = np.zeros(len(blocks), dtype=np.uint64)
        offs[1:] = np.cumsum(lens + 3)[:-1]
        offs += shift
        buf = bytearray(int(offs[-1] + lens[-1]) + 64)
        for o, b in zip(offs, blocks):
            buf[int(o):int(o) + len(b)] = b
        arena = torch.from_numpy(np.frombuffer(bytes(buf), dtype=np.uint8).copy()).cuda()
        r = coder.encode_batch(arena, offs, lens, [order] * len(blocks))
        torch.cuda.synchronize()
        assert np.all(r.status.cpu().numpy() == 0)
        ol = r.out_len.cpu().numpy()
        host = r.dst.cpu().numpy()
        for j, b in enumerate(blocks):
            got = host[int(r.dst_off[j]):int(r.dst_off[j]) + int(ol[j])].tobytes()
            assert got == oracle.encode(b, order), (order, len(b))


def test_order1_wide_malformed_same_error(coder, oracle):
    rng = np.random.default_rng(77)
    data = rng.integers(0, 180, 20000).astype(np.uint8).tobytes()
    ref = bytearray(oracle.encode(data, 1))
    cases = [bytes(ref[:len(ref) - 3]), bytes(ref[:len(ref) // 2])]
    for _ in range(40):
        m = bytearray(ref)
        i = int(rng.integers(0, len(m)))
        m[i] ^= int(rng.integers(1, 256))
        cases.append(bytes(m))
    checked = 0
    for c in cases:
        n = oracle.peek_len(c)
        if n is not None and n > len(data) * 4:  # a mutated length: skip the huge allocation
            continue
        rc_o, out_o, _ = oracle.decode_ex(c)
        assert _err_code(lambda: coder.decode(c)) == rc_o
        if rc_o == 0:
            assert coder.decode(c) == out_o
        checked += 1
    assert checked > 30


def test_decode_with_every_backend(coder, oracle, fixture_streams):
    """lib.rs:241-255 `decode_with`: each device form (the single-block launch
    and the batched kernel fed a one-job queue) decodes byte-identically to the
    oracle, on the edge cases and the reference-written fixture streams, and
    fails malformed streams with the oracle's error; Backend::detect names the
    form plain decode uses."""
    import fqxv_amd
    from fqxv_amd import Backend
    assert Backend.detect() == Backend.Gfx950
    streams = [oracle.encode(d, o) for d in _cases() for o in (0, 1)] + [s for _, s in fixture_streams]
    for s in streams:
        ref = oracle.decode(s)
        for be in (Backend.Gfx950, Backend.Gfx950Batched):
            assert coder.decode_with(s, be) == ref, (len(s), be)
    assert fqxv_amd.decode_with(streams[3], Backend.Gfx950Batched) == oracle.decode(streams[3])
    r = np.random.default_rng(78)
    checked = 0
    for it in range(120):
        data = r.integers(0, 30, r.integers(0, 2000)).astype(np.uint8).tobytes()
        enc = bytearray(oracle.encode(data, it & 1))
        for _ in range(r.integers(1, 6)):
            if enc:
                enc[r.integers(0, len(enc))] = r.integers(0, 256)
        b = bytes(enc)
        n = oracle.peek_len(b)
        if n is not None and n > max(len(data), 1) * 4:
            continue
        rc, out, _ = oracle.decode_ex(b)
        for be in (Backend.Gfx950, Backend.Gfx950Batched):
            assert _err_code(lambda: coder.decode_with(b, be)) == rc, (it, be)
        checked += 1
    assert checked > 60
