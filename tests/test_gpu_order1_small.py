"""Order-1 decode of blocks whose contexts and symbols together number at
most eight (the register path: DNA-like blocks), against the oracle.

The decoder composes the round's per-slot maps in registers for these
blocks, so they get their own cases: every code-space size 1..8, context 0
inside and outside the symbol set, round counts around the 8-round blocks
and the 32-symbol rounds, i.i.d. and strongly correlated symbols, and
mutated or truncated streams (the first error must be the oracle's,
including absent contexts reached mid-round).
"""
import numpy as np
import pytest

from conftest import rng_bytes

pytestmark = pytest.mark.gpu

SIZES = (1, 2, 31, 32, 33, 255, 256, 257, 1000, 4109, 70001)


def _markov(rng, alphabet, n, stay):
    """Symbols from `alphabet`; each repeats the previous one with
    probability `stay`, else is drawn uniformly."""
    a = np.asarray(alphabet, dtype=np.uint8)
    pick = rng.integers(0, len(a), n)
    keep = rng.random(n) < stay
    idx = np.empty(n, dtype=np.int64)
    cur = pick[0] if n else 0
    for i in range(n):
        if not keep[i] or i == 0:
            cur = pick[i]
        idx[i] = cur
    return a[idx].tobytes()


def _blocks():
    rng = np.random.default_rng(8)
    out = []
    for a in range(1, 9):
        with_zero = list(range(a))                 # context 0 is a symbol: nu = a
        offset = [65 + 3 * i for i in range(a)]    # context 0 is extra: nu = a + 1
        for alpha in (with_zero, offset):
            if len(alpha) + (alpha[0] != 0) > 8:
                continue
            for n in SIZES:
                out.append(_markov(rng, alpha, n, 0.0))
            out.append(_markov(rng, alpha, 30011, 0.9))
            out.append(_markov(rng, alpha, 65536, 0.5))
    from fqxv_amd import synth
    for i, n in enumerate((4096, 100003, 1 << 20)):
        out.append(synth.fill(synth.K_DNA, n, 700 + i).tobytes())
    # A symbol that is no context (only at the end) and a skewed table.
    out.append(b"A" * 5000 + b"C" * 3000 + b"T")
    out.append(bytes([7]) * 9000 + bytes([1, 2, 3]) * 17)
    return out


def _nu(data):
    syms = set(data)
    ctx = {0} | set(data[:-1])
    return len(syms | ctx)


def test_small_code_space_round_trips(coder, oracle):
    import torch
    blocks = _blocks()
    assert all(_nu(b) <= 8 for b in blocks)
    streams = [oracle.encode(b, 1) for b in blocks]
    for b, s in zip(blocks, streams):
        assert coder.decode(s) == b, (len(b), _nu(b))
    # The same streams in one batched launch.
    sizes = np.array([len(b) for b in blocks], dtype=np.uint64)
    s_len = np.array([len(s) for s in streams], dtype=np.uint64)
    s_off = np.zeros(len(streams), dtype=np.uint64)
    s_off[1:] = np.cumsum(s_len)[:-1]
    src = torch.from_numpy(np.frombuffer(b"".join(streams) + bytes(64), dtype=np.uint8).copy()).cuda()
    d = coder.decode_batch(src, s_off, s_len, sizes)
    torch.cuda.synchronize()
    assert np.all(d.status.cpu().numpy() == 0)
    host = d.dst.cpu().numpy()
    for j, b in enumerate(blocks):
        assert host[int(d.dst_off[j]):int(d.dst_off[j]) + len(b)].tobytes() == b, j


def _err_code(fn):
    import fqxv_amd
    try:
        fn()
        return 0
    except fqxv_amd.Malformed as e:
        return e.code


def test_small_code_space_mutants_same_error(coder, oracle):
    from fqxv_amd import _lib
    rng = np.random.default_rng(88)
    bases = []
    for a in (2, 4, 5, 6):
        alpha = [65 + i for i in range(a)]
        for n in (40, 333, 4000, 20000):
            d = bytearray(_markov(rng, alpha, n, 0.6))
            d[-1] = 90  # 'Z' decodes only at the end: no context of its own
            bases.append(bytes(d))
    bases.append(rng_bytes(5, 6000, 6))
    seen = {}
    checked = 0
    for it in range(600):
        data = bases[it % len(bases)]
        ref = bytearray(oracle.encode(data, 1))
        m = bytearray(ref)
        kind = it % 3
        if kind == 0:      # states and words
            k = sum(bin(b).count("1") for b in m[9:41])
            for _ in range(int(rng.integers(1, 4))):
                i = int(rng.integers(41 + 512 * k, len(m)))
                m[i] ^= int(rng.integers(1, 256))
        elif kind == 1:    # anywhere
            i = int(rng.integers(0, len(m)))
            m[i] ^= int(rng.integers(1, 256))
        else:              # truncation
            del m[int(rng.integers(0, len(m))):]
        c = bytes(m)
        n = oracle.peek_len(c)
        if n is not None and n > len(data) * 4:
            continue
        rc_o, out_o, _ = oracle.decode_ex(c)
        rc_g = _err_code(lambda: coder.decode(c))
        assert rc_g == rc_o, (it, rc_o, rc_g)
        if rc_o == 0:
            assert coder.decode(c) == out_o, it
        seen[rc_o] = seen.get(rc_o, 0) + 1
        checked += 1
    assert checked > 400
    # The cases reach the path's own exits: clean decodes, truncation and
    # absent contexts met inside a round.
    for code in (0, _lib.ERR_TRUNCATED, _lib.ERR_ABSENT_CONTEXT):
        assert seen.get(code, 0) > 5, seen
