"""Batched encode with the histogram pre-pass: the long jobs (2 MiB and up,
first in the queue) have their byte and pair counts taken by many waves in
slices before the encode kernel runs.  The streams must be the oracle's
byte for byte: slices that split a symbol pair, the first slice's context 0,
an alphabet too wide for the dense pair table (counted by the job's own
wave), order-0 long jobs (byte counts only), unaligned block offsets, and a
queue in index order, where the long jobs are not a prefix of the queue.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MIB = 1 << 20


def _blocks():
    from fqxv_amd import synth
    rng = np.random.default_rng(2024)
    blocks, orders = [], []

    def add(b, o):
        blocks.append(b)
        orders.append(o)
    add(synth.fill(synth.K_MARKOV, 2 * MIB + 12345, 41).tobytes(), 1)
    add(synth.fill(synth.K_DNA, 3 * MIB + 7, 42).tobytes(), 1)
    add(synth.fill(synth.K_NAMES, 2 * MIB + 1, 43).tobytes(), 0)
    add(rng.integers(0, 200, 2 * MIB + 333).astype(np.uint8).tobytes(), 1)   # 200 symbols: not dense
    add(synth.fill(synth.K_MARKOV, 2 * MIB, 44).tobytes(), 0)
    add(bytes([9]) * (2 * MIB + 5), 1)                                        # one symbol
    add(rng.choice([0, 1, 2, 250], 2 * MIB + 77, p=[0.7, 0.2, 0.09, 0.01]).astype(np.uint8).tobytes(), 1)
    for j in range(12):
        n = int(rng.integers(1000, 300000))
        add(synth.fill((synth.K_MARKOV, synth.K_DNA, synth.K_NAMES)[j % 3], n, 60 + j).tobytes(), j & 1)
    return blocks, np.array(orders, dtype=np.int32)


def _arena(blocks):
    import torch
    # Unaligned starts: one to three bytes of padding before each block.
    pad = [1 + (j % 3) for j in range(len(blocks))]
    offs, parts, at = [], [], 0
    for p, b in zip(pad, blocks):
        parts.append(bytes(p))
        at += p
        offs.append(at)
        parts.append(b)
        at += len(b)
    arena = torch.from_numpy(np.frombuffer(b"".join(parts) + bytes(64), dtype=np.uint8).copy()).cuda()
    return arena, np.array(offs, dtype=np.uint64), np.array([len(b) for b in blocks], dtype=np.uint64)


@pytest.mark.parametrize("lpt", [True, False])
def test_encode_with_prepass_matches_oracle(coder, oracle, lpt):
    import torch
    blocks, orders = _blocks()
    arena, offs, sizes = _arena(blocks)
    r = coder.encode_batch(arena, offs, sizes, orders, lpt=lpt)
    torch.cuda.synchronize()
    assert np.all(r.status.cpu().numpy() == 0)
    ol = r.out_len.cpu().numpy()
    host = r.dst.cpu().numpy()
    for j, b in enumerate(blocks):
        got = host[int(r.dst_off[j]):int(r.dst_off[j]) + int(ol[j])].tobytes()
        assert got == oracle.encode(b, int(orders[j])), (lpt, j, len(b))


def test_encode_best_with_prepass_matches_oracle(coder, oracle):
    import torch
    blocks, orders = _blocks()
    arena, offs, sizes = _arena(blocks)
    ties = (np.arange(len(blocks)) % 2).astype(np.int32)
    r = coder.encode_best_batch(arena, offs, sizes, ties)
    torch.cuda.synchronize()
    assert np.all(r.status.cpu().numpy() == 0)
    ol = r.out_len.cpu().numpy()
    host = r.dst.cpu().numpy()
    for j, b in enumerate(blocks):
        got = host[int(r.dst_off[j]):int(r.dst_off[j]) + int(ol[j])].tobytes()
        assert got == oracle.encode_best(b, int(ties[j])), (j, len(b))


def test_prepass_graph_replays_match_eager(coder, oracle):
    """A captured encode (pre-pass kernels included) replayed three times,
    with eager work between the replays, gives the eager call's streams."""
    import torch
    blocks, orders = _blocks()
    arena, offs, sizes = _arena(blocks)
    eager = coder.encode_batch(arena, offs, sizes, orders)
    torch.cuda.synchronize()
    caps = coder.encode_bounds(sizes, orders)
    dst_off = np.zeros_like(caps)
    dst_off[1:] = np.cumsum(caps)[:-1]
    jobs = coder.jobs_array(offs, sizes, dst_off, caps, order=orders)
    dst = torch.zeros(int(caps.sum()) + 64, dtype=torch.uint8, device="cuda")
    from fqxv_amd import tune
    g = coder.graph("encode", arena, dst, jobs, tune.queue_order(coder, "lpt-cost", sizes, orders, "encode"))
    ref = eager.dst.cpu().numpy()
    rl = eager.out_len.cpu().numpy()
    for rep in range(3):
        dst.fill_(0xEE)
        g.launch()
        torch.cuda.synchronize()
        assert np.all(g.status.cpu().numpy() == 0)
        host = dst.cpu().numpy()
        for j in range(len(blocks)):
            a, n = int(dst_off[j]), int(rl[j])
            assert host[a:a + n].tobytes() == ref[int(eager.dst_off[j]):int(eager.dst_off[j]) + n].tobytes(), (rep, j)
        # eager work between replays
        coder.encode_batch(arena, offs[:3], sizes[:3], orders[:3])
        torch.cuda.synchronize()
    g.close()
