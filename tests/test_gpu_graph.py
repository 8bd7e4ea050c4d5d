"""GPU: the batched call captured into a hipGraph, and the grid setter.

Replays run several times with eager device work (fills, other kernels,
synchronisations) between them; every replay must reproduce the eager
batched result byte for byte, and the eager result is pinned to the oracle.
"""
import numpy as np
import pytest

from conftest import rng_bytes

pytestmark = pytest.mark.gpu


def _blocks():
    rng = np.random.default_rng(3)
    out = []
    for i in range(40):
        n = int(rng.integers(0, 60000)) if i % 5 else int(rng.integers(200000, 400000))
        out.append(rng_bytes(900 + i, n, int(rng.integers(2, 120))))
    return out


def _arena(blocks):
    import torch
    sizes = np.array([len(b) for b in blocks], dtype=np.uint64)
    offs = np.zeros(len(blocks), dtype=np.uint64)
    offs[1:] = np.cumsum(sizes)[:-1]
    arena = torch.from_numpy(np.frombuffer(b"".join(blocks) + bytes(64), dtype=np.uint8).copy()).cuda()
    return arena, offs, sizes


def test_decode_graph_replays_match_eager_and_oracle(coder, oracle):
    import torch
    blocks = _blocks()
    orders = [i & 1 for i in range(len(blocks))]
    arena, offs, sizes = _arena(blocks)
    enc = coder.encode_batch(arena, offs, sizes, orders)
    torch.cuda.synchronize()
    assert np.all(enc.status.cpu().numpy() == 0)
    ol = enc.out_len.cpu().numpy().astype(np.uint64)
    host = enc.dst.cpu().numpy()
    for j in range(0, len(blocks), 3):
        s = host[int(enc.dst_off[j]):int(enc.dst_off[j] + ol[j])].tobytes()
        assert s == oracle.encode(blocks[j], orders[j]), j

    dst_off = np.zeros_like(sizes)
    dst_off[1:] = np.cumsum(sizes)[:-1]
    dst = torch.empty(int(sizes.sum()) + 64, dtype=torch.uint8, device="cuda")
    jobs = coder.jobs_array(enc.dst_off, ol, dst_off, sizes)
    g = coder.graph("decode", enc.dst, dst, jobs, coder.lpt_order(sizes))
    n = int(sizes.sum())
    junk = torch.zeros(1 << 20, device="cuda")
    for rep in range(4):
        dst.fill_(0xA5 + rep)
        g.status.fill_(-1)
        g.out_len.zero_()
        junk.add_(1.0)
        torch.cuda.synchronize()
        g.launch()
        junk.mul_(0.5)
        torch.cuda.synchronize()
        assert np.all(g.status.cpu().numpy() == 0), rep
        assert np.array_equal(g.out_len.cpu().numpy(), sizes.astype(np.int64)), rep
        assert torch.equal(dst[:n], arena[:n]), rep
        # an eager batched call between replays shares the context
        e = coder.decode_batch(enc.dst, enc.dst_off, ol, sizes)
        torch.cuda.synchronize()
        assert torch.equal(e.dst[:n], arena[:n]), rep
    g.close()


def test_encode_graph_replays_match_eager(coder):
    import torch
    blocks = _blocks()[:24]
    orders = [(i >> 1) & 1 for i in range(len(blocks))]
    arena, offs, sizes = _arena(blocks)
    ref = coder.encode_batch(arena, offs, sizes, orders)
    torch.cuda.synchronize()
    ref_host = ref.dst.cpu().numpy()
    ref_len = ref.out_len.cpu().numpy()
    caps = np.array([coder.encode_bound(int(n), o) for n, o in zip(sizes, orders)], dtype=np.uint64)
    dst_off = np.zeros_like(caps)
    dst_off[1:] = np.cumsum(caps)[:-1]
    dst = torch.empty(int(caps.sum()) + 64, dtype=torch.uint8, device="cuda")
    jobs = coder.jobs_array(offs, sizes, dst_off, caps, order=np.array(orders, dtype=np.int32))
    g = coder.graph("encode", arena, dst, jobs, None)
    for rep in range(3):
        dst.fill_(rep)
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        assert np.all(g.status.cpu().numpy() == 0)
        got_len = g.out_len.cpu().numpy()
        assert np.array_equal(got_len, ref_len)
        h = dst.cpu().numpy()
        for j in range(len(blocks)):
            a = h[int(dst_off[j]):int(dst_off[j]) + int(got_len[j])]
            b = ref_host[int(ref.dst_off[j]):int(ref.dst_off[j]) + int(ref_len[j])]
            assert np.array_equal(a, b), (rep, j)
    g.close()


@pytest.mark.parametrize("waves", [1, 3, 64, 0])
def test_grid_size_does_not_change_output(coder, oracle, waves):
    import torch
    blocks = _blocks()[:20]
    orders = [i & 1 for i in range(len(blocks))]
    arena, offs, sizes = _arena(blocks)
    try:
        coder.set_grid(waves)
        enc = coder.encode_batch(arena, offs, sizes, orders, lpt=bool(waves & 1))
        torch.cuda.synchronize()
        assert np.all(enc.status.cpu().numpy() == 0)
        ol = enc.out_len.cpu().numpy()
        host = enc.dst.cpu().numpy()
        for j in (0, 1, 5, 19):
            s = host[int(enc.dst_off[j]):int(enc.dst_off[j]) + int(ol[j])].tobytes()
            assert s == oracle.encode(blocks[j], orders[j]), j
        d = coder.decode_batch(enc.dst, enc.dst_off, ol, sizes, lpt=not (waves & 1))
        torch.cuda.synchronize()
        assert np.all(d.status.cpu().numpy() == 0)
        n = int(sizes.sum())
        assert torch.equal(d.dst[:n], arena[:n])
    finally:
        coder.set_grid(0)


def test_autotune_picks_a_measured_candidate_and_caches(coder, tmp_path, monkeypatch):
    import torch
    from fqxv_amd import tune
    monkeypatch.setenv("FQXV_RANS_TUNE_CACHE", str(tmp_path / "tune.json"))
    blocks = _blocks()
    orders = np.array([i & 1 for i in range(len(blocks))], dtype=np.int32)
    arena, offs, sizes = _arena(blocks)
    enc = coder.encode_batch(arena, offs, sizes, orders)
    torch.cuda.synchronize()
    ol = enc.out_len.cpu().numpy().astype(np.uint64)
    dst_off = np.zeros_like(sizes)
    dst_off[1:] = np.cumsum(sizes)[:-1]
    dst = torch.empty(int(sizes.sum()) + 64, dtype=torch.uint8, device="cuda")
    jobs = coder.jobs_array(enc.dst_off, ol, dst_off, sizes)
    try:
        t = tune.autotune(coder, "decode", enc.dst, dst, jobs, sizes, orders, grids=[40, 8], reps=1)
        assert (t.grid, t.policy, t.staging, t.ms) in [tuple(r) for r in t.table]
        assert len(t.table) == 2 * len(tune.POLICIES) * len(tune.STAGINGS)
        assert {r[2] for r in t.table} == set(tune.STAGINGS)
        assert (tmp_path / "tune.json").exists()
        again = tune.autotune(coder, "decode", enc.dst, dst, jobs, sizes, orders, grids=[40, 8], reps=1)
        assert (again.grid, again.policy, again.staging) == (t.grid, t.policy, t.staging)
        g = coder.graph("decode", enc.dst, dst, t.jobs(jobs), t.order(coder, sizes, orders))
        dst.zero_()
        g.launch()
        torch.cuda.synchronize()
        n = int(sizes.sum())
        assert np.all(g.status.cpu().numpy() == 0)
        assert torch.equal(dst[:n], arena[:n])
        g.close()
    finally:
        coder.set_grid(0)
