"""GPU: the serving engine (random-access decode from a resident archive).

Requests arrive as a Poisson process and overlap on the device; each result
must be the same bytes alone and in the mix, and equal to the blocks that
were encoded (whose encoding is pinned to the oracle).
"""
import numpy as np
import pytest

from conftest import rng_bytes

pytestmark = pytest.mark.gpu


def _archive(coder, oracle):
    import torch
    rng = np.random.default_rng(55)
    blocks = []
    for i in range(48):
        n = int(np.exp(rng.uniform(np.log(64), np.log(300000))))
        blocks.append(rng_bytes(5000 + i, n, int(rng.integers(2, 60))))
    orders = np.array([int(rng.integers(0, 2)) for _ in blocks], dtype=np.int32)
    sizes = np.array([len(b) for b in blocks], dtype=np.uint64)
    offs = np.zeros(len(blocks), dtype=np.uint64)
    offs[1:] = np.cumsum(sizes)[:-1]
    arena = torch.from_numpy(np.frombuffer(b"".join(blocks) + bytes(64), dtype=np.uint8).copy()).cuda()
    enc = coder.encode_batch(arena, offs, sizes, orders)
    torch.cuda.synchronize()
    assert np.all(enc.status.cpu().numpy() == 0)
    ol = enc.out_len.cpu().numpy().astype(np.uint64)
    host = enc.dst.cpu().numpy()
    for j in range(0, len(blocks), 5):
        assert host[int(enc.dst_off[j]):int(enc.dst_off[j] + ol[j])].tobytes() == oracle.encode(blocks[j], orders[j])
    return blocks, enc, ol


def _server(enc, ol, blocks, budget, slots=3, extra_bad=False, mode="slots"):
    from fqxv_amd.serve import Server
    src_off = list(enc.dst_off)
    src_len = list(ol)
    n = [len(b) for b in blocks]
    if extra_bad:  # a truncated copy of block 0's stream
        src_off.append(enc.dst_off[0])
        src_len.append(5)
        n.append(len(blocks[0]))
    return Server(enc.dst, src_off, src_len, n, budget=budget, slots=slots, max_batch=256, mode=mode,
                  ring=256, idle_ms=50)


MODES = ["slots", "persistent"]


@pytest.mark.parametrize("mode", MODES)
def test_serving_alone_equals_in_mix(coder, oracle, mode):
    from fqxv_amd.serve import poisson_requests, run_load
    blocks, enc, ol = _archive(coder, oracle)
    nb = len(blocks)
    arrivals, reqs = poisson_requests(5, 60, nb, rate=400.0, kmax=16)
    sample = [0, 7, 13, 29, 44, 59]

    alone = {}
    srv = _server(enc, ol, blocks, budget=16 << 20, mode=mode)
    for i in sample:
        def grab(_, c, ids=reqs[i]):
            alone[i] = [t.cpu().numpy().tobytes() for t in srv.output(c, ids)]
        recs = run_load(srv, [0.0], [reqs[i]], on_done=grab)
        assert recs[0].status == 0
    srv.close()

    seen = {}
    srv = _server(enc, ol, blocks, budget=6 << 20, mode=mode)  # small budget: admission must wait

    def check(i, c):
        outs = srv.output(c, reqs[i])
        seen[i] = [t.cpu().numpy().tobytes() for t in outs]

    recs = run_load(srv, arrivals, reqs, on_done=check)
    assert srv.max_in_flight > 1, "requests never overlapped on the device"
    for i, r in enumerate(recs):
        assert r.status == 0, i
        assert r.t_arrival <= r.t_admit <= r.t_done
        assert seen[i] == [blocks[b] for b in reqs[i]], i
    for i in sample:
        assert alone[i] == seen[i], i
    srv.close()


@pytest.mark.parametrize("mode", MODES)
def test_serving_errors_stay_per_request(coder, oracle, mode):
    import fqxv_amd
    from fqxv_amd.serve import run_load
    blocks, enc, ol = _archive(coder, oracle)
    nb = len(blocks)
    srv = _server(enc, ol, blocks, budget=32 << 20, extra_bad=True, mode=mode)
    reqs = [np.array([1, 2, 3], np.uint32), np.array([4, nb], np.uint32), np.array([5, 6], np.uint32)]
    outs = {}

    def grab(i, c):
        if c.status == 0:
            outs[i] = [t.cpu().numpy().tobytes() for t in srv.output(c, reqs[i])]

    recs = run_load(srv, [0.0, 0.0, 0.0], reqs, on_done=grab)
    assert recs[0].status == 0 and recs[2].status == 0
    assert recs[1].status == fqxv_amd.rans._lib.ERR_TRUNCATED
    assert outs[0] == [blocks[b] for b in reqs[0]]
    assert outs[2] == [blocks[b] for b in reqs[2]]
    with pytest.raises(ValueError):
        srv.submit([nb + 1])
    with pytest.raises(ValueError):
        srv.submit([])
    srv.close()
    sizes = [len(b) for b in blocks]
    big = int(np.argmax(sizes))
    with pytest.raises(ValueError):  # a block larger than the whole budget
        _server(enc, ol, blocks, budget=1024, mode=mode)
    small = _server(enc, ol, blocks, budget=(sizes[big] + 255) // 256 * 256, mode=mode)
    second = int(np.argsort(sizes)[-2])
    with pytest.raises(ValueError):  # a request larger than the whole budget
        small.submit([big, second])
    small.close()


def test_persistent_grid_leaves_when_idle_and_comes_back(coder, oracle):
    """The resident grid drains after idle_ms without work; later requests
    relaunch it; destroy returns (every wave has an exit)."""
    import time
    from fqxv_amd.serve import run_load
    blocks, enc, ol = _archive(coder, oracle)
    srv = _server(enc, ol, blocks, budget=32 << 20, mode="persistent")
    for rnd in range(3):
        ids = np.array([rnd, 10 + rnd, 20 + rnd, 47], np.uint32)
        got = {}

        def grab(_, c, ids=ids):
            got[0] = [t.cpu().numpy().tobytes() for t in srv.output(c, ids)]

        recs = run_load(srv, [0.0], [ids], on_done=grab)
        assert recs[0].status == 0
        assert got[0] == [blocks[b] for b in ids]
        time.sleep(0.2)  # > idle_ms: the grid leaves
    t0 = time.perf_counter()
    srv.close()
    assert time.perf_counter() - t0 < 5.0


def test_persistent_ring_backpressure(coder, oracle):
    """A ring far smaller than the offered work: jobs wait for ring slots,
    slots are reused many times, every output is still the original."""
    from fqxv_amd.serve import Server, run_load
    blocks, enc, ol = _archive(coder, oracle)
    nb = len(blocks)
    srv = Server(enc.dst, list(enc.dst_off), list(ol), [len(b) for b in blocks], budget=64 << 20,
                 mode="persistent", ring=16, idle_ms=50)
    rng = np.random.default_rng(8)
    reqs = [np.sort(rng.choice(nb, int(rng.integers(1, 9)), replace=False)).astype(np.uint32) for _ in range(40)]
    seen = {}

    def grab(i, c):
        seen[i] = [t.cpu().numpy().tobytes() for t in srv.output(c, reqs[i])]

    recs = run_load(srv, [0.0] * len(reqs), reqs, on_done=grab)
    for i, r in enumerate(recs):
        assert r.status == 0, i
        assert seen[i] == [blocks[b] for b in reqs[i]], i
    with pytest.raises(ValueError):  # more blocks than the ring holds
        srv.submit(np.arange(17, dtype=np.uint32))
    srv.close()
