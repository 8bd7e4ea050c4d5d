"""GPU: fqxv_rans_decode_batch_noalloc, the batched decode that allocates
nothing (its counter and workspaces live in a fixed device array), called
through the C ABI and checked against fqxv_rans_decode_batch and the input:
two calls in a row, the second on a smaller input, so the second call runs
on scratch the first left behind."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _corpus(seed, nblocks):
    from fqxv_amd import synth
    rng = np.random.default_rng(seed)
    kinds = rng.choice([synth.K_MARKOV, synth.K_DNA, synth.K_NAMES], nblocks)
    sizes = np.exp(rng.uniform(np.log(64), np.log(600_000), nblocks)).astype(np.uint64)
    sizes[0] = 0
    orders = rng.integers(0, 2, nblocks).astype(np.int32)
    return synth.make(kinds, sizes, orders, seed, False, f"noalloc{seed}")


def _call_noalloc(L, src, dst, jobs, order, status, out_len, waves=0):
    import torch
    return L.fqxv_rans_decode_batch_noalloc(
        ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(jobs.data_ptr()),
        len(status), ctypes.c_void_p(order.data_ptr()) if order is not None else None,
        ctypes.c_void_p(status.data_ptr()), ctypes.c_void_p(out_len.data_ptr()), waves,
        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))


def test_noalloc_decode_matches_batched_decode(coder, oracle):
    import torch
    from fqxv_amd import _lib
    L = _lib.load()
    assert L.fqxv_rans_noalloc_max_waves() == 32
    prepared = []
    for seed, nb in ((61, 300), (62, 40)):  # the second input is smaller
        c = _corpus(seed, nb)
        arena = torch.from_numpy(c.data).cuda()
        enc = coder.encode_batch(arena, c.offs, c.sizes, c.orders)
        torch.cuda.synchronize()
        assert int((enc.status != 0).sum()) == 0
        comp = enc.out_len.cpu().numpy().astype(np.uint64)
        ref = coder.decode_batch(enc.dst, enc.dst_off, comp, c.sizes)
        torch.cuda.synchronize()
        assert int((ref.status != 0).sum()) == 0
        jobs_np = coder.jobs_array(enc.dst_off, comp, ref.dst_off, c.sizes)
        jobs = torch.from_numpy(jobs_np.view(np.uint8).copy()).cuda()
        order = torch.from_numpy(coder.lpt_order(c.sizes).view(np.int32)).cuda()
        dst = torch.full_like(ref.dst, 0xEE)
        status = torch.full((nb,), -1, dtype=torch.int32, device="cuda")
        out_len = torch.zeros(nb, dtype=torch.int64, device="cuda")
        prepared.append((c, arena, enc, ref, jobs, order, dst, status, out_len))
    for c, arena, enc, ref, jobs, order, dst, status, out_len in prepared:
        print(f"noalloc: {len(c.sizes)} blocks, {c.total} bytes", flush=True)
        rc = _call_noalloc(L, enc.dst, dst, jobs, order, status, out_len)
        assert rc == 0, _lib.strerror(rc)
        torch.cuda.synchronize()
        assert torch.equal(status, ref.status)
        assert torch.equal(out_len, ref.out_len)
        n = c.total
        assert torch.equal(dst[:n], ref.dst[:n])
        assert torch.equal(dst[:n], arena[:n])
        assert bool((dst[n:] == 0xEE).all()), "wrote past the last block"
        j = int(np.argmax(c.sizes))
        o, m = int(enc.dst_off[j]), int(enc.out_len[j])
        a, k = int(ref.dst_off[j]), int(c.sizes[j])
        assert dst[a:a + k].cpu().numpy().tobytes() == oracle.decode(enc.dst[o:o + m].cpu().numpy().tobytes())
    # Over the fixed scratch: refused before any work, nothing written.
    c, arena, enc, ref, jobs, order, dst, status, out_len = prepared[1]
    status.fill_(-7)
    rc = _call_noalloc(L, enc.dst, dst, jobs, order, status, out_len, waves=33)
    assert rc == _lib.ERR_SCRATCH_LIMIT
    assert _lib.strerror(rc) == "request exceeds the fixed scratch of the no-allocation call"
    torch.cuda.synchronize()
    assert bool((status == -7).all())
