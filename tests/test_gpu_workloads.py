"""GPU: the full-size workloads the benchmark reports, under test.

* The uneven 10,000-block corpus (mixed_10k: 6.02 GB, log-uniform 4 KiB-4 MiB,
  order-0/1 50/50, Markov quality / DNA / names, seed 4) decoded in full
  through the autotuned, captured batched decode into a sentinel-filled
  buffer: every byte must come from the replay, equal the input, and a
  sample of blocks must equal the oracle's decode of the same streams.
* The serving workload: the Markov order-1 corpus (markov_64x4mib) resident
  compressed in HBM, Poisson arrivals of 1-64 block requests (seed 5) under a
  32 GiB output budget: each request's output alone == in the mix == the
  original blocks.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SENTINEL = 0xEE


def _encode_corpus(coder, oracle, corpus, sample):
    import torch
    arena = torch.from_numpy(corpus.data).cuda()
    enc = coder.encode_batch(arena, corpus.offs, corpus.sizes, corpus.orders)
    torch.cuda.synchronize()
    assert int((enc.status != 0).sum()) == 0
    comp_len = enc.out_len.cpu().numpy().astype(np.uint64)
    for j in sample:
        o, n = int(enc.dst_off[j]), int(comp_len[j])
        got = enc.dst[o:o + n].cpu().numpy().tobytes()
        assert got == oracle.encode(corpus.block(j), int(corpus.orders[j])), j
    return arena, enc, comp_len


def test_mixed_10k_full_tuned_graph_into_poisoned_buffer(coder, oracle):
    import torch
    from fqxv_amd import synth, tune
    corpus = synth.workload("mixed_10k")
    nb, n_total = len(corpus.sizes), corpus.total
    assert nb == 10000 and n_total > 6_000_000_000
    o1 = [j for j in range(nb) if corpus.orders[j] == 1]
    big_o1 = max(o1, key=lambda j: corpus.sizes[j])
    sample = sorted({0, nb - 1, int(np.argmax(corpus.sizes)), big_o1, *range(17, nb, 997)})
    arena, enc, comp_len = _encode_corpus(coder, oracle, corpus, sample[:6])
    dst = torch.full((n_total + 64,), SENTINEL, dtype=torch.uint8, device="cuda")
    dst_off = corpus.offs
    jobs = coder.jobs_array(enc.dst_off, comp_len, dst_off, corpus.sizes)
    tuned = tune.autotune(coder, "decode", enc.dst, dst, jobs, corpus.sizes, corpus.orders,
                          policies=("lpt-cost", "lpt-size"), use_cache=False)
    g = coder.graph("decode", enc.dst, dst, tuned.jobs(jobs), tuned.order(coder, corpus.sizes, corpus.orders))
    try:
        for rep in range(2):
            dst.fill_(SENTINEL)
            torch.cuda.synchronize()
            g.launch()
            torch.cuda.synchronize()
            assert int((g.status != 0).sum()) == 0, rep
            assert torch.equal(g.out_len.cpu(), torch.from_numpy(corpus.sizes.astype(np.int64))), rep
            assert torch.equal(dst[:n_total], arena[:n_total]), f"replay {rep} differs from the input"
            assert bool((dst[n_total:] == SENTINEL).all()), "decode wrote past the last block"
    finally:
        g.close()
    for j in sample:
        o, n = int(enc.dst_off[j]), int(comp_len[j])
        ref = oracle.decode(enc.dst[o:o + n].cpu().numpy().tobytes())
        a, m = int(dst_off[j]), int(corpus.sizes[j])
        assert dst[a:a + m].cpu().numpy().tobytes() == ref, j
    coder.set_grid(0)


@pytest.mark.parametrize("mode", ["persistent", "slots"])
def test_serving_markov_corpus_alone_equals_in_mix(coder, oracle, mode):
    import torch
    from fqxv_amd import synth
    from fqxv_amd.serve import Server, poisson_requests, run_load
    corpus = synth.workload("markov_64x4mib")
    nb = len(corpus.sizes)
    blk = int(corpus.sizes[0])
    arena, enc, comp_len = _encode_corpus(coder, oracle, corpus, [0, nb - 1])
    budget = 32 << 30
    # Offered load about half of the measured batched peak (~43 GB/s).
    rate = 0.5 * 43e9 / (32.5 * blk)
    arrivals, reqs = poisson_requests(5, 24, nb, rate, kmax=64)
    srv = Server(enc.dst, enc.dst_off, comp_len, corpus.sizes, budget=budget, slots=4, max_batch=4096,
                 mode=mode, idle_ms=100)
    try:
        alone = {}
        for i in (0, 5, 11, 23):
            def grab(_, c, i=i):
                alone[i] = [t.clone() for t in srv.output(c, reqs[i])]
            r = run_load(srv, [0.0], [reqs[i]], on_done=grab)
            assert r[0].status == 0
        bad = []

        def check(i, c):
            outs = srv.output(c, reqs[i])
            for t, b in zip(outs, reqs[i]):
                o = int(corpus.offs[b])
                if not torch.equal(t, arena[o:o + blk]):
                    bad.append((i, int(b), "original"))
            if i in alone:
                for t, a, b in zip(outs, alone[i], reqs[i]):
                    if not torch.equal(t, a):
                        bad.append((i, int(b), "alone"))

        recs = run_load(srv, arrivals, reqs, on_done=check)
        assert all(r.status == 0 for r in recs)
        assert not bad, bad[:5]
        assert srv.max_in_flight > 1, "requests never overlapped on the device"
    finally:
        srv.close()
