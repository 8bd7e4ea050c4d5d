#!/usr/bin/env python3
"""rANS Nx16 decode throughput on one MI355X: batched vs per-block launch.

Headline workload (``value``): ``mixed_10k`` -- 10,000 blocks, sizes
log-uniform on [4 KiB, 4 MiB] (6.02 GB), order-0/1 at 50/50 per block,
streams split in thirds between Markov quality strings, DNA (ACGT + 0.1% N)
and ':'-separated decimal name fields; seed 4.  The blocks are generated on
the host, copied to HBM and encoded on the device (byte identity against the
oracle on a sample of blocks, device round trip on all of them).  One step =
one replay of a captured hipGraph holding the batched decode of the whole
corpus (one persistent-kernel launch over all blocks from a device work
queue, grid and queue order chosen by the autotuner), inputs resident in HBM;
GB/s = 1e9 bytes of uncompressed output per second.

Also reported on the same JSON line:

* ``perblock``: the plain form -- one kernel launch per block, back to back
  on one stream -- on an every-k-th-block sample, with per-launch p50/p99.
* ``configs``: 1 MiB binned order-0 (encode+decode round trip); 64 x 1 MiB
  normal-quality order-0 and 64 x 4 MiB Markov order-1, each in the plain
  per-block form (encode and decode) and batched, with the per-round
  critical-path time of one block.
* ``serving``: the Markov order-1 corpus resident compressed in HBM, Poisson
  arrivals of 1-64 block requests (seed 5) swept from 10% to 120% of the
  batched peak under a 32 GiB output budget: p50/p99 latency and GB/s, for
  the persistent engine (one resident grid fed block by block) and the
  launch-per-admission engine (4 batched launches in flight).
* ``roofline``: HBM-bound; algorithmic bytes = compressed bytes read +
  uncompressed bytes written, over the decode kernel's measured duration;
  ``traffic`` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE runs.
* ``cpu_baseline``: the oracle (plain-C scalar restatement of fqxv-rans)
  decoding a sample of the same streams on the host's cores, all cores
  (block-parallel, like the reference's rayon dispatch) and one thread.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "rANS Nx16 o0/o1 decode GB/s, batched vs per-block launch; p99 block latency"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_k_rans_decode_mixed_10k.json")
SECTIONS = ("c4", "configs", "serve", "cpu")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def events(n):
    import torch
    return [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(n)]


def slices_to_host(t, offs, lens):
    """Copy device byte ranges to one compact host array (+ new offsets)."""
    lens = np.asarray(lens, dtype=np.uint64)
    o2 = np.zeros(len(lens), dtype=np.uint64)
    if len(lens):
        o2[1:] = np.cumsum(lens)[:-1]
    out = np.empty(int(lens.sum()) + 64, dtype=np.uint8)
    for a, b, n in zip(offs, o2, lens):
        out[int(b):int(b) + int(n)] = t[int(a):int(a) + int(n)].cpu().numpy()
    return out, o2


def encode_corpus(coder, corpus, dev, reps=2):
    """Batched device encode; returns (BatchResult, comp_len, best seconds)."""
    import torch
    arena = torch.from_numpy(corpus.data).to(dev)
    stream = torch.cuda.current_stream()
    enc = coder.encode_batch(arena, corpus.offs, corpus.sizes, corpus.orders)
    torch.cuda.synchronize()
    assert int((enc.status != 0).sum()) == 0, f"{corpus.name}: device encode failed"
    times = []
    for _ in range(reps):
        (a, b), = events(1)
        a.record(stream)
        enc = coder.encode_batch(arena, corpus.offs, corpus.sizes, corpus.orders, stream=stream)
        b.record(stream)
        torch.cuda.synchronize()
        times.append(a.elapsed_time(b) / 1e3)
    return arena, enc, enc.out_len.cpu().numpy().astype(np.uint64), min(times)


def check_sample(oracle, corpus, enc, comp_len, sample):
    for j in sample:
        o, n = int(enc.dst_off[j]), int(comp_len[j])
        got = enc.dst[o:o + n].cpu().numpy().tobytes()
        ref = oracle.encode(corpus.block(j), int(corpus.orders[j]))
        assert got == ref, f"{corpus.name} block {j}: device encode differs from the oracle"


def perblock(coder, src, src_off, src_len, dst, dst_off, caps, idx, decode=True, orders=None):
    """The plain form: one launch per block, back to back on one stream."""
    import torch
    from fqxv_amd import _lib
    stream = torch.cuda.current_stream()
    st1 = torch.zeros(1, dtype=torch.int32, device=src.device)
    ol1 = torch.zeros(1, dtype=torch.int64, device=src.device)
    jobs = [_lib.Job(int(src_off[j]), int(src_len[j]), int(dst_off[j]), int(caps[j]), 2**64 - 1,
                     0 if orders is None else int(orders[j]), 0) for j in idx]
    fn = coder.decode_one_device if decode else coder.encode_one_device
    for jb in jobs[:2]:
        fn(src, dst, jb, st1, ol1, stream=stream)
    torch.cuda.synchronize()
    ev = events(len(jobs))
    t0 = time.perf_counter()
    for (a, b), jb in zip(ev, jobs):
        a.record(stream)
        fn(src, dst, jb, st1, ol1, stream=stream)
        b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    assert int(st1.item()) == 0
    lat = np.array([a.elapsed_time(b) for a, b in ev])
    return wall, lat


def perblock_streams(src, src_off, src_len, dst, dst_off, caps, idx, nstreams=4):
    """The plain form spread over `nstreams` streams (one context each: a
    context's single-block calls share its workspace), launches dealt round
    robin, so the per-block form is not held to one queue."""
    import torch
    import fqxv_amd
    from fqxv_amd import _lib
    coders = [fqxv_amd.Coder(0) for _ in range(nstreams)]
    streams = [torch.cuda.Stream() for _ in range(nstreams)]
    st = torch.zeros(nstreams, dtype=torch.int32, device=src.device)
    ol = torch.zeros(nstreams, dtype=torch.int64, device=src.device)
    jobs = [_lib.Job(int(src_off[j]), int(src_len[j]), int(dst_off[j]), int(caps[j]), 2**64 - 1, 0, 0) for j in idx]
    for k, jb in enumerate(jobs[:nstreams]):
        coders[k].decode_one_device(src, dst, jb, st[k:k + 1], ol[k:k + 1], stream=streams[k])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i, jb in enumerate(jobs):
        k = i % nstreams
        coders[k].decode_one_device(src, dst, jb, st[k:k + 1], ol[k:k + 1], stream=streams[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    assert int((st != 0).sum()) == 0
    for c in coders:
        c.close()
    return wall


def bench_percall(coder, oracle):
    """The drop-in single call (host buffers, PCIe included), as the
    reference's callers make it: fixture-sized streams (names sub-streams of
    n <= 1,300 bytes, crates/fqxv-tokenizer/src/lib.rs:98-110) and the 1 MiB
    binned-quality block, per call mean over repeated calls."""
    from fqxv_amd import synth
    res = {}
    cases = {"names_1000b_o0": (synth.fill(synth.K_NAMES, 1000, 11).tobytes(), 0),
             "names_1000b_o1": (synth.fill(synth.K_NAMES, 1000, 11).tobytes(), 1),
             "binned_1mib_o0": (synth.workload("binned_1mib").block(0).tobytes(), 0)}
    for name, (data, order) in cases.items():
        reps = 200 if len(data) < 4096 else 10
        enc = coder.encode(data, order)
        assert enc == oracle.encode(data, order)
        assert coder.decode(enc) == data
        t0 = time.perf_counter()
        for _ in range(reps):
            coder.encode(data, order)
        te = (time.perf_counter() - t0) / reps
        t0 = time.perf_counter()
        for _ in range(reps):
            coder.decode(enc)
        td = (time.perf_counter() - t0) / reps
        res[name] = {"n": len(data), "encode_us": round(te * 1e6, 1), "decode_us": round(td * 1e6, 1),
                     "decode_mbs": round(len(data) / td / 1e6, 2)}
    log("[percall] " + ", ".join(f"{k}: enc {v['encode_us']} us / dec {v['decode_us']} us" for k, v in res.items()))
    return {"form": "host-buffer fqxv_rans_encode / fqxv_rans_decode, one call per stream, PCIe included",
            "reference_context": "fqxv-rans AVX-512 order-0 decode ~817 MiB/s on one Cascade Lake core "
                                 "(crates/fqxv-rans/src/lib.rs:176-178; published, other hardware)",
            **res}


def batched_decode_time(coder, enc, comp_len, sizes, dst, reps=3):
    import torch
    stream = torch.cuda.current_stream()
    dst_off = np.zeros_like(sizes)
    dst_off[1:] = np.cumsum(sizes)[:-1]
    jobs = coder.jobs_array(enc.dst_off, comp_len, dst_off, sizes)
    g = coder.graph("decode", enc.dst, dst, jobs, coder.lpt_order(sizes))
    g.launch(stream)
    ev = events(reps)
    for a, b in ev:
        a.record(stream)
        g.launch(stream)
        b.record(stream)
    torch.cuda.synchronize()
    assert int((g.status != 0).sum()) == 0
    g.close()
    return min(a.elapsed_time(b) for a, b in ev) / 1e3


# --------------------------------------------------------------------------- C4
def bench_c4(args, coder, dev, oracle):
    import torch
    from fqxv_amd import synth, tune
    t0 = time.time()
    corpus = synth.workload(args.workload, blocks=args.blocks or None)
    nb, n_total = len(corpus.sizes), corpus.total
    log(f"[c4] generated {args.workload}: {nb} blocks, {n_total / 1e9:.3f} GB in {time.time() - t0:.1f}s")
    arena, enc, comp_len, t_enc = encode_corpus(coder, corpus, dev)
    c_total = int(comp_len.sum())
    log(f"[c4] batched encode {n_total / t_enc / 1e9:.1f} GB/s, ratio {n_total / c_total:.3f}")

    rng = np.random.default_rng(0)
    sample = sorted(set(rng.choice(nb, min(24, nb), replace=False).tolist()) | {int(np.argmax(corpus.sizes))})
    check_sample(oracle, corpus, enc, comp_len, sample)
    enc_tuned = bench_encode_tuned(coder, corpus, arena, enc, comp_len, args)
    best = bench_encode_best(coder, corpus, arena, oracle, sample[:6]) if args.encode_best else None
    dec = coder.decode_batch(enc.dst, enc.dst_off, comp_len, corpus.sizes)
    torch.cuda.synchronize()
    assert int((dec.status != 0).sum()) == 0, "device decode failed"
    assert torch.equal(dec.dst[:n_total], arena[:n_total]), "device round trip differs"
    log(f"[c4] parity ok ({len(sample)} blocks vs oracle, all {nb} round-tripped)")
    eager_ms = batched_decode_eager_ms(coder, enc, comp_len, corpus.sizes)
    crc = bench_crc(coder, oracle, enc, comp_len, args.workload) if args.crc else None

    out = dec.dst
    jobs_np = coder.jobs_array(enc.dst_off, comp_len, dec.dst_off, corpus.sizes)
    if args.grid:
        coder.set_grid(args.grid)
        tuned = tune.Tuned(args.grid, args.policy, float("nan"), [])
    else:
        tuned = tune.autotune(coder, "decode", enc.dst, out, jobs_np, corpus.sizes, corpus.orders,
                              use_cache=False, log=log)
    log(f"[c4] tuned grid={tuned.grid} policy={tuned.policy} staging={tuned.staging}")
    g = coder.graph("decode", enc.dst, out, tuned.jobs(jobs_np), tuned.order(coder, corpus.sizes, corpus.orders))
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        g.launch(stream)
    torch.cuda.synchronize()
    ev = events(args.steps)
    # The output region is filled with a sentinel before the timed replays,
    # so the comparison with the input afterwards proves the timed replays
    # wrote every byte (the eager decode and the tuner wrote it before).
    out.fill_(0xEE)
    g.status.fill_(-1)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for a, b in ev:
        a.record(stream)
        g.launch(stream)
        b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    kern = [a.elapsed_time(b) / 1e3 for a, b in ev]
    assert int((g.status != 0).sum()) == 0
    assert torch.equal(out[:n_total], arena[:n_total]), "timed replays did not reproduce the input"
    t_step = wall / args.steps
    value = n_total / t_step / 1e9
    t_kern = float(np.mean(kern))
    log(f"[c4] graph-replayed batched decode {value:.1f} GB/s ({t_step * 1e3:.2f} ms/step, "
        f"kernel {t_kern * 1e3:.2f} ms)")

    # Latency floor: one wave decodes a block as ceil(n/32) dependent rounds,
    # so the largest order-1 block bounds the makespan.
    big1 = [j for j in range(nb) if corpus.orders[j] == 1]
    jbig = max(big1, key=lambda j: corpus.sizes[j]) if big1 else int(np.argmax(corpus.sizes))
    _, lat_big = perblock(coder, enc.dst, enc.dst_off, comp_len, out, dec.dst_off, corpus.sizes, [jbig])
    # Per-round time of one block alone, per stream kind and order (the
    # largest block of each class).
    kind_rounds = {}
    for kd in sorted(set(int(k) for k in corpus.kinds)):
        for od in (0, 1):
            cand = np.nonzero((corpus.kinds == kd) & (corpus.orders == od))[0]
            if len(cand) == 0:
                continue
            j = int(cand[np.argmax(corpus.sizes[cand])])
            _, lat = perblock(coder, enc.dst, enc.dst_off, comp_len, out, dec.dst_off, corpus.sizes, [j])
            rounds = (int(corpus.sizes[j]) + 31) // 32
            kind_rounds[f"{synth.KIND_NAMES[kd]}_o{od}"] = {
                "block": j, "n": int(corpus.sizes[j]), "ms": round(float(lat[0]), 3),
                "ns_per_round": round(float(lat[0]) * 1e6 / rounds, 1)}
    log("[c4] ns/round alone: " + ", ".join(f"{k} {v['ns_per_round']}" for k, v in kind_rounds.items()))

    pb = None
    if not args.no_perblock:
        idx = list(range(0, nb, args.perblock_every))
        wall_pb, lat = perblock(coder, enc.dst, enc.dst_off, comp_len, out, dec.dst_off, corpus.sizes, idx)
        sb = int(corpus.sizes[idx].sum())
        pb_gbs = sb / wall_pb / 1e9
        pb = {"value": round(pb_gbs, 4), "unit": "GB/s",
              "form": "one kernel launch per block, back to back on one stream",
              "sample": f"every {args.perblock_every}th block ({len(idx)} launches, {sb / 1e9:.3f} GB)",
              "speedup_batched": round(value / pb_gbs, 1),
              "p50_block_ms": round(float(np.percentile(lat, 50)), 3),
              "p99_block_ms": round(float(np.percentile(lat, 99)), 3)}
        log(f"[c4] per-block {pb_gbs:.3f} GB/s, batched speedup {value / pb_gbs:.1f}x, "
            f"p50/p99 {pb['p50_block_ms']}/{pb['p99_block_ms']} ms")
        wall4 = perblock_streams(enc.dst, enc.dst_off, comp_len, out, dec.dst_off, corpus.sizes, idx)
        pb["four_streams"] = {"value": round(sb / wall4 / 1e9, 4), "unit": "GB/s",
                              "form": "one launch per block, dealt round robin over 4 streams (4 contexts)",
                              "speedup_batched": round(value / (sb / wall4 / 1e9), 1)}
        log(f"[c4] per-block on 4 streams {sb / wall4 / 1e9:.3f} GB/s")
        torch.cuda.synchronize()
        assert torch.equal(out[:n_total], arena[:n_total])
    xxh = bench_xxh3(coder, arena, corpus.offs, corpus.sizes, args.workload) if args.crc else None
    # Fingerprint of the generated corpus (one XXH3-64 over all of it, on the
    # device), so a result can be tied to the exact input bytes.
    fp = "%016x" % (int(coder.xxh3_batch(arena, [0], [n_total]).cpu().numpy().view(np.uint64)[0]))
    pcie = bench_pcie(coder, enc, comp_len, corpus, arena) if args.pcie else None

    # CPU-baseline sample (host copies of a subset of the streams).
    cpu_idx = np.arange(0, nb, args.cpu_every)
    host_c, host_off = slices_to_host(enc.dst, enc.dst_off[cpu_idx], comp_len[cpu_idx])
    res = {
        "crc32c": crc,
        "xxh3": xxh,
        "fingerprint": fp,
        "pcie": pcie,
        "encode_best": best,
        "corpus": corpus, "value": value, "t_step": t_step, "t_kern": t_kern, "n_total": n_total,
        "c_total": c_total, "nb": nb, "encode_gbs": n_total / t_enc / 1e9, "encode_tuned": enc_tuned, "perblock": pb,
        "tuned": tuned, "eager_ms": eager_ms, "graph_kernel_ms": t_kern * 1e3,
        "kind_rounds": kind_rounds,
        "largest_o1": {"block": int(jbig), "n": int(corpus.sizes[jbig]), "ms": round(float(lat_big[0]), 3),
                       "rounds": int((int(corpus.sizes[jbig]) + 31) // 32),
                       "ns_per_round": round(float(lat_big[0]) * 1e6 / ((int(corpus.sizes[jbig]) + 31) // 32),
                                             1)},
        "cpu_sample": (host_c, host_off, comp_len[cpu_idx], corpus.sizes[cpu_idx], corpus.orders[cpu_idx]),
    }
    g.close()
    del arena, enc, dec, out
    torch.cuda.empty_cache()
    return res


def bench_encode_tuned(coder, corpus, arena, enc, comp_len, args):
    """The batched encode as the decode is measured: the autotuner picks the
    persistent grid and queue order by timing graph replays, then the tuned
    graph is replayed with HIP events around each launch.  Every block's
    length, and the bytes of a sample of blocks, equal the eager call's
    (whose sample was checked against the oracle)."""
    import torch
    from fqxv_amd import tune
    sizes, orders = corpus.sizes, corpus.orders
    caps = coder.encode_bounds(sizes, orders)
    d_off = np.zeros_like(caps)
    d_off[1:] = np.cumsum(caps)[:-1]
    jobs = coder.jobs_array(corpus.offs, sizes, d_off, caps, order=orders)
    dst = torch.empty(int(caps.sum()) + 64, dtype=torch.uint8, device=arena.device)
    top = coder.max_waves()
    grids = sorted({min(top, g) for g in (2048, 1792, 1536, 1280)}, reverse=True)
    t = tune.autotune(coder, "encode", arena, dst, jobs, sizes, orders, grids=grids,
                      policies=("lpt-cost", "lpt-size"), use_cache=False, log=log)
    g = coder.graph("encode", arena, dst, t.jobs(jobs), t.order(coder, sizes, orders, "encode"))
    stream = torch.cuda.current_stream()
    g.launch(stream)
    torch.cuda.synchronize()
    ev = events(3)
    dst.fill_(0xEE)
    for a, b in ev:
        a.record(stream)
        g.launch(stream)
        b.record(stream)
    torch.cuda.synchronize()
    ms = min(a.elapsed_time(b) for a, b in ev)
    assert int((g.status != 0).sum()) == 0
    ol = g.out_len.cpu().numpy().astype(np.uint64)
    assert np.array_equal(ol, comp_len), "tuned encode lengths differ from the eager call's"
    rng = np.random.default_rng(7)
    for j in sorted(set(rng.choice(len(sizes), 16, replace=False).tolist()) | {int(np.argmax(sizes))}):
        a0, b0, n = int(d_off[j]), int(enc.dst_off[j]), int(ol[j])
        assert torch.equal(dst[a0:a0 + n], enc.dst[b0:b0 + n]), f"tuned encode block {j} differs"
    g.close()
    del dst
    coder.set_grid(0)
    gbs = corpus.total / (ms / 1e3) / 1e9
    log(f"[c4] tuned batched encode {gbs:.1f} GB/s ({ms:.2f} ms, grid {t.grid}, {t.policy})")
    return {"gbs": gbs, "ms": ms, "grid": t.grid, "policy": t.policy, "table": t.table}


def bench_encode_best(coder, corpus, arena, oracle, sample):
    """Both orders per block, the shorter kept (ties to order-0, as the
    tokenizer does): batched GB/s, and how many blocks order-1 wins."""
    import torch
    stream = torch.cuda.current_stream()
    r = coder.encode_best_batch(arena, corpus.offs, corpus.sizes, 0, stream=stream)
    torch.cuda.synchronize()
    assert int((r.status != 0).sum()) == 0
    times = []
    for _ in range(2):
        # The previous result is freed first, so the call reuses its cached
        # arena instead of allocating a new one inside the timed region.
        del r
        (a, b), = events(1)
        a.record(stream)
        r = coder.encode_best_batch(arena, corpus.offs, corpus.sizes, 0, stream=stream)
        b.record(stream)
        torch.cuda.synchronize()
        times.append(a.elapsed_time(b) / 1e3)
    t = min(times)
    ol = r.out_len.cpu().numpy()
    for j in sample:
        o = int(r.dst_off[j])
        assert r.dst[o:o + int(ol[j])].cpu().numpy().tobytes() == oracle.encode_best(corpus.block(j), 0), j
    tags = r.dst[torch.from_numpy(r.dst_off.astype(np.int64)).to(r.dst.device)].cpu().numpy()
    n = corpus.total
    res = {"value": round(n / t / 1e9, 3), "unit": "GB/s", "ties": "order-0",
           "compressed_bytes": int(ol.sum()), "order1_chosen": int(tags.sum()),
           "checked_vs_oracle": len(sample)}
    log(f"[c4] batched encode-best {res['value']} GB/s, order-1 kept for {res['order1_chosen']} of {len(ol)} blocks")
    del r
    torch.cuda.empty_cache()
    return res


def batched_decode_eager_ms(coder, enc, comp_len, sizes, reps=2):
    """The batched call launched eagerly (no graph), for comparison."""
    import torch
    stream = torch.cuda.current_stream()
    d = coder.decode_batch(enc.dst, enc.dst_off, comp_len, sizes, stream=stream)
    ev = events(reps)
    for a, b in ev:
        a.record(stream)
        d = coder.decode_batch(enc.dst, enc.dst_off, comp_len, sizes, stream=stream, dst=d.dst)
        b.record(stream)
    torch.cuda.synchronize()
    assert int((d.status != 0).sum()) == 0
    ms = min(a.elapsed_time(b) for a, b in ev)
    del d
    return round(ms, 3)


# ------------------------------------------------------------------ C1, C2, C3
def bench_config(coder, dev, oracle, name, check):
    import torch
    from fqxv_amd import synth
    corpus = synth.workload(name)
    nb, n_total = len(corpus.sizes), corpus.total
    arena, enc, comp_len, t_enc_b = encode_corpus(coder, corpus, dev)
    check_sample(oracle, corpus, enc, comp_len, check)
    c_total = int(comp_len.sum())
    idx = list(range(nb))
    # plain form: encode then decode, one launch per block
    caps = coder.encode_bounds(corpus.sizes, corpus.orders)
    e_off = np.zeros_like(caps)
    e_off[1:] = np.cumsum(caps)[:-1]
    edst = torch.empty(int(caps.sum()) + 64, dtype=torch.uint8, device=dev)
    w_enc, lat_enc = perblock(coder, arena, corpus.offs, corpus.sizes, edst, e_off, caps, idx, decode=False,
                              orders=corpus.orders)
    for j in check[:2]:
        o = int(e_off[j])
        assert edst[o:o + int(comp_len[j])].cpu().numpy().tobytes() == oracle.encode(
            corpus.block(j), int(corpus.orders[j]))
    del edst
    dst = torch.empty(n_total + 64, dtype=torch.uint8, device=dev)
    d_off = corpus.offs
    w_dec, lat_dec = perblock(coder, enc.dst, enc.dst_off, comp_len, dst, d_off, corpus.sizes, idx)
    torch.cuda.synchronize()
    assert torch.equal(dst[:n_total], arena[:n_total]), f"{name}: per-block round trip differs"
    dst.zero_()
    t_b = batched_decode_time(coder, enc, comp_len, corpus.sizes, dst)
    assert torch.equal(dst[:n_total], arena[:n_total]), f"{name}: batched round trip differs"
    n0 = int(corpus.sizes[0])
    rounds = (n0 + 31) // 32
    r = {"blocks": nb, "uncompressed_bytes": n_total, "compressed_bytes": c_total,
         "order": int(corpus.orders[0]), "bits_per_symbol": round(8 * c_total / n_total, 3),
         "perblock_encode_gbs": round(n_total / w_enc / 1e9, 3),
         "perblock_decode_gbs": round(n_total / w_dec / 1e9, 3),
         "batched_encode_gbs": round(n_total / t_enc_b / 1e9, 3),
         "batched_decode_gbs": round(n_total / t_b / 1e9, 3),
         "block_decode_ms_p50": round(float(np.percentile(lat_dec, 50)), 3),
         "block_decode_ms_p99": round(float(np.percentile(lat_dec, 99)), 3),
         "block_encode_ms_p50": round(float(np.percentile(lat_enc, 50)), 3),
         "decode_ns_per_round": round(float(np.median(lat_dec)) * 1e6 / rounds, 1),
         "rounds_per_block": rounds,
         "decode_roofline_frac": round((n_total + c_total) / t_b / 1e9 / HBM_PEAK_GBS, 5)}
    log(f"[{name}] " + ", ".join(f"{k}={v}" for k, v in r.items()))
    state = (corpus, arena, enc, comp_len)
    return r, state


# ------------------------------------------------------------ permutation planes
def bench_perm(coder, dev, oracle, n=1 << 26):
    """reorder.rs:597-607 / 1089-1100 transposes on the device: u32 indices
    <-> four byte planes.  HBM-bound: 8n algorithmic bytes per launch."""
    import torch
    from fqxv_amd import perm
    stream = torch.cuda.current_stream()
    p = torch.randint(-2**31, 2**31 - 1, (n,), dtype=torch.int32, device=dev)
    planes = perm.split(p, coder=coder)
    back = perm.join(planes, n, coder=coder)
    torch.cuda.synchronize()
    assert torch.equal(back, p)
    k = 4096
    assert np.array_equal(planes[:4 * n].view(4, n)[:, :k].cpu().numpy().T.reshape(-1).view("<u4"),
                          p[:k].cpu().numpy().view(np.uint32))
    res = {}
    for name, fn in (("split", lambda: perm.split(p, planes, coder=coder)),
                     ("join", lambda: perm.join(planes, n, back, coder=coder))):
        fn()
        ev = events(5)
        for a, b in ev:
            a.record(stream)
            fn()
            b.record(stream)
        torch.cuda.synchronize()
        t = min(a.elapsed_time(b) for a, b in ev) / 1e3
        res[name] = {"gbs": round(8 * n / t / 1e9, 1), "ms": round(t * 1e3, 3),
                     "roofline_frac": round(8 * n / t / 1e9 / HBM_PEAK_GBS, 3)}
    assert torch.equal(back, p)
    log(f"[perm] split {res['split']['gbs']} GB/s, join {res['join']['gbs']} GB/s (n = {n})")
    del p, planes, back
    torch.cuda.empty_cache()
    return {"indices": n, "bytes_per_launch": 8 * n, **res}


# ------------------------------------------------------------------------ CRC
def bench_crc(coder, oracle, enc, comp_len, corpus_name):
    """Container checksums (crates/fqxv/src/crc.rs) on the device: CRC-32C of
    every compressed block (one range each) and of the whole compressed
    arena as one range.  HBM-bound: algorithmic bytes = bytes read."""
    import torch
    stream = torch.cuda.current_stream()
    offs = np.asarray(enc.dst_off, dtype=np.uint64)
    lens = np.asarray(comp_len, dtype=np.uint64)
    got = coder.crc32c_batch(enc.dst, offs, lens)
    torch.cuda.synchronize()
    idx = list(range(0, len(offs), max(1, len(offs) // 16)))
    for j in idx:
        o, n = int(offs[j]), int(lens[j])
        ref = oracle.crc32c(enc.dst[o:o + n].cpu().numpy())
        assert int(got[j]) & 0xFFFFFFFF == ref, j
    res = {}
    total = int(lens.sum())
    span = int(offs[-1] + lens[-1])
    for name, (o, l, nbytes) in {"per_block": (offs, lens, total),
                                 "whole_arena": (np.zeros(1, np.uint64), np.array([span], np.uint64), span)}.items():
        coder.crc32c_batch(enc.dst, o, l)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            coder.crc32c_batch(enc.dst, o, l)
        t = (time.perf_counter() - t0) / reps
        res[name] = {"ranges": len(o), "bytes": nbytes, "gbs": round(nbytes / t / 1e9, 1),
                     "roofline_frac": round(nbytes / t / 1e9 / HBM_PEAK_GBS, 3)}
    log(f"[crc] per-block {res['per_block']['gbs']} GB/s, whole arena {res['whole_arena']['gbs']} GB/s")
    return {"corpus": f"compressed blocks of {corpus_name}", "checked_vs_oracle": len(idx),
            "timing": "host wall clock per call (the call includes its host-side planning and returns "
                      "when the checksums are written)", **res}


def bench_pcie(coder, enc, comp_len, corpus, arena):
    """The same decode from host buffers: the compressed streams in pinned
    host memory are copied to HBM, decoded in one batched launch, and the
    output copied back to pinned host memory, each phase timed with events
    on one stream.  Not the headline (inputs there are resident in HBM)."""
    import torch
    n_total = corpus.total
    # The streams packed back to back (the encode regions carry slack).
    lens = np.asarray(comp_len, dtype=np.uint64)
    offs = np.zeros(len(lens), dtype=np.uint64)
    offs[1:] = np.cumsum(lens)[:-1]
    span = int(lens.sum())
    packed = torch.cat([enc.dst[int(o):int(o) + int(n)] for o, n in zip(enc.dst_off, lens)])
    h_in = torch.empty(span, dtype=torch.uint8, pin_memory=True)
    h_in.copy_(packed)
    del packed
    h_out = torch.empty(n_total, dtype=torch.uint8, pin_memory=True)
    d_in = torch.empty(span + 64, dtype=torch.uint8, device=arena.device)
    stream = torch.cuda.current_stream()
    best = None
    for _ in range(2):
        e = events(3)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e[0][0].record(stream)
        d_in[:span].copy_(h_in, non_blocking=True)
        e[0][1].record(stream)
        e[1][0].record(stream)
        d = coder.decode_batch(d_in, offs, lens, corpus.sizes, stream=stream)
        e[1][1].record(stream)
        e[2][0].record(stream)
        h_out.copy_(d.dst[:n_total], non_blocking=True)
        e[2][1].record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ms = [a.elapsed_time(b) for a, b in e]
        if best is None or wall < best[0]:
            best = (wall, ms)
        del d
    k = 1 << 20
    assert torch.equal(h_out[:k], arena[:k].cpu()) and torch.equal(h_out[-k:], arena[n_total - k:n_total].cpu())
    wall, ms = best
    res = {"gbs": round(n_total / wall / 1e9, 2), "ms": round(wall * 1e3, 1), "h2d_ms": round(ms[0], 1),
           "decode_ms": round(ms[1], 1), "d2h_ms": round(ms[2], 1),
           "h2d_gbs": round(span / ms[0] / 1e6, 1), "d2h_gbs": round(n_total / ms[2] / 1e6, 1),
           "form": "pinned host -> HBM, batched decode (eager, LPT order), HBM -> pinned host, one stream"}
    log(f"[pcie] host-to-host decode {res['gbs']} GB/s ({res['h2d_ms']} + {res['decode_ms']} + {res['d2h_ms']} ms)")
    del h_in, h_out, d_in
    torch.cuda.empty_cache()
    return res


def bench_xxh3(coder, arena, offs, lens, corpus_name):
    """Content digests (XXH3-64, block.rs:164-176) on the device over the
    decoded blocks, one range each, and over the whole arena as one range.
    The block-sum kernel is HBM-bound (1 KiB read per 64 B of lane sums);
    timing is host wall clock per call, which includes the host-side plan
    and returns when the digests are written."""
    import torch
    import xxhash
    offs = np.asarray(offs, dtype=np.uint64)
    lens = np.asarray(lens, dtype=np.uint64)
    got = coder.xxh3_batch(arena, offs, lens).cpu().numpy().view(np.uint64)
    idx = list(range(0, len(offs), max(1, len(offs) // 8)))
    for j in idx:
        o, n = int(offs[j]), int(lens[j])
        assert int(got[j]) == xxhash.xxh3_64_intdigest(arena[o:o + n].cpu().numpy().tobytes()), j
    res = {}
    total = int(lens.sum())
    # One range the size of a block's largest decoded sequence stream
    # (MAX_BLOCK_SEQ_BYTES = 256 MiB, crates/fqxv/src/container/format.rs:11):
    # the digest's scramble chain is serial over 1 KiB blocks, so a single
    # range is latency-bound.
    one = min(256 << 20, int(offs[-1] + lens[-1]))
    for name, (o, l, nbytes) in {"per_block": (offs, lens, total),
                                 "one_256mib_range": (np.zeros(1, np.uint64), np.array([one], np.uint64), one)}.items():
        coder.xxh3_batch(arena, o, l)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            coder.xxh3_batch(arena, o, l)
        t = (time.perf_counter() - t0) / reps
        res[name] = {"ranges": len(o), "bytes": nbytes, "gbs": round(nbytes / t / 1e9, 1),
                     "roofline_frac": round(nbytes / t / 1e9 / HBM_PEAK_GBS, 3)}
    log(f"[xxh3] per-block {res['per_block']['gbs']} GB/s, one 256 MiB range {res['one_256mib_range']['gbs']} GB/s")
    return {"corpus": f"decoded blocks of {corpus_name}", "checked_vs_xxhash": len(idx),
            "timing": "host wall clock per call", **res}


# ---------------------------------------------------------------------- names
def bench_names(coder, blocks=32, per_block=40000, distinct=8):
    """Read-name streams (crates/fqxv-tokenizer): ``blocks`` blocks of
    ``per_block`` names (``distinct`` generated lists -- paired Illumina,
    Illumina, SRA -- repeated), encoded and decoded through the batched
    device path from host buffers (framing and tokenizing on the host, every
    sub-stream of every block in one device launch), then the names digests
    on the device.  Wall clock per call, PCIe transfers included."""
    from fqxv_amd import names, synth
    kinds = ("paired", "illumina", "sra")
    lists = [synth.make_names(kinds[i % 3], per_block, 500 + i) for i in range(distinct)]
    recs = [names.records(lists[i % distinct]) for i in range(blocks)]
    raw = sum(len(r) for r in recs)
    n_names = blocks * per_block
    streams = names.encode_batch_records(recs, coder)
    out = names.decode_batch_records(streams, coder)
    assert out == recs
    comp = sum(len(s) for s in streams)
    res = {"blocks": blocks, "names": n_names, "record_bytes": raw, "stream_bytes": comp,
           "ratio": round(raw / comp, 2)}
    for name, fn in (("encode", lambda: names.encode_batch_records(recs, coder)),
                     ("decode", lambda: names.decode_batch_records(streams, coder)),
                     ("decode_and_digest", lambda: names.decode_batch_records(streams, coder, digests=True))):
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        t = min(ts)
        res[name] = {"ms": round(t * 1e3, 2), "mnames_s": round(n_names / t / 1e6, 2),
                     "record_mbs": round(raw / t / 1e6, 1)}
    log(f"[names] {n_names} names in {blocks} blocks: encode {res['encode']['mnames_s']} M names/s, "
        f"decode {res['decode']['mnames_s']} M names/s, +digest {res['decode_and_digest']['mnames_s']} M names/s")
    return res


# ------------------------------------------------------------------------ C5
def serve_peak(dev, enc, comp_len, corpus, arena, waves, peak_blocks):
    """Batched decode rate of this corpus on `waves` decode waves: one graph
    launch over copies of every block, on a context of its own."""
    import torch
    import fqxv_amd
    nb = len(corpus.sizes)
    blk = int(corpus.sizes[0])
    reps = max(1, peak_blocks // nb)
    k = nb * reps
    pc = fqxv_amd.Coder(dev.index or 0)
    pc.set_grid(waves)
    d_off = np.arange(k, dtype=np.uint64) * np.uint64(blk)
    dst = torch.empty(k * blk + 64, dtype=torch.uint8, device=dev)
    jobs = pc.jobs_array(np.tile(enc.dst_off, reps), np.tile(comp_len, reps), d_off, np.tile(corpus.sizes, reps))
    g = pc.graph("decode", enc.dst, dst, jobs, None)
    g.launch()
    stream = torch.cuda.current_stream()
    ev = events(2)
    for a, b in ev:
        a.record(stream)
        g.launch(stream)
        b.record(stream)
    torch.cuda.synchronize()
    assert int((g.status != 0).sum()) == 0
    assert torch.equal(dst[(k - 1) * blk:k * blk], arena[(nb - 1) * blk:nb * blk])
    t = min(a.elapsed_time(b) for a, b in ev) / 1e3
    g.close()
    del dst, pc
    torch.cuda.empty_cache()
    return k * blk / t, k, t


def bench_serve(args, coder, dev, state):
    import torch
    from fqxv_amd.serve import Server, poisson_requests, run_load, summarize
    corpus, arena, enc, comp_len = state
    nb = len(corpus.sizes)
    blk = int(corpus.sizes[0])
    budget = args.serve_budget_gib << 30
    mean_req = 32.5 * blk
    bad = []
    modes, peaks, single = {}, {}, {}
    for mode in args.serve_modes:
        srv = Server(enc.dst, enc.dst_off, comp_len, corpus.sizes, budget=budget, slots=args.serve_slots,
                     max_batch=4096, mode=mode, waves=args.serve_waves)
        waves = srv.waves
        # The offered-load axis is relative to a batched decode on the same
        # number of decode waves as this engine runs.
        peak, k, t_peak = serve_peak(dev, enc, comp_len, corpus, arena, waves, args.peak_blocks)
        log(f"[serve:{mode}] {waves} decode waves; batched peak on them {peak / 1e9:.2f} GB/s "
            f"({k} blocks in {t_peak * 1e3:.1f} ms)")
        # One-block requests on an idle engine: the latency floor of a request.
        one = []
        for b in range(min(8, nb)):
            recs = run_load(srv, [0.0], [np.array([b], np.uint32)])
            one.append(recs[0].latency * 1e3)
        single[mode] = round(float(np.median(one)), 3)
        peaks[mode] = {"waves": waves, "gbs": round(peak / 1e9, 3), "blocks": k, "ms": round(t_peak * 1e3, 2)}
        points = []
        for load in args.serve_loads:
            rate = load * peak / mean_req
            arrivals, reqs = poisson_requests(5, args.serve_requests, nb, rate, kmax=64)

            def verify(i, c, reqs=reqs, srv=srv, load=load, mode=mode):
                if i % 8:
                    return
                for t, b in zip(srv.output(c, reqs[i]), reqs[i]):
                    if not torch.equal(t, arena[int(corpus.offs[b]):int(corpus.offs[b]) + blk]):
                        bad.append((mode, load, i, int(b)))

            recs = run_load(srv, arrivals, reqs, on_done=verify)
            s = summarize(recs, lambda i, reqs=reqs: len(reqs[i]) * blk)
            s.update({"offered_load": load, "offered_gbs": round(rate * mean_req / 1e9, 3),
                      "p99_over_single_block": round(s["p99_ms"] / single[mode], 3),
                      "gbs_over_same_grid_peak": round(s["gbs"] * 1e9 / peak, 3),
                      "max_in_flight": srv.max_in_flight, "max_queued": srv.max_queued})
            points.append(s)
            log(f"[serve:{mode}] load {load:.0%}: p50 {s['p50_ms']} ms, p99 {s['p99_ms']} ms "
                f"({s['p99_over_single_block']}x one block), {s['gbs']} GB/s "
                f"({s['gbs_over_same_grid_peak']:.0%} of peak), in flight <= {srv.max_in_flight}, "
                f"queued <= {srv.max_queued}")
        srv.close()
        modes[mode] = points
    assert not bad, f"serving outputs differ: {bad[:5]}"
    assert all(p["errors"] == 0 for pts in modes.values() for p in pts)
    first = args.serve_modes[0]
    return {"corpus": "markov_64x4mib, resident compressed in HBM", "batched_peak_gbs": peaks[first]["gbs"],
            "peak": "batched decode (one graph launch over copies of every block) on the engine's own decode waves",
            "peaks": peaks, "single_block_ms": single,
            "requests_per_point": args.serve_requests, "request_blocks": "uniform 1-64 distinct ids, seed 5",
            "budget_gib": args.serve_budget_gib, "mode": first, "points": modes[first],
            "modes": {"persistent": "one resident grid fed block by block from a host-appended ring",
                      "slots": f"admitted requests packed into batched launches, {args.serve_slots} in flight"},
            "by_mode": modes,
            "verified": "every 8th request's blocks compared on device with the original bytes"}


# ------------------------------------------------------------------------ CPU
def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def bench_cpu(args, oracle, sample, workload):
    host_c, host_off, lens, sizes, orders = sample
    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    # The oracle compiled for the highest x86-64 level this host supports.
    blib, bflags = oracle.baseline_lib()

    def run(idx, nthreads, min_s):
        nbytes, t, passes = 0, 0.0, 0
        while t < min_s:
            t0 = time.perf_counter()
            _, _, ol, st = oracle.decode_many(host_c, host_off[idx], lens[idx], sizes[idx], nthreads, use=blib)
            t += time.perf_counter() - t0
            assert np.all(st == 0)
            nbytes += int(sizes[idx].sum())
            passes += 1
        return nbytes / t / 1e9, passes, t

    allidx = np.arange(len(sizes))
    gbs_all, p_all, t_all = run(allidx, threads, args.cpu_seconds)
    sub = allidx[::10]
    gbs_1, p_1, t_1 = run(sub, 1, args.cpu_seconds / 2)
    log(f"[cpu] oracle decode {gbs_all:.3f} GB/s on {threads} threads, {gbs_1:.4f} GB/s on 1 thread")
    return {"value": round(gbs_all, 4), "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"{len(allidx)} blocks of {workload} (every {args.cpu_every}th, "
                      f"{int(sizes.sum()) / 1e9:.3f} GB, order-1 {int(orders.sum())}), decoded {p_all}x "
                      f"block-parallel over {threads} threads in {t_all:.1f}s",
            "single_thread": {"value": round(gbs_1, 4), "unit": "GB/s",
                              "sample": f"every 10th block of that sample ({len(sub)} blocks), {p_1}x, "
                                        f"{t_1:.1f}s"},
            "build": f"oracle/rans_oracle.c, gcc {bflags}", "host_cpu": cpu_model(), "nproc": os.cpu_count()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="mixed_10k")
    ap.add_argument("--blocks", type=int, default=0, help="truncate the workload (0 = all)")
    ap.add_argument("--sections", default=",".join(SECTIONS))
    ap.add_argument("--grid", type=int, default=0, help="skip the autotuner: persistent grid")
    ap.add_argument("--policy", default="lpt-cost", help="with --grid: queue order policy")
    ap.add_argument("--perblock-every", type=int, default=20)
    ap.add_argument("--no-perblock", action="store_true")
    ap.add_argument("--encode-best", action=argparse.BooleanOptionalAction, default=True)
    ap.add_argument("--crc", action=argparse.BooleanOptionalAction, default=True)
    ap.add_argument("--pcie", action=argparse.BooleanOptionalAction, default=True)
    ap.add_argument("--cpu-every", type=int, default=10)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    ap.add_argument("--peak-blocks", type=int, default=2048)
    ap.add_argument("--serve-requests", type=int, default=500)
    ap.add_argument("--serve-loads", type=float, nargs="+", default=[0.1, 0.25, 0.5, 0.75, 1.0, 1.2])
    ap.add_argument("--serve-budget-gib", type=int, default=32)
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-waves", type=int, default=0, help="persistent serving grid (0 = occupancy max)")
    ap.add_argument("--serve-modes", nargs="+", default=["persistent", "slots"],
                    choices=["persistent", "slots"], help="serving engines to sweep (the first is reported)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run a 1 GiB read + 1 GiB write kernel (counter calibration under rocprofv3 --pmc)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the rANS blocks are independent and the "
              "multi-GPU form (static LPT sharding of the block list, no collectives) is "
              "not part of this build; run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    sections = set(args.sections.split(","))

    import torch
    import fqxv_amd
    from oracle import oracle

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    dev = torch.device("cuda", 0)
    coder = fqxv_amd.Coder(0)

    if args.copy_calib:
        # Known traffic for counter calibration: an elementwise kernel that
        # reads 1 GiB and writes 1 GiB (fill_ writes 1 GiB).
        a = torch.empty(1 << 28, dtype=torch.int32, device=dev).fill_(1)
        b = torch.empty_like(a)
        torch.add(a, 1, out=b)
        torch.cuda.synchronize()
        del a, b

    c4 = bench_c4(args, coder, dev, oracle)
    configs = {}
    serving = None
    if "configs" in sections or "serve" in sections:
        coder.set_grid(0)
        if "configs" in sections:
            r1, _ = bench_config(coder, dev, oracle, "binned_1mib", [0])
            r2, _ = bench_config(coder, dev, oracle, "normal_64x1mib", [0, 21, 63])
            configs = {"binned_1mib": r1, "normal_64x1mib": r2, "perm_planes": bench_perm(coder, dev, oracle),
                       "names_streams": bench_names(coder)}
        r3, st3 = bench_config(coder, dev, oracle, "markov_64x4mib", [0, 63])
        if "configs" in sections:
            configs["markov_64x4mib"] = r3
        if "serve" in sections:
            serving = bench_serve(args, coder, dev, st3)
        del st3
    cpu = bench_cpu(args, oracle, c4["cpu_sample"], args.workload) if "cpu" in sections else None
    percall = bench_percall(coder, oracle) if "configs" in sections else None

    n_total, c_total, t_kern = c4["n_total"], c4["c_total"], c4["t_kern"]
    achieved = (n_total + c_total) / t_kern / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            tf = json.load(f)
        if tf.get("uncompressed_bytes") == n_total and tf.get("compressed_bytes") == c_total:
            traffic = tf["hbm_bytes_per_launch"]
    corpus = c4["corpus"]
    tuned = c4["tuned"]
    result = {
        "metric": METRIC, "value": round(c4["value"], 3), "unit": "GB/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(c4["t_step"] * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
        "data": "synthetic (host generator, fixed seeds; byte-identical to the oracle on a sample, "
                "device round trip on every block)",
        "config": {"workload": f"{args.workload}: {c4['nb']} blocks, log-uniform 4 KiB-4 MiB, "
                               f"order-0/1 50/50, Markov quality / DNA / names, seed 4",
                   "blocks": c4["nb"], "uncompressed_bytes": n_total, "compressed_bytes": c_total,
                   "order1_blocks": int(corpus.orders.sum()),
                   "launch": "hipGraph replay of one persistent-kernel launch over all blocks",
                   "checked": "output filled with a sentinel before the timed replays, equal to the input after",
                   "grid": tuned.grid, "queue_policy": tuned.policy,
                   "generator": "counter-based splitmix64 (fqxv_amd/csrc/synth.c), seeds 1-5",
                   "corpus_xxh3": c4["fingerprint"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "kernel": "k_rans_decode", "kernel_ms": round(t_kern * 1e3, 3)},
        "cpu_baseline": cpu,
        "perblock": c4["perblock"],
        "batched_eager_ms": c4["eager_ms"],
        "latency_floor": c4["largest_o1"],
        "ns_per_round_alone": c4["kind_rounds"],
        "percall": percall,
        "tuner": {"grid": tuned.grid, "policy": tuned.policy, "staging": tuned.staging,
                  "dimensions": ["grid", "policy", "staging"], "table": tuned.table},
        "encode": {"value": round(c4["encode_tuned"]["gbs"], 3), "unit": "GB/s",
                   "form": "graph replay of one batched launch (tuned grid and queue order), HIP events",
                   "ms": round(c4["encode_tuned"]["ms"], 3), "grid": c4["encode_tuned"]["grid"],
                   "policy": c4["encode_tuned"]["policy"], "tuner": c4["encode_tuned"]["table"],
                   "eager_call": {"value": round(c4["encode_gbs"], 3), "unit": "GB/s",
                                  "form": "fqxv_rans_encode_batch at the default grid, host job setup included"}},
        "encode_best": c4["encode_best"],
        "crc32c": c4["crc32c"],
        "xxh3": c4["xxh3"],
        "pcie_inclusive": c4["pcie"],
        "configs": configs,
        "serving": serving,
    }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
