#!/usr/bin/env python3
"""rANS Nx16 decode throughput on one MI355X: batched vs per-block launch.

Workload (default ``mixed_10k``): 10,000 blocks, sizes log-uniform on
[4 KiB, 4 MiB] (6.02 GB), order-0/1 at 50/50 per block, streams split in
thirds between Markov quality strings, DNA (ACGT + 0.1% N) and ':'-separated
decimal name fields; seed 4.  The blocks are generated on the host, copied to
HBM, encoded on the device (byte identity against the oracle is checked on a
sample of blocks, and the device round trip on all of them), then:

* ``value``: one step = one batched decode of the whole corpus (one
  persistent-kernel launch over all blocks, inputs resident in HBM);
  GB/s = 1e9 bytes of uncompressed output per second.
* ``perblock``: the plain form -- one kernel launch per block -- timed on an
  every-k-th-block sample, with per-launch p50/p99 latency.
* ``roofline``: HBM-bound; algorithmic bytes = compressed bytes read +
  uncompressed bytes written, over the decode kernel's measured duration.
* ``cpu_baseline``: the oracle (plain-C scalar restatement of fqxv-rans)
  decoding a sample of the same streams on the host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "rANS Nx16 o0/o1 decode GB/s, batched vs per-block launch; p99 block latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="mixed_10k")
    ap.add_argument("--blocks", type=int, default=0, help="truncate the workload (0 = all)")
    ap.add_argument("--perblock-every", type=int, default=20)
    ap.add_argument("--cpu-every", type=int, default=10)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-perblock", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the rANS blocks are independent and the "
              "multi-GPU form (static LPT sharding of the block list, no collectives) is "
              "not part of this build; run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    import fqxv_amd
    from fqxv_amd import synth

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    dev = torch.device("cuda", 0)
    coder = fqxv_amd.Coder(0)

    t0 = time.time()
    corpus = synth.workload(args.workload, blocks=args.blocks or None)
    n_blocks = len(corpus.sizes)
    n_total = corpus.total
    log(f"[bench] generated {args.workload}: {n_blocks} blocks, {n_total / 1e9:.3f} GB "
        f"in {time.time() - t0:.1f}s")
    arena = torch.from_numpy(corpus.data).to(dev)
    torch.cuda.synchronize()

    # ---- encode (batched, one persistent launch) ---------------------------
    stream = torch.cuda.current_stream()
    enc = coder.encode_batch(arena, corpus.offs, corpus.sizes, corpus.orders)
    torch.cuda.synchronize()
    assert int((enc.status != 0).sum()) == 0, "device encode failed"
    enc_times = []
    for _ in range(2):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        enc = coder.encode_batch(arena, corpus.offs, corpus.sizes, corpus.orders, stream=stream)
        e1.record(stream)
        torch.cuda.synchronize()
        enc_times.append(e0.elapsed_time(e1) / 1e3)
    comp_len = enc.out_len.cpu().numpy().astype(np.uint64)
    c_total = int(comp_len.sum())
    log(f"[bench] encode {n_total / min(enc_times) / 1e9:.1f} GB/s, ratio {n_total / c_total:.3f}")

    # ---- parity before timing ---------------------------------------------
    from oracle import oracle
    host_c = enc.dst.cpu().numpy()
    rng = np.random.default_rng(0)
    sample = sorted(set(rng.choice(n_blocks, min(24, n_blocks), replace=False).tolist()) |
                    {int(np.argmax(corpus.sizes))})
    for j in sample:
        got = host_c[int(enc.dst_off[j]):int(enc.dst_off[j]) + int(comp_len[j])].tobytes()
        ref = oracle.encode(corpus.block(j), int(corpus.orders[j]))
        assert got == ref, f"block {j}: device encode differs from the oracle"
    dec = coder.decode_batch(enc.dst, enc.dst_off, comp_len, corpus.sizes)
    torch.cuda.synchronize()
    assert int((dec.status != 0).sum()) == 0, "device decode failed"
    assert torch.equal(dec.dst[:n_total], arena[:n_total]), "device round trip differs"
    log(f"[bench] parity ok ({len(sample)} blocks vs oracle, all {n_blocks} round-tripped)")

    # ---- batched decode: the timed steps -----------------------------------
    order_np = coder.lpt_order(corpus.sizes)
    jobs_np = coder.jobs_array(enc.dst_off, comp_len, dec.dst_off, corpus.sizes)
    import ctypes
    jobs_t = torch.from_numpy(jobs_np.view(np.uint8).copy()).to(dev)
    ord_t = torch.from_numpy(order_np.view(np.int32).copy()).to(dev)
    status = torch.zeros(n_blocks, dtype=torch.int32, device=dev)
    olen = torch.zeros(n_blocks, dtype=torch.int64, device=dev)
    out = dec.dst
    L = coder._L

    def launch():
        rc = L.fqxv_rans_decode_batch(coder._h, ctypes.c_void_p(enc.dst.data_ptr()),
                                      ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(jobs_t.data_ptr()),
                                      n_blocks, ctypes.c_void_p(ord_t.data_ptr()),
                                      ctypes.c_void_p(status.data_ptr()), ctypes.c_void_p(olen.data_ptr()),
                                      ctypes.c_void_p(stream.cuda_stream))
        assert rc == 0

    for _ in range(args.warmup):
        launch()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for a, b in evs:
        a.record(stream)
        launch()
        b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    kern = [a.elapsed_time(b) / 1e3 for a, b in evs]
    assert int((status != 0).sum()) == 0
    assert torch.equal(out[:n_total], arena[:n_total])
    t_step = wall / args.steps
    value = n_total / t_step / 1e9
    t_kern = float(np.mean(kern))
    log(f"[bench] batched decode {value:.1f} GB/s ({t_step * 1e3:.2f} ms/step, kernel {t_kern * 1e3:.2f} ms)")

    # ---- per-block launches (plain form) on a sample ------------------------
    perblock = None
    if not args.no_perblock:
        from fqxv_amd import _lib
        idx = list(range(0, n_blocks, args.perblock_every))
        st1 = torch.zeros(1, dtype=torch.int32, device=dev)
        ol1 = torch.zeros(1, dtype=torch.int64, device=dev)
        jobs = [_lib.Job(int(enc.dst_off[j]), int(comp_len[j]), int(dec.dst_off[j]), int(corpus.sizes[j]),
                         2**64 - 1, 0, 0) for j in idx]
        for jb in jobs[:3]:
            coder.decode_one_device(enc.dst, out, jb, st1, ol1, stream=stream)
        torch.cuda.synchronize()
        pe = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in idx]
        tp0 = time.perf_counter()
        for (a, b), jb in zip(pe, jobs):
            a.record(stream)
            coder.decode_one_device(enc.dst, out, jb, st1, ol1, stream=stream)
            b.record(stream)
        torch.cuda.synchronize()
        tp = time.perf_counter() - tp0
        lat = np.array([a.elapsed_time(b) for a, b in pe])
        sb = int(corpus.sizes[idx].sum())
        pb_gbs = sb / tp / 1e9
        perblock = {"value": round(pb_gbs, 3), "unit": "GB/s", "sample": f"every {args.perblock_every}th block "
                    f"({len(idx)} launches, {sb / 1e9:.3f} GB)", "speedup_batched": round(value / pb_gbs, 2),
                    "p50_block_ms": round(float(np.percentile(lat, 50)), 3),
                    "p99_block_ms": round(float(np.percentile(lat, 99)), 3)}
        log(f"[bench] per-block {pb_gbs:.2f} GB/s, batched speedup {value / pb_gbs:.1f}x, "
            f"p99 {perblock['p99_block_ms']} ms")

    # ---- CPU baseline: the oracle on a sample ------------------------------
    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        idx = np.arange(0, n_blocks, args.cpu_every)
        tc0 = time.perf_counter()
        dst, doff, ol, stc = oracle.decode_many(host_c, enc.dst_off[idx], comp_len[idx], corpus.sizes[idx],
                                                threads)
        tc = time.perf_counter() - tc0
        assert np.all(stc == 0)
        sbytes = int(corpus.sizes[idx].sum())
        cpu = {"value": round(sbytes / tc / 1e9, 4), "unit": "GB/s", "cores": threads, "kind": "port",
               "sample": f"every {args.cpu_every}th block of {args.workload} ({len(idx)} blocks, "
                         f"{sbytes / 1e9:.3f} GB), block-parallel over {threads} threads, {tc:.1f}s"}
        log(f"[bench] cpu oracle {cpu['value']} GB/s on {threads} threads")

    achieved = (n_total + c_total) / t_kern / 1e9
    result = {
        "metric": METRIC, "value": round(value, 3), "unit": "GB/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t_step * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {n_blocks} blocks, log-uniform 4 KiB-4 MiB, "
                               f"order-0/1 50/50, Markov quality / DNA / names, seed 4",
                   "blocks": n_blocks, "uncompressed_bytes": n_total, "compressed_bytes": c_total,
                   "order1_blocks": int(corpus.orders.sum())},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None},
        "cpu_baseline": cpu,
        "perblock": perblock,
        "encode": {"value": round(n_total / min(enc_times) / 1e9, 3), "unit": "GB/s",
                   "form": "batched, one persistent launch"},
    }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
