"""Autotuner for the batched (persistent-kernel) encode/decode launch.

The batched kernels run one 64-lane wave per workgroup and pull block ids from
a device work queue; what is left to choose per workload is

* ``grid``: how many persistent waves the launch starts (the occupancy
  maximum, or fewer: fewer waves hold fewer LDS tables and workspace slices
  resident and contend less in L2 when blocks are few or large);
* ``policy``: the order in which waves take blocks from the queue --
  ``"index"`` (as given), ``"lpt-size"`` (largest first) or ``"lpt-cost"``
  (largest estimated decode time first, order-1 bytes weighted by
  :data:`O1_WEIGHT`, the measured per-byte cost ratio of the two decoders);
* ``staging`` (decode): where an order-1 block's distinct code rows live --
  ``"lds"`` (behind the block's (context x symbol) table in LDS when they
  fit, the workspace otherwise) or ``"workspace"`` (always the wave's device
  workspace, read through L1/L2), set per job with
  ``FQXV_RANS_JOB_O1_ROWS_WORKSPACE``.

Two shapes the kernels fix instead of searching: one block per wave (the
order-1 fast path gives every lane a context, so a second block would need
its own wave anyway; only order-0 leaves lanes 32..63 idle, and order-0 is
the smaller share of a mixed batch's wave time) and one wave per workgroup
(each wave decodes its own block out of its own LDS block, so a larger
workgroup shares nothing and only coarsens the LDS allocation, which sets
the occupancy either way).

Every candidate is captured into a hipGraph and timed by replay with HIP
events, exactly the form a caller then replays.  Results are cached as JSON
(``$FQXV_RANS_TUNE_CACHE``, default ``~/.cache/fqxv_amd/tune.json``) under a
key made of the device name, direction and workload shape.
"""
from __future__ import annotations

import json
import os
from dataclasses import asdict, dataclass

import numpy as np

POLICIES = ("lpt-cost", "lpt-size", "index")
STAGINGS = ("lds", "workspace")
# Per-byte decode time of an order-1 block over an order-0 block (gfx950,
# ~650 vs ~190 ns per 32-symbol round under load); encode is closer to 1.4.
O1_WEIGHT = {"decode": 3.4, "encode": 1.4}


def staging_flags(staging: str) -> int:
    """Job flags of a staging choice (fqxv_rans_job.flags)."""
    from . import _lib
    if staging == "lds":
        return 0
    if staging == "workspace":
        return _lib.JOB_O1_ROWS_WORKSPACE
    raise ValueError(f"unknown table staging {staging!r}")


def with_staging(jobs_np: np.ndarray, staging: str) -> np.ndarray:
    """A copy of a job array with every job's staging flag set."""
    j = jobs_np.copy()
    j["flags"] = staging_flags(staging)
    return j


@dataclass
class Tuned:
    grid: int
    policy: str
    ms: float
    table: list  # (grid, policy, staging, ms) for every candidate measured
    staging: str = "lds"

    def order(self, coder, sizes, orders, direction: str = "decode") -> np.ndarray | None:
        return queue_order(coder, self.policy, sizes, orders, direction)

    def jobs(self, jobs_np: np.ndarray) -> np.ndarray:
        """The job array with the chosen staging flag."""
        return with_staging(jobs_np, self.staging)


def queue_order(coder, policy: str, sizes, orders, direction: str = "decode") -> np.ndarray | None:
    sizes = np.asarray(sizes, dtype=np.uint64)
    if policy == "index":
        return None
    if policy == "lpt-size":
        return coder.lpt_order(sizes)
    if policy == "lpt-cost":
        w = np.where(np.asarray(orders) == 1, O1_WEIGHT[direction], 1.0)
        return coder.lpt_order((sizes.astype(np.float64) * w).astype(np.uint64))
    raise ValueError(f"unknown queue policy {policy!r}")


def _cache_path() -> str:
    return os.environ.get("FQXV_RANS_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "fqxv_amd", "tune.json"))


def _key(direction, sizes, orders) -> str:
    import torch
    sizes = np.asarray(sizes, dtype=np.uint64)
    o1 = float(np.asarray(orders).mean()) if len(orders) else 0.0
    name = torch.cuda.get_device_name(0).replace(" ", "_")
    return f"{name}|{direction}|{len(sizes)}|{int(sizes.sum())}|{int(sizes.max()) if len(sizes) else 0}|{o1:.2f}"


def _load_cache() -> dict:
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(c: dict) -> None:
    p = _cache_path()
    try:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        tmp = p + ".tmp"
        with open(tmp, "w") as f:
            json.dump(c, f, indent=1, sort_keys=True)
        os.replace(tmp, p)
    except OSError:
        pass


def default_grids(max_waves: int, count: int) -> list[int]:
    g = {max_waves, max_waves * 3 // 4, max_waves // 2, max_waves // 4, 256}
    return sorted({min(max(1, x), count) for x in g}, reverse=True)


def autotune(coder, direction: str, src, dst, jobs_np: np.ndarray, sizes, orders, grids=None,
             policies=POLICIES, stagings=None, reps: int = 2, use_cache: bool = True, log=None) -> Tuned:
    """Pick (grid, queue policy, table staging) for one batched workload by
    timing graph replays.

    ``sizes`` are the per-block costs' byte counts (uncompressed lengths);
    ``orders`` the per-block orders.  Leaves the coder's grid at the winner;
    apply the staging with :meth:`Tuned.jobs`.
    """
    import torch
    key = _key(direction, sizes, orders)
    if use_cache:
        hit = _load_cache().get(key)
        if hit:
            t = Tuned(**hit)
            coder.set_grid(t.grid)
            return t
    count = len(jobs_np)
    grids = grids or default_grids(coder.max_waves(), count)
    if stagings is None:
        stagings = STAGINGS if direction == "decode" else ("lds",)
    table = []
    stream = torch.cuda.current_stream()
    for stg in stagings:
        jobs_s = with_staging(jobs_np, stg)
        for grid in grids:
            for pol in policies:
                coder.set_grid(grid)
                g = coder.graph(direction, src, dst, jobs_s, queue_order(coder, pol, sizes, orders, direction))
                g.launch(stream)  # warm
                evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                       for _ in range(reps)]
                for a, b in evs:
                    a.record(stream)
                    g.launch(stream)
                    b.record(stream)
                torch.cuda.synchronize()
                if int((g.status != 0).sum()) != 0:
                    raise RuntimeError(f"autotune candidate grid={grid} policy={pol} staging={stg} "
                                       "reported block errors")
                ms = min(a.elapsed_time(b) for a, b in evs)
                g.close()
                table.append((grid, pol, stg, round(ms, 3)))
                if log:
                    log(f"[tune] {direction} grid={grid} policy={pol} staging={stg}: {ms:.2f} ms")
    grid, pol, stg, ms = min(table, key=lambda r: r[3])
    t = Tuned(grid=grid, policy=pol, ms=ms, table=table, staging=stg)
    coder.set_grid(grid)
    if use_cache:
        c = _load_cache()
        c[key] = asdict(t)
        _store_cache(c)
    return t
