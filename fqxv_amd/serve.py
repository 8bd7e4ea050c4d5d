"""Request loop for random-access block decode (the library's serving engine).

A :class:`Server` wraps ``fqxv_rans_server_*`` (include/fqxv_rans_hip.h): a
compressed archive resident in HBM, an output pool under a byte budget,
arrival-order admission, and several batched decodes in flight at once.
:func:`run_load` drives it with a Poisson arrival process and returns one
record per request (arrival, admission and completion times in the engine's
clock), from which :func:`summarize` derives p50/p99 latency and GB/s.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass

import numpy as np

from . import _lib
from .rans import Coder, _raise


@dataclass
class Completed:
    ticket: int
    out_off: int
    out_bytes: int
    t_arrival: float
    t_admit: float
    t_done: float
    status: int
    nblocks: int

    @property
    def latency(self) -> float:
        return self.t_done - self.t_arrival


class Server:
    """Decode service over one resident archive.

    ``archive``: torch uint8 device tensor of compressed streams;
    ``src_off``/``src_len``: each block's stream; ``n``: each block's decoded
    length.  ``budget``: bytes of decoded output that may be held at once.
    ``mode``: ``"slots"`` (admitted requests packed into batched launches, up
    to ``slots`` at once) or ``"persistent"`` (one resident grid of ``waves``
    decode waves, 0 = half of the chip, fed block by block through a ``ring``
    of jobs; the grid leaves after ``idle_ms`` without new work).
    """

    def __init__(self, archive, src_off, src_len, n, budget: int, slots: int = 4, max_batch: int = 4096,
                 device: int = 0, mode: str = "slots", ring: int = 16384, waves: int = 0, idle_ms: int = 200):
        import torch
        self._L = _lib.load()
        self.archive = archive
        self.n = np.asarray(n, dtype=np.uint64)
        jobs = Coder.jobs_array(np.asarray(src_off, np.uint64), np.asarray(src_len, np.uint64),
                                np.zeros(len(self.n), np.uint64), self.n)
        self._jobs = np.ascontiguousarray(jobs)
        self.budget = int(budget)
        self.pool = torch.empty(self.budget, dtype=torch.uint8, device=archive.device)
        h = ctypes.c_void_p()
        self.mode = mode
        if mode == "persistent":
            rc = self._L.fqxv_rans_server_create_persistent(
                device, ctypes.c_void_p(archive.data_ptr()), ctypes.c_void_p(self._jobs.ctypes.data), len(self.n),
                ctypes.c_void_p(self.pool.data_ptr()), self.budget, ring, waves, idle_ms, ctypes.byref(h))
        elif mode == "slots":
            rc = self._L.fqxv_rans_server_create(device, ctypes.c_void_p(archive.data_ptr()),
                                                 ctypes.c_void_p(self._jobs.ctypes.data), len(self.n),
                                                 ctypes.c_void_p(self.pool.data_ptr()), self.budget,
                                                 slots, max_batch, ctypes.byref(h))
        else:
            raise ValueError(f"unknown serving mode {mode!r}")
        _raise(rc)
        self._h = h
        self._buf = (_lib.Done * 256)()

    def close(self):
        if getattr(self, "_h", None):
            self._L.fqxv_rans_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def now(self) -> float:
        t = ctypes.c_double()
        _raise(self._L.fqxv_rans_server_now(self._h, ctypes.byref(t)))
        return t.value

    def submit(self, ids, t_arrival: float = -1.0) -> int:
        a = np.ascontiguousarray(ids, dtype=np.uint32)
        tk = ctypes.c_uint64()
        _raise(self._L.fqxv_rans_server_submit(self._h, ctypes.c_void_p(a.ctypes.data), len(a),
                                               float(t_arrival), ctypes.byref(tk)))
        return int(tk.value)

    def poll(self) -> tuple[int, int]:
        q, f = ctypes.c_uint32(), ctypes.c_uint32()
        _raise(self._L.fqxv_rans_server_poll(self._h, ctypes.byref(q), ctypes.byref(f)))
        return int(q.value), int(f.value)

    def take(self) -> list[Completed]:
        out = []
        while True:
            got = ctypes.c_uint32()
            _raise(self._L.fqxv_rans_server_take(self._h, self._buf, len(self._buf), ctypes.byref(got)))
            for i in range(got.value):
                d = self._buf[i]
                out.append(Completed(d.ticket, d.out_off, d.out_bytes, d.t_arrival, d.t_admit, d.t_done,
                                     d.status, d.nblocks))
            if got.value < len(self._buf):
                return out

    def release(self, ticket: int) -> None:
        _raise(self._L.fqxv_rans_server_release(self._h, int(ticket)))

    @property
    def waves(self) -> int:
        """Decode waves the service runs (resident grid, or each batched launch)."""
        w = ctypes.c_uint32()
        _raise(self._L.fqxv_rans_server_waves(self._h, ctypes.byref(w)))
        return int(w.value)

    def block_offsets(self, ids) -> np.ndarray:
        """Offsets of a request's blocks inside its output region."""
        r = (self.n[np.asarray(ids, dtype=np.int64)] + 255) // 256 * 256
        o = np.zeros(len(r), dtype=np.uint64)
        if len(r) > 1:
            o[1:] = np.cumsum(r)[:-1]
        return o

    def output(self, done: Completed, ids):
        """Device views (torch) of a completed request's decoded blocks."""
        offs = self.block_offsets(ids)
        return [self.pool[done.out_off + int(o): done.out_off + int(o) + int(self.n[i])]
                for o, i in zip(offs, ids)]


def poisson_requests(seed: int, count: int, nblocks: int, rate: float, kmax: int = 64):
    """Arrival times (s, from 0) of a Poisson process of `rate` requests/s,
    each a uniform random set of 1..kmax distinct block ids."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, count)
    t = np.cumsum(gaps) - gaps[0]
    reqs = []
    for _ in range(count):
        k = int(rng.integers(1, min(kmax, nblocks) + 1))
        reqs.append(np.sort(rng.choice(nblocks, k, replace=False)).astype(np.uint32))
    return t, reqs


def run_load(server: Server, arrivals, reqs, on_done=None, keep=()) -> list[Completed]:
    """Submit each request at its arrival time (offset from now), run the
    engine until all complete, release outputs (after ``on_done`` sees them;
    tickets of request indices in ``keep`` are not released)."""
    base = server.now() + 0.002
    t_sub = base + np.asarray(arrivals, dtype=np.float64)
    nxt = 0
    done: dict[int, Completed] = {}
    idx_of: dict[int, int] = {}
    keep = set(keep)
    total = len(reqs)
    server.max_in_flight = server.max_queued = 0
    while len(done) < total:
        now = server.now()
        while nxt < total and t_sub[nxt] <= now:
            tk = server.submit(reqs[nxt], t_arrival=float(t_sub[nxt]))
            idx_of[tk] = nxt
            nxt += 1
        queued, in_flight = server.poll()
        server.max_in_flight = max(server.max_in_flight, in_flight)
        server.max_queued = max(server.max_queued, queued)
        for c in server.take():
            i = idx_of[c.ticket]
            done[i] = c
            if on_done is not None:
                on_done(i, c)
            if i not in keep:
                server.release(c.ticket)
        if nxt < total and len(done) < total:
            wait = t_sub[nxt] - server.now()
            if wait > 0.0005:
                time.sleep(min(wait, 0.0005))
    return [done[i] for i in range(total)]


def summarize(records: list[Completed], nbytes_of) -> dict:
    lat = np.array([r.latency for r in records]) * 1e3
    first = min(r.t_arrival for r in records)
    last = max(r.t_done for r in records)
    tot = sum(nbytes_of(i) for i in range(len(records)))
    return {"requests": len(records), "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3), "mean_ms": round(float(lat.mean()), 3),
            "gbs": round(tot / (last - first) / 1e9, 3), "span_s": round(last - first, 3),
            "errors": int(sum(r.status != 0 for r in records))}
