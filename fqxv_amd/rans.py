"""Host-side mirror of the fqxv-rans API over the gfx950 library.

Reference surface (crates/fqxv-rans/src/lib.rs):

    pub enum Order { Zero, One }                          lib.rs:58-64
    pub enum Backend { Scalar, Sse42, Avx2, Avx512 }      lib.rs:67-97
    pub enum Error { InputTooLarge, Malformed, NotImplemented }  lib.rs:40-52
    pub fn encode(src, order) -> Result<Vec<u8>>          lib.rs:111-127
    pub fn decode(src) -> Result<Vec<u8>>                 lib.rs:186-200
    pub fn decode_bounded(src, max_len) -> Result<...>    lib.rs:226-231
    pub fn decode_with(src, backend) -> Result<...>       lib.rs:241-255

Same names, argument meaning and error behaviour: every malformed stream
raises :class:`Malformed` carrying the reference's message; nothing here
falls back to a CPU coder.  On top of that, :class:`Coder` exposes the
device-resident batched forms (one persistent launch over many blocks, or
one launch per block) used by the benchmark and the serving loop.
"""
from __future__ import annotations

import ctypes
import enum
import threading
from dataclasses import dataclass

import numpy as np

from . import _lib, tune

# Declared lengths above this are treated like a failed `try_reserve_exact`
# (model.rs:199-205): the decode still validates tables and states first.
ALLOC_LIMIT = 1 << 36


class Order(enum.IntEnum):
    """lib.rs:58-64 -- the stream's order tag byte."""
    Zero = 0
    One = 1


class Backend(enum.Enum):
    """Which device form performs the coding (lib.rs:67-97 analogue).

    ``Gfx950`` is the single-block launch, ``Gfx950Batched`` routes the call
    through the persistent batched kernel.  Output is byte-identical.
    """
    Gfx950 = "gfx950"
    Gfx950Batched = "gfx950-batched"

    @staticmethod
    def detect() -> "Backend":
        """lib.rs:80-96: the form plain ``decode`` uses (fqxv_rans_backend_detect)."""
        code = _lib.load().fqxv_rans_backend_detect()
        return Backend.Gfx950Batched if code == _lib.BACKEND_GFX950_BATCHED else Backend.Gfx950


class Error(Exception):
    """lib.rs:40-52."""


class InputTooLarge(Error):
    pass


class Malformed(Error):
    """`Error::Malformed(&'static str)`; ``str(e)`` is the reference message."""

    def __init__(self, msg: str, code: int = 0):
        super().__init__(msg)
        self.code = code


class NotImplemented_(Error):
    pass


class DeviceError(Error):
    """The HIP runtime failed (no device, out of device memory, ...)."""


def _raise(code: int):
    if code == _lib.OK:
        return
    msg = _lib.strerror(code)
    if code in (_lib.ERR_DEVICE,):
        raise DeviceError(msg)
    if code in (_lib.ERR_INVALID_ARG, _lib.ERR_DST_TOO_SMALL):
        raise ValueError(msg)
    raise Malformed(msg, code)


def _as_bytes(src) -> bytes:
    if isinstance(src, (bytes, bytearray, memoryview)):
        return bytes(src)
    if isinstance(src, np.ndarray):
        return np.ascontiguousarray(src, dtype=np.uint8).tobytes()
    return bytes(src)


def peek_output_len(src) -> int | None:
    """lib.rs:207-210."""
    b = _as_bytes(src)
    if len(b) < 9:
        return None
    return int.from_bytes(b[1:9], "little")


@dataclass
class BatchResult:
    dst: object          # torch uint8 device tensor holding the outputs
    dst_off: np.ndarray  # per-job output offsets (host)
    out_len: object      # torch int64 device tensor
    status: object       # torch int32 device tensor
    # The device job and order arrays the launch reads: held as long as the
    # result, so the caching allocator cannot hand them out while the
    # launch (on whatever stream it was given) may still read them.
    inputs: tuple = ()


class Coder:
    """One context on one HIP device (the library serialises its calls)."""

    def __init__(self, device: int = 0):
        self._L = _lib.load()
        h = ctypes.c_void_p()
        rc = self._L.fqxv_rans_ctx_create(device, ctypes.byref(h))
        if rc:
            raise DeviceError(f"fqxv_rans_ctx_create({device}): {_lib.strerror(rc)}")
        self._h = h
        self.device = device

    def close(self):
        if getattr(self, "_h", None):
            self._L.fqxv_rans_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------ reference API (host)
    def encode(self, src, order: Order | int) -> bytes:
        b = _as_bytes(src)
        order = int(order)
        if order not in (0, 1):
            raise ValueError("order must be Order.Zero or Order.One")
        cap = int(self._L.fqxv_rans_encode_bound(len(b), order))
        dst = ctypes.create_string_buffer(cap)
        out = ctypes.c_size_t(0)
        sbuf = ctypes.create_string_buffer(b, len(b)) if b else None
        rc = self._L.fqxv_rans_encode(self._h, sbuf, len(b), order, dst, cap, ctypes.byref(out))
        _raise(rc)
        return dst.raw[: out.value]

    def decode_bounded(self, src, max_len: int) -> bytes:
        b = _as_bytes(src)
        n = peek_output_len(b)
        if n is not None and n > max_len:
            raise Malformed(_lib.strerror(_lib.ERR_BOUND), _lib.ERR_BOUND)
        cap = n if (n is not None and n <= ALLOC_LIMIT) else 0
        dst = ctypes.create_string_buffer(max(cap, 1))
        out = ctypes.c_size_t(0)
        sbuf = ctypes.create_string_buffer(b, len(b)) if b else None
        rc = self._L.fqxv_rans_decode_bounded(self._h, sbuf, len(b), max_len, dst, cap,
                                              ctypes.byref(out))
        _raise(rc)
        return dst.raw[: out.value]

    def decode(self, src) -> bytes:
        return self.decode_bounded(src, (1 << 64) - 1)

    def decode_with(self, src, backend: Backend) -> bytes:
        """lib.rs:241-255: decode pinned to one device form."""
        code = {Backend.Gfx950: _lib.BACKEND_GFX950, Backend.Gfx950Batched: _lib.BACKEND_GFX950_BATCHED}[backend]
        b = _as_bytes(src)
        n = peek_output_len(b)
        cap = n if (n is not None and n <= ALLOC_LIMIT) else 0
        dst = ctypes.create_string_buffer(max(cap, 1))
        out = ctypes.c_size_t(0)
        sbuf = ctypes.create_string_buffer(b, len(b)) if b else None
        _raise(self._L.fqxv_rans_decode_with(self._h, code, sbuf, len(b), dst, cap, ctypes.byref(out)))
        return dst.raw[: out.value]

    # ------------------------------------------------ device-resident forms
    @staticmethod
    def _stream(stream):
        import torch
        if stream is None:
            stream = torch.cuda.current_stream()
        return ctypes.c_void_p(stream.cuda_stream)

    @staticmethod
    def jobs_array(src_off, src_len, dst_off, dst_cap, max_len=None, order=None, flags=0) -> np.ndarray:
        k = len(src_off)
        a = np.zeros(k, dtype=np.dtype(_lib.JOB_DTYPE_FIELDS))
        a["src_off"] = src_off
        a["src_len"] = src_len
        a["dst_off"] = dst_off
        a["dst_cap"] = dst_cap
        a["max_len"] = (np.uint64(2**64 - 1) if max_len is None else max_len)
        a["order"] = 0 if order is None else order
        a["flags"] = flags
        return a

    def _run_batch(self, fn, src, dst, jobs_np, order_np, stream):
        import torch
        dev = src.device
        jobs = torch.from_numpy(jobs_np.view(np.uint8).copy()).to(dev)
        k = len(jobs_np)
        status = torch.full((k,), -1, dtype=torch.int32, device=dev)
        out_len = torch.zeros((k,), dtype=torch.int64, device=dev)
        ordt = None
        if order_np is not None:
            ordt = torch.from_numpy(np.ascontiguousarray(order_np, dtype=np.uint32).view(np.int32)).to(dev)
        if stream is not None:
            # Allocated on the current stream, read on `stream`.
            s = torch.cuda.ExternalStream(stream.cuda_stream) if not isinstance(stream, torch.cuda.Stream) else stream
            for t in (jobs, status, out_len, ordt):
                if t is not None:
                    t.record_stream(s)
        rc = fn(self._h, ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()),
                ctypes.c_void_p(jobs.data_ptr()), k,
                ctypes.c_void_p(ordt.data_ptr()) if ordt is not None else None,
                ctypes.c_void_p(status.data_ptr()), ctypes.c_void_p(out_len.data_ptr()),
                self._stream(stream))
        _raise(rc)
        return status, out_len, (jobs, ordt)

    def encode_bound(self, n: int, order: int) -> int:
        return int(self._L.fqxv_rans_encode_bound(int(n), int(order)))

    @staticmethod
    def encode_bounds(src_len, orders) -> np.ndarray:
        """``fqxv_rans_encode_bound`` of many blocks at once (the same
        formula: header bound -- order-0 9 + 512 + 128, order-1
        9 + 32 + 512 * min(n, 256) + 128 -- plus 2n + 16; 9 for an empty
        block), so a batch of thousands of blocks needs no per-block call."""
        n = np.asarray(src_len, dtype=np.uint64)
        o = np.asarray(orders)
        k = np.minimum(n, np.uint64(256))
        hdr = np.where(o == 0, np.uint64(9 + 512 + 128), np.uint64(9 + 32 + 128) + np.uint64(512) * k)
        return np.where(n == 0, np.uint64(9), hdr + np.uint64(2) * n + np.uint64(16)).astype(np.uint64)

    @staticmethod
    def encode_best_bounds(src_len) -> np.ndarray:
        """``fqxv_rans_encode_best_bound`` of many blocks at once."""
        n = np.asarray(src_len, dtype=np.uint64)
        return (Coder.encode_bounds(n, np.zeros(len(n), np.int32)) +
                Coder.encode_bounds(n, np.ones(len(n), np.int32))).astype(np.uint64)

    def lpt_order(self, sizes) -> np.ndarray:
        s = np.ascontiguousarray(sizes, dtype=np.uint64)
        o = np.empty(len(s), dtype=np.uint32)
        self._L.fqxv_rans_lpt_order(ctypes.c_void_p(s.ctypes.data), len(s), ctypes.c_void_p(o.ctypes.data))
        return o

    def encode_batch(self, src, src_off, src_len, orders, lpt: bool = True, stream=None) -> BatchResult:
        """Encode every block of a device arena in one persistent launch."""
        import torch
        src_off = np.asarray(src_off, dtype=np.uint64)
        src_len = np.asarray(src_len, dtype=np.uint64)
        orders = np.asarray(orders, dtype=np.int32)
        caps = self.encode_bounds(src_len, orders)
        dst_off = np.zeros_like(caps)
        if len(caps):
            dst_off[1:] = np.cumsum(caps)[:-1]
        dst = torch.empty(int(caps.sum()) + 64, dtype=torch.uint8, device=src.device)
        jobs = self.jobs_array(src_off, src_len, dst_off, caps, order=orders)
        # Largest estimated encode time first (order-1 bytes cost more).
        order_np = tune.queue_order(self, "lpt-cost", src_len, orders, "encode") if lpt else None
        status, out_len, keep = self._run_batch(self._L.fqxv_rans_encode_batch, src, dst, jobs, order_np, stream)
        return BatchResult(dst, dst_off, out_len, status, keep)

    def encode_best_bound(self, n: int) -> int:
        return int(self._L.fqxv_rans_encode_best_bound(int(n)))

    def encode_best_batch(self, src, src_off, src_len, tie_order=0, lpt: bool = True, stream=None) -> BatchResult:
        """Encode every block at both orders and keep the smaller stream.

        ``tie_order`` (scalar or per block) is the order kept when both
        streams have the same length: 0 as the tokenizer and the permutation
        planes do, 1 as the long-read codec does."""
        import torch
        src_off = np.asarray(src_off, dtype=np.uint64)
        src_len = np.asarray(src_len, dtype=np.uint64)
        ties = np.broadcast_to(np.asarray(tie_order, dtype=np.int32), src_len.shape).copy()
        caps = self.encode_best_bounds(src_len)
        dst_off = np.zeros_like(caps)
        if len(caps):
            dst_off[1:] = np.cumsum(caps)[:-1]
        dst = torch.empty(int(caps.sum()) + 64, dtype=torch.uint8, device=src.device)
        jobs = self.jobs_array(src_off, src_len, dst_off, caps, order=ties)
        order_np = self.lpt_order(src_len) if lpt else None
        status, out_len, keep = self._run_batch(self._L.fqxv_rans_encode_best_batch, src, dst, jobs, order_np,
                                                stream)
        return BatchResult(dst, dst_off, out_len, status, keep)

    def decode_batch(self, src, src_off, src_len, out_len, max_len=None, lpt: bool = True,
                     stream=None, dst=None, dst_off=None) -> BatchResult:
        """Decode every stream of a device arena in one persistent launch.

        ``out_len`` are the output capacities (normally each stream's n)."""
        import torch
        src_off = np.asarray(src_off, dtype=np.uint64)
        src_len = np.asarray(src_len, dtype=np.uint64)
        caps = np.asarray(out_len, dtype=np.uint64)
        if dst_off is None:
            dst_off = np.zeros_like(caps)
            if len(caps):
                dst_off[1:] = np.cumsum(caps)[:-1]
        if dst is None:
            dst = torch.empty(int(caps.sum()) + 64, dtype=torch.uint8, device=src.device)
        jobs = self.jobs_array(src_off, src_len, dst_off, caps, max_len=max_len)
        order_np = self.lpt_order(caps) if lpt else None
        status, ol, keep = self._run_batch(self._L.fqxv_rans_decode_batch, src, dst, jobs, order_np, stream)
        return BatchResult(dst, dst_off, ol, status, keep)

    def _one(self, fn, src, dst, job: _lib.Job, status, out_len, stream):
        rc = fn(self._h, ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()),
                ctypes.byref(job), ctypes.c_void_p(status.data_ptr()),
                ctypes.c_void_p(out_len.data_ptr()), self._stream(stream))
        _raise(rc)

    def decode_one_device(self, src, dst, job: _lib.Job, status, out_len, stream=None):
        """The plain form: one kernel launch for one block (device pointers)."""
        self._one(self._L.fqxv_rans_decode_one_device, src, dst, job, status, out_len, stream)

    def encode_one_device(self, src, dst, job: _lib.Job, status, out_len, stream=None):
        self._one(self._L.fqxv_rans_encode_one_device, src, dst, job, status, out_len, stream)

    def max_waves(self) -> int:
        w = ctypes.c_uint32(0)
        _raise(self._L.fqxv_rans_max_waves(self._h, ctypes.byref(w)))
        return int(w.value)

    def crc32c_batch(self, src, offs, lens, stream=None):
        """CRC-32C (crates/fqxv/src/crc.rs) of byte ranges of a device arena;
        returns a device uint32 (int32-typed) tensor."""
        import torch
        offs = np.ascontiguousarray(offs, dtype=np.uint64)
        lens = np.ascontiguousarray(lens, dtype=np.uint64)
        out = torch.zeros(max(len(offs), 1), dtype=torch.int32, device=src.device)
        _raise(self._L.fqxv_rans_crc32c_batch(self._h, ctypes.c_void_p(src.data_ptr()),
                                              ctypes.c_void_p(offs.ctypes.data), ctypes.c_void_p(lens.ctypes.data),
                                              len(offs), ctypes.c_void_p(out.data_ptr()), self._stream(stream)))
        return out[: len(offs)]

    def xxh3_batch(self, src, offs, lens, stream=None):
        """XXH3-64 (seed 0; the reference's xxhash-rust ``Xxh3``) of byte
        ranges of a device arena; returns a device int64 tensor holding the
        u64 digests' bits."""
        import torch
        offs = np.ascontiguousarray(offs, dtype=np.uint64)
        lens = np.ascontiguousarray(lens, dtype=np.uint64)
        out = torch.zeros(max(len(offs), 1), dtype=torch.int64, device=src.device)
        _raise(self._L.fqxv_rans_xxh3_batch(self._h, ctypes.c_void_p(src.data_ptr()),
                                            ctypes.c_void_p(offs.ctypes.data), ctypes.c_void_p(lens.ctypes.data),
                                            len(offs), ctypes.c_void_p(out.data_ptr()), self._stream(stream)))
        return out[: len(offs)]

    def xxh3_concat(self, src, segments, stream=None):
        """XXH3-64 of concatenated byte ranges of a device arena:
        ``segments[i]`` is a list of (offset, length) pairs whose bytes, in
        order, form string i (e.g. ``[n_reads u64][lens u32...][bytes]``,
        the sequence/quality digest layout of block.rs:164-196).  Returns a
        device int64 tensor holding the u64 digests' bits."""
        import torch
        first = np.zeros(len(segments) + 1, dtype=np.uint32)
        first[1:] = np.cumsum([len(s) for s in segments])
        flat = [p for s in segments for p in s]
        offs = np.ascontiguousarray([p[0] for p in flat] or [0], dtype=np.uint64)
        lens = np.ascontiguousarray([p[1] for p in flat] or [0], dtype=np.uint64)
        out = torch.zeros(max(len(segments), 1), dtype=torch.int64, device=src.device)
        _raise(self._L.fqxv_rans_xxh3_concat(self._h, ctypes.c_void_p(src.data_ptr()),
                                             ctypes.c_void_p(offs.ctypes.data), ctypes.c_void_p(lens.ctypes.data),
                                             ctypes.c_void_p(first.ctypes.data), len(segments),
                                             ctypes.c_void_p(out.data_ptr()), self._stream(stream)))
        return out[: len(segments)]

    def set_grid(self, waves: int) -> None:
        """Persistent-grid size of later batched calls (0 = occupancy max)."""
        _raise(self._L.fqxv_rans_ctx_set_grid(self._h, int(waves)))

    def graph(self, direction: str, src, dst, jobs_np: np.ndarray, order_np=None) -> "Graph":
        """Capture one batched call (``"decode"``, ``"encode"`` or ``"encode_best"``) over fixed
        device buffers into a hipGraph; :meth:`Graph.launch` replays it."""
        import torch
        d = {"decode": 0, "encode": 1, "encode_best": 2}[direction]
        dev = src.device
        jobs = torch.from_numpy(jobs_np.view(np.uint8).copy()).to(dev)
        k = len(jobs_np)
        status = torch.full((k,), -1, dtype=torch.int32, device=dev)
        out_len = torch.zeros((k,), dtype=torch.int64, device=dev)
        ordt = None
        if order_np is not None:
            ordt = torch.from_numpy(np.ascontiguousarray(order_np, dtype=np.uint32).view(np.int32)).to(dev)
        torch.cuda.synchronize(dev)
        h = ctypes.c_void_p()
        rc = self._L.fqxv_rans_graph_create(
            self._h, d, ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()),
            ctypes.c_void_p(jobs.data_ptr()), k, ctypes.c_void_p(ordt.data_ptr()) if ordt is not None else None,
            ctypes.c_void_p(status.data_ptr()), ctypes.c_void_p(out_len.data_ptr()), ctypes.byref(h))
        _raise(rc)
        return Graph(self, h, src, dst, jobs, ordt, status, out_len)


class Graph:
    """A captured batched call; buffers stay referenced while it lives."""

    def __init__(self, coder: Coder, h, src, dst, jobs, ordt, status, out_len):
        self._coder = coder
        self._h = h
        self._keep = (src, jobs, ordt)
        self.dst = dst
        self.status = status
        self.out_len = out_len

    def launch(self, stream=None) -> None:
        _raise(self._coder._L.fqxv_rans_graph_launch(self._h, Coder._stream(stream)))

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._coder._L.fqxv_rans_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_tls = threading.local()


def default_coder() -> Coder:
    c = getattr(_tls, "coder", None)
    if c is None:
        c = Coder(0)
        _tls.coder = c
    return c


def encode(src, order: Order | int) -> bytes:
    """lib.rs:111-127."""
    return default_coder().encode(src, order)


def decode(src) -> bytes:
    """lib.rs:186-200."""
    return default_coder().decode(src)


def decode_bounded(src, max_len: int) -> bytes:
    """lib.rs:226-231."""
    return default_coder().decode_bounded(src, max_len)


def decode_with(src, backend: Backend) -> bytes:
    """lib.rs:241-255: decode through the named device form
    (``fqxv_rans_decode_with``); output is identical."""
    return default_coder().decode_with(src, backend)
