"""Host-side mirror of the read-name tokenizer's decode over the gfx950 library.

Reference surface (crates/fqxv-tokenizer/src/lib.rs):

    pub enum Error { Malformed(&'static str), Rans(fqxv_rans::Error) }  lib.rs:40-49
    pub fn decode(src: &[u8]) -> Result<Vec<Vec<u8>>>                  lib.rs:543-650

``decode`` keeps that signature; ``decode_batch`` is the batched form the
container's block loop would call (crates/fqxv/src/container/block.rs:1060,
one names stream per block): the framing of every stream is walked on the
host, all rANS sub-streams of all streams are decoded in ONE device launch
(``fqxv_names_decode_batch``), and the names are rebuilt natively.  Errors
carry the reference's messages; there is no CPU fallback behind this module.
"""
from __future__ import annotations

import ctypes
import struct

import numpy as np

from . import _lib
from .rans import Coder, DeviceError, default_coder


class Error(Exception):
    """lib.rs:40-49.  ``str(e)`` is the reference's message (for a rANS
    sub-stream failure, fqxv-rans's message)."""

    def __init__(self, msg: str, code: int):
        super().__init__(msg)
        self.code = code


class Malformed(Error):
    """``Error::Malformed`` -- a framing or token-level check failed."""


class Rans(Error):
    """``Error::Rans`` -- a sub-stream failed in the entropy coder."""


class Corrupt(Exception):
    """The container's ``Error::Corrupt { what: "block names digest" }``
    (crates/fqxv/src/container/block.rs:1097-1109): the names decoded, but
    not to the bytes the block's stored digest covers."""


def _error(code: int) -> Error:
    msg = _lib.strerror(code)
    return Malformed(msg, code) if code >= _lib.NAMES_ERR_FIRST else Rans(msg, code)


def parse_records(rec: bytes) -> list[bytes]:
    """[u64 n] then n x [u32 len][bytes] -> the list of names."""
    n = struct.unpack_from("<Q", rec, 0)[0]
    out, p = [], 8
    for _ in range(n):
        ln = struct.unpack_from("<I", rec, p)[0]
        out.append(rec[p + 4:p + 4 + ln])
        p += 4 + ln
    return out


def decode_batch_records(streams, coder: Coder | None = None, digests: bool = False):
    """Decode many names streams in one batched device pass.

    Returns one entry per stream: the names in the digest layout of
    crates/fqxv/src/container/block.rs:164-176 (``bytes``), or the
    :class:`Error` the reference would return for that stream.  With
    ``digests=True`` returns ``(records, digests)``, the digests (xxh3-64 of
    each stream's records, 0 for a failed stream) computed on the device.
    """
    c = coder or default_coder()
    L = c._L
    parts = [bytes(s) for s in streams]
    lens = np.array([len(s) for s in parts], dtype=np.uint64)
    offs = np.zeros(len(parts), dtype=np.uint64)
    if len(parts) > 1:
        offs[1:] = np.cumsum(lens[:-1])
    blob = b"".join(parts) or b"\0"
    sbuf = ctypes.create_string_buffer(blob, len(blob))
    h = ctypes.c_void_p()
    rc = L.fqxv_names_decode_batch(c._h, sbuf, ctypes.c_void_p(offs.ctypes.data), ctypes.c_void_p(lens.ctypes.data),
                                   len(parts), ctypes.byref(h))
    if rc:
        if rc == _lib.ERR_DEVICE:
            raise DeviceError(_lib.strerror(rc))
        raise ValueError(_lib.strerror(rc))
    try:
        out = []
        for i in range(len(parts)):
            p = ctypes.c_void_p()
            size = ctypes.c_uint64()
            n = ctypes.c_uint64()
            st = L.fqxv_names_records(h, i, ctypes.byref(p), ctypes.byref(size), ctypes.byref(n))
            if st:
                out.append(_error(st))
            else:
                out.append(ctypes.string_at(p, size.value))
        if not digests:
            return out
        dg = np.zeros(max(len(parts), 1), dtype=np.uint64)
        rc = L.fqxv_names_digests(c._h, h, ctypes.c_void_p(dg.ctypes.data))
        if rc:
            raise DeviceError(_lib.strerror(rc))
        return out, [int(x) for x in dg[: len(parts)]]
    finally:
        L.fqxv_names_batch_destroy(h)


def decode_batch(streams, coder: Coder | None = None, expected_digests=None):
    """One list of names (or an :class:`Error`) per stream.  With
    ``expected_digests`` (one stored names digest per stream, or None to
    skip one), each decoded stream is checked on the device as
    block.rs:1097-1109 checks it; a mismatch gives :class:`Corrupt`."""
    if expected_digests is None:
        return [r if isinstance(r, Error) else parse_records(r) for r in decode_batch_records(streams, coder)]
    recs, got = decode_batch_records(streams, coder, digests=True)
    out = []
    for r, d, want in zip(recs, got, expected_digests):
        if isinstance(r, Error):
            out.append(r)
        elif want is not None and d != int(want):
            out.append(Corrupt("block names digest"))
        else:
            out.append(parse_records(r))
    return out


def records(names) -> bytes:
    """The records layout ([u64 n] then per name [u32 len][bytes]) of a
    list of names: the input of :func:`encode_batch`."""
    out = [struct.pack("<Q", len(names))]
    for nm in names:
        b = bytes(nm)
        out.append(struct.pack("<I", len(b)))
        out.append(b)
    return b"".join(out)


def encode_batch_records(recs, coder: Coder | None = None):
    """lib.rs:455-540 ``encode`` of many blocks given as records buffers:
    host tokenizing, every sub-stream of every block encoded at both orders
    in one device launch (the smaller kept, order-0 on a tie), host framing.
    Returns one names stream (``bytes``) per block."""
    c = coder or default_coder()
    L = c._L
    parts = [bytes(r) for r in recs]
    lens = np.array([len(p) for p in parts], dtype=np.uint64)
    offs = np.zeros(len(parts), dtype=np.uint64)
    if len(parts) > 1:
        offs[1:] = np.cumsum(lens[:-1])
    blob = b"".join(parts) or b"\0"
    sbuf = ctypes.create_string_buffer(blob, len(blob))
    h = ctypes.c_void_p()
    rc = L.fqxv_names_encode_batch(c._h, sbuf, ctypes.c_void_p(offs.ctypes.data), ctypes.c_void_p(lens.ctypes.data),
                                   len(parts), ctypes.byref(h))
    if rc:
        if rc == _lib.ERR_DEVICE:
            raise DeviceError(_lib.strerror(rc))
        raise ValueError(_lib.strerror(rc))
    try:
        out = []
        for i in range(len(parts)):
            p = ctypes.c_void_p()
            size = ctypes.c_uint64()
            st = L.fqxv_names_output(h, i, ctypes.byref(p), ctypes.byref(size))
            if st == _lib.ERR_DEVICE:
                raise DeviceError(_lib.strerror(st))
            if st:
                raise ValueError(f"block {i}: {_lib.strerror(st)}")
            out.append(ctypes.string_at(p, size.value))
        return out
    finally:
        L.fqxv_names_batch_destroy(h)


def encode_batch(lists, coder: Coder | None = None) -> list[bytes]:
    """One names stream per list of names."""
    return encode_batch_records([records(n) for n in lists], coder)


def encode(names, coder: Coder | None = None) -> bytes:
    """lib.rs:455-540 ``encode(names)``."""
    return encode_batch([names], coder)[0]


def decode(src, coder: Coder | None = None) -> list[bytes]:
    """lib.rs:543-650 ``decode(src)``."""
    r = decode_batch([src], coder)[0]
    if isinstance(r, (Error, Corrupt)):
        raise r
    return r
