"""ctypes front-end to the C oracle (oracle/rans_oracle.c) -- TEST INFRASTRUCTURE.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module, and only as the checker / CPU baseline.  The product
package ``fqxv_amd`` never imports it.

The functions mirror the reference API in crates/fqxv-rans/src/lib.rs:
``encode`` (lib.rs:111-127), ``decode`` (lib.rs:186-200), ``decode_bounded``
(lib.rs:226-231), plus ``normalize`` (model.rs:213-264) and the ``EncSym``
map (model.rs:70-133) for the model-level tests.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "librans_oracle.so")

OK = 0
ERR_TRUNCATED = 1
ERR_UNKNOWN_ORDER = 2
ERR_FREQ_EXCEEDS = 3
ERR_FREQ_SUM = 4
ERR_OUTPUT_ALLOC = 5
ERR_ABSENT_CONTEXT = 6
ERR_BOUND = 7
ERR_DST_TOO_SMALL = 8

MESSAGES = {
    ERR_TRUNCATED: "truncated stream",
    ERR_UNKNOWN_ORDER: "unknown order tag",
    ERR_FREQ_EXCEEDS: "rANS freq table exceeds TOTFREQ",
    ERR_FREQ_SUM: "rANS freq table does not sum to TOTFREQ",
    ERR_OUTPUT_ALLOC: "rANS output length too large to allocate",
    ERR_ABSENT_CONTEXT: "symbol references absent context",
    ERR_BOUND: "rANS output length exceeds caller bound",
}


class OracleError(Exception):
    def __init__(self, code: int):
        self.code = code
        super().__init__(MESSAGES.get(code, f"oracle error {code}"))


def build(force: bool = False) -> str:
    """Compile the oracle with the recipe in oracle/Makefile."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def _host_level() -> int:
    """Highest x86-64 micro-architecture level of the running CPU (2, 3 or 4)."""
    try:
        flags = set()
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    flags = set(line.split(":", 1)[1].split())
                    break
    except OSError:
        return 2
    v3 = {"avx", "avx2", "bmi1", "bmi2", "f16c", "fma", "movbe", "xsave"}
    v4 = {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"}
    if not v3 <= flags:
        return 2
    return 4 if v4 <= flags else 3


def baseline_build() -> tuple[str, str]:
    """The oracle build the CPU baseline times: the one compiled for the
    highest level this host supports (path, compiler flags)."""
    level = _host_level()
    while level > 2:
        path = os.path.join(HERE, "_build", f"librans_oracle_v{level}.so")
        if os.path.exists(path):
            return path, f"-O3 -march=x86-64-v{level}"
        level -= 1
    return LIB_PATH, "-O3 -march=x86-64-v2"


def _declare(L):
    c_sz = ctypes.c_size_t
    p8 = ctypes.c_void_p
    L.orc_encode_bound.restype = c_sz
    L.orc_encode_bound.argtypes = [c_sz, ctypes.c_int]
    L.orc_encode.argtypes = [p8, c_sz, ctypes.c_int, p8, c_sz, ctypes.POINTER(c_sz)]
    L.orc_decode.argtypes = [p8, c_sz, p8, c_sz, ctypes.POINTER(c_sz), ctypes.POINTER(c_sz)]
    L.orc_decode_bounded.argtypes = [p8, c_sz, c_sz, p8, c_sz, ctypes.POINTER(c_sz)]
    L.orc_peek_len.argtypes = [p8, c_sz, ctypes.POINTER(ctypes.c_uint64)]
    L.orc_normalize.argtypes = [p8, p8]
    L.orc_validate_freqs.argtypes = [p8]
    L.orc_encsym_apply.restype = ctypes.c_uint32
    L.orc_encsym_apply.argtypes = [p8, ctypes.c_uint32]
    L.orc_encsym_new.argtypes = [ctypes.c_uint32, ctypes.c_uint32, p8]
    for fn in (L.orc_encode_many, L.orc_decode_many):
        fn.restype = ctypes.c_int
    L.orc_encode_many.argtypes = [p8, p8, p8, p8, c_sz, p8, p8, p8, p8, p8, ctypes.c_int]
    L.orc_decode_many.argtypes = [p8, p8, p8, c_sz, p8, p8, p8, p8, p8, ctypes.c_int]
    L.orc_crc32c.restype = ctypes.c_uint32
    L.orc_crc32c.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
    L.orc_crc32c_combine.restype = ctypes.c_uint32
    L.orc_crc32c_combine.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64]
    L.orc_crc32c_many.restype = None
    L.orc_crc32c_many.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_uint32, ctypes.c_void_p]
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _declare(ctypes.CDLL(LIB_PATH))
    return _lib


_baseline = None


def baseline_lib():
    """The oracle built for this host's highest x86-64 level (CPU baseline)."""
    global _baseline
    if _baseline is None:
        path, flags = baseline_build()
        _baseline = (_declare(ctypes.CDLL(path)), flags)
    return _baseline


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


def _as_u8(data) -> np.ndarray:
    if isinstance(data, np.ndarray):
        return np.ascontiguousarray(data, dtype=np.uint8)
    return np.frombuffer(bytes(data), dtype=np.uint8)


def encode_bound(n: int, order: int) -> int:
    return int(lib().orc_encode_bound(n, order))


def encode(data, order: int) -> bytes:
    src = _as_u8(data)
    cap = encode_bound(src.size, order)
    dst = np.empty(cap, dtype=np.uint8)
    out = ctypes.c_size_t(0)
    rc = lib().orc_encode(_ptr(src) if src.size else None, src.size, order, _ptr(dst), cap,
                          ctypes.byref(out))
    if rc:
        raise OracleError(rc)
    return dst[: out.value].tobytes()


def encode_best(data, tie_order: int = 0) -> bytes:
    """Both orders, the smaller kept; ``tie_order`` wins a tie.  Follows the
    callers: crates/fqxv-tokenizer/src/lib.rs:91-95 and
    crates/fqxv/src/container/reorder.rs:597-607 (order-0 on a tie),
    crates/fqxv-lroverlap/src/codec.rs:292-303 (order-1 on a tie)."""
    e0 = encode(data, 0)
    e1 = encode(data, 1)
    if tie_order == 1:
        return e1 if len(e1) <= len(e0) else e0
    return e1 if len(e1) < len(e0) else e0


def decode_ex(data, cap: int | None = None):
    """Return (status, output bytes, consumed stream bytes)."""
    src = _as_u8(data)
    n = peek_len(src)
    if cap is None:
        cap = n if (n is not None and n <= (1 << 31)) else 0
    dst = np.empty(max(cap, 1), dtype=np.uint8)
    out = ctypes.c_size_t(0)
    used = ctypes.c_size_t(0)
    rc = lib().orc_decode(_ptr(src) if src.size else None, src.size, _ptr(dst), cap,
                          ctypes.byref(out), ctypes.byref(used))
    return rc, dst[: out.value].tobytes(), used.value


def decode(data, cap: int | None = None) -> bytes:
    rc, out, _ = decode_ex(data, cap)
    if rc:
        raise OracleError(rc)
    return out


def decode_bounded(data, max_len: int) -> bytes:
    src = _as_u8(data)
    n = peek_len(src)
    if n is not None and n > max_len:
        raise OracleError(ERR_BOUND)
    return decode(src)


def peek_len(data) -> int | None:
    src = _as_u8(data)
    v = ctypes.c_uint64(0)
    rc = lib().orc_peek_len(_ptr(src) if src.size else None, src.size, ctypes.byref(v))
    return None if rc else int(v.value)


def normalize(counts) -> np.ndarray:
    c = np.ascontiguousarray(counts, dtype=np.uint32)
    assert c.shape == (256,)
    f = np.zeros(256, dtype=np.uint16)
    lib().orc_normalize(_ptr(c), _ptr(f))
    return f


def validate_freqs(freq) -> int:
    f = np.ascontiguousarray(freq, dtype=np.uint16)
    return int(lib().orc_validate_freqs(_ptr(f)))


class _EncSym(ctypes.Structure):
    _fields_ = [("x_max", ctypes.c_uint64), ("rcp_freq", ctypes.c_uint32),
                ("rcp_shift", ctypes.c_uint32), ("bias", ctypes.c_uint32),
                ("cmpl_freq", ctypes.c_uint32), ("freq", ctypes.c_uint32),
                ("cum", ctypes.c_uint32), ("use_div", ctypes.c_int)]


def encsym(freq: int, cum: int) -> _EncSym:
    e = _EncSym()
    lib().orc_encsym_new(freq, cum, ctypes.byref(e))
    return e


def encsym_apply(e: _EncSym, v: int) -> int:
    return int(lib().orc_encsym_apply(ctypes.byref(e), v))


def encode_many(arena: np.ndarray, offs, lens, orders, threads: int):
    """Encode many blocks over `threads` pthreads; returns (dst, dst_off, out_len, status)."""
    offs = np.ascontiguousarray(offs, dtype=np.uint64)
    lens = np.ascontiguousarray(lens, dtype=np.uint64)
    orders = np.ascontiguousarray(orders, dtype=np.int32)
    caps = np.array([encode_bound(int(n), int(o)) for n, o in zip(lens, orders)], dtype=np.uint64)
    doff = np.zeros_like(caps)
    if caps.size:
        doff[1:] = np.cumsum(caps)[:-1]
    dst = np.empty(int(caps.sum()) + 1, dtype=np.uint8)
    out_len = np.zeros(len(lens), dtype=np.uint64)
    status = np.zeros(len(lens), dtype=np.int32)
    lib().orc_encode_many(_ptr(arena), _ptr(offs), _ptr(lens), _ptr(orders), len(lens), _ptr(dst),
                          _ptr(doff), _ptr(caps), _ptr(out_len), _ptr(status), threads)
    return dst, doff, out_len, status


def decode_many(arena: np.ndarray, offs, lens, out_caps, threads: int, use=None):
    offs = np.ascontiguousarray(offs, dtype=np.uint64)
    lens = np.ascontiguousarray(lens, dtype=np.uint64)
    caps = np.ascontiguousarray(out_caps, dtype=np.uint64)
    doff = np.zeros_like(caps)
    if caps.size:
        doff[1:] = np.cumsum(caps)[:-1]
    dst = np.empty(int(caps.sum()) + 1, dtype=np.uint8)
    out_len = np.zeros(len(lens), dtype=np.uint64)
    status = np.zeros(len(lens), dtype=np.int32)
    (use or lib()).orc_decode_many(_ptr(arena), _ptr(offs), _ptr(lens), len(lens), _ptr(dst), _ptr(doff),
                          _ptr(caps), _ptr(out_len), _ptr(status), threads)
    return dst, doff, out_len, status


# ---- reorder permutation byte planes (crates/fqxv/src/container/reorder.rs) ----
class PermError(Exception):
    """The container's `Error::Malformed` for a permutation payload."""


def perm_split(perm) -> np.ndarray:
    """reorder.rs:597-604: plane k (bytes [k*n, (k+1)*n)) = byte k of each index."""
    p = np.ascontiguousarray(perm, dtype="<u4")
    return p.view(np.uint8).reshape(-1, 4).T.copy().reshape(-1)


def perm_join(planes, n: int) -> np.ndarray:
    """reorder.rs:1097-1099: index i = LE bytes (planes[i], [n+i], [2n+i], [3n+i])."""
    b = np.ascontiguousarray(planes, dtype=np.uint8)[: 4 * n].reshape(4, n).astype(np.uint32)
    return b[0] | (b[1] << 8) | (b[2] << 16) | (b[3] << 24)


def encode_permutation(perm) -> bytes:
    """reorder.rs:597-607: planes, then the smaller rANS order (order-0 on a tie)."""
    return encode_best(perm_split(perm).tobytes(), 0)


def decode_permutation(stream: bytes, n: int) -> np.ndarray:
    """reorder.rs:1089-1100: decode bounded by 4n, exact length, join."""
    try:
        pb = decode_bounded(stream, 4 * n)
    except OracleError:
        raise PermError("bad permutation")
    if len(pb) != 4 * n:
        raise PermError("permutation length mismatch")
    return perm_join(np.frombuffer(pb, dtype=np.uint8), n)


# ---- CRC-32C (crates/fqxv/src/crc.rs) ----------------------------------------
def crc32c(data) -> int:
    b = _as_u8(data)
    return int(lib().orc_crc32c(_ptr(b) if b.size else None, b.size))


def crc32c_combine(crc_a: int, crc_b: int, len_b: int) -> int:
    return int(lib().orc_crc32c_combine(crc_a, crc_b, len_b))


def crc32c_many(arena: np.ndarray, offs, lens) -> np.ndarray:
    offs = np.ascontiguousarray(offs, dtype=np.uint64)
    lens = np.ascontiguousarray(lens, dtype=np.uint64)
    out = np.zeros(len(offs), dtype=np.uint32)
    lib().orc_crc32c_many(_ptr(arena), _ptr(offs), _ptr(lens), len(offs), _ptr(out))
    return out
