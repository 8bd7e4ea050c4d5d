"""ctypes binding of libfqxv_rans_hip.so (include/fqxv_rans_hip.h).

The library is built in-tree by ``python -m fqxv_amd.build_lib`` (or
``__graft_entry__.build()``).  Loading fails loudly when it is missing: there
is no CPU fallback behind this package.
"""
from __future__ import annotations

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libfqxv_rans_hip.so")

OK = 0
ERR_TRUNCATED = 1
ERR_UNKNOWN_ORDER = 2
ERR_FREQ_EXCEEDS = 3
ERR_FREQ_SUM = 4
ERR_OUTPUT_ALLOC = 5
ERR_ABSENT_CONTEXT = 6
ERR_BOUND = 7
ERR_DST_TOO_SMALL = 8
ERR_INVALID_ARG = 9
JOB_O1_ROWS_WORKSPACE = 1  # fqxv_rans_job.flags: order-1 code rows in the workspace, not LDS
ERR_DEVICE = 10
ERR_SCRATCH_LIMIT = 11
BACKEND_GFX950 = 0
BACKEND_GFX950_BATCHED = 1
NAMES_ERR_FIRST = 32

# Every symbol include/fqxv_rans_hip.h declares.
EXPORTS = [
    "fqxv_rans_strerror", "fqxv_rans_encode_bound", "fqxv_rans_peek_len",
    "fqxv_rans_ctx_create", "fqxv_rans_ctx_destroy", "fqxv_rans_encode",
    "fqxv_rans_decode", "fqxv_rans_decode_bounded", "fqxv_rans_decode_batch",
    "fqxv_rans_encode_batch", "fqxv_rans_decode_one_device",
    "fqxv_rans_encode_one_device", "fqxv_rans_lpt_order", "fqxv_rans_max_waves",
    "fqxv_rans_graph_create", "fqxv_rans_graph_launch", "fqxv_rans_graph_destroy",
    "fqxv_rans_ctx_set_grid", "fqxv_rans_server_create", "fqxv_rans_server_destroy",
    "fqxv_rans_server_now", "fqxv_rans_server_submit",
    "fqxv_rans_server_poll", "fqxv_rans_server_take", "fqxv_rans_server_release", "fqxv_rans_server_waves",
    "fqxv_rans_encode_best_batch", "fqxv_rans_encode_best_bound",
    "fqxv_rans_perm_split", "fqxv_rans_perm_join", "fqxv_rans_crc32c_batch",
    "fqxv_names_decode_batch", "fqxv_names_status", "fqxv_names_records",
    "fqxv_names_batch_destroy", "fqxv_names_digests", "fqxv_rans_xxh3_batch",
    "fqxv_names_encode_batch", "fqxv_names_output", "fqxv_rans_server_create_persistent",
    "fqxv_rans_xxh3_concat", "fqxv_rans_decode_with", "fqxv_rans_backend_detect",
    "fqxv_rans_decode_batch_noalloc", "fqxv_rans_noalloc_max_waves",
]


class Job(ctypes.Structure):
    """fqxv_rans_job (48 bytes)."""
    _fields_ = [("src_off", ctypes.c_uint64), ("src_len", ctypes.c_uint64),
                ("dst_off", ctypes.c_uint64), ("dst_cap", ctypes.c_uint64),
                ("max_len", ctypes.c_uint64), ("order", ctypes.c_int32),
                ("flags", ctypes.c_uint32)]


class Done(ctypes.Structure):
    """fqxv_rans_done (56 bytes)."""
    _fields_ = [("ticket", ctypes.c_uint64), ("out_off", ctypes.c_uint64),
                ("out_bytes", ctypes.c_uint64), ("t_arrival", ctypes.c_double),
                ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double),
                ("status", ctypes.c_int32), ("nblocks", ctypes.c_uint32)]


JOB_DTYPE_FIELDS = [("src_off", "<u8"), ("src_len", "<u8"), ("dst_off", "<u8"),
                    ("dst_cap", "<u8"), ("max_len", "<u8"), ("order", "<i4"),
                    ("flags", "<u4")]

_lib = None
_lock = threading.Lock()


def load():
    """Load the library (raises OSError when it has not been built)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is not built; run python -m fqxv_amd.build_lib")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        sz = ctypes.c_size_t
        L.fqxv_rans_strerror.restype = ctypes.c_char_p
        L.fqxv_rans_strerror.argtypes = [ctypes.c_int]
        L.fqxv_rans_encode_bound.restype = sz
        L.fqxv_rans_encode_bound.argtypes = [sz, ctypes.c_int]
        L.fqxv_rans_peek_len.argtypes = [vp, sz, ctypes.POINTER(ctypes.c_uint64)]
        L.fqxv_rans_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
        L.fqxv_rans_ctx_destroy.argtypes = [vp]
        L.fqxv_rans_ctx_destroy.restype = None
        L.fqxv_rans_encode.argtypes = [vp, vp, sz, ctypes.c_int, vp, sz, ctypes.POINTER(sz)]
        L.fqxv_rans_decode.argtypes = [vp, vp, sz, vp, sz, ctypes.POINTER(sz)]
        L.fqxv_rans_decode_bounded.argtypes = [vp, vp, sz, sz, vp, sz, ctypes.POINTER(sz)]
        L.fqxv_rans_decode_with.argtypes = [vp, ctypes.c_int, vp, sz, vp, sz, ctypes.POINTER(sz)]
        L.fqxv_rans_backend_detect.argtypes = []
        L.fqxv_rans_decode_batch_noalloc.argtypes = [vp, vp, vp, ctypes.c_uint32, vp, vp, vp, ctypes.c_uint32, vp]
        L.fqxv_rans_noalloc_max_waves.argtypes = []
        L.fqxv_rans_noalloc_max_waves.restype = ctypes.c_uint32
        L.fqxv_rans_encode_best_bound.restype = sz
        L.fqxv_rans_encode_best_bound.argtypes = [sz]
        for fn in (L.fqxv_rans_decode_batch, L.fqxv_rans_encode_batch, L.fqxv_rans_encode_best_batch):
            fn.argtypes = [vp, vp, vp, vp, ctypes.c_uint32, vp, vp, vp, vp]
        for fn in (L.fqxv_rans_decode_one_device, L.fqxv_rans_encode_one_device):
            fn.argtypes = [vp, vp, vp, ctypes.POINTER(Job), vp, vp, vp]
        L.fqxv_rans_lpt_order.argtypes = [vp, ctypes.c_uint32, vp]
        L.fqxv_rans_lpt_order.restype = None
        L.fqxv_rans_max_waves.argtypes = [vp, ctypes.POINTER(ctypes.c_uint32)]
        L.fqxv_rans_graph_create.argtypes = [vp, ctypes.c_int, vp, vp, vp, ctypes.c_uint32, vp, vp, vp,
                                             ctypes.POINTER(vp)]
        L.fqxv_rans_graph_launch.argtypes = [vp, vp]
        L.fqxv_rans_graph_destroy.argtypes = [vp]
        L.fqxv_rans_graph_destroy.restype = None
        L.fqxv_rans_ctx_set_grid.argtypes = [vp, ctypes.c_uint32]
        L.fqxv_rans_perm_split.argtypes = [vp, ctypes.c_uint64, vp, vp]
        L.fqxv_rans_perm_join.argtypes = [vp, ctypes.c_uint64, vp, vp]
        L.fqxv_rans_crc32c_batch.argtypes = [vp, vp, vp, vp, ctypes.c_uint32, vp, vp]
        u32 = ctypes.c_uint32
        L.fqxv_rans_server_create.argtypes = [ctypes.c_int, vp, vp, u32, vp, ctypes.c_uint64, u32, u32,
                                              ctypes.POINTER(vp)]
        L.fqxv_rans_server_create_persistent.argtypes = [ctypes.c_int, vp, vp, u32, vp, ctypes.c_uint64, u32, u32,
                                                         u32, ctypes.POINTER(vp)]
        L.fqxv_rans_server_destroy.argtypes = [vp]
        L.fqxv_rans_server_destroy.restype = None
        L.fqxv_rans_server_now.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
        L.fqxv_rans_server_submit.argtypes = [vp, vp, u32, ctypes.c_double, ctypes.POINTER(ctypes.c_uint64)]
        L.fqxv_rans_server_poll.argtypes = [vp, ctypes.POINTER(u32), ctypes.POINTER(u32)]
        L.fqxv_rans_server_take.argtypes = [vp, ctypes.POINTER(Done), u32, ctypes.POINTER(u32)]
        L.fqxv_rans_server_release.argtypes = [vp, ctypes.c_uint64]
        L.fqxv_rans_server_waves.argtypes = [vp, ctypes.POINTER(u32)]
        L.fqxv_names_decode_batch.argtypes = [vp, vp, vp, vp, u32, ctypes.POINTER(vp)]
        L.fqxv_names_status.argtypes = [vp, u32]
        L.fqxv_names_status.restype = ctypes.c_int32
        L.fqxv_names_records.argtypes = [vp, u32, ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_uint64),
                                         ctypes.POINTER(ctypes.c_uint64)]
        L.fqxv_names_batch_destroy.argtypes = [vp]
        L.fqxv_names_batch_destroy.restype = None
        L.fqxv_names_digests.argtypes = [vp, vp, vp]
        L.fqxv_names_encode_batch.argtypes = [vp, vp, vp, vp, u32, ctypes.POINTER(vp)]
        L.fqxv_names_output.argtypes = [vp, u32, ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_uint64)]
        L.fqxv_rans_xxh3_batch.argtypes = [vp, vp, vp, vp, u32, vp, vp]
        L.fqxv_rans_xxh3_concat.argtypes = [vp, vp, vp, vp, vp, u32, vp, vp]
        _lib = L
        return L


def strerror(code: int) -> str:
    return load().fqxv_rans_strerror(code).decode()
