"""The reorder permutation's byte planes, on the device.

crates/fqxv/src/container/reorder.rs stores the global permutation (n u32
read indices) as four byte planes of n bytes -- plane k holds byte k of every
index -- and rANS-codes the 4n bytes at the smaller order (order-0 on a tie,
reorder.rs:597-607); decoding is bounded by 4n and must give exactly 4n
bytes (reorder.rs:1089-1100).  Here the transposes are the library's
``fqxv_rans_perm_split`` / ``_join`` kernels and the coding is the batched
device coder; errors carry the container's messages.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .rans import Coder, Malformed, _raise, default_coder


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def split(perm, planes=None, stream=None, coder: Coder | None = None):
    """Device u32 indices (torch int32/uint32 tensor) -> 4n-byte planes."""
    import torch
    c = coder or default_coder()
    n = perm.numel()
    if planes is None:
        planes = torch.empty(4 * n + 16, dtype=torch.uint8, device=perm.device)
    _raise(c._L.fqxv_rans_perm_split(_ptr(perm), n, _ptr(planes), Coder._stream(stream)))
    return planes


def join(planes, n: int, perm=None, stream=None, coder: Coder | None = None):
    """4n-byte planes (device) -> u32 indices as a torch int32 tensor."""
    import torch
    c = coder or default_coder()
    if perm is None:
        perm = torch.empty(max(n, 1), dtype=torch.int32, device=planes.device)
    _raise(c._L.fqxv_rans_perm_join(_ptr(planes), n, _ptr(perm), Coder._stream(stream)))
    return perm[:n]


def encode_permutation(perm, coder: Coder | None = None) -> bytes:
    """reorder.rs:597-607 on the device: planes, then the shorter order
    (order-0 on a tie).  ``perm``: host or device u32 indices."""
    import torch
    c = coder or default_coder()
    p = perm if isinstance(perm, torch.Tensor) else torch.from_numpy(
        np.ascontiguousarray(perm, dtype=np.uint32).view(np.int32))
    p = p.to(torch.device("cuda", c.device)).contiguous()
    n = p.numel()
    planes = split(p, coder=c)
    r = c.encode_best_batch(planes, [0], [4 * n], 0)
    torch.cuda.synchronize()
    _raise(int(r.status[0]))
    return r.dst[: int(r.out_len[0])].cpu().numpy().tobytes()


def decode_permutation(stream: bytes, n: int, coder: Coder | None = None):
    """reorder.rs:1089-1100 on the device; returns the indices as a device
    int32 tensor.  Any decode error is "bad permutation"; a stream that is
    not exactly 4n bytes long is "permutation length mismatch"."""
    import torch
    c = coder or default_coder()
    dev = torch.device("cuda", c.device)
    b = bytes(stream)
    s = torch.frombuffer(bytearray(b + bytes(8)), dtype=torch.uint8).to(dev)
    r = c.decode_batch(s, [0], [len(b)], [4 * n], max_len=[4 * n])
    torch.cuda.synchronize()
    if int(r.status[0]) != _lib.OK:
        raise Malformed("bad permutation")
    if int(r.out_len[0]) != 4 * n:
        raise Malformed("permutation length mismatch")
    return join(r.dst, n, coder=c)
