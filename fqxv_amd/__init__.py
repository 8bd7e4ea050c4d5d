"""fqxv_amd -- MI355X (gfx950) rANS Nx16 coder, a drop-in for fqxv-rans.

The product is the C-ABI library ``fqxv_amd/lib/libfqxv_rans_hip.so`` built
from ``fqxv_amd/csrc`` (hand-written HIP kernels for gfx950); this package is
its Python front-end mirroring crates/fqxv-rans/src/lib.rs.
"""
from .rans import (  # noqa: F401
    ALLOC_LIMIT, Backend, BatchResult, Coder, DeviceError, Error, Graph, InputTooLarge,
    Malformed, Order, decode, decode_bounded, decode_with, default_coder, encode,
    peek_output_len,
)

__all__ = [
    "Backend", "BatchResult", "Coder", "DeviceError", "Error", "Graph", "InputTooLarge",
    "Malformed", "Order", "decode", "decode_bounded", "decode_with", "encode",
    "peek_output_len", "default_coder", "ALLOC_LIMIT",
]
