"""The C-ABI library loads and exports what include/fqxv_rans_hip.h declares
(no device calls: these run without a GPU)."""
import ctypes
import os
import re
import struct

import pytest

from conftest import ROOT


def header_functions():
    text = open(os.path.join(ROOT, "include", "fqxv_rans_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fqxv_(?:rans|names)_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def L():
    from fqxv_amd import _lib
    return _lib.load()


def test_exports_every_declared_symbol(L):
    from fqxv_amd import _lib
    declared = header_functions()
    assert declared == sorted(_lib.EXPORTS)
    for name in declared:
        assert hasattr(L, name), name


def test_job_struct_layout():
    from fqxv_amd import _lib
    import numpy as np
    assert ctypes.sizeof(_lib.Job) == 48
    assert np.dtype(_lib.JOB_DTYPE_FIELDS).itemsize == 48


def test_strerror_matches_reference_messages(L, oracle):
    from fqxv_amd import _lib
    for code, msg in oracle.MESSAGES.items():
        assert _lib.strerror(code) == msg


def test_encode_bound(L, oracle):
    for n in (0, 1, 31, 32, 1000, 1 << 20):
        for order in (0, 1):
            b = L.fqxv_rans_encode_bound(n, order)
            assert b >= oracle.encode_bound(n, order)
            assert b == (9 if n == 0 else oracle.encode_bound(n, order) + 16)


def test_peek_len(L):
    v = ctypes.c_uint64(0)
    buf = ctypes.create_string_buffer(b"\x01" + struct.pack("<Q", 12345), 9)
    assert L.fqxv_rans_peek_len(buf, 9, ctypes.byref(v)) == 0 and v.value == 12345
    assert L.fqxv_rans_peek_len(buf, 8, ctypes.byref(v)) == 1


def test_lpt_order(L):
    import numpy as np
    from fqxv_amd.rans import Coder
    s = np.array([5, 9, 1, 9, 3], dtype=np.uint64)
    o = np.empty(5, dtype=np.uint32)
    L.fqxv_rans_lpt_order(ctypes.c_void_p(s.ctypes.data), 5, ctypes.c_void_p(o.ctypes.data))
    assert list(o) == [1, 3, 0, 4, 2]


def test_product_never_imports_oracle():
    """The shipped package must not reach the checker."""
    pkg = os.path.join(ROOT, "fqxv_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h", ".c")):
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in text.lower().replace("the oracle", ""), f


def test_vectorised_encode_bounds_match_the_library(L):
    """Coder.encode_bounds / encode_best_bounds (numpy, used for whole
    batches) equal fqxv_rans_encode_bound / _encode_best_bound per block."""
    import numpy as np
    from fqxv_amd.rans import Coder
    rng = np.random.default_rng(5)
    ns = [0, 1, 2, 8, 9, 255, 256, 257, 1000, 1 << 20, (1 << 32) - 1, 1 << 33]
    ns += [int(x) for x in rng.integers(0, 1 << 26, 200)]
    for order in (0, 1):
        got = Coder.encode_bounds(ns, [order] * len(ns))
        assert [int(v) for v in got] == [int(L.fqxv_rans_encode_bound(n, order)) for n in ns]
    mixed = [int(o) for o in rng.integers(0, 2, len(ns))]
    got = Coder.encode_bounds(ns, mixed)
    assert [int(v) for v in got] == [int(L.fqxv_rans_encode_bound(n, o)) for n, o in zip(ns, mixed)]
    got = Coder.encode_best_bounds(ns)
    assert [int(v) for v in got] == [int(L.fqxv_rans_encode_best_bound(n)) for n in ns]


C_DRIVER = os.path.join(ROOT, "tests", "c_driver", "_build", "abi_driver")


def test_c_driver_links_against_the_library():
    """The plain-C driver (tests/c_driver, built by __graft_entry__.build())
    links libfqxv_rans_hip.so and the HIP runtime and resolves the host-only
    entry points without a device."""
    import subprocess
    assert os.path.exists(C_DRIVER), "tests/c_driver is not built (run __graft_entry__.build())"
    r = subprocess.run([C_DRIVER, "--probe"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "probe ok" in r.stdout, r.stdout + r.stderr
