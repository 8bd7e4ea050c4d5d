"""GPU: the C ABI driven from plain C (tests/c_driver/abi_driver.c), the way
the Rust FFI crate would call it -- no ctypes, no torch in the process that
codes.  Cases: the reference-written fixture streams (re-encode identity and
decode), the reference's edge cases, and the 1 MiB binned-quality order-0
block (seed 1) at both orders, each with the oracle's stream."""
import os
import struct
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

C_DRIVER = os.path.join(ROOT, "tests", "c_driver", "_build", "abi_driver")


def test_c_driver_encodes_and_decodes_fixture_and_workload_streams(oracle, fixture_streams, tmp_path):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from fqxv_amd import synth
    cases = [(e["order"], oracle.decode(s), s) for e, s in fixture_streams]
    for data in (b"", b"Q", b"AB", b"A" * 1000, bytes(range(256)) * 40, bytes(range(33))):
        for order in (0, 1):
            cases.append((order, data, oracle.encode(data, order)))
    c1 = synth.workload("binned_1mib").block(0).tobytes()
    for order in (0, 1):
        cases.append((order, c1, oracle.encode(c1, order)))
    path = tmp_path / "cases.bin"
    with open(path, "wb") as f:
        f.write(b"FQXVDRV1" + struct.pack("<I", len(cases)))
        for order, raw, enc in cases:
            f.write(struct.pack("<IQ", order, len(raw)) + raw + struct.pack("<Q", len(enc)) + enc)
    print(f"c driver: {len(cases)} cases", flush=True)
    r = subprocess.run([C_DRIVER, str(path)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"{len(cases)} cases ok" in r.stdout
