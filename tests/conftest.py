import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X / gfx950)")
    config.addinivalue_line("markers", "slow: long-running parity case")


@pytest.fixture(scope="session")
def fixture_streams():
    """rANS streams written by the reference encoder, extracted from the
    reference's golden archives (tests/golden/make_fixture_streams.py)."""
    index = json.load(open(os.path.join(GOLDEN, "fixture_streams.json")))
    blob = open(os.path.join(GOLDEN, "fixture_streams.bin"), "rb").read()
    out = []
    for e in index:
        s = blob[e["blob_offset"]:e["blob_offset"] + e["stream_len"]]
        out.append((e, s))
    return out


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o
    o.lib()
    return o


@pytest.fixture(scope="session")
def coder():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import fqxv_amd
    return fqxv_amd.Coder(0)


def rng_bytes(seed: int, n: int, alphabet: int = 256) -> bytes:
    r = np.random.default_rng(seed)
    return r.integers(0, alphabet, n, dtype=np.int64).astype(np.uint8).tobytes()
