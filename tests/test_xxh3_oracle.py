"""The XXH3-64 restatement (oracle/xxh3_oracle.py) against the xxhash
package and the names digests stored in the reference's golden archives --
CPU only."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

xxhash = pytest.importorskip("xxhash")


@pytest.fixture(scope="module")
def xo():
    from oracle import xxh3_oracle
    return xxh3_oracle


def test_every_short_form_and_long_boundaries(xo):
    """All of 0..1300 (the 0-3, 4-8, 9-16, 17-128, 129-240 short forms and
    the long loop's stripe and block edges), then multi-block sizes."""
    rng = np.random.default_rng(5)
    lens = list(range(0, 1301)) + [2047, 2048, 2049, 3072, 4097, 65536, 100003]
    for n in lens:
        b = rng.integers(0, 256, n).astype(np.uint8).tobytes()
        assert xo.xxh3_64(b) == xxhash.xxh3_64_intdigest(b), n


def test_fixture_names_digests(xo, oracle):
    """The names digest of each golden block (block.rs:164-176) over the
    names the names oracle decodes."""
    from oracle import names_oracle as no
    v = json.load(open(os.path.join(GOLDEN, "names_vectors.json")))
    blob = open(os.path.join(GOLDEN, "names_streams.bin"), "rb").read()
    for e in v:
        names = no.decode(blob[e["off"]:e["off"] + e["len"]])
        assert "%016x" % xo.xxh3_64(no.digest_input(names)) == e["names_digest"]


def test_plants_are_caught(xo):
    """A wrong secret byte or a dropped scramble must change the digest."""
    b = bytes(range(256)) * 20
    good = xo.xxh3_64(b)
    sec = xo.SECRET
    try:
        xo.SECRET = sec[:130] + bytes([sec[130] ^ 1]) + sec[131:]
        assert xo.xxh3_64(b) != good
    finally:
        xo.SECRET = sec
    assert xo.xxh3_64(b) == good
