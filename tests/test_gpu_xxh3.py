"""XXH3-64 on the device (fqxv_rans_xxh3_batch, fqxv_names_digests) against
the xxhash package and the golden archives' names digests."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu
xxhash = pytest.importorskip("xxhash")


def test_ranges_every_length_and_alignment(coder):
    import torch
    rng = np.random.default_rng(9)
    arena = rng.integers(0, 256, 6 << 20).astype(np.uint8)
    lens = list(range(0, 1301)) + [2047, 2048, 2049, 3072, 4097, 65536, 100003, 1 << 20, (3 << 20) + 5]
    offs = [int(rng.integers(0, len(arena) - n)) for n in lens]
    d = torch.from_numpy(arena).cuda()
    out = coder.xxh3_batch(d, offs, lens).cpu().numpy().view(np.uint64)
    for o, n, h in zip(offs, lens, out):
        assert int(h) == xxhash.xxh3_64_intdigest(arena[o:o + n].tobytes()), (o, n)


def test_large_single_range(coder):
    import torch
    rng = np.random.default_rng(10)
    arena = rng.integers(0, 256, (64 << 20) + 777).astype(np.uint8)
    d = torch.from_numpy(arena).cuda()
    h = coder.xxh3_batch(d, [3], [len(arena) - 3]).cpu().numpy().view(np.uint64)[0]
    assert int(h) == xxhash.xxh3_64_intdigest(arena[3:].tobytes())


def test_names_digests_match_fixture_blocks(coder):
    from fqxv_amd import names
    v = json.load(open(os.path.join(GOLDEN, "names_vectors.json")))
    blob = open(os.path.join(GOLDEN, "names_streams.bin"), "rb").read()
    streams = [blob[e["off"]:e["off"] + e["len"]] for e in v]
    recs, dg = names.decode_batch_records(streams, coder, digests=True)
    assert ["%016x" % x for x in dg] == [e["names_digest"] for e in v]
    want = [int(e["names_digest"], 16) for e in v]
    ok = names.decode_batch(streams, coder, expected_digests=want)
    assert all(isinstance(r, list) and len(r) == e["n_reads"] for r, e in zip(ok, v))
    bad = names.decode_batch(streams, coder, expected_digests=[want[0] ^ 1, None, want[2]])
    assert isinstance(bad[0], names.Corrupt) and str(bad[0]) == "block names digest"
    assert isinstance(bad[1], list) and isinstance(bad[2], list)


def test_names_digests_of_failed_and_large_streams(coder, oracle):
    from fqxv_amd import names, synth
    from oracle import names_oracle as no
    big = synth.make_names("illumina", 50000, 3)
    streams = [no.encode(big), b"\x02", no.encode([])]
    recs, dg = names.decode_batch_records(streams, coder, digests=True)
    assert dg[0] == no.names_digest(big)
    assert dg[1] == 0 and isinstance(recs[1], names.Error)
    assert dg[2] == no.names_digest([])


def test_concat_digest_block_layout(coder):
    """fqxv_rans_xxh3_concat on the sequence/quality digest layout of
    crates/fqxv/src/container/block.rs:164-196: XXH3-64 over
    [n_reads u64][lens u32...][bytes], the parts placed apart in the arena
    (header, lengths and bytes at unaligned offsets), checked against the
    xxhash package on the same bytes; plus empty segments, an empty string,
    segments longer than one 64 KiB staging piece, and many strings in one
    call."""
    import torch
    rng = np.random.default_rng(31)
    pieces, cases = [], []
    pos = 0

    def put(b):
        nonlocal pos
        pos += int(rng.integers(0, 13))  # misalign
        o = pos
        pieces.append((o, b))
        pos += len(b)
        return o

    for n_reads in (0, 1, 7, 1000, 25000):
        lens = rng.integers(0, 301, n_reads).astype("<u4")
        seq = rng.integers(0, 256, int(lens.sum())).astype(np.uint8).tobytes()
        hdr = np.array([n_reads], dtype="<u8").tobytes()
        segs = [(put(hdr), 8), (put(lens.tobytes()), 4 * n_reads), (put(seq), len(seq))]
        cases.append((segs, hdr + lens.tobytes() + seq))
    big = rng.integers(0, 256, (1 << 20) + 3).astype(np.uint8).tobytes()
    cases.append(([(put(big), len(big)), (put(b""), 0), (put(b"xyz"), 3)], big + b"xyz"))
    cases.append(([], b""))
    one = rng.integers(0, 256, 5000).astype(np.uint8).tobytes()
    cases.append(([(put(one), len(one))], one))
    arena = np.zeros(pos + 64, dtype=np.uint8)
    for o, b in pieces:
        arena[o:o + len(b)] = np.frombuffer(b, dtype=np.uint8)
    d = torch.from_numpy(arena).cuda()
    out = coder.xxh3_concat(d, [c[0] for c in cases]).cpu().numpy().view(np.uint64)
    for (segs, ref), h in zip(cases, out):
        assert int(h) == xxhash.xxh3_64_intdigest(ref), [s[1] for s in segs]
    # The same strings as single ranges through the plain batch entry point.
    single = [c for c in cases if len(c[0]) == 1]
    got = coder.xxh3_batch(d, [c[0][0][0] for c in single], [c[0][0][1] for c in single]).cpu().numpy()
    assert [int(x) for x in got.view(np.uint64)] == [xxhash.xxh3_64_intdigest(c[1]) for c in single]
