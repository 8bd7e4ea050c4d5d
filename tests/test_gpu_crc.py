"""GPU: batched CRC-32C against the checksums stored in reference-written
archives and against the oracle on random ranges."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_crc32c_batch_reproduces_archive_checksums(coder):
    import torch
    vecs = json.load(open(os.path.join(GOLDEN, "crc_vectors.json")))
    for name in sorted({e["archive"] for e in vecs}):
        b = open(os.path.join(GOLDEN, f"crc_{name}.bin"), "rb").read()
        ev = [e for e in vecs if e["archive"] == name]
        arena = torch.from_numpy(np.frombuffer(b, dtype=np.uint8).copy()).cuda()
        got = coder.crc32c_batch(arena, [e["off"] for e in ev], [e["len"] for e in ev])
        torch.cuda.synchronize()
        assert got.cpu().numpy().view(np.uint32).tolist() == [e["crc"] for e in ev], name


def test_crc32c_batch_matches_oracle_on_random_ranges(coder, oracle):
    import torch
    rng = np.random.default_rng(21)
    data = rng.integers(0, 256, 3 << 20, dtype=np.uint8)
    arena = torch.from_numpy(data).cuda()
    offs, lens = [], []
    for n in [0, 1, 2, 7, 8, 9, 1023, 1024, 1025, 65535, 65536, 65537, 131072 + 5, 1 << 20, (3 << 20) - 64]:
        for o in (0, 1, 3, 8, 61):
            if o + n <= data.size:
                offs.append(o)
                lens.append(n)
    for _ in range(200):
        n = int(rng.integers(0, 300000))
        offs.append(int(rng.integers(0, data.size - n)))
        lens.append(n)
    got = coder.crc32c_batch(arena, offs, lens)
    torch.cuda.synchronize()
    ref = oracle.crc32c_many(data, offs, lens)
    assert np.array_equal(got.cpu().numpy().view(np.uint32), ref)


def test_crc32c_batch_large_single_range(coder, oracle):
    """A range of many chunks (the whole-file checksum shape), and a batch
    mixing it with small ranges."""
    import torch
    rng = np.random.default_rng(8)
    data = rng.integers(0, 256, (64 << 20) + 12345, dtype=np.uint8)
    arena = torch.from_numpy(data).cuda()
    offs = [0, 3, 17, 5]
    lens = [data.size, data.size - 3, 100, (1 << 22) + 1]
    got = coder.crc32c_batch(arena, offs, lens)
    torch.cuda.synchronize()
    assert np.array_equal(got.cpu().numpy().view(np.uint32), oracle.crc32c_many(data, offs, lens))


def test_crc32c_batch_two_level_fold(coder, oracle):
    """Ranges longer than 4096 chunks (256 MiB) fold their 256 MiB segments
    first and the segment values second; a range of exactly 4096 chunks, one
    of 4097 (a one-byte second segment) and small ranges share the batch."""
    import torch
    rng = np.random.default_rng(9)
    data = rng.integers(0, 256, (260 << 20) + 777, dtype=np.uint8)
    arena = torch.from_numpy(data).cuda()
    offs = [0, 5, 1000, 3, 7, 0]
    lens = [data.size, data.size - 5, 333, (256 << 20) + 1, 256 << 20, 0]
    got = coder.crc32c_batch(arena, offs, lens)
    torch.cuda.synchronize()
    assert np.array_equal(got.cpu().numpy().view(np.uint32), oracle.crc32c_many(data, offs, lens))
