"""Extract the rANS streams embedded in the reference's golden archives.

The reference keeps four archives written by an earlier fqxv release under
crates/fqxv/tests/fixtures/*.fqxv (pinned by crates/fqxv/tests/golden.rs:62-103
and manifest.tsv).  Their names sub-streams (tokenizer lib.rs:98-110), the
reorder side streams and the long-read overlap streams are fqxv-rans streams
written by the reference encoder.

This script scans every byte offset for a well-formed rANS Nx16 header
([u8 order][u64 n] + frequency tables that sum to 4096, scalar.rs:8-20),
decodes the candidate with the oracle, and keeps it only if re-encoding the
decoded bytes reproduces the stored stream byte for byte.  A kept stream
therefore pins both directions of the coder against bytes the reference
wrote.  Output: fixture_streams.bin (the streams, concatenated) and
fixture_streams.json (offsets, orders, lengths, CRC-32 of the decoded bytes).

Run from the repository root:  python tests/golden/make_fixture_streams.py
"""
import json
import os
import sys
import zlib

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle  # noqa: E402

FIXDIR = "/root/reference/crates/fqxv/tests/fixtures"
NAMES = ["plain_se", "paired_labels", "reorder", "longread"]
MAX_N = 1 << 26


def window_sums(data: bytes, start: int) -> dict:
    """For every p, the sum of the 256 little-endian u16 values at data[p+start:]."""
    out = {}
    for a in (0, 1):
        buf = data[a:]
        m = len(buf) // 2
        arr = np.frombuffer(buf[: 2 * m], dtype="<u2").astype(np.int64)
        cs = np.concatenate([[0], np.cumsum(arr)])
        if m >= 256:
            out[a] = cs[256:] - cs[:-256]  # out[a][k] = sum arr[k:k+256]
        else:
            out[a] = np.zeros(0, dtype=np.int64)
    return out


def scan(data: bytes):
    found = []
    ws = window_sums(data, 0)
    L = len(data)
    p = 0
    while p + 9 <= L:
        order = data[p]
        if order not in (0, 1):
            p += 1
            continue
        n = int.from_bytes(data[p + 1:p + 9], "little")
        if not (1 <= n <= MAX_N):
            p += 1
            continue
        t = p + 9 if order == 0 else p + 9 + 32
        if order == 1:
            if t > L:
                p += 1
                continue
            bm = data[p + 9:p + 41]
            if not any(bm):
                p += 1
                continue
        a = t % 2
        k = (t - a) // 2
        if k >= len(ws[a]) or ws[a][k] != 4096:
            p += 1
            continue
        rc, out, used = oracle.decode_ex(data[p:], cap=n)
        if rc != 0 or len(out) != n:
            p += 1
            continue
        stream = data[p:p + used]
        if oracle.encode(out, order) != stream:
            p += 1
            continue
        found.append((p, order, n, stream, out))
        p += used
    return found


def main():
    blobs = []
    index = []
    off = 0
    for name in NAMES:
        data = open(os.path.join(FIXDIR, name + ".fqxv"), "rb").read()
        for (p, order, n, stream, out) in scan(data):
            k = 0
            if order == 1:
                k = sum(bin(b).count("1") for b in stream[9:41])
            hdr = 9 + (512 if order == 0 else 32 + 512 * k) + 128
            index.append({
                "fixture": name, "archive_offset": p, "order": order, "n": n,
                "stream_len": len(stream), "words": (len(stream) - hdr) // 2,
                "contexts": k, "decoded_crc32": zlib.crc32(out) & 0xFFFFFFFF,
                "blob_offset": off,
            })
            blobs.append(stream)
            off += len(stream)
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "fixture_streams.bin"), "wb") as f:
        f.write(b"".join(blobs))
    with open(os.path.join(here, "fixture_streams.json"), "w") as f:
        json.dump(index, f, indent=1)
    by = {}
    for e in index:
        by.setdefault((e["fixture"], e["order"]), 0)
        by[(e["fixture"], e["order"])] += 1
    print(f"{len(index)} streams, {off} bytes: {by}")


if __name__ == "__main__":
    main()
