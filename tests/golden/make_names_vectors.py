"""Extract the names streams and their stored digests from the reference's
golden archives.

Each plain-layout block payload (crates/fqxv/src/container/mod.rs:80-97)
starts with three xxh3-64 content digests (names, sequence, quality) of the
block's DECODED streams, then [4] n_reads and [4] names_len + the names
stream in fqxv-tokenizer format (crates/fqxv-tokenizer/src/lib.rs).  The
names digest covers [n_reads u64][len u32 + name bytes]* of the decoded names
(crates/fqxv/src/container/block.rs:164-176), so it pins a names decoder end
to end.  The reorder fixture discards read order and stores no names, so it
contributes none.

Writes names_streams.bin (the streams back to back) and names_vectors.json
([{archive, block, off, len, n_reads, names_digest}]).  Run from the
repository root in a container that has the reference; the outputs are
committed, so tests never read the reference.
"""
import json
import os
import struct

REF = "/root/reference/crates/fqxv/tests/fixtures"
HERE = os.path.dirname(os.path.abspath(__file__))
ARCHIVES = ["plain_se", "paired_labels", "longread"]


def blocks(b):
    assert b[:4] == b"FQXV"
    ext_len = struct.unpack_from("<H", b, 19)[0]
    p = 21 + ext_len + 4  # header + header CRC
    foot = struct.unpack_from("<Q", b, len(b) - 12)[0]
    while p < foot:
        assert b[p:p + 4] == b"FQXB"
        ln = struct.unpack_from("<Q", b, p + 4)[0]
        if ln == 0:
            return
        yield b[p + 16:p + 16 + ln]
        p += 16 + ln


def main():
    blob = bytearray()
    vecs = []
    for name in ARCHIVES:
        b = open(os.path.join(REF, name + ".fqxv"), "rb").read()
        for i, pay in enumerate(blocks(b)):
            nd = struct.unpack_from("<Q", pay, 0)[0]
            n_reads, nl = struct.unpack_from("<II", pay, 24)
            vecs.append({"archive": name, "block": i, "off": len(blob), "len": nl, "n_reads": n_reads,
                         "names_digest": "%016x" % nd})
            blob += pay[32:32 + nl]
    with open(os.path.join(HERE, "names_streams.bin"), "wb") as f:
        f.write(blob)
    with open(os.path.join(HERE, "names_vectors.json"), "w") as f:
        json.dump(vecs, f, indent=0)
    print(len(vecs), "names streams,", len(blob), "bytes")


if __name__ == "__main__":
    main()
