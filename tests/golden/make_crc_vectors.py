"""Extract CRC-32C golden vectors from the reference's fixture archives.

The reference's golden archives (crates/fqxv/tests/fixtures/*.fqxv, written
by an earlier fqxv release) store CRC-32C values computed by
crates/fqxv/src/crc.rs over byte ranges of the archive itself.  This script
walks the container layout documented in crates/fqxv/src/container/mod.rs:1-97
(header + header_crc, [FQXB][len][crc][payload] frames, the footer's
per-stream (offset, len, crc) triples, whole_file_crc and footer_crc) and
writes, for the two small archives:

  crc_<name>.bin   -- the archive bytes (a data file the reference's tests hold)
  crc_vectors.json -- [{archive, what, off, len, crc}] with the stored CRCs

Run from the repository root in a container that has the reference; the
outputs are committed, so tests never read the reference.
"""
import json
import os
import struct

REF = "/root/reference/crates/fqxv/tests/fixtures"
HERE = os.path.dirname(os.path.abspath(__file__))
ARCHIVES = ["plain_se", "paired_labels"]


def u32(b, o):
    return struct.unpack_from("<I", b, o)[0]


def u64(b, o):
    return struct.unpack_from("<Q", b, o)[0]


def vectors(name, b):
    out = []
    assert b[:4] == b"FQXV"
    ext_len = struct.unpack_from("<H", b, 19)[0]
    hdr = 21 + ext_len
    out.append({"archive": name, "what": "header", "off": 0, "len": hdr, "crc": u32(b, hdr)})
    # Trailer: [8] footer_offset [4] "FQXF".
    assert b[-4:] == b"FQXF"
    foot = u64(b, len(b) - 12)
    p = hdr + 4
    nblk = 0
    while p < foot:
        assert b[p:p + 4] == b"FQXB", (name, p)
        ln = u64(b, p + 4)
        if ln == 0:
            break
        out.append({"archive": name, "what": f"frame{nblk}", "off": p + 16, "len": ln, "crc": u32(b, p + 12)})
        p += 16 + ln
        nblk += 1
    q = foot
    ngroups = u32(b, q)
    q += 4
    for g in range(ngroups):
        q += 12  # byte_offset, read_count
        for sname in ("names", "sequence", "quality"):
            off, ln, crc = u64(b, q), u32(b, q + 8), u32(b, q + 12)
            out.append({"archive": name, "what": f"group{g}.{sname}", "off": off, "len": ln, "crc": crc})
            q += 16
    q += 8  # total_reads
    out.append({"archive": name, "what": "whole_file", "off": 0, "len": q, "crc": u32(b, q)})
    out.append({"archive": name, "what": "footer", "off": foot, "len": q + 4 - foot, "crc": u32(b, q + 4)})
    return out


def main():
    allv = []
    for name in ARCHIVES:
        b = open(os.path.join(REF, name + ".fqxv"), "rb").read()
        with open(os.path.join(HERE, f"crc_{name}.bin"), "wb") as f:
            f.write(b)
        allv += vectors(name, b)
    with open(os.path.join(HERE, "crc_vectors.json"), "w") as f:
        json.dump(allv, f, indent=0)
    print(len(allv), "vectors")


if __name__ == "__main__":
    main()
