"""The names oracle (oracle/names_oracle.py) against the reference's golden
archives and the tokenizer's own test cases (crates/fqxv-tokenizer/src/lib.rs
tests module) -- CPU only."""
import json
import os

import pytest

from conftest import GOLDEN

pytest.importorskip("xxhash")


@pytest.fixture(scope="module")
def no(oracle):
    from oracle import names_oracle
    return names_oracle


def fixture_names_streams():
    v = json.load(open(os.path.join(GOLDEN, "names_vectors.json")))
    blob = open(os.path.join(GOLDEN, "names_streams.bin"), "rb").read()
    return [(e, blob[e["off"]:e["off"] + e["len"]]) for e in v]


@pytest.mark.parametrize("idx", range(3))
def test_fixture_names_digest_and_reencode(no, idx):
    """Decoded names digest to the block's stored names digest
    (crates/fqxv/src/container/block.rs:164-176), and re-encoding them gives
    the stored stream byte for byte (pins encode and every sub-stream)."""
    e, s = fixture_names_streams()[idx]
    names = no.decode(s)
    assert len(names) == e["n_reads"]
    assert "%016x" % no.names_digest(names) == e["names_digest"]
    assert no.encode(names) == s


REF_LISTS = [
    [],                                                                     # roundtrip_empty
    [b"INST:1:FC:1:1101:1000:2000", b"INST:1:FC:1:1101:1005:2050", b"INST:1:FC:1:1101:1010:2100"],
    [b"SRR453566.1 HWI-ST167:4:1101:0042:1986 length=101", b"SRR453566.2 HWI-ST167:4:1101:0043:1990 length=101"],
    [b"X:007:0000:5", b"X:008:0001:5", b"X:009:0002:5", b"X:010:1000:5"],
    [b"a", b"bb99", b"", b"12", b"x1y2z3"],
    [bytes([0x80, 0x0a]), bytes([0xff]), b"read\x80\xa5\xffname", bytes([0, 0x80, 0x7f, 0xff, 0x0a]), b"plain"],
    [bytes(b"a1"[j % 2] for j in range(62))],
    [bytes(b"a1"[j % 2] for j in range(63))],
    [bytes(b"a1"[j % 2] for j in range(64))],
    [b"n1234567890123456789x", b"n1234567890123456790x", b"n99999999999999999x"],  # > 18 digits -> string
]


@pytest.mark.parametrize("names", REF_LISTS)
def test_reference_roundtrips(no, names):
    assert no.decode(no.encode(names)) == names


def test_tokenize_max_col(no):
    # roundtrip_at_max_col_boundary's construction.
    assert len(no.tokenize(bytes(b"a1"[j % 2] for j in range(63)))) == 63
    assert len(no.tokenize(bytes(b"a1"[j % 2] for j in range(64)))) == 64


def test_past_max_col_rejected_like_the_reference(no):
    """Tokens past column 63 all land in column 63, so a record with 65+
    tokens puts more ops there than there are records and the reference's
    per-column guard (lib.rs:232-237) rejects its own encoding."""
    names = [bytes(b"a1"[j % 2] for j in range(66))]
    enc = no.encode(names)
    with pytest.raises(no.NamesError, match="^op run too long$"):
        no.decode(enc)


def test_incrementing_names_compress_well(no):
    names = [b"INST:1:FC:1:1101:%d:%d" % (1000 + i, 2000 + 2 * i) for i in range(20000)]
    enc = no.encode(names)
    assert len(enc) < sum(len(n) + 1 for n in names) // 10
    assert no.decode(enc) == names


def test_corrupt_record_count(no):
    """corrupt_run_length_cannot_drive_huge_allocation (lib.rs tests)."""
    out = bytearray([1])
    no.write_varint(out, (1 << 64) // 2 - 1)
    counts = bytearray()
    no.write_varint(counts, 1)
    no._push_stream(out, counts)
    for _ in range(4):
        no._push_stream(out, b"")
    opc = bytearray()
    for c in range(64):
        no.write_varint(opc, 1 if c == 0 else 0)
    no._push_stream(out, opc)
    no._push_stream(out, bytes([0]))
    lens = bytearray()
    no.write_varint(lens, 1 << 30)
    no._push_stream(out, lens)
    with pytest.raises(no.NamesError, match="^record count exceeds token stream$"):
        no.decode(bytes(out))


def test_overflowing_take(no):
    """take_rejects_overflowing_length_without_panicking (lib.rs tests)."""
    src = bytes([0x01, 0x01] + [0xFF] * 9 + [0x37])
    with pytest.raises(no.NamesError, match="^truncated$"):
        no.decode(src)


def test_varint_semantics(no):
    r = no.Reader(bytes([0x80]))
    with pytest.raises(no.NamesError, match="^varint too long$"):
        r.varint()
    assert no.unzigzag(no.zigzag(-5)) == -5 and no.zigzag(-1) == 1 and no.zigzag(1) == 2


@pytest.mark.parametrize("kind", ["illumina", "paired", "sra", "arbitrary", "highbit", "maxcol"])
def test_synthetic_roundtrip(no, kind):
    from fqxv_amd import synth
    names = synth.make_names(kind, 300, 11)
    assert no.decode(no.encode(names)) == names
