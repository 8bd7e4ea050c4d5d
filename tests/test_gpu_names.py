"""Names streams on the device: fqxv_names_decode_batch against the names
oracle and the reference's golden archives (crates/fqxv-tokenizer/src/lib.rs
decode, crates/fqxv/src/container/block.rs:164-176 digests)."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu
xxhash = pytest.importorskip("xxhash")


@pytest.fixture(scope="module")
def no(oracle):
    from oracle import names_oracle
    return names_oracle


@pytest.fixture(scope="module")
def nm(coder):
    from fqxv_amd import names
    return names


def fixture_names_streams():
    v = json.load(open(os.path.join(GOLDEN, "names_vectors.json")))
    blob = open(os.path.join(GOLDEN, "names_streams.bin"), "rb").read()
    return [(e, blob[e["off"]:e["off"] + e["len"]]) for e in v]


def expect(no, s):
    try:
        return no.decode(s)
    except no.NamesError as e:
        return ("err", str(e))


def got(nm, r):
    return ("err", str(r)) if isinstance(r, nm.Error) else r


def test_fixture_digests(nm, coder):
    """The stored names digest of each golden block, computed over the
    device path's output records."""
    fx = fixture_names_streams()
    recs = nm.decode_batch_records([s for _, s in fx], coder)
    for (e, _), rec in zip(fx, recs):
        assert isinstance(rec, bytes), rec
        assert "%016x" % xxhash.xxh3_64_intdigest(rec) == e["names_digest"], e["archive"]
        assert len(nm.parse_records(rec)) == e["n_reads"]


def test_reference_lists(nm, no, coder):
    from test_names_oracle import REF_LISTS
    streams = [no.encode(n) for n in REF_LISTS]
    out = nm.decode_batch(streams, coder)
    for names, r in zip(REF_LISTS, out):
        assert r == names


@pytest.mark.parametrize("kind", ["illumina", "paired", "sra", "arbitrary", "highbit", "maxcol"])
def test_synthetic_batch(nm, no, coder, kind):
    from fqxv_amd import synth
    lists = [synth.make_names(kind, n, 100 + i) for i, n in enumerate([0, 1, 2, 97, 1000, 4000])]
    streams = [no.encode(n) for n in lists]
    out = nm.decode_batch(streams, coder)
    for names, r in zip(lists, out):
        assert r == names


def test_large_block(nm, no, coder):
    from fqxv_amd import synth
    names = synth.make_names("paired", 120000, 5)
    s = no.encode(names)
    r = nm.decode(s, coder)
    assert r == names
    assert xxhash.xxh3_64_intdigest(nm.decode_batch_records([s], coder)[0]) == no.names_digest(names)


def corrupt(data: bytes, rng) -> bytes:
    """Substitutions, an optional 0xFF window and an optional truncation, as
    the tokenizer's decode_survives_mutation property test applies them."""
    d = bytearray(data)
    for _ in range(int(rng.integers(0, 12))):
        if d:
            d[int(rng.integers(0, len(d)))] = int(rng.integers(0, 256))
    if d and rng.random() < 0.5:
        i = int(rng.integers(0, len(d)))
        for k in range(i, min(len(d), i + int(rng.integers(1, 9)))):
            d[k] = 0xFF
    if d and rng.random() < 0.5:
        d = d[: int(rng.integers(0, len(d)))]
    return bytes(d)


def test_mutations_match_oracle(nm, no, coder):
    """Every mutated stream gives the oracle's result: the same names or
    the same first error."""
    from fqxv_amd import synth
    rng = np.random.default_rng(2024)
    bases = [no.encode(synth.make_names(k, 60, 7)) for k in ("illumina", "paired", "sra", "arbitrary")]
    muts = [corrupt(bases[i % len(bases)], rng) for i in range(240)]
    out = nm.decode_batch(muts, coder)
    for m, r in zip(muts, out):
        assert got(nm, r) == expect(no, m), m.hex()


def test_garbage_matches_oracle(nm, no, coder):
    rng = np.random.default_rng(77)
    gs = [rng.integers(0, 256, int(rng.integers(0, 256))).astype(np.uint8).tobytes() for _ in range(200)]
    gs += [bytes([1]) + g for g in gs[:100]]  # past the version check
    out = nm.decode_batch(gs, coder)
    for g, r in zip(gs, out):
        assert got(nm, r) == expect(no, g), g.hex()


def test_reference_corrupt_cases(nm, no, coder):
    from test_names_oracle import test_corrupt_record_count  # noqa: F401  (same construction)
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
    cases = [bytes(out), bytes([0x01, 0x01] + [0xFF] * 9 + [0x37]), b"", bytes([2]),
             no.encode([bytes(b"a1"[j % 2] for j in range(66))])]
    res = nm.decode_batch(cases, coder)
    msgs = [str(r) for r in res]
    assert msgs == ["record count exceeds token stream", "truncated", "truncated", "unsupported version",
                    "op run too long"]
    for c, r in zip(cases, res):
        assert got(nm, r) == expect(no, c)
        assert isinstance(r, nm.Malformed)


def test_rans_error_kind(nm, no, coder):
    """A sub-stream that fails inside the coder is Error::Rans."""
    s = bytearray(no.encode([b"INST:1:FC:1:1101:1000:2000"] * 3))
    # [ver][varint n = 3][varint len, 2 bytes]: the first sub-stream's order
    # tag is byte 4.
    s[4] = 7
    r = nm.decode_batch([bytes(s)], coder)[0]
    assert isinstance(r, nm.Rans)
    assert got(nm, r) == expect(no, bytes(s))


def structured_mutants(no, rng, count: int):
    """Valid framing around corrupted RAW sub-streams (counts, ops, pads,
    planes, record count ...), so decoding walks past the entropy coder into
    the token-level checks (lib.rs:202-245, 590-650)."""
    from fqxv_amd import synth
    bases = [no.encode_parts(synth.make_names(k, 40, 9 + i))
             for i, k in enumerate(("illumina", "paired", "sra", "arbitrary", "maxcol"))]
    out = []
    for i in range(count):
        n, flat, planes = bases[i % len(bases)]
        flat = [bytearray(x) for x in flat]
        planes = [[bytearray(p) for p in col] for col in planes]
        what = int(rng.integers(0, 6))
        if what == 0:
            n = max(0, n + int(rng.integers(-3, 4)))
        elif what in (1, 2):
            k = int(rng.integers(0, 8))
            st = flat[k]
            if st and rng.random() < 0.6:
                st[int(rng.integers(0, len(st)))] = int(rng.integers(0, 256))
            elif st and rng.random() < 0.5:
                del st[int(rng.integers(0, len(st))):]
            else:
                st += bytes(rng.integers(0, 256, int(rng.integers(1, 4))).astype(np.uint8))
        elif what == 3:
            cols = [c for c in range(64) if planes[c]]
            if cols:
                c = cols[int(rng.integers(0, len(cols)))]
                p = planes[c][int(rng.integers(0, len(planes[c])))]
                if rng.random() < 0.5 and p:
                    p[int(rng.integers(0, len(p)))] ^= 0xFF
                elif p:
                    del p[-1]
                else:
                    p.append(1)
        elif what == 4:
            c = int(rng.integers(0, 64))
            planes[c].append(bytearray(rng.integers(0, 4, int(rng.integers(0, 50))).astype(np.uint8)))
        else:
            c = int(rng.integers(0, 64))
            planes[c] = []
        out.append(no.assemble(n, [bytes(x) for x in flat], [[bytes(p) for p in col] for col in planes]))
    return out


def test_structured_mutations_match_oracle(nm, no, coder):
    rng = np.random.default_rng(31)
    muts = structured_mutants(no, rng, 300)
    out = nm.decode_batch(muts, coder)
    kinds = set()
    for m, r in zip(muts, out):
        e = expect(no, m)
        kinds.add(e[1] if isinstance(e, tuple) else "ok")
        assert got(nm, r) == e, m.hex()
    assert len(kinds) >= 8, kinds


def test_encode_matches_oracle(nm, no, coder):
    """fqxv_names_encode_batch: host tokenizing + one batched encode-best
    launch + host framing gives the reference's bytes."""
    from fqxv_amd import synth
    from test_names_oracle import REF_LISTS
    lists = list(REF_LISTS) + [synth.make_names(k, n, 40 + n) for k in synth.NAME_KINDS for n in (0, 1, 9, 800)]
    lists.append([bytes(b"a1"[j % 2] for j in range(70))] * 3)  # past the column cap: still encodes
    out = nm.encode_batch(lists, coder)
    for names, s in zip(lists, out):
        assert s == no.encode(names)


def test_encode_reproduces_fixture_streams(nm, no, coder):
    """Re-encoding each golden block's names on the device gives the stored
    names stream byte for byte."""
    streams = [s for _, s in fixture_names_streams()]
    names = nm.decode_batch(streams, coder)
    assert nm.encode_batch(names, coder) == streams


def test_encode_decode_large_batch(nm, no, coder):
    from fqxv_amd import synth
    lists = [synth.make_names(k, 30000, 60 + i) for i, k in enumerate(("illumina", "paired", "sra", "illumina"))]
    streams = nm.encode_batch(lists, coder)
    assert streams[2] == no.encode(lists[2])
    recs, dg = nm.decode_batch_records(streams, coder, digests=True)
    for names, r, d in zip(lists, recs, dg):
        assert nm.parse_records(r) == names
        assert d == no.names_digest(names)


def test_encode_batch_with_a_malformed_records_buffer(nm, no, coder):
    """A block whose records buffer is truncated fails alone: the others in
    the same launch still encode (their streams come back packed), and the
    context serves the next batch normally."""
    names = [b"r1:1:AC", b"r1:2:AC", b"r1:10:GT"]
    good = nm.records(names)
    bad = good[:-1]  # the last name is one byte short
    with pytest.raises(ValueError, match="block 1"):
        nm.encode_batch_records([good, bad, good], coder)
    assert nm.encode_batch_records([good, good], coder) == [no.encode(names)] * 2
