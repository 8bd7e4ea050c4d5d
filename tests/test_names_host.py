"""CPU: the names-stream host code (fqxv_amd/csrc/rans_names.cpp) built with
AddressSanitizer and UBSan, its device batch replaced by the oracle
(tests/names_host), against the names oracle: the framing walk, the host
header checks, the two device phases, the replay and the encode framing run
here on every case, with every access checked.  Includes streams whose
sub-streams declare outputs near their bound (the allocation budget case)."""
import os
import struct
import subprocess

import numpy as np
import pytest

from conftest import ROOT

BIN = os.path.join(ROOT, "tests", "names_host", "_build", "names_host")


def _pack(items) -> bytes:
    return struct.pack("<I", len(items)) + b"".join(struct.pack("<Q", len(b)) + b for b in items)


def _run(mode, items, tmp_path, env=None):
    assert os.path.exists(BIN), "tests/names_host is not built (run __graft_entry__.build())"
    src, dst = tmp_path / f"{mode}.in", tmp_path / f"{mode}.out"
    src.write_bytes(_pack(items))
    e = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1")
    e.update(env or {})
    r = subprocess.run([BIN, mode, str(src), str(dst)], capture_output=True, text=True, timeout=600, env=e)
    assert r.returncode == 0, r.stderr[-4000:]
    out, blob, p = [], dst.read_bytes(), 0
    for _ in items:
        st, size = struct.unpack_from("<iQ", blob, p)
        p += 12
        out.append((st, blob[p:p + size]))
        p += size
    return out


def _check_decode(no, streams, results):
    from fqxv_amd import _lib, names
    for s, (st, rec) in zip(streams, results):
        try:
            ref = ("ok", names.records(no.decode(s)))
        except no.NamesError as e:
            ref = ("err", str(e))
        got = ("ok", rec) if st == 0 else ("err", _lib.strerror(st))
        assert got == ref, s[:64].hex()


@pytest.fixture(scope="module")
def no(oracle):
    from oracle import names_oracle
    return names_oracle


def test_decode_matches_oracle_under_sanitizers(no, tmp_path):
    from fqxv_amd import synth
    from test_gpu_names import corrupt, fixture_names_streams, structured_mutants
    from test_names_oracle import REF_LISTS
    rng = np.random.default_rng(5)
    streams = [s for _, s in fixture_names_streams()]
    streams += [no.encode(n) for n in REF_LISTS]
    streams += [no.encode(synth.make_names(k, n, 7)) for k in synth.NAME_KINDS for n in (0, 1, 300)]
    bases = [no.encode(synth.make_names(k, 60, 7)) for k in ("illumina", "paired", "sra", "arbitrary")]
    streams += [corrupt(bases[i % len(bases)], rng) for i in range(300)]
    streams += [rng.integers(0, 256, int(rng.integers(0, 200))).astype(np.uint8).tobytes() for _ in range(100)]
    streams += structured_mutants(no, rng, 300)
    _check_decode(no, streams, _run("decode", streams, tmp_path))


def _hostile(no, header_kind: str, declared_of, pad: int = 0) -> bytes:
    """A names stream with valid framing whose eight flat sub-streams each
    declare `declared_of(max_out)` output bytes: 9-byte headers only
    ("short"), or complete order-0 headers (valid table, 32 states) with no
    renormalisation words ("full"), which decode as far as the first word.
    `pad` extra bytes in the second sub-stream raise the stream's max_out."""
    def sub(n):
        if header_kind == "short":
            return b"\x00" + struct.pack("<Q", n)
        freq = [0] * 256
        freq[65], freq[66] = 2048, 2048
        return b"\x00" + struct.pack("<Q", n) + struct.pack("<256H", *freq) + struct.pack("<32I", *[1 << 16] * 32)
    # Two passes: the stream length fixes max_out (len * 65536, at least 1 MiB).
    length = 0
    for _ in range(2):
        max_out = max(length * 65536, 1 << 20)
        out = bytearray([1])
        no.write_varint(out, 3)
        for k in range(8):
            s = sub(declared_of(max_out)) + (bytes(pad) if k == 1 else b"")
            no.write_varint(out, len(s))
            out += s
        out += bytes(64)  # every column width 0
        length = len(out)
    return bytes(out)


def test_streams_declaring_large_outputs(no, tmp_path):
    """Sub-streams that declare their stream's whole bound: each fails like the
    reference (truncated or the bound), next to valid streams in the same
    batch, and nothing is sized to the sum of what they declare."""
    from fqxv_amd import synth
    good = [no.encode(synth.make_names("illumina", 500, 3)), no.encode(synth.make_names("sra", 200, 4))]
    bad = [_hostile(no, "short", lambda m: m), _hostile(no, "full", lambda m: m),
           _hostile(no, "full", lambda m: m + 1), _hostile(no, "short", lambda m: 0)]
    streams = [good[0]] + bad + [good[1]]
    _check_decode(no, streams, _run("decode", streams, tmp_path))


def test_oversized_job_fails_alone(no, tmp_path):
    """A sub-stream declaring more than one launch's budget (1 GiB) gets a
    launch of its own; when its region cannot be allocated (the stub's device
    budget set below it) that stream fails with the reference's allocation
    error and the rest of the batch decodes."""
    from fqxv_amd import _lib, names, synth
    big = _hostile(no, "full", lambda m: (1 << 30) + 4096, pad=20000)
    good = no.encode(synth.make_names("paired", 300, 8))
    res = _run("decode", [good, big, good], tmp_path, env={"FQXV_STUB_DEVICE_BYTES": str(1 << 30)})
    assert res[0] == (0, names.records(no.decode(good))) and res[2] == res[0]
    assert _lib.strerror(res[1][0]) == "rANS output length too large to allocate"
    # With room for it, the same stream fails as the reference does.
    _check_decode(no, [big], _run("decode", [big], tmp_path))


def test_encode_matches_oracle_under_sanitizers(no, tmp_path):
    from fqxv_amd import names, synth
    from test_names_oracle import REF_LISTS
    lists = list(REF_LISTS) + [synth.make_names(k, n, 3) for k in synth.NAME_KINDS for n in (0, 1, 250)]
    res = _run("encode", [names.records(n) for n in lists], tmp_path)
    for n, (st, s) in zip(lists, res):
        assert st == 0 and s == no.encode(n)
