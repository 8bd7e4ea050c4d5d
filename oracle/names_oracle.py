"""CPU restatement of the read-name tokenizer codec -- TEST INFRASTRUCTURE.

Follows crates/fqxv-tokenizer/src/lib.rs (format version 1) on top of the
rANS oracle (oracle/oracle.py); only tests/, ``__graft_entry__.smoke()`` and
bench.py's CPU legs use it, as the checker.  Pure Python loops: meant for
fixture-sized and test-sized name lists.

Pinned by the names streams of the reference's golden archives
(crates/fqxv/tests/fixtures/{plain_se,paired_labels,longread}.fqxv): the
decoded names must digest (xxh3-64 over the layout of
crates/fqxv/src/container/block.rs:164-176) to the value each block stores,
and re-encoding them must give the stored stream byte for byte.

Errors carry the reference's ``Error::Malformed`` strings (lib.rs) or the
rANS oracle's (lib.rs of fqxv-rans), in the reference's check order.
"""
from __future__ import annotations

import struct

from . import oracle as ro

FORMAT_VERSION = 1                 # lib.rs:55
MATCH, STR, NUM, DELTA = 0, 1, 2, 3  # lib.rs:58-61
MAX_NUM_DIGITS = 18                # lib.rs:63
MAX_COL = 63                       # lib.rs:66
MAX_STREAM_EXPANSION = 1 << 16     # lib.rs:75
MIN_STREAM_BOUND = 1 << 20         # lib.rs:76
U64 = (1 << 64) - 1


class NamesError(Exception):
    """A malformed names stream; ``str()`` is the reference's message."""


def _i64(v: int) -> int:
    v &= U64
    return v - (1 << 64) if v >> 63 else v


# ---- fqxv-bytes primitives (crates/fqxv-bytes/src/lib.rs) -----------------

def write_varint(out: bytearray, v: int) -> None:
    """LEB128 (fqxv-bytes lib.rs:14-24)."""
    v &= U64
    while True:
        b = v & 0x7F
        v >>= 7
        if v == 0:
            out.append(b)
            return
        out.append(b | 0x80)


def zigzag(d: int) -> int:
    """fqxv-bytes lib.rs:53-55 on i64."""
    d = _i64(d)
    return ((d << 1) ^ (d >> 63)) & U64


def unzigzag(z: int) -> int:
    """fqxv-bytes lib.rs:58-60."""
    z &= U64
    return _i64((z >> 1) ^ (-(z & 1) & U64))


class Reader:
    """fqxv-bytes ``Reader`` with the tokenizer's error strings
    (lib.rs:700-712): truncation of a varint reports "varint too long"
    because ``read_varint`` returns one None for both (fqxv-bytes
    lib.rs:31-47, 120-123)."""

    def __init__(self, buf: bytes):
        self.buf = buf
        self.pos = 0

    def u8(self) -> int:
        if self.pos >= len(self.buf):
            raise NamesError("truncated")
        b = self.buf[self.pos]
        self.pos += 1
        return b

    def varint(self) -> int:
        v, shift = 0, 0
        while True:
            if self.pos >= len(self.buf):
                raise NamesError("varint too long")
            b = self.buf[self.pos]
            self.pos += 1
            v |= (b & 0x7F) << shift
            if not b & 0x80:
                return v & U64
            shift += 7
            if shift >= 64:
                raise NamesError("varint too long")

    def take(self, n: int) -> bytes:
        if n > len(self.buf) - self.pos:
            raise NamesError("truncated")
        s = self.buf[self.pos:self.pos + n]
        self.pos += n
        return s


# ---- tokens (lib.rs:654-698, 253-263) -------------------------------------

def tokenize(name: bytes):
    """Maximal digit / non-digit runs (lib.rs:654-681): (is_num, bytes, value)."""
    toks = []
    i = 0
    while i < len(name):
        dig = 48 <= name[i] <= 57
        s = i
        while i < len(name) and (48 <= name[i] <= 57) == dig:
            i += 1
        b = bytes(name[s:i])
        is_num = dig and len(b) <= MAX_NUM_DIGITS
        toks.append((is_num, b, int(b) if is_num else 0))
    return toks


def natural_digits(value: int) -> int:
    """lib.rs:253-263."""
    return 1 if value <= 0 else len(str(value))


def num_tok(value: int, width: int):
    """lib.rs:684-698: i64 ``to_string`` left-padded with '0' to width."""
    s = str(value).encode()
    if len(s) < width:
        s = b"0" * (width - len(s)) + s
    return (True, s, value)


# ---- encode (lib.rs:84-200, 455-540) --------------------------------------

def _push_stream(out: bytearray, stream: bytes) -> None:
    """lib.rs:98-103 with encode_best (lib.rs:88-95, order-0 on a tie)."""
    c = ro.encode_best(bytes(stream), 0)
    write_varint(out, len(c))
    out += c


def encode_parts(names):
    """The raw (pre-rANS) sub-streams of lib.rs:455-540: (n_records,
    [counts, str_lens, str_data, num_vals, pads, op_counts, op_syms,
    op_lens], per-column lists of delta byte planes)."""
    counts = bytearray()
    op_cols = [bytearray() for _ in range(MAX_COL + 1)]
    str_lens, str_data, num_vals, pads = bytearray(), bytearray(), bytearray(), bytearray()
    delta_vals = [[] for _ in range(MAX_COL + 1)]
    prev = []
    for name in names:
        toks = tokenize(bytes(name))
        for i, (is_num, b, value) in enumerate(toks):
            col = min(i, MAX_COL)
            p = prev[i] if i < len(prev) else None
            if p is not None and p[1] == b:
                op_cols[col].append(MATCH)
            elif is_num:
                pad = len(b) - natural_digits(value)
                if p is not None and p[0]:
                    op_cols[col].append(DELTA)
                    delta_vals[col].append(zigzag(value - p[2]))
                else:
                    op_cols[col].append(NUM)
                    write_varint(num_vals, value)
                pads.append(pad & 0xFF)
            else:
                op_cols[col].append(STR)
                write_varint(str_lens, len(b))
                str_data += b
        write_varint(counts, len(toks))
        prev = toks
    # Op RLE (lib.rs:175-200).
    rc, syms, lens = bytearray(), bytearray(), bytearray()
    for col in op_cols:
        runs, i = 0, 0
        while i < len(col):
            s, run = col[i], 1
            while i + run < len(col) and col[i + run] == s:
                run += 1
            syms.append(s)
            write_varint(lens, run)
            runs += 1
            i += run
        write_varint(rc, runs)
    # Delta byte planes (lib.rs:118-135).
    planes = []
    for col in delta_vals:
        if not col:
            planes.append([])
            continue
        width = max(1, (max(col).bit_length() + 7) // 8)
        planes.append([bytes((v >> (8 * k)) & 0xFF for v in col) for k in range(width)])
    flat = [bytes(x) for x in (counts, str_lens, str_data, num_vals, pads, rc, syms, lens)]
    return len(names), flat, planes


def assemble(n_records: int, flat, planes) -> bytes:
    """The framing of lib.rs:515-537 around raw sub-streams."""
    out = bytearray([FORMAT_VERSION])
    write_varint(out, n_records)
    for st in flat:
        _push_stream(out, st)
    for col in planes:
        write_varint(out, len(col))
        for pl in col:
            _push_stream(out, pl)
    return bytes(out)


def encode(names) -> bytes:
    """lib.rs:455-540."""
    return assemble(*encode_parts(names))


# ---- decode (lib.rs:105-110, 137-168, 202-245, 543-650) -------------------

def _read_stream(r: Reader, max_out: int) -> bytes:
    """lib.rs:105-110: [varint len][bytes], rANS decode_bounded."""
    n = r.varint()
    s = r.take(n)
    try:
        return ro.decode_bounded(s, max_out)
    except ro.OracleError as e:
        raise NamesError(str(e)) from None


def _read_op_rle(r: Reader, n_records: int, max_out: int):
    """lib.rs:202-245."""
    counts = _read_stream(r, max_out)
    syms = _read_stream(r, max_out)
    lens = _read_stream(r, max_out)
    cc, cl = Reader(counts), Reader(lens)
    sym_pos, total = 0, 0
    cols = []
    for _ in range(MAX_COL + 1):
        n_runs = cc.varint()
        col = bytearray()
        for _ in range(n_runs):
            if sym_pos >= len(syms):
                raise NamesError("op sym underrun")
            s = syms[sym_pos]
            sym_pos += 1
            run = cl.varint()
            if len(col) + run > n_records:
                raise NamesError("op run too long")
            total += run
            if total > max_out:
                raise NamesError("op columns exceed the stream bound")
            col += bytes([s]) * run
        cols.append(bytes(col))
    return cols


def _read_delta_planes(r: Reader, n_delta):
    """lib.rs:137-168 (a plane index >= 8 shifts by 8*plane mod 64, the
    release-build meaning of Rust's ``<<``)."""
    cols = []
    for n in n_delta:
        width = r.varint()
        if width == 0:
            cols.append([])
            continue
        vals = [0] * n
        for plane in range(width):
            ln = r.varint()
            b = r.take(ln)
            try:
                dec = ro.decode_bounded(b, n)
            except ro.OracleError as e:
                raise NamesError(str(e)) from None
            if len(dec) != n:
                raise NamesError("delta plane length mismatch")
            sh = (8 * plane) & 63
            for k, v in enumerate(dec):
                vals[k] = (vals[k] | (v << sh)) & U64
        cols.append(vals)
    return cols


def decode_streams(src: bytes):
    """The framing half of lib.rs:543-588: every sub-stream decoded, checks
    in the reference's order.  Returns (n_records, streams dict)."""
    r = Reader(bytes(src))
    if r.u8() != FORMAT_VERSION:
        raise NamesError("unsupported version")
    n_records = r.varint()
    max_out = max(min(len(src) * MAX_STREAM_EXPANSION, U64), MIN_STREAM_BOUND)
    counts = _read_stream(r, max_out)
    if n_records > len(counts):
        raise NamesError("record count exceeds token stream")
    st = {"counts": counts}
    for k in ("str_lens", "str_data", "num_vals", "pads"):
        st[k] = _read_stream(r, max_out)
    ops = _read_op_rle(r, n_records, max_out)
    st["ops"] = ops
    st["deltas"] = _read_delta_planes(r, [c.count(DELTA) for c in ops])
    return n_records, st


def decode(src: bytes):
    """lib.rs:543-650: the list of names."""
    n_records, st = decode_streams(src)
    ops, deltas, pads = st["ops"], st["deltas"], st["pads"]
    op_cur = [Reader(c) for c in ops]
    delta_pos = [0] * (MAX_COL + 1)
    c_counts, c_sl, c_num = Reader(st["counts"]), Reader(st["str_lens"]), Reader(st["num_vals"])
    str_data = st["str_data"]
    str_pos = pad_pos = 0
    names, prev = [], []
    for _ in range(n_records):
        n_toks = c_counts.varint()
        cur = []
        for _ in range(n_toks):
            col = min(len(cur), MAX_COL)
            op = op_cur[col].u8()
            if op == MATCH:
                if len(cur) >= len(prev):
                    raise NamesError("MATCH without prior token")
                cur.append(prev[len(cur)])
            elif op == STR:
                ln = c_sl.varint()
                if str_pos + ln > len(str_data):
                    raise NamesError("string data underrun")
                cur.append((False, str_data[str_pos:str_pos + ln], 0))
                str_pos += ln
            elif op == NUM:
                value = _i64(c_num.varint())
                if pad_pos >= len(pads):
                    raise NamesError("pad underrun")
                width = natural_digits(value) + pads[pad_pos]
                pad_pos += 1
                cur.append(num_tok(value, width))
            elif op == DELTA:
                if delta_pos[col] >= len(deltas[col]):
                    raise NamesError("delta underrun")
                z = deltas[col][delta_pos[col]]
                delta_pos[col] += 1
                d = unzigzag(z)
                p = prev[len(cur)] if len(cur) < len(prev) else None
                if p is None or not p[0]:
                    raise NamesError("DELTA without numeric prior")
                value = _i64(p[2] + d)
                if pad_pos >= len(pads):
                    raise NamesError("pad underrun")
                width = natural_digits(value) + pads[pad_pos]
                pad_pos += 1
                cur.append(num_tok(value, width))
            else:
                raise NamesError("unknown op")
        names.append(b"".join(t[1] for t in cur))
        prev = cur
    return names


# ---- block digest (crates/fqxv/src/container/block.rs:164-176) ------------

def digest_input(names) -> bytes:
    """The bytes the block's names digest covers: [n u64 LE] then per name
    [len u32 LE][bytes]."""
    out = bytearray(struct.pack("<Q", len(names)))
    for nm in names:
        out += struct.pack("<I", len(nm)) + bytes(nm)
    return bytes(out)


def names_digest(names) -> int:
    """xxh3-64 (seed 0) of ``digest_input`` -- the checker uses the xxhash
    package's XXH3_64bits, the algorithm of the reference's xxhash-rust
    0.8.18 (Cargo.lock)."""
    import xxhash
    return xxhash.xxh3_64_intdigest(digest_input(names))
