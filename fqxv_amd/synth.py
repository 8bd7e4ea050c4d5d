"""Synthetic workloads for the rANS benchmarks (host generator in csrc/synth.c).

Named workloads (sizes in bytes of uncompressed data):

- ``binned_1mib``  : 1 block of 1 MiB, i.i.d. over {35,45,56,70} with
  p = .02/.05/.08/.85, seed 1, order-0.
- ``normal_64x1mib``: 64 blocks x 1 MiB, 33 + round(N(30, 8^2)) clipped to
  2..41, seed 2, order-0.
- ``markov_64x4mib``: 64 blocks x 4 MiB cut from one continuous stream of
  150-symbol Markov quality strings, seed 3, order-1.
- ``mixed_10k``    : 10,000 blocks, sizes log-uniform on [4 KiB, 4 MiB],
  order 0/1 at 50/50, kind uniform over {Markov quality, DNA, names},
  seed 4.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libfqxv_synth.so")

K_BINNED, K_NORMAL, K_MARKOV, K_DNA, K_NAMES = 0, 1, 2, 3, 4
KIND_NAMES = {K_BINNED: "binned", K_NORMAL: "normal", K_MARKOV: "markov", K_DNA: "dna",
              K_NAMES: "names"}
MIB = 1 << 20


def build() -> str:
    os.makedirs(os.path.dirname(LIB_PATH), exist_ok=True)
    tmp = LIB_PATH + ".tmp"
    subprocess.run(["gcc", "-O3", "-fPIC", "-shared", "-std=c11", "-Wall", "-o", tmp,
                    os.path.join(HERE, "csrc", "synth.c"), "-lm", "-lpthread"], check=True)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        u64 = ctypes.c_uint64
        L.synth_fill.argtypes = [ctypes.c_int, vp, u64, u64, u64, u64]
        L.synth_fill.restype = None
        L.synth_mixed_layout.argtypes = [u64, ctypes.c_uint32, u64, u64, vp, vp, vp]
        L.synth_mixed_layout.restype = None
        L.synth_fill_blocks.argtypes = [vp, vp, vp, vp, ctypes.c_uint32, u64, vp, ctypes.c_int]
        L.synth_fill_blocks.restype = None
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def fill(kind: int, n: int, seed: int, stream: int = 0, start: int = 0) -> np.ndarray:
    out = np.empty(n, dtype=np.uint8)
    lib().synth_fill(kind, _p(out), n, seed, stream, start)
    return out


@dataclass
class Corpus:
    name: str
    data: np.ndarray        # all blocks back to back
    offs: np.ndarray        # u64 block offsets
    sizes: np.ndarray       # u64 block sizes
    orders: np.ndarray      # i32 order per block
    kinds: np.ndarray       # i32 generator kind per block

    @property
    def total(self) -> int:
        return int(self.sizes.sum())

    def block(self, j: int) -> np.ndarray:
        o, n = int(self.offs[j]), int(self.sizes[j])
        return self.data[o:o + n]


def _offsets(sizes: np.ndarray) -> np.ndarray:
    offs = np.zeros(len(sizes), dtype=np.uint64)
    if len(sizes):
        offs[1:] = np.cumsum(sizes)[:-1]
    return offs


def make(kinds, sizes, orders, seed: int, continuous: bool, name: str, threads: int = 0) -> Corpus:
    kinds = np.ascontiguousarray(kinds, dtype=np.int32)
    sizes = np.ascontiguousarray(sizes, dtype=np.uint64)
    orders = np.ascontiguousarray(orders, dtype=np.int32)
    offs = _offsets(sizes)
    data = np.empty(int(sizes.sum()) + 64, dtype=np.uint8)
    starts = offs.copy() if continuous else None
    if threads <= 0:
        threads = min(16, os.cpu_count() or 1)
    lib().synth_fill_blocks(_p(kinds), _p(sizes), _p(offs), _p(starts), len(sizes), seed, _p(data),
                            threads)
    return Corpus(name, data, offs, sizes, orders, kinds)


def mixed_layout(count: int, seed: int = 4, lo: int = 4096, hi: int = 4 * MIB):
    sizes = np.empty(count, dtype=np.uint64)
    orders = np.empty(count, dtype=np.int32)
    kinds = np.empty(count, dtype=np.int32)
    lib().synth_mixed_layout(seed, count, lo, hi, _p(sizes), _p(orders), _p(kinds))
    return sizes, orders, kinds


def workload(name: str, blocks: int | None = None) -> Corpus:
    """A named workload; ``blocks`` truncates it to its first blocks."""
    if name == "binned_1mib":
        c = make([K_BINNED], [MIB], [0], 1, False, name)
    elif name == "normal_64x1mib":
        k = blocks or 64
        c = make([K_NORMAL] * k, [MIB] * k, [0] * k, 2, False, name)
    elif name == "markov_64x4mib":
        k = blocks or 64
        c = make([K_MARKOV] * k, [4 * MIB] * k, [1] * k, 3, True, name)
    elif name == "mixed_10k":
        sizes, orders, kinds = mixed_layout(10000, 4)
        if blocks:
            sizes, orders, kinds = sizes[:blocks], orders[:blocks], kinds[:blocks]
        c = make(kinds, sizes, orders, 4, False, name)
    else:
        raise KeyError(name)
    return c


# ---- read-name lists for the names-stream path ------------------------------

NAME_KINDS = ("illumina", "paired", "sra", "arbitrary", "highbit", "maxcol")


def make_names(kind: str, n: int, seed: int) -> list[bytes]:
    """Synthetic FASTQ read names of one style:

    - ``illumina``: ``INST7:41:HFXXX:<lane>:<tile>:<x>:<y>``, tiles rising,
      x/y coordinates drifting upward within a tile;
    - ``paired``: the same, mates interleaved with `` 1:N:0:<barcode>`` /
      `` 2:N:0:<barcode>`` comments (a G=2 block);
    - ``sra``: ``SRR<run>.<i> <i> length=<len>`` (a counter column);
    - ``arbitrary``: 0-40 bytes drawn from ``AB:._-0129 `` (the alphabet of
      the tokenizer's own property tests);
    - ``highbit``: arbitrary bytes 0x00-0xff;
    - ``maxcol``: alternating letter/digit names of 58-64 single-byte tokens,
      up to the 64-column cap.  (Past it, the last column collects every
      further token, and the reference's decoder rejects a column holding
      more ops than there are records: crates/fqxv-tokenizer/src/lib.rs:
      232-237.)
    """
    r = np.random.default_rng(seed)
    out: list[bytes] = []
    if kind in ("illumina", "paired"):
        m = n // 2 + 1 if kind == "paired" else n
        tile, x, y = 1101, 1000, 2000
        lane = int(r.integers(1, 9))
        steps = r.integers(0, 40, size=(m, 2))
        newtile = r.random(m) < 0.001
        codes = [b"ACGTAC", b"TTAGGC", b"CGATGT"]
        for i in range(m):
            if newtile[i]:
                tile += 1
                x, y = 1000, 2000
            x += int(steps[i, 0])
            y += int(steps[i, 1]) * 3
            base = b"INST7:41:HFXXX:%d:%d:%d:%d" % (lane, tile, x, y)
            if kind == "paired":
                bc = codes[i % 3]
                out.append(base + b" 1:N:0:" + bc)
                out.append(base + b" 2:N:0:" + bc)
            else:
                out.append(base)
        return out[:n]
    if kind == "sra":
        run = int(r.integers(100000, 999999))
        lens = r.choice([100, 101, 150, 151], size=n, p=[0.1, 0.2, 0.6, 0.1])
        return [b"SRR%d.%d %d length=%d" % (run, i + 1, i + 1, int(lens[i])) for i in range(n)]
    if kind == "arbitrary":
        alpha = np.frombuffer(b"AB:._-0129 ", dtype=np.uint8)
        ls = r.integers(0, 41, size=n)
        return [alpha[r.integers(0, len(alpha), size=int(k))].tobytes() for k in ls]
    if kind == "highbit":
        ls = r.integers(0, 24, size=n)
        return [r.integers(0, 256, size=int(k), dtype=np.int64).astype(np.uint8).tobytes() for k in ls]
    if kind == "maxcol":
        ls = r.integers(58, 65, size=n)
        return [bytes(b"a1"[j % 2] for j in range(int(k))) for k in ls]
    raise ValueError(kind)
