"""CPU restatement of XXH3-64 (seed 0, default secret) -- TEST INFRASTRUCTURE.

The reference hashes decoded block streams with xxhash-rust 0.8.18
(Cargo.lock; crates/fqxv/src/container/mod.rs:146 ``Xxh3``, block.rs:164-176
``stream_digests``), a dependency absent from the reference tree.  This is
the published XXH3_64bits algorithm of xxHash 0.8 (the streaming ``update``
form hashes the concatenation of its inputs, so one call over the digest
layout equals the reference's running hash): short inputs (0-3, 4-8, 9-16,
17-128, 129-240 bytes) and the long-input loop of 1 KiB blocks of 16 stripes
with an accumulator scramble after each block.  Only tests/ and bench.py use
it, as the checker.

Pinned by the xxhash package's XXH3_64bits on every length 0..2099 and
larger sizes (tests/test_xxh3_oracle.py), and by the names digests stored in
the reference's golden archives.
"""
import struct

M = (1 << 64) - 1
SECRET = bytes([
0xb8,0xfe,0x6c,0x39,0x23,0xa4,0x4b,0xbe,0x7c,0x01,0x81,0x2c,0xf7,0x21,0xad,0x1c,
0xde,0xd4,0x6d,0xe9,0x83,0x90,0x97,0xdb,0x72,0x40,0xa4,0xa4,0xb7,0xb3,0x67,0x1f,
0xcb,0x79,0xe6,0x4e,0xcc,0xc0,0xe5,0x78,0x82,0x5a,0xd0,0x7d,0xcc,0xff,0x72,0x21,
0xb8,0x08,0x46,0x74,0xf7,0x43,0x24,0x8e,0xe0,0x35,0x90,0xe6,0x81,0x3a,0x26,0x4c,
0x3c,0x28,0x52,0xbb,0x91,0xc3,0x00,0xcb,0x88,0xd0,0x65,0x8b,0x1b,0x53,0x2e,0xa3,
0x71,0x64,0x48,0x97,0xa2,0x0d,0xf9,0x4e,0x38,0x19,0xef,0x46,0xa9,0xde,0xac,0xd8,
0xa8,0xfa,0x76,0x3f,0xe3,0x9c,0x34,0x3f,0xf9,0xdc,0xbb,0xc7,0xc7,0x0b,0x4f,0x1d,
0x8a,0x51,0xe0,0x4b,0xcd,0xb4,0x59,0x31,0xc8,0x9f,0x7e,0xc9,0xd9,0x78,0x73,0x64,
0xea,0xc5,0xac,0x83,0x34,0xd3,0xeb,0xc3,0xc5,0x81,0xa0,0xff,0xfa,0x13,0x63,0xeb,
0x17,0x0d,0xdd,0x51,0xb7,0xf0,0xda,0x49,0xd3,0x16,0x55,0x26,0x29,0xd4,0x68,0x9e,
0x2b,0x16,0xbe,0x58,0x7d,0x47,0xa1,0xfc,0x8f,0xf8,0xb8,0xd1,0x7a,0xd0,0x31,0xce,
0x45,0xcb,0x3a,0x8f,0x95,0x16,0x04,0x28,0xaf,0xd7,0xfb,0xca,0xbb,0x4b,0x40,0x7e])
assert len(SECRET) == 192
P32_1, P32_2, P32_3 = 0x9E3779B1, 0x85EBCA77, 0xC2B2AE3D
P64_1, P64_2, P64_3, P64_4, P64_5 = 0x9E3779B185EBCA87, 0xC2B2AE3D27D4EB4F, 0x165667B19E3779F9, 0x85EBCA77C2B2AE63, 0x27D4EB2F165667C5
MX1, MX2 = 0x165667919E3779F9, 0x9FB21C651E98DF25
r64 = lambda b, o: struct.unpack_from("<Q", b, o)[0]
r32 = lambda b, o: struct.unpack_from("<I", b, o)[0]
def rotl(x, r): return ((x << r) | (x >> (64 - r))) & M
def fold(a, b): p = a * b; return (p ^ (p >> 64)) & M
def av64(h):
    h ^= h >> 33; h = h * P64_2 & M; h ^= h >> 29; h = h * P64_3 & M; return h ^ (h >> 32)
def av3(h): h ^= h >> 37; h = h * MX1 & M; return h ^ (h >> 32)
def rrmxmx(h, n):
    h ^= rotl(h, 49) ^ rotl(h, 24); h = h * MX2 & M; h ^= (h >> 35) + n; h = h * MX2 & M; return h ^ (h >> 28)
def mix16(b, o, so): return fold(r64(b, o) ^ r64(SECRET, so), r64(b, o + 8) ^ r64(SECRET, so + 8))
def bswap64(x): return int.from_bytes(x.to_bytes(8, 'little'), 'big')
def xxh3_64(b) -> int:
    """XXH3_64bits(b), seed 0."""
    b = bytes(b)
    n = len(b)
    if n == 0: return av64(r64(SECRET, 56) ^ r64(SECRET, 64))
    if n <= 3:
        c = (b[0] << 16) | (b[n >> 1] << 24) | b[n - 1] | (n << 8)
        return av64(c ^ ((r32(SECRET, 0) ^ r32(SECRET, 4))))
    if n <= 8:
        i1, i2 = r32(b, 0), r32(b, n - 4)
        return rrmxmx(((i2 + (i1 << 32)) & M) ^ (r64(SECRET, 8) ^ r64(SECRET, 16)), n)
    if n <= 16:
        lo = r64(b, 0) ^ (r64(SECRET, 24) ^ r64(SECRET, 32)); hi = r64(b, n - 8) ^ (r64(SECRET, 40) ^ r64(SECRET, 48))
        return av3((n + bswap64(lo) + hi + fold(lo, hi)) & M)
    if n <= 128:
        acc = n * P64_1 & M
        if n > 32:
            if n > 64:
                if n > 96:
                    acc += mix16(b, 48, 96) + mix16(b, n - 64, 112)
                acc += mix16(b, 32, 64) + mix16(b, n - 48, 80)
            acc += mix16(b, 16, 32) + mix16(b, n - 32, 48)
        acc += mix16(b, 0, 0) + mix16(b, n - 16, 16)
        return av3(acc & M)
    if n <= 240:
        acc = n * P64_1 & M
        for i in range(8): acc += mix16(b, 16 * i, 16 * i)
        acc = av3(acc & M)
        for i in range(8, n // 16): acc += mix16(b, 16 * i, 16 * (i - 8) + 3)
        acc += mix16(b, n - 16, 136 - 17)
        return av3(acc & M)
    acc = [P32_3, P64_1, P64_2, P64_3, P64_4, P32_2, P64_5, P32_1]
    def a512(o, so):
        for i in range(8):
            dv = r64(b, o + 8 * i); dk = dv ^ r64(SECRET, so + 8 * i)
            acc[i ^ 1] = (acc[i ^ 1] + dv) & M
            acc[i] = (acc[i] + (dk & 0xffffffff) * (dk >> 32)) & M
    def scr():
        for i in range(8):
            a = acc[i]; a ^= a >> 47; a ^= r64(SECRET, 128 + 8 * i); acc[i] = a * P32_1 & M
    nbs = 16; bl = 1024; nb = (n - 1) // bl
    for k in range(nb):
        for s in range(nbs): a512(k * bl + 64 * s, 8 * s)
        scr()
    ns = ((n - 1) - bl * nb) // 64
    for s in range(ns): a512(nb * bl + 64 * s, 8 * s)
    a512(n - 64, 192 - 64 - 7)
    res = n * P64_1 & M
    for i in range(4): res += fold(acc[2 * i] ^ r64(SECRET, 11 + 16 * i), acc[2 * i + 1] ^ r64(SECRET, 11 + 16 * i + 8))
    return av3(res & M)
