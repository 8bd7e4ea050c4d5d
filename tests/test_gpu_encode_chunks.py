"""Batched encode at the staging chunk's edges, against the oracle.

The encoder's reverse walk stages the input 2,048 positions (plus the 4
bytes before them) at a time.  Chunks wholly inside the input are fetched
with wide aligned loads and realigned when staged; the block's first chunk
and a chunk below a top chunk of fewer than 4 positions take the byte-exact
path.  These cases put every block size around the chunk boundaries (the
top chunk 0..5 positions long), every start alignment, both orders and
encode-best, and an alphabet too wide for the dense order-1 table.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CH = 2048


def _blocks():
    rng = np.random.default_rng(77)
    blocks, orders = [], []
    for k in (1, 2, 3):
        for extra in range(6):
            n = k * CH + extra
            for kind in range(3):
                if kind == 0:    # skewed, few symbols: dense tables
                    b = rng.choice([3, 9, 27, 81], n, p=[0.6, 0.25, 0.1, 0.05]).astype(np.uint8)
                elif kind == 1:  # correlated (order-1 structure)
                    b = np.cumsum(rng.integers(-2, 3, n)).astype(np.uint8) % 40
                    b = b.astype(np.uint8)
                else:            # wide alphabet: global order-1 tables
                    b = rng.integers(0, 256, n).astype(np.uint8)
                blocks.append(b.tobytes())
                orders.append((k + extra + kind) & 1)
    # Sizes just below and above a chunk with both orders on each.
    for n in (CH - 1, CH, CH + 1, 4 * CH - 3, 4 * CH + 3):
        b = rng.integers(0, 12, n).astype(np.uint8).tobytes()
        blocks += [b, b]
        orders += [0, 1]
    return blocks, np.array(orders, dtype=np.int32)


def _arena(blocks, shift):
    import torch
    offs, parts, at = [], [], 0
    for j, b in enumerate(blocks):
        pad = (shift + j) % 4
        parts.append(bytes(pad))
        at += pad
        offs.append(at)
        parts.append(b)
        at += len(b)
    arena = torch.from_numpy(np.frombuffer(b"".join(parts) + bytes(64), dtype=np.uint8).copy()).cuda()
    return arena, np.array(offs, dtype=np.uint64), np.array([len(b) for b in blocks], dtype=np.uint64)


@pytest.mark.parametrize("shift", [0, 1, 2, 3])
def test_encode_chunk_edges_match_oracle(coder, oracle, shift):
    import torch
    blocks, orders = _blocks()
    arena, offs, sizes = _arena(blocks, shift)
    r = coder.encode_batch(arena, offs, sizes, orders)
    torch.cuda.synchronize()
    assert np.all(r.status.cpu().numpy() == 0)
    ol = r.out_len.cpu().numpy()
    host = r.dst.cpu().numpy()
    for j, b in enumerate(blocks):
        got = host[int(r.dst_off[j]):int(r.dst_off[j]) + int(ol[j])].tobytes()
        assert got == oracle.encode(b, int(orders[j])), (shift, j, len(b), int(orders[j]))


def test_encode_best_chunk_edges_match_oracle(coder, oracle):
    import torch
    blocks, _ = _blocks()
    arena, offs, sizes = _arena(blocks, 1)
    ties = (np.arange(len(blocks)) % 2).astype(np.int32)
    r = coder.encode_best_batch(arena, offs, sizes, ties)
    torch.cuda.synchronize()
    assert np.all(r.status.cpu().numpy() == 0)
    ol = r.out_len.cpu().numpy()
    host = r.dst.cpu().numpy()
    for j, b in enumerate(blocks):
        got = host[int(r.dst_off[j]):int(r.dst_off[j]) + int(ol[j])].tobytes()
        assert got == oracle.encode_best(b, int(ties[j])), (j, len(b))
