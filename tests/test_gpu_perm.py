"""GPU: the reorder permutation's byte planes (split/join kernels, and the
device encode/decode of a permutation) against the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [0, 1, 3, 4, 5, 1000, (1 << 20) + 3, 1 << 20])
@pytest.mark.parametrize("misalign", [False, True])
def test_split_join_match_oracle(coder, oracle, n, misalign):
    import torch
    from fqxv_amd import perm
    rng = np.random.default_rng(n + 7 * misalign)
    q = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    base = torch.from_numpy(np.concatenate([np.zeros(1, np.uint32), q]).view(np.int32)).cuda()
    p = base[1:] if misalign else base[1:].clone()
    planes_buf = torch.full((4 * n + 32,), 0xEE, dtype=torch.uint8, device="cuda")
    planes = planes_buf[1:] if misalign else planes_buf
    perm.split(p, planes, coder=coder)
    torch.cuda.synchronize()
    got = planes[:4 * n].cpu().numpy()
    assert np.array_equal(got, oracle.perm_split(q))
    assert np.all(planes_buf[4 * n + (1 if misalign else 0):].cpu().numpy() == 0xEE)  # nothing past 4n
    back = perm.join(planes, n, coder=coder)
    torch.cuda.synchronize()
    assert np.array_equal(back.cpu().numpy().view(np.uint32), q)


def test_permutation_encode_decode_match_oracle(coder, oracle):
    import torch
    import fqxv_amd
    from fqxv_amd import perm
    rng = np.random.default_rng(4)
    for n in (1, 1000, 100000):
        p = rng.permutation(n).astype(np.uint32)
        s = perm.encode_permutation(p, coder=coder)
        assert s == oracle.encode_permutation(p), n
        d = perm.decode_permutation(s, n, coder=coder)
        assert np.array_equal(d.cpu().numpy().view(np.uint32), p)
    s = oracle.encode_permutation(rng.permutation(5000).astype(np.uint32))
    for bad, n, msg in ((s[:20], 5000, "bad permutation"), (s, 4999, "bad permutation"),
                        (s, 5001, "permutation length mismatch")):
        with pytest.raises(fqxv_amd.Malformed, match=f"^{msg}$"):
            perm.decode_permutation(bad, n, coder=coder)
        with pytest.raises(oracle.PermError, match=f"^{msg}$"):
            oracle.decode_permutation(bad, n)
