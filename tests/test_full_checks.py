"""The full-size property checks of test_gpu_full.py, run on small CPU
tensors: each must accept the true sorted output and reject a planted
mistake (an element replaced by its neighbour, two elements swapped, a value
paired with the wrong key)."""
import pytest

torch = pytest.importorskip("torch")

from test_gpu_full import SIGN64, _nondecreasing, _same_multiset_u32, _u32  # noqa: E402


def _data(n=5000, seed=3):
    g = torch.Generator().manual_seed(seed)
    return torch.randint(-(1 << 31), (1 << 31) - 1, (n,), dtype=torch.int32, generator=g)


def test_int32_checks_accept_the_sorted_output():
    x = _data()
    y = torch.sort(x).values
    assert _nondecreasing(y)
    assert _same_multiset_u32(_u32(x), _u32(y))


def test_int32_checks_reject_a_duplicated_element():
    x = _data()
    y = torch.sort(x).values.clone()
    y[100] = y[101]  # still sorted, one element lost
    assert _nondecreasing(y)
    assert not _same_multiset_u32(_u32(x), _u32(y))


def test_int32_checks_reject_a_swap():
    x = _data()
    y = torch.sort(x).values.clone()
    y[[7, 8]] = y[[8, 7]]
    assert _same_multiset_u32(_u32(x), _u32(y))
    assert not _nondecreasing(y)


def test_multiset_hash_sees_a_low_bit_change():
    # same top-16-bit histogram and an equal plain sum (and equal sums of any
    # mix that is linear in the low bits): the non-linear terms differ
    x = torch.tensor([0x10000001, 0x10000004], dtype=torch.int64)
    y = torch.tensor([0x10000002, 0x10000003], dtype=torch.int64)
    assert int(x.sum()) == int(y.sum())
    assert not _same_multiset_u32(x, y)


def test_pair_checks():
    n = 4096
    g = torch.Generator().manual_seed(5)
    k = torch.randint(0, 64, (n,), dtype=torch.int64, generator=g) - (1 << 62)  # duplicates
    order = sorted(range(n), key=lambda i: ((int(k[i]) ^ SIGN64), i))
    idx = torch.tensor(order, dtype=torch.int64)
    ks = k[idx]

    def ok(ks, idx):
        perm = torch.equal(torch.bincount(idx, minlength=n), torch.ones(n, dtype=torch.int64))
        s = ks ^ SIGN64
        strictly = bool(((s[:-1] < s[1:]) | ((s[:-1] == s[1:]) & (idx[:-1] < idx[1:]))).all())
        return perm and torch.equal(ks, k[idx]) and strictly

    assert ok(ks, idx)
    bad = idx.clone()
    bad[[10, 11]] = bad[[11, 10]]  # values swapped, keys kept
    assert not ok(ks, bad)
