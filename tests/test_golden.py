"""Golden vectors (tests/golden/vectors.npz, written by
tests/golden/make_golden.py). The reference ships no vectors of its own, so
these are pinned here without the oracle: every expected output must be a
permutation of its input that is ordered under the comparator contract
(src/cmp.hxx:21-73), checked with numpy's own sorts and an independently
written float total-order key. Then the oracle must reproduce the vectors
from its seeds, and (GPU) the sort must reproduce the outputs."""
import os

import numpy as np
import pytest

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "vectors.npz"), allow_pickle=False)

KEYED = ["i32_full", "i32_cfg1", "i32_dup", "i32_glibc", "u64"]


def _f32_key(bits):
    b = bits.astype(np.uint64)
    return np.where(b >> 31 == 1, (~b) & 0xFFFFFFFF, b | 0x80000000)


def _f64_key(bits):
    b = bits.astype(np.uint64)
    return np.where(b >> np.uint64(63) == 1, ~b, b | np.uint64(1 << 63))


@pytest.mark.parametrize("name", KEYED)
def test_integer_vectors_are_sorted_permutations(name):
    a, w = G[name + "_in"], G[name + "_out"]
    assert np.array_equal(np.sort(a), w)  # numpy's sort, not the oracle
    assert (w[1:] >= w[:-1]).all()


def test_float_vectors_follow_total_order():
    for nm, key in (("f32", _f32_key), ("f64", _f64_key)):
        a, w = G[nm + "_in_bits"], G[nm + "_out_bits"]
        assert np.array_equal(np.sort(a), np.sort(w))      # same bit patterns
        kw = key(w)
        assert (kw[1:] >= kw[:-1]).all()                   # ordered by the total-order key
    w32 = G["f32_out_bits"]
    # -0.0 sorts before +0.0, -NaN first, +NaN last
    pos_m0 = np.flatnonzero(w32 == 0x80000000)
    pos_p0 = np.flatnonzero(w32 == 0x00000000)
    assert pos_m0.max() < pos_p0.min()
    assert w32[0] == 0xFFFFFFFF and w32[-1] == 0x7FFFFFFF


@pytest.mark.parametrize("d", range(4))
def test_pair_vectors_follow_key_value_order(d):
    k, v = G[f"pairs{d}_k_in"], G[f"pairs{d}_v_in"]
    order = np.lexsort((v, k))
    assert np.array_equal(k[order], G[f"pairs{d}_k_out"])
    assert np.array_equal(v[order], G[f"pairs{d}_v_out"])
    assert np.array_equal(v, np.arange(v.size, dtype=np.uint32))


def test_segment_vectors():
    offs, a, w = G["seg_offsets"], G["seg_in"], G["seg_out"]
    for s in range(offs.size - 1):
        assert np.array_equal(np.sort(a[offs[s]:offs[s + 1]]), w[offs[s]:offs[s + 1]])
    head = G["seglens_head"]
    assert ((head >= 1 << 10) & (head < 1 << 22)).all()


def test_oracle_reproduces_vectors(oracle):
    assert np.array_equal(oracle.gen_i32(0x6D6D7301, 4099, 32), G["i32_full_in"])
    assert np.array_equal(oracle.gen_i32(0x6D6D7300, 5000, 34), G["i32_cfg1_in"])
    assert np.array_equal(oracle.glibc_rand_list(0, 2048, 1024), G["i32_glibc_in"])
    assert np.array_equal(oracle.gen_u64(0x6D6D7306, 3000), G["u64_in"])
    f = oracle.gen_f32_normal(0x6D6D7304, 3000).view(np.uint32)
    assert np.array_equal(f[200:], G["f32_in_bits"][200:])
    k, v = oracle.gen_pairs(0x6D6D7313, 2500, 3, 1.1, 1 << 20)
    assert np.array_equal(k, G["pairs3_k_in"]) and np.array_equal(v, G["pairs3_v_in"])
    lens = oracle.gen_seglens(0x6D6D7303, 8192, 10, 22)
    assert np.array_equal(lens[:64], G["seglens_head"]) and lens.sum() == G["seglens_sum"][0]
    for name in KEYED:
        assert np.array_equal(oracle.sort(G[name + "_in"]), G[name + "_out"])
    assert np.array_equal(oracle.sort(G["f32_in_bits"].view(np.float32)).view(np.uint32), G["f32_out_bits"])
    for d in range(4):
        wk, wv = oracle.sort_pairs(G[f"pairs{d}_k_in"], G[f"pairs{d}_v_in"])
        assert np.array_equal(wk, G[f"pairs{d}_k_out"]) and np.array_equal(wv, G[f"pairs{d}_v_out"])


@pytest.mark.gpu
def test_gpu_sort_reproduces_vectors(gmms, dev):
    import torch

    def run(a):
        return gmms.sort(torch.from_numpy(np.ascontiguousarray(a)).to(dev)).cpu().numpy()

    for name in KEYED:
        assert np.array_equal(run(G[name + "_in"]), G[name + "_out"]), name
    assert np.array_equal(run(G["f32_in_bits"].view(np.float32)).view(np.uint32), G["f32_out_bits"])
    assert np.array_equal(run(G["f64_in_bits"].view(np.float64)).view(np.uint64), G["f64_out_bits"])
    for d in range(4):
        k = torch.from_numpy(G[f"pairs{d}_k_in"].view(np.int64)).to(dev)
        v = torch.from_numpy(G[f"pairs{d}_v_in"].view(np.int32)).to(dev)
        ks, vs = gmms.sort(k, v)
        assert np.array_equal(ks.cpu().numpy().view(np.uint64), G[f"pairs{d}_k_out"])
        assert np.array_equal(vs.cpu().numpy().view(np.uint32), G[f"pairs{d}_v_out"])
    got = gmms.segmented_sort(torch.from_numpy(G["seg_in"]).to(dev), G["seg_offsets"]).cpu().numpy()
    assert np.array_equal(got, G["seg_out"])
