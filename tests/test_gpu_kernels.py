"""GPU parity of every K-way merge kernel -- pipeline=0 (tile merge, one
merger per wavefront; the default), 1 (the pipelined minBlockHeap) and 2
(tile merge, one merger per workgroup) -- against the oracle, bit-exact:
sizes, branching factors, merger counts, duplicate-heavy, presorted and
clustered inputs (which drive the tile merges' galloping run copies), float
total order, 64-bit keys and (key, value) pairs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[0, 1, 2], ids=["wave", "heap", "workgroup"])
def kern(request):
    return request.param


def _to_dev(a, dev):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _sort(gmms, a, dev, kern, **kw):
    return gmms.sort(_to_dev(a, dev), params=gmms.make_params(pipeline=kern, **kw)).cpu().numpy()


@pytest.mark.parametrize("n", [1025, 5000, 65537, 300_001, (1 << 20) + 9000])
def test_tmerge_int32(gmms, oracle, dev, kern, n):
    print(f"kern={kern} tmerge int32 n={n}", flush=True)
    a = oracle.gen_i32(0x6D6D7300 + n, n, 32)
    assert np.array_equal(_sort(gmms, a, dev, kern), oracle.sort(a))


@pytest.mark.parametrize("K", [2, 4, 8, 16, 32, 64])
def test_tmerge_branching_factors(gmms, oracle, dev, kern, K):
    n = 200_003
    print(f"kern={kern} tmerge K={K}", flush=True)
    a = oracle.gen_i32(77 + K, n, 34)
    assert np.array_equal(_sort(gmms, a, dev, kern, K=K), oracle.sort(a))


@pytest.mark.parametrize("kind", ["dup1024", "allsame", "sorted", "reverse", "extremes", "blocks"])
def test_tmerge_patterns(gmms, oracle, dev, kern, kind):
    n = 700_001
    print(f"kern={kern} tmerge pattern {kind}", flush=True)
    if kind == "dup1024":
        a = oracle.gen_i32(3, n, 54)
    elif kind == "allsame":
        a = np.full(n, -5, np.int32)
    elif kind == "sorted":
        a = np.sort(oracle.gen_i32(4, n, 32))
    elif kind == "reverse":
        a = np.sort(oracle.gen_i32(4, n, 32))[::-1].copy()
    elif kind == "extremes":
        a = oracle.gen_i32(6, n, 32)
        a[::5] = np.iinfo(np.int32).max
        a[1::7] = np.iinfo(np.int32).min
    else:  # sorted runs of 3000 at random offsets: long single-list stretches
        a = np.concatenate([np.sort(oracle.gen_i32(i, 3000, 40)) + i * 7 for i in range(n // 3000 + 1)])[:n]
        a = a.astype(np.int32)
    assert np.array_equal(_sort(gmms, a, dev, kern), oracle.sort(a))


@pytest.mark.parametrize("mergers", [1, 7, 512, 3000])
def test_tmerge_merger_counts(gmms, oracle, dev, kern, mergers):
    n = 1 << 20
    print(f"kern={kern} tmerge mergers={mergers}", flush=True)
    a = oracle.gen_i32(5, n, 54)
    assert np.array_equal(_sort(gmms, a, dev, kern, K=16, mergers=mergers), oracle.sort(a))


def test_tmerge_float_and_64bit(gmms, oracle, dev, kern):
    import torch

    print("tmerge float32/uint64", flush=True)
    x = oracle.gen_f32_normal(8, 262_147)
    x[:10] = [-0.0, 0.0, np.inf, -np.inf, np.nan, -np.nan, 1e-45, -1e-45, 3.0, -3.0]
    got = _sort(gmms, x, dev, kern)
    assert np.array_equal(got.view(np.uint32), oracle.sort(x).view(np.uint32))
    u = oracle.gen_u64(21, 150_001)
    t = torch.from_numpy(u.view(np.int64).copy()).to(dev)
    alt = torch.empty_like(t)
    res = gmms.sort_inplace(t, alt, dtype=gmms.UINT64, params=gmms.make_params(pipeline=kern))
    assert np.array_equal((alt if res else t).cpu().numpy().view(np.uint64), oracle.sort(u))


@pytest.mark.parametrize("dist", [0, 1, 2, 3])
def test_tmerge_pairs(gmms, oracle, dev, kern, dist):
    import torch

    n = 300_000
    print(f"kern={kern} tmerge pairs dist={dist}", flush=True)
    k, v = oracle.gen_pairs(2000 + dist, n, dist, 1.1, 1 << 20)
    p = gmms.make_params(pipeline=kern)
    ks, vs = gmms.sort(torch.from_numpy(k.view(np.int64)).to(dev), torch.from_numpy(v.view(np.int32)).to(dev),
                       params=p)
    wk, wv = oracle.sort_pairs(k, v)
    assert np.array_equal(ks.cpu().numpy().view(np.uint64), wk)
    assert np.array_equal(vs.cpu().numpy().view(np.uint32), wv)
