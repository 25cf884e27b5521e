"""GPU parity: libgmms.so on cuda:0 against the oracle (std::sort with the
reference comparator contract), bit-exact. Sizes stay small enough for the
oracle to finish in seconds; full-size properties live in test_gpu_full.py."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 63, 64, 1000, 1023, 1024, 1025, 4096, 5000, 65536, 65537, 300_001]


def _to_dev(a, dev):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _host(t):
    return t.cpu().numpy()


def _gpu_sort(gmms, a, dev, **kw):
    params = gmms.make_params(**kw) if kw else None
    out = gmms.sort(_to_dev(a, dev), params=params)
    return _host(out)


def test_library_is_the_hip_build(gmms):
    assert "gfx950" in gmms.version()


@pytest.mark.parametrize("n", [1000, 1024, 5000, 300_001])
def test_squaresort_tiles_match_oracle(gmms, oracle, dev, n):
    print(f"squaresort n={n}", flush=True)
    a = oracle.gen_i32(10 + n, n, 32)
    t = _to_dev(a, dev)
    gmms.squaresort(t)
    assert np.array_equal(_host(t), oracle.basecase_i32(a))


@pytest.mark.parametrize("n", SIZES)
def test_int32_full_range(gmms, oracle, dev, n):
    print(f"int32 full range n={n}", flush=True)
    a = oracle.gen_i32(0x6D6D7300 + n, n, 32)
    assert np.array_equal(_gpu_sort(gmms, a, dev), oracle.sort(a))


@pytest.mark.parametrize("n", [1 << 20, (1 << 20) + 9000])
def test_int32_config1_and_race_size(gmms, oracle, dev, n):
    # config 1: N=2^20 in [0, 2^30); (1<<20)+9000 is scripts/raceConditions.sh's size
    print(f"int32 [0,2^30) n={n}", flush=True)
    a = oracle.gen_i32(0x6D6D7300, n, 34)
    assert np.array_equal(_gpu_sort(gmms, a, dev), oracle.sort(a))


@pytest.mark.parametrize("K", [2, 4, 8, 16, 32, 64])
def test_branching_factors(gmms, oracle, dev, K):
    n = 200_003
    print(f"K={K} n={n}", flush=True)
    a = oracle.gen_i32(77 + K, n, 34)
    assert np.array_equal(_gpu_sort(gmms, a, dev, K=K), oracle.sort(a))


@pytest.mark.parametrize("mergers", [1, 3, 37, 512, 4096])
def test_merger_counts(gmms, oracle, dev, mergers):
    n = 1 << 20
    print(f"mergers={mergers}", flush=True)
    a = oracle.gen_i32(5, n, 54)  # duplicate heavy: ties across every cut
    assert np.array_equal(_gpu_sort(gmms, a, dev, K=16, mergers=mergers), oracle.sort(a))


@pytest.mark.parametrize("kind", ["dup1024", "allsame", "sorted", "reverse", "extremes", "glibc"])
def test_int32_patterns(gmms, oracle, dev, kind):
    n = 400_000
    print(f"int32 pattern {kind}", flush=True)
    if kind == "dup1024":
        a = oracle.gen_i32(3, n, 54)
    elif kind == "allsame":
        a = np.full(n, 7, np.int32)
    elif kind == "sorted":
        a = np.sort(oracle.gen_i32(4, n, 32))
    elif kind == "reverse":
        a = np.sort(oracle.gen_i32(4, n, 32))[::-1].copy()
    elif kind == "extremes":
        a = oracle.gen_i32(6, n, 32)
        a[::5] = np.iinfo(np.int32).max  # the reference's MAXVAL sentinel as a real key
        a[1::7] = np.iinfo(np.int32).min
    else:
        a = oracle.glibc_rand_list(0, n, 1024)  # the reference harness input
    assert np.array_equal(_gpu_sort(gmms, a, dev), oracle.sort(a))


@pytest.mark.parametrize("dt", [np.uint32, np.int64, np.uint64])
def test_integer_dtypes(gmms, oracle, dev, dt):
    import torch

    n = 150_001
    print(f"dtype {dt.__name__}", flush=True)
    raw = oracle.gen_u64(21, n)
    a = raw.astype(np.uint32) if dt == np.uint32 else raw.view(dt)
    a = a.copy()
    a[::11] = np.iinfo(dt).max
    a[1::13] = np.iinfo(dt).min
    want = oracle.sort(a)
    t = torch.from_numpy(a.view(np.int32 if dt == np.uint32 else np.int64).copy()).to(dev)
    tdt = {np.uint32: gmms.UINT32, np.int64: gmms.INT64, np.uint64: gmms.UINT64}[dt]
    alt = torch.empty_like(t)
    res = gmms.sort_inplace(t, alt, dtype=tdt)
    got = _host(alt if res else t).view(dt)
    assert np.array_equal(got, want)


def test_float32_total_order(gmms, oracle, dev):
    n = 262_147
    print("float32", flush=True)
    x = oracle.gen_f32_normal(8, n)
    bits = x.view(np.uint32)
    specials = np.array([0x80000000, 0x00000000, 0x7F800000, 0xFF800000, 0x7FC00000, 0xFFC00000,
                         0x00000001, 0x80000001, 0x7FFFFFFF, 0xFFFFFFFF], np.uint32)
    bits[: len(specials) * 50] = np.tile(specials, 50)
    got = _gpu_sort(gmms, x, dev)
    assert np.array_equal(got.view(np.uint32), oracle.sort(x).view(np.uint32))


def test_float64_total_order(gmms, oracle, dev):
    import torch

    n = 70_001
    print("float64", flush=True)
    x = oracle.gen_u64(9, n).view(np.float64).copy()  # every bit pattern, NaNs included
    x[::17] = -0.0
    x[1::17] = 0.0
    got = _host(gmms.sort(torch.from_numpy(x).to(dev)))
    assert np.array_equal(got.view(np.uint64), oracle.sort(x).view(np.uint64))


@pytest.mark.parametrize("dist", [0, 1, 2, 3])
def test_pairs(gmms, oracle, dev, dist):
    import torch

    n = 200_000
    print(f"pairs dist={dist}", flush=True)
    k, v = oracle.gen_pairs(1000 + dist, n, dist, 1.1, 1 << 20)
    ks, vs = gmms.sort(torch.from_numpy(k.view(np.int64)).to(dev), torch.from_numpy(v.view(np.int32)).to(dev))
    wk, wv = oracle.sort_pairs(k, v)
    assert np.array_equal(_host(ks).view(np.uint64), wk)
    assert np.array_equal(_host(vs).view(np.uint32), wv)


@pytest.mark.parametrize("n", [1, 63, 64, 1023, 1025, 4095, 4097, 65537, 300_001, (1 << 20) + 3])
def test_pairs_sizes(gmms, oracle, dev, n):
    """Pairs across the tile, run and round boundaries (partial store rows of
    the two-pass transpose, lists of every length), Zipf keys (ties broken by
    value)."""
    import torch

    print(f"pairs n={n}", flush=True)
    k, v = oracle.gen_pairs(4000 + n, n, 3, 1.1, 1 << 12)
    ks, vs = gmms.sort(torch.from_numpy(k.view(np.int64)).to(dev), torch.from_numpy(v.view(np.int32)).to(dev))
    wk, wv = oracle.sort_pairs(k, v)
    assert np.array_equal(_host(ks).view(np.uint64), wk)
    assert np.array_equal(_host(vs).view(np.uint32), wv)


@pytest.mark.parametrize("distinct,n,mergers,desc", [(64, 3_000_017, 0, False), (4096, 2_000_003, 101, False),
                                                     (1 << 20, 1 << 21, 0, False), (64, 3_000_017, 0, True)])
def test_pairs_value_is_position_handover(gmms, oracle, dev, distinct, n, mergers, desc):
    """Zipf(1.1) keys over few or many distinct values, value = input
    position: equal keys are ordered by value, so each list's copies of a
    key follow the previous list's, and a merger hands over from one list
    to the next (the elements one list takes before the next list's head
    are written ahead of that list's gallop, and gallops chain from list to
    list). Split tasks (mergers = 101) start inside such runs; descending
    order reverses the lists' order within each key."""
    import torch

    print(f"pairs handover distinct={distinct} n={n} mergers={mergers} desc={desc}", flush=True)
    k, v = oracle.gen_pairs(77 + distinct, n, 3, 1.1, distinct)
    p = gmms.make_params(K=16, mergers=mergers, descending=desc) if (mergers or desc) else None
    ks, vs = gmms.sort(torch.from_numpy(k.view(np.int64)).to(dev), torch.from_numpy(v.view(np.int32)).to(dev),
                       params=p)
    wk, wv = oracle.sort_pairs(k, v)
    if desc:
        wk, wv = wk[::-1], wv[::-1]
    assert np.array_equal(_host(ks).view(np.uint64), wk)
    assert np.array_equal(_host(vs).view(np.uint32), wv)


@pytest.mark.parametrize("n,mergers", [(600_001, 37), (3_000_017, 101), (3_000_017, 3072)])
def test_pairs_split_tasks(gmms, oracle, dev, n, mergers):
    """Tasks split over several mergers each (the merger count fills whole
    generations of `mergers` resident mergers, so a task of 600_001 over 37
    gets 7 mergers), on duplicate keys whose ties are broken by value."""
    import torch

    print(f"pairs n={n} mergers={mergers}", flush=True)
    k, v = oracle.gen_pairs(31 + mergers, n, 0)
    k = k >> np.uint64(48)
    p = gmms.make_params(K=16, mergers=mergers)
    ks, vs = gmms.sort(torch.from_numpy(k.view(np.int64)).to(dev), torch.from_numpy(v.view(np.int32)).to(dev),
                       params=p)
    wk, wv = oracle.sort_pairs(k, v)
    assert np.array_equal(_host(ks).view(np.uint64), wk)
    assert np.array_equal(_host(vs).view(np.uint32), wv)


def test_find_partitions_cuts_are_consistent(gmms, oracle, dev):
    n, L, K = 1 << 18, 1 << 14, 16
    print("find_partitions", flush=True)
    a = oracle.gen_i32(13, n, 56)  # [0,256): heavy ties
    b = oracle.basecase_i32(a, L)  # runs of L sorted
    # internal order of int32 is the sign-flipped bits (unsigned order)
    internal = (b.view(np.uint32) ^ np.uint32(0x80000000))
    t2 = _to_dev(internal.view(np.int32), dev)
    p = gmms.make_params(K=K, mergers=64, slack=100)
    cuts, skips, m = gmms.find_partitions(t2, L, dtype=gmms.UINT32, params=p)
    cuts, skips = _host(cuts), _host(skips)
    assert m > 1
    T = K * L
    lists = [internal[k * L:(k + 1) * L] for k in range(K)]
    for j in range(1, m):
        r = j * T // m
        c = cuts[j]
        assert 0 <= skips[j] <= 100 and c.sum() == r - skips[j]
        # consistency: max of the taken part precedes min of the rest under
        # the stable order (key, list, position)
        taken = [(int(lists[k][c[k] - 1]), k, c[k] - 1) for k in range(K) if c[k] > 0]
        rest = [(int(lists[k][c[k]]), k, c[k]) for k in range(K) if c[k] < L]
        if taken and rest:
            assert max(taken) < min(rest)


def test_merge_round_matches_oracle(gmms, oracle, dev):
    import torch

    n, K = 300_000, 8
    print("merge_round", flush=True)
    a = oracle.gen_i32(17, n, 40)
    b = oracle.basecase_i32(a)
    internal = (b.view(np.uint32) ^ np.uint32(0x80000000)).view(np.int32)
    tin = _to_dev(internal, dev)
    tout = torch.empty_like(tin)
    gmms.merge_round(tin, tout, 1024, dtype=gmms.INT32, params=gmms.make_params(K=K), to_external=True)
    assert np.array_equal(_host(tout), oracle.merge_round_i32(b, K, 1024))


@pytest.mark.parametrize("L", [1000, 2048])
def test_merge_round_pairs_any_list_length(gmms, oracle, dev, L):
    """A merge round over lists of L elements: L = 2048 lets the partition
    search the 256-stride sample array (keys and values), L = 1000 does not
    (lists do not start at multiples of 256). Reference: each task of K
    lists sorted by (key, value) on the CPU."""
    import torch

    n, K = 150_007, 8
    print(f"merge_round pairs L={L}", flush=True)
    k, v = oracle.gen_pairs(23 + L, n, 0)
    k = k >> np.uint64(44)  # duplicate keys: ties broken by value
    for s0 in range(0, n, L):  # sorted input lists
        o = np.lexsort((v[s0:s0 + L], k[s0:s0 + L]))
        k[s0:s0 + L] = k[s0:s0 + L][o]
        v[s0:s0 + L] = v[s0:s0 + L][o]
    want_k, want_v = k.copy(), v.copy()
    for s0 in range(0, n, K * L):
        o = np.lexsort((v[s0:s0 + K * L], k[s0:s0 + K * L]))
        want_k[s0:s0 + K * L] = k[s0:s0 + K * L][o]
        want_v[s0:s0 + K * L] = v[s0:s0 + K * L][o]
    tk, tv = _to_dev(k.view(np.int64), dev), _to_dev(v.view(np.int32), dev)
    ok_, ov_ = torch.empty_like(tk), torch.empty_like(tv)
    gmms.merge_round(tk, ok_, L, tv, ov_, dtype=gmms.PAIR_U64_U32, params=gmms.make_params(K=K))
    assert np.array_equal(_host(ok_).view(np.uint64), want_k)
    assert np.array_equal(_host(ov_).view(np.uint32), want_v)


def test_multimergesort_dropin_returns_result_buffer(gmms, oracle, dev):
    import torch

    for n, rounds in ((1000, 0), (5000, 1), (1 << 17, 2)):
        print(f"multimergesort n={n}", flush=True)
        a = oracle.gen_i32(31, n, 34)
        inp = _to_dev(a, dev)
        out = torch.empty_like(inp)
        res = gmms.multimergesort(inp, out, thread_blocks=482)
        assert gmms.sort_rounds(gmms.INT32, n) == rounds
        assert (res is out) == (rounds % 2 == 1)
        assert np.array_equal(_host(res), oracle.sort(a))


def test_buffers_off_16_byte_alignment(gmms, oracle, dev):
    """Key buffers that start 4 bytes past a 16-byte boundary (views at
    element offset 1): the 16-byte loads and stores of the base case and the
    merge rounds need only element alignment. The guard elements around the
    views stay untouched."""
    import torch

    n = 3 * 65536 + 4097
    print("unaligned buffers", flush=True)
    a = oracle.gen_i32(41, n, 34)
    big = torch.full((n + 2,), -7, dtype=torch.int32, device=dev)
    alt = torch.full((n + 2,), -9, dtype=torch.int32, device=dev)
    big[1:n + 1] = _to_dev(a, dev)
    k, ka = big[1:n + 1], alt[1:n + 1]
    assert k.data_ptr() % 16 == 4
    which = gmms.sort_inplace(k, ka)
    torch.cuda.synchronize()
    got = _host(ka if which else k)
    assert np.array_equal(got, oracle.sort(a))
    for t, guard in ((big, -7), (alt, -9)):
        h = _host(t)
        assert h[0] == guard and h[-1] == guard


def test_pad_writes_maximum(gmms, dev):
    import torch

    t = torch.zeros(100, dtype=torch.int32, device=dev)
    gmms.pad(t, 37)
    h = _host(t)
    assert (h[:37] == np.iinfo(np.int32).max).all() and (h[37:] == 0).all()


def test_repeat_is_bit_identical(gmms, oracle, dev):
    # the reference's scripts/raceConditions.sh: many runs at N = 2^20 + 9000
    n = (1 << 20) + 9000
    print("repeat", flush=True)
    a = oracle.gen_i32(99, n, 54)
    want = oracle.sort(a)
    src = _to_dev(a, dev)
    for _ in range(20):
        got = gmms.sort(src, params=gmms.make_params(mergers=997))
        assert np.array_equal(_host(got), want)


@pytest.mark.parametrize("run", [1024, 4096])
def test_base_run_lengths(gmms, oracle, dev, run):
    # the base case leaves runs of one tile (the reference's M) or of four
    # tiles merged in a workgroup; ragged ends leave some wavefronts idle
    import torch

    p = gmms.make_params(base_run=run)
    for n in (1, 1000, 1024, 1025, 3 * 1024 + 7, 4095, 4096, 4097, 5000, 9000, 65541, 300_001, (1 << 20) + 3):
        print(f"base_run={run} n={n}", flush=True)
        a = oracle.gen_i32(7 * n + run, n, 32)
        assert np.array_equal(_gpu_sort(gmms, a, dev, base_run=run), oracle.sort(a)), n
    k, v = oracle.gen_pairs(5, 70_001, 3, 1.1, 1 << 20)
    ks, vs = gmms.sort(torch.from_numpy(k.view(np.int64)).to(dev), torch.from_numpy(v.view(np.int32)).to(dev),
                       params=p)
    wk, wv = oracle.sort_pairs(k, v)
    assert np.array_equal(_host(ks).view(np.uint64), wk) and np.array_equal(_host(vs).view(np.uint32), wv)
    x = oracle.gen_u64(6, 50_003).view(np.float64).copy()
    got = _host(gmms.sort(torch.from_numpy(x).to(dev), params=p))
    assert np.array_equal(got.view(np.uint64), oracle.sort(x).view(np.uint64))


@pytest.mark.parametrize("run", [1024, 4096])
def test_descending(gmms, oracle, dev, run):
    # greater<T> (cmp(a, b) = a > b): the exact reverse of the ascending order
    import torch

    p = gmms.make_params(descending=True, base_run=run)
    for n in (1, 1000, 4097, 300_001):
        print(f"descending int32 n={n} base_run={run}", flush=True)
        a = oracle.gen_i32(n + 17, n, 32)
        got = _host(gmms.sort(_to_dev(a, dev), params=p))
        assert np.array_equal(got, oracle.sort(a)[::-1]), n
    x = oracle.gen_f32_normal(3, 70_001)
    x.view(np.uint32)[::41] = 0x80000000
    x.view(np.uint32)[1::41] = 0xFFC00000
    got = _host(gmms.sort(_to_dev(x, dev), params=p))
    assert np.array_equal(got.view(np.uint32), oracle.sort(x)[::-1].view(np.uint32))
    u = oracle.gen_u64(4, 50_001).view(np.int64)  # int64 tensor: signed order
    got = _host(gmms.sort(torch.from_numpy(u).to(dev), params=p))
    assert np.array_equal(got, oracle.sort(u)[::-1])
    k, v = oracle.gen_pairs(9, 60_000, 3, 1.1, 1 << 20)
    ks, vs = gmms.sort(torch.from_numpy(k.view(np.int64)).to(dev), torch.from_numpy(v.view(np.int32)).to(dev),
                       params=p)
    wk, wv = oracle.sort_pairs(k, v)
    assert np.array_equal(_host(ks).view(np.uint64), wk[::-1]) and np.array_equal(_host(vs).view(np.uint32), wv[::-1])


def test_descending_segments(gmms, oracle, dev):
    offs = np.array([0, 1, 1000, 5000, 70_000, 70_000, 200_003], np.int64)
    a = oracle.gen_i32(21, int(offs[-1]), 32)
    got = _host(gmms.segmented_sort(_to_dev(a, dev), offs, params=gmms.make_params(descending=True)))
    for s in range(len(offs) - 1):
        x, y = int(offs[s]), int(offs[s + 1])
        assert np.array_equal(got[x:y], oracle.sort(a[x:y])[::-1]), s
