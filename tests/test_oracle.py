"""Pins of the oracle (CPU only). Each check is independent of the oracle's
own formula: published known answers, glibc itself, numpy's sort, brute
force, or the sizes quoted for the reference."""
import ctypes
import heapq
import itertools
import math
import struct

import numpy as np
import pytest


def test_mt19937_64_known_answer(oracle):
    # C++ standard [rand.predef]: the 10000th invocation of a default-constructed
    # mt19937_64 (seed 5489) produces 9981545732273789042.
    assert oracle.lib.oracle_mt64_nth(5489, 10000) == 9981545732273789042


def test_glibc_rand_matches_libc(oracle):
    # The reference harness generator: srand(0); data[i] = rand() % RANGE with a
    # second rand() consumed per element (src/testMM.cu:156, src/buildData.h:54-65).
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(0)
    expect = []
    for _ in range(3000):
        expect.append(libc.rand() % 1024)
        libc.rand()
    got = oracle.glibc_rand_list(0, 3000, 1024)
    assert got.tolist() == expect


def test_generators_bit_layout(oracle):
    raw = oracle.gen_u64(7, 1000)
    assert np.array_equal(oracle.gen_i32(7, 1000, 32), (raw >> np.uint64(32)).astype(np.uint32).view(np.int32))
    assert np.array_equal(oracle.gen_i32(7, 1000, 34), (raw >> np.uint64(34)).astype(np.int32))
    small = oracle.gen_i32(7, 1000, 54)
    assert small.min() >= 0 and small.max() < 1024
    lens = oracle.gen_seglens(3, 5000, 10.0, 22.0)
    assert lens.min() >= 1024 and lens.max() < (1 << 22)


@pytest.mark.parametrize("dt", [np.int32, np.uint32, np.int64, np.uint64])
def test_sort_matches_numpy(oracle, dt):
    rng = np.random.default_rng(11)
    info = np.iinfo(dt)
    for n in (0, 1, 2, 7, 1000, 4097):
        a = rng.integers(info.min, info.max, size=n, dtype=dt, endpoint=True)
        assert np.array_equal(oracle.sort(a), np.sort(a, kind="stable"))
    dup = rng.integers(0, 3, size=5000).astype(dt)
    assert np.array_equal(oracle.sort(dup), np.sort(dup))


def test_sort_brute_force_small(oracle):
    # every permutation of a multiset sorts to the same sequence
    base = np.array([3, -1, 3, 0, 2147483647, -2147483648, 0], np.int32)
    want = sorted(base.tolist())
    for perm in itertools.islice(itertools.permutations(range(7)), 0, 5040, 7):
        assert oracle.sort(base[list(perm)]).tolist() == want


def _f32(bits):
    return struct.unpack("<f", struct.pack("<I", bits))[0]


def test_float_total_order_special_values(oracle):
    # IEEE 754 totalOrder on bits: -NaN < -inf < -1 < -0 < +0 < 1 < +inf < +NaN
    order = [0xFFC00001, 0xFFC00000, 0xFF800000, 0xBF800000, 0x80000001, 0x80000000,
             0x00000000, 0x00000001, 0x3F800000, 0x7F800000, 0x7FC00000, 0x7FC00001]
    keys = [oracle.lib.oracle_f32_order_key(b) for b in order]
    assert keys == sorted(keys) and len(set(keys)) == len(keys)
    rng = np.random.default_rng(5)
    a = np.array(order * 3, np.uint32)[rng.permutation(len(order) * 3)].view(np.float32)
    got = oracle.sort(a).view(np.uint32).tolist()
    assert got == [b for b in order for _ in range(3)]
    # non-NaN values agree with numeric order, -0 before +0
    x = oracle.gen_f32_normal(9, 20000)
    x[::97] = -0.0
    x[1::97] = 0.0
    s = oracle.sort(x)
    assert np.array_equal(s, np.sort(x))
    z = s.view(np.uint32)
    zeros = z[(s == 0)]
    assert (zeros[: (zeros == 0x80000000).sum()] == 0x80000000).all()


def test_f64_order_key_monotone(oracle):
    vals = [-math.inf, -1e300, -1.0, -5e-324, -0.0, 0.0, 5e-324, 1.0, 1e300, math.inf]
    bits = [struct.unpack("<Q", struct.pack("<d", v))[0] for v in vals]
    keys = [oracle.lib.oracle_f64_order_key(b) for b in bits]
    assert keys == sorted(keys) and len(set(keys)) == len(keys)


def test_pairs_match_lexsort(oracle):
    for dist in range(4):
        k, v = oracle.gen_pairs(100 + dist, 20000, dist, 1.1, 1 << 12)
        if dist == 1:
            assert np.all(k[:-1] <= k[1:])
        if dist == 2:
            assert np.all(k[:-1] >= k[1:])
        ks, vs = oracle.sort_pairs(k, v)
        order = np.lexsort((v, k))
        assert np.array_equal(ks, k[order]) and np.array_equal(vs, v[order])
    # zipf: heavy duplicates
    k, _ = oracle.gen_pairs(5, 20000, 3, 1.1, 1 << 20)
    assert len(np.unique(k)) < 20000 // 2


def test_schedule_closed_forms(oracle):
    # sizes quoted for the reference defaults (K=64, M=1024)
    assert oracle.lib.oracle_rounds(1 << 28, 64, 1024) == 3
    assert oracle.lib.oracle_rounds(10**8 + 17, 64, 1024) == 3
    assert oracle.lib.oracle_rounds(1 << 20, 64, 1024) == 2
    assert oracle.lib.oracle_rounds(1024, 64, 1024) == 0
    assert oracle.lib.oracle_rounds(1025, 64, 1024) == 1
    for n in (1, 1000, 1024, 5000, 65536, 65537, 10**7 + 3):
        for K in (2, 4, 16, 64):
            r = oracle.lib.oracle_rounds(n, K, 1024)
            assert r == (0 if n <= 1024 else math.ceil(math.log(n / 1024, K) - 1e-12))
            L = 1024
            for _ in range(r):
                t = oracle.lib.oracle_tasks(n, K, L)
                e = oracle.lib.oracle_edge_task_size(n, K, L)
                assert (t - 1) * K * L + e == n and 0 < e <= K * L
                L *= K
    assert [oracle.lib.oracle_merger_target(j, 1000, 7) for j in range(8)] == [j * 1000 // 7 for j in range(8)]


def test_corank_against_heapq(oracle):
    rng = np.random.default_rng(3)
    for trial in range(60):
        K = int(rng.integers(1, 9))
        stride = int(rng.integers(1, 40))
        lens = [int(rng.integers(0, stride + 1)) for _ in range(K)]
        flat = np.zeros(K * stride, np.int32)
        for k in range(K):
            flat[k * stride:k * stride + lens[k]] = np.sort(rng.integers(0, 6, lens[k]))
        tags = heapq.merge(*[[(int(flat[k * stride + p]), k, p) for p in range(lens[k])] for k in range(K)])
        tags = list(tags)
        for r in range(0, sum(lens) + 1, max(1, sum(lens) // 7)):
            s = oracle.corank_i32(flat, stride, lens, r)
            want = [sum(1 for t in tags[:r] if t[1] == k) for k in range(K)]
            assert s.tolist() == want


def test_basecase_and_round_oracle(oracle):
    a = oracle.gen_i32(1, 5000, 54)
    b = oracle.basecase_i32(a)
    for t in range(0, 5000, 1024):
        assert np.array_equal(b[t:t + 1024], np.sort(a[t:t + 1024]))
    r1 = oracle.merge_round_i32(b, 4, 1024)
    assert np.array_equal(r1[:4096], np.sort(a[:4096])) and np.array_equal(r1[4096:], np.sort(a[4096:]))


@pytest.mark.parametrize("threads", [1, 4, 0])
def test_psort_i32_matches_numpy(oracle, threads):
    """oracle_psort_i32 (libstdc++ parallel sort): the bit-exact checker of
    the 2^28 benchmark output and the CPU baseline. Pinned by numpy's sort
    on the same bits, with 1, 4 and all host threads."""
    rng = np.random.default_rng(21 + threads)
    cases = [np.zeros(0, np.int32), np.array([7], np.int32),
             rng.integers(-2**31, 2**31, 4097, dtype=np.int64).astype(np.int32),
             oracle.gen_i32(0x6D6D7301, 1 << 20, 32),
             rng.integers(0, 3, 1 << 20).astype(np.int32),  # duplicate-heavy
             np.full(100_000, -2**31, np.int32)]
    for a in cases:
        b = a.copy()
        oracle.lib.oracle_psort_i32(b.ctypes.data, b.size, threads)
        assert np.array_equal(b, np.sort(a)), (a.size, threads)


@pytest.mark.parametrize("threads", [1, 4, 0])
def test_psort_u64pairs_matches_lexsort(oracle, threads):
    """oracle_psort_u64pairs: (key, value) lexicographic, pinned by numpy's
    lexsort, including keys with many equal copies and equal values."""
    rng = np.random.default_rng(31 + threads)
    for n in (0, 1, 4097, 1 << 20):
        k = rng.integers(0, 2**64, n, dtype=np.uint64)
        k[::3] = k[0] if n else k[::3]  # a third of the keys equal
        v = rng.integers(0, 8, n, dtype=np.uint64).astype(np.uint32)
        k2, v2 = k.copy(), v.copy()
        oracle.lib.oracle_psort_u64pairs(k2.ctypes.data, v2.ctypes.data, n, threads)
        order = np.lexsort((v, k))
        assert np.array_equal(k2, k[order]) and np.array_equal(v2, v[order]), (n, threads)


def test_time_sort_reports_positive_medians(oracle):
    a = oracle.gen_i32(3, 1 << 16, 32)
    for threads in (1, 2):
        t = oracle.lib.oracle_time_sort_i32(a.ctypes.data, a.size, threads, 3)
        assert 0.0 < t < 10.0


def test_normal_generator_distribution(oracle):
    """gen_f32_normal is normal(0, 1): Kolmogorov-Smirnov against the
    standard normal CDF, sample mean and variance within 5 standard errors
    on 10^6 draws."""
    from scipy import stats

    n = 10**6
    x = oracle.gen_f32_normal(0x6D6D7304, n).astype(np.float64)
    assert np.isfinite(x).all()
    assert abs(x.mean()) < 5 / math.sqrt(n)
    assert abs(x.var() - 1.0) < 5 * math.sqrt(2.0 / n)
    ks = stats.kstest(x, "norm")
    assert ks.pvalue > 1e-3, ks
    # tails: P(|x| > 3) = 0.0027
    assert abs((np.abs(x) > 3).mean() - 0.0026998) < 5 * math.sqrt(0.0027 / n)


def test_zipf_generator_rank_frequency(oracle):
    """Zipf(1.1) keys over 2^20 values: the log-log rank-frequency slope
    over the 1000 most frequent keys is -1.1, and the most frequent key
    takes the share 1 / H(2^20, 1.1) of the draws."""
    n = 1 << 20
    k, v = oracle.gen_pairs(0x6D6D7313, n, 3, 1.1, 1 << 20)
    assert np.array_equal(v, np.arange(n, dtype=np.uint32))
    _, counts = np.unique(k, return_counts=True)
    counts = np.sort(counts)[::-1].astype(np.float64)
    r = np.arange(1, 1001, dtype=np.float64)
    slope = np.polyfit(np.log(r), np.log(counts[:1000]), 1)[0]
    assert abs(slope + 1.1) < 0.05, slope
    h = float(np.sum(np.arange(1, (1 << 20) + 1, dtype=np.float64) ** -1.1))
    share = counts[0] / n
    assert abs(share - 1.0 / h) < 5 * math.sqrt(share / n), (share, 1.0 / h)


def test_sorted_and_reverse_pairs_are_the_uniform_draw(oracle):
    """Distributions 1 and 2 are the uniform draw (distribution 0, same
    seed) sorted ascending and descending; values are the indexes."""
    n = 100_000
    k0, v0 = oracle.gen_pairs(77, n, 0, 1.1, 1 << 20)
    k1, v1 = oracle.gen_pairs(77, n, 1, 1.1, 1 << 20)
    k2, v2 = oracle.gen_pairs(77, n, 2, 1.1, 1 << 20)
    assert np.array_equal(k1, np.sort(k0)) and np.array_equal(k2, np.sort(k0)[::-1])
    for v in (v0, v1, v2):
        assert np.array_equal(v, np.arange(n, dtype=np.uint32))


def test_seglens_log_uniform(oracle):
    """gen_seglens: log2 of the lengths is uniform on [10, 22) (KS test),
    the lengths are integers in [2^10, 2^22)."""
    from scipy import stats

    lens = oracle.gen_seglens(0x6D6D7303, 8192, 10.0, 22.0)
    assert lens.min() >= 1 << 10 and lens.max() < 1 << 22
    # floor() shifts log2 down by < 1/len: negligible against the KS bound
    ks = stats.kstest((np.log2(lens) - 10.0) / 12.0, "uniform")
    assert ks.pvalue > 1e-3, ks
    # the literal 8192-segment batch: about 4.06e9 keys
    assert 4.0e9 < lens.sum() < 4.12e9


def test_uniform_int32_generator_spreads(oracle):
    """(int32)(u64 >> 32) is uniform over the full int32 range: a chi-square
    test over 256 buckets of the top byte on 2^20 draws; shift 34 stays in
    [0, 2^30) and shift 54 in [0, 1024) with every value present."""
    from scipy import stats

    a = oracle.gen_i32(0x6D6D7301, 1 << 20, 32)
    top = (a.view(np.uint32) >> 24).astype(np.int64)
    chi = stats.chisquare(np.bincount(top, minlength=256))
    assert chi.pvalue > 1e-3, chi
    b = oracle.gen_i32(0x6D6D7302, 1 << 20, 34)
    assert b.min() >= 0 and b.max() < 1 << 30 and b.max() > (1 << 30) - (1 << 14)
    c = oracle.gen_i32(0x6D6D7302, 1 << 20, 54)
    assert np.array_equal(np.unique(c), np.arange(1024, dtype=np.int32))


@pytest.mark.parametrize("threads", [1, 0])
def test_psort_u64_and_f32(oracle, threads):
    """oracle_psort_u64 against numpy's sort; oracle_psort_f32 against the
    single-thread total-order sort and an independent numpy key (negative
    floats reversed, -0 before +0, NaNs by sign and payload)."""
    rng = np.random.default_rng(41)
    u = rng.integers(0, 2**64, 1 << 18, dtype=np.uint64)
    u[::5] = 7
    assert np.array_equal(oracle.psort(u, threads), np.sort(u))
    bits = rng.integers(0, 2**32, 1 << 18, dtype=np.uint64).astype(np.uint32)
    bits[:4] = [0x80000000, 0, 0x7FC00000, 0xFFC00000]
    f = bits.view(np.float32)
    got = oracle.psort(f, threads).view(np.uint32)
    assert np.array_equal(got, oracle.sort(f).view(np.uint32))
    key = np.where(bits >> 31 == 1, ~bits, bits | np.uint32(0x80000000)).astype(np.uint32)
    assert np.array_equal(got, bits[np.argsort(key, kind="stable")])


def test_time_sort_runs(oracle):
    a = oracle.gen_i32(5, 1 << 16, 32)
    secs = np.zeros(3)
    assert oracle.lib.oracle_time_sort_i32_runs(a.ctypes.data, a.size, 1, 1, 3, secs.ctypes.data) == 0
    assert (secs > 0).all() and (secs < 10).all()
