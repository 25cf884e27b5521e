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
