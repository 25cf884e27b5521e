"""GPU parity of the batched segmented sort: every segment of a batch equals
the oracle's sort of that segment, bit for bit, and equals what the plain
per-array sort (gmms_sort on the segment alone) returns."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _to_dev(a, dev):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _offsets(lens, start=0):
    return np.concatenate([[start], start + np.cumsum(np.asarray(lens, np.int64))]).astype(np.int64)


def _bits(x):
    return x.view({4: np.uint32, 8: np.uint64}[x.dtype.itemsize])


def _check_segments(oracle, got, src, offs):
    # compared as bit patterns: NaNs compare equal to themselves only so
    for s in range(len(offs) - 1):
        a, b = int(offs[s]), int(offs[s + 1])
        want = oracle.sort(src[a:b])
        if not np.array_equal(_bits(got[a:b]), _bits(want)):
            raise AssertionError(f"segment {s} [{a},{b}) differs from the oracle")


MIXED = [0, 1, 2, 63, 1023, 1024, 1025, 4096, 5000, 0, 65537, 300_001, 1 << 20, 7, 16385]


def test_int32_mixed_lengths(gmms, oracle, dev):
    print("segments int32 mixed lengths", flush=True)
    offs = _offsets(MIXED)
    a = oracle.gen_i32(0x5E6, int(offs[-1]), 32)
    got = gmms.segmented_sort(_to_dev(a, dev), offs).cpu().numpy()
    _check_segments(oracle, got, a, offs)


def test_log_uniform_batch_equals_per_array(gmms, oracle, dev):
    import torch

    print("segments log-uniform 2^10..2^17 x 300 vs per-array", flush=True)
    lens = oracle.gen_seglens(0x6D6D7303, 300, 10, 17)
    offs = _offsets(lens)
    a = oracle.gen_i32(0x6D6D7303, int(offs[-1]), 33)  # [0, 2^31)
    t = _to_dev(a, dev)
    got = gmms.segmented_sort(t, offs).cpu().numpy()
    _check_segments(oracle, got, a, offs)
    per = torch.empty_like(t)
    for s in range(len(lens)):
        x, y = int(offs[s]), int(offs[s + 1])
        per[x:y] = gmms.sort(t[x:y])
    assert np.array_equal(per.cpu().numpy(), got)


def test_nonzero_first_offset(gmms, oracle, dev):
    import torch

    print("segments starting at 777", flush=True)
    offs = _offsets([5000, 3, 40000, 1024], start=777)
    n = int(offs[-1])
    a = oracle.gen_i32(31, n, 32)
    k = _to_dev(a, dev)
    ka = torch.full_like(k, -5)
    plan = gmms.SegPlan(gmms.INT32, offs)
    alt = plan.sort(k, ka)
    torch.cuda.synchronize()
    res = (ka if alt else k).cpu().numpy()
    _check_segments(oracle, res, a, offs)
    # elements before offsets[0] are neither read nor written
    assert np.array_equal(k.cpu().numpy()[:777], a[:777])
    assert (ka.cpu().numpy()[:777] == -5).all()


@pytest.mark.parametrize("dt", ["uint32", "int64", "uint64"])
def test_integer_dtypes(gmms, oracle, dev, dt):
    print(f"segments {dt}", flush=True)
    offs = _offsets([3000, 70000, 1, 0, 129_000, 2048])
    n = int(offs[-1])
    u = oracle.gen_u64(500 + len(dt), n)
    a = {"uint32": (u >> 32).astype(np.uint32), "int64": u.view(np.int64), "uint64": u}[dt]
    got = gmms.segmented_sort(_to_dev(a, dev), offs).cpu().numpy()
    _check_segments(oracle, got, a, offs)


def test_float_total_order(gmms, oracle, dev):
    print("segments float32/float64 with specials", flush=True)
    offs = _offsets([20000, 1500, 99_999, 1])
    n = int(offs[-1])
    x = oracle.gen_f32_normal(8, n)
    bits = x.view(np.uint32)
    specials = np.array([0x80000000, 0x00000000, 0x7F800000, 0xFF800000, 0x7FC00000, 0xFFC00000,
                         0x00000001, 0x80000001, 0x7FFFFFFF, 0xFFFFFFFF], np.uint32)
    bits[::97] = np.resize(specials, bits[::97].size)
    got = gmms.segmented_sort(_to_dev(x, dev), offs).cpu().numpy()
    _check_segments(oracle, got, x, offs)
    y = oracle.gen_u64(9, n).view(np.float64).copy()
    y[::13] = -0.0
    got = gmms.segmented_sort(_to_dev(y, dev), offs).cpu().numpy()
    for s in range(len(offs) - 1):
        a, b = int(offs[s]), int(offs[s + 1])
        assert np.array_equal(got[a:b].view(np.uint64), oracle.sort(y[a:b]).view(np.uint64))


@pytest.mark.parametrize("dist", [0, 3])
def test_pairs(gmms, oracle, dev, dist):
    print(f"segments pairs dist={dist}", flush=True)
    offs = _offsets([40000, 5, 1024, 150_000, 333])
    n = int(offs[-1])
    k, v = oracle.gen_pairs(700 + dist, n, dist, 1.1, 1 << 20)
    ks, vs = gmms.segmented_sort(_to_dev(k.view(np.int64), dev), offs, _to_dev(v.view(np.int32), dev))
    ks, vs = ks.cpu().numpy().view(np.uint64), vs.cpu().numpy().view(np.uint32)
    for s in range(len(offs) - 1):
        a, b = int(offs[s]), int(offs[s + 1])
        wk, wv = oracle.sort_pairs(k[a:b], v[a:b])
        assert np.array_equal(ks[a:b], wk) and np.array_equal(vs[a:b], wv), f"segment {s}"


def test_pairs_value_is_position_handover(gmms, oracle, dev):
    """Zipf keys over 64 values, value = input position, in segments long
    enough for several merge rounds: the persistent merger's hand-overs and
    chained gallops (each list's copies of a key follow the previous
    list's)."""
    print("segments pairs handover", flush=True)
    offs = _offsets([1_000_003, 70_000, 2_000_000, 4097])
    n = int(offs[-1])
    k, v = oracle.gen_pairs(711, n, 3, 1.1, 64)
    for desc in (False, True):
        p = gmms.make_params(descending=True) if desc else None
        ks, vs = gmms.segmented_sort(_to_dev(k.view(np.int64), dev), offs, _to_dev(v.view(np.int32), dev), params=p)
        ks, vs = ks.cpu().numpy().view(np.uint64), vs.cpu().numpy().view(np.uint32)
        for s in range(len(offs) - 1):
            a, b = int(offs[s]), int(offs[s + 1])
            wk, wv = oracle.sort_pairs(k[a:b], v[a:b])
            if desc:
                wk, wv = wk[::-1], wv[::-1]
            assert np.array_equal(ks[a:b], wk) and np.array_equal(vs[a:b], wv), f"segment {s} descending={desc}"


@pytest.mark.parametrize("K", [2, 4, 64])
def test_forced_branching_factor(gmms, oracle, dev, K):
    print(f"segments K={K}", flush=True)
    offs = _offsets([200_003, 9000, 1, 70_000])
    a = oracle.gen_i32(90 + K, int(offs[-1]), 34)
    got = gmms.segmented_sort(_to_dev(a, dev), offs, params=gmms.make_params(K=K)).cpu().numpy()
    _check_segments(oracle, got, a, offs)


def test_many_small_and_duplicates(gmms, oracle, dev):
    print("segments 4000 small, keys in [0,4)", flush=True)
    rng = np.random.default_rng(5)
    lens = rng.integers(0, 3000, 4000)
    offs = _offsets(lens)
    a = (oracle.gen_u64(77, int(offs[-1])) >> 62).astype(np.int32)
    got = gmms.segmented_sort(_to_dev(a, dev), offs).cpu().numpy()
    _check_segments(oracle, got, a, offs)


def test_tuning_table_per_class(gmms, oracle, dev):
    print("segments tuning table", flush=True)
    kk0, mm0 = gmms.seg_tuning_get(gmms.INT32)
    try:
        kk, mm = list(kk0), list(mm0)
        kk[13], mm[13] = 4, 8192     # 4097..8192
        kk[17], mm[17] = 64, 0       # 65537..131072
        kk[20], mm[20] = 8, 1 << 15  # up to 2^20
        gmms.seg_tuning_set(gmms.INT32, kk, mm)
        offs = _offsets([8000, 100_000, 1 << 20, 700_000, 5000])
        plan = gmms.SegPlan(gmms.INT32, offs)
        # long segments start from 4096-element runs
        assert plan.segment(0) == (4, 1)
        assert plan.segment(1) == (64, 1)
        assert plan.segment(2) == (8, 3)
        a = oracle.gen_i32(4242, int(offs[-1]), 32)
        got = gmms.segmented_sort(_to_dev(a, dev), offs).cpu().numpy()
        _check_segments(oracle, got, a, offs)
    finally:
        gmms.seg_tuning_set(gmms.INT32, kk0, mm0)


def test_graph_replays_match_eager(gmms, oracle, dev):
    import torch

    print("segments hip graph replays", flush=True)
    offs = _offsets([30000, 1, 250_000, 4096, 0, 77_777])
    n = int(offs[-1])
    plan = gmms.SegPlan(gmms.INT32, offs)
    k = torch.empty(n, dtype=torch.int32, device=dev)
    ka = torch.empty_like(k)
    g = plan.graph(k, ka)
    for rep in range(3):
        a = oracle.gen_i32(1000 + rep, n, 32)
        # eager sort of the same input between replays
        e, ea = _to_dev(a, dev), torch.empty_like(k)
        ealt = plan.sort(e, ea)
        torch.cuda.synchronize()
        eager = (ea if ealt else e).cpu().numpy()
        k.copy_(_to_dev(a, dev))
        alt = g.launch()
        torch.cuda.synchronize()
        got = (ka if alt else k).cpu().numpy()
        assert np.array_equal(got, eager), f"replay {rep}"
        _check_segments(oracle, got, a, offs)
    g.close()


def test_plan_reuse_smaller_second_batch(gmms, oracle, dev):
    print("segments plan reuse, workspace reuse", flush=True)
    ws = gmms.Workspace()
    import torch

    for lens, seed in (([100_000, 3, 50_000], 1), ([2000, 1], 2)):
        offs = _offsets(lens)
        a = oracle.gen_i32(seed, int(offs[-1]), 32)
        plan = gmms.SegPlan(gmms.INT32, offs)
        k = _to_dev(a, dev)
        ka = torch.empty_like(k)
        alt = plan.sort(k, ka, workspace=ws)
        torch.cuda.synchronize()
        _check_segments(oracle, (ka if alt else k).cpu().numpy(), a, offs)


def test_autotune_installs_valid_choices(gmms, oracle, dev):
    print("segments autotune classes 11..13", flush=True)
    kk0, mm0 = gmms.seg_tuning_get(gmms.INT32)
    try:
        res = gmms.seg_autotune(gmms.INT32, 11, 13, batch_elems=1 << 19, reps=1)
        kk, mm = gmms.seg_tuning_get(gmms.INT32)
        for c, (K, m, ms) in res.items():
            assert K == 16 and ms > 0
            assert kk[c] == K and mm[c] == m
        offs = _offsets([1500, 3000, 8000, 6000])
        a = oracle.gen_i32(3, int(offs[-1]), 32)
        got = gmms.segmented_sort(_to_dev(a, dev), offs).cpu().numpy()
        _check_segments(oracle, got, a, offs)
    finally:
        gmms.seg_tuning_set(gmms.INT32, kk0, mm0)
