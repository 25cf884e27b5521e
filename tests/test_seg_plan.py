"""Host logic of the batched segmented sort (no GPU): the planner's round
alignment, tile and merger counts against the reference schedule's closed
forms (src/multimergesort.hxx:87-90: rounds = ceil(log_K(n/1024)),
tasks = ceil(n / (K*listSize))), the tuning table and argument checks."""
import numpy as np
import pytest


def _offsets(lens, start=0):
    return np.concatenate([[start], start + np.cumsum(np.asarray(lens, np.int64))]).astype(np.int64)


def _run(n):
    # segments longer than one 1024-element tile start from 4096-element runs
    return 4096 if n > 1024 else 1024


def _rounds(n, K):
    r, L = 0, _run(n)
    while L < n:
        r, L = r + 1, L * K
    return r


LENS = [0, 1, 1023, 1024, 1025, 5000, 70_000, 1 << 20, 3, (1 << 22) + 17]


@pytest.mark.parametrize("K", [4, 16, 64])
def test_plan_matches_schedule_closed_forms(gmms, K):
    offs = _offsets(LENS)
    # mergers=1: every task is one merger, so merge items = tasks over all rounds
    plan = gmms.SegPlan(gmms.INT32, offs, gmms.make_params(K=K, mergers=1))
    info = plan.info
    assert info.n_seg == len(LENS)
    assert info.elements == sum(LENS) == info.extent
    assert info.tiles == sum((n + 1023) // 1024 for n in LENS)
    assert info.rounds == max(_rounds(n, K) for n in LENS)
    tasks = 0
    for n in LENS:
        L = _run(n)
        for _ in range(_rounds(n, K)):
            tasks += -(-n // (K * L))
            L *= K
    assert info.merge_items == tasks
    assert info.partition_items == 0
    # setup, one base-case launch per run length present, one merge per round
    assert info.launches == 1 + 2 + info.rounds
    for s, n in enumerate(LENS):
        assert plan.segment(s) == (K, _rounds(n, K))


def test_split_tasks_get_partition_items(gmms):
    offs = _offsets([1 << 22, 5000])
    plan = gmms.SegPlan(gmms.INT32, offs, gmms.make_params(K=16, mergers=512))
    info = plan.info
    # round volumes are 2^22+5000, then 2^22 twice: each is split over ~512 mergers
    assert info.merge_items >= 3 * 400
    assert info.partition_items == info.merge_items - sum(
        -(-n // (16 * 4096 * 16 ** r)) for n in (1 << 22, 5000) for r in range(_rounds(n, 16)))
    assert info.launches == 2 + 2 * info.rounds


def test_empty_batch_and_empty_segments(gmms):
    plan = gmms.SegPlan(gmms.INT32, [0])
    assert plan.info.rounds == 0 and plan.info.launches == 0 and plan.info.elements == 0
    plan = gmms.SegPlan(gmms.INT32, [5, 5, 5])
    assert plan.info.n_seg == 2 and plan.info.tiles == 0 and plan.info.launches == 0


def test_pairs_k_limit(gmms):
    # 64-bit and pair elements use 512-element windows: at most 32 lists
    assert gmms.SegPlan(gmms.PAIR_U64_U32, _offsets([1 << 20]), gmms.make_params(K=32)).segment(0) == (32, 2)
    with pytest.raises(gmms.GmmsError) as e:
        gmms.SegPlan(gmms.PAIR_U64_U32, _offsets([1 << 20]), gmms.make_params(K=64))
    assert e.value.status == gmms.ERROR_UNSUPPORTED


def test_invalid_arguments(gmms):
    with pytest.raises(gmms.GmmsError):
        gmms.SegPlan(gmms.INT32, [0, 10, 5])          # decreasing offsets
    with pytest.raises(gmms.GmmsError):
        gmms.SegPlan(gmms.INT32, [-1, 10])            # negative start
    with pytest.raises(gmms.GmmsError):
        gmms.SegPlan(99, [0, 10])                     # bad dtype
    with pytest.raises(gmms.GmmsError) as e:
        gmms.SegPlan(gmms.INT32, [0, 10], gmms.make_params(pipeline=1))
    assert e.value.status == gmms.ERROR_UNSUPPORTED


def test_tuning_table_roundtrip_and_validation(gmms):
    kk0, mm0 = gmms.seg_tuning_get(gmms.FLOAT32)
    try:
        kk, mm = list(kk0), list(mm0)
        kk[14], mm[14] = 8, 12345
        gmms.seg_tuning_set(gmms.FLOAT32, kk, mm)
        assert gmms.seg_tuning_get(gmms.FLOAT32) == (kk, mm)
        plan = gmms.SegPlan(gmms.FLOAT32, _offsets([10_000, 20_000]))
        assert plan.segment(0) == (8, 1)      # class 14: 8193..16384
        assert plan.segment(1) == (16, 1)     # class 15: default K
        for bad in (3, 128, -2):
            kb = list(kk)
            kb[5] = bad
            with pytest.raises(gmms.GmmsError):
                gmms.seg_tuning_set(gmms.FLOAT32, kb, mm)
        with pytest.raises(gmms.GmmsError):
            gmms.seg_tuning_set(gmms.UINT64, [64] * gmms.SEG_CLASSES, [0] * gmms.SEG_CLASSES)  # > 32 lists
    finally:
        gmms.seg_tuning_set(gmms.FLOAT32, kk0, mm0)
