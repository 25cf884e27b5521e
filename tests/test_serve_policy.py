"""Host logic of the sort server (no GPU): the request footprint, the batching
policy, and a discrete-event simulation of the admission loop showing that
the HBM budget is never exceeded, every request is served, and the oldest
pending request is always in the next batch."""
import heapq

import numpy as np
import pytest


def _al(x):
    return (x + 255) // 256 * 256


@pytest.mark.parametrize("n", [0, 1, 1000, 1 << 20, 12345677])
def test_request_bytes(gmms, n):
    assert gmms.server_request_bytes(gmms.INT32, n) == 2 * _al(4 * n)
    assert gmms.server_request_bytes(gmms.FLOAT64, n) == 2 * _al(8 * n)
    assert gmms.server_request_bytes(gmms.PAIR_U64_U32, n) == 2 * _al(8 * n) + 2 * _al(4 * n)


def test_policy_takes_head_then_same_dtype_in_order(gmms):
    d = [0, 2, 0, 0, 6, 0]
    n = [100, 5, 10**6, 10, 7, 1]
    assert gmms.server_batch_policy(d, n, 10**7) == [0, 2, 3, 5]
    # a different head dtype batches its own kind
    assert gmms.server_batch_policy([6, 0, 6], [10, 10, 10], 10**7) == [0, 2]


def test_policy_head_must_fit(gmms):
    # the oldest request does not fit: nothing is admitted, younger ones wait
    assert gmms.server_batch_policy([0, 0], [10**7, 10], 10**7) == []


def test_policy_stops_at_first_misfit(gmms):
    b = gmms.server_request_bytes(gmms.INT32, 1000)
    # room for two: the third same-dtype request stops the batch, the fourth is not skipped ahead
    assert gmms.server_batch_policy([0, 0, 0, 0], [1000, 1000, 10**6, 1000], 2 * b + 100) == [0, 1]


def test_policy_limits(gmms):
    d, n = [0] * 10, [1000] * 10
    assert gmms.server_batch_policy(d, n, 10**9, max_batch_requests=3) == [0, 1, 2]
    assert gmms.server_batch_policy(d, n, 10**9, max_batch_elems=4500) == [0, 1, 2, 3]
    assert gmms.server_batch_policy([], [], 10**9) == []


def test_simulated_admission_never_exceeds_budget(gmms):
    """Poisson arrivals, sizes log-uniform in [2^10, 2^26], dtypes 60% int32 /
    25% float32 / 15% pairs; up to 3 batches in flight; a batch takes time
    proportional to its elements. The admission loop mirrors the server's:
    retire finished batches, then admit while a slot is free and the policy
    returns a batch for the free bytes."""
    rng = np.random.default_rng(11)
    budget = 2 * 10**9
    nreq = 600
    gaps = rng.exponential(0.002, nreq)
    arrivals = np.cumsum(gaps)
    sizes = np.floor(2.0 ** rng.uniform(10, 26, nreq)).astype(np.int64)
    dts = rng.choice([gmms.INT32, gmms.FLOAT32, gmms.PAIR_U64_U32], nreq, p=[0.6, 0.25, 0.15])
    need = [gmms.server_request_bytes(int(d), int(s)) for d, s in zip(dts, sizes)]
    assert max(need) <= budget
    t, i = 0.0, 0
    pending, inflight, done = [], [], set()
    reserved, peak, slots = 0, 0, 3
    heads = []
    while len(done) < nreq:
        # next event: an arrival or a batch completion
        t_arr = arrivals[i] if i < nreq else np.inf
        t_fin = inflight[0][0] if inflight else np.inf
        t = min(t_arr, t_fin)
        if t == t_arr:
            pending.append(i)
            i += 1
        else:
            _, bid, members, nbytes = heapq.heappop(inflight)
            reserved -= nbytes
            done.update(members)
        while len(inflight) < slots and pending:
            ch = gmms.server_batch_policy([int(dts[k]) for k in pending], [int(sizes[k]) for k in pending],
                                          budget - reserved)
            if not ch:
                break
            members = [pending[c] for c in ch]
            heads.append(members[0] == min(pending))
            nbytes = sum(need[m] for m in members)
            reserved += nbytes
            peak = max(peak, reserved)
            assert reserved <= budget
            pending = [k for k in pending if k not in set(members)]
            dur = sum(int(sizes[m]) for m in members) / 20e9 + 20e-6
            heapq.heappush(inflight, (t + dur, len(heads), members, nbytes))
    assert all(heads)                 # the oldest pending request is always served next
    assert peak <= budget and peak > budget // 4  # the budget was actually the constraint at times
