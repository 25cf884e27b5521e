"""GPU tests of the sort server: every result is bit-identical alone and in a
mixed batch (and equal to the oracle's sort), the HBM budget bounds the
device memory in use, batches overlap, over-budget requests are refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pinned(a):
    import torch

    t = torch.from_numpy(np.ascontiguousarray(a))
    return t.pin_memory()


def _bits(x):
    return x.view({4: np.uint32, 8: np.uint64}[x.dtype.itemsize])


def _requests(oracle):
    reqs = []
    for i, n in enumerate([1, 1000, 4096, 70_001, 300_000, 1 << 20]):
        reqs.append(("i32", oracle.gen_i32(900 + i, n, 32), None))
    for i, n in enumerate([777, 65_536, 250_001]):
        x = oracle.gen_f32_normal(950 + i, n)
        x.view(np.uint32)[::61] = 0x80000000  # -0
        x.view(np.uint32)[1::61] = 0x7FC00000  # NaN
        reqs.append(("f32", x, None))
    for i, n in enumerate([5000, 131_073]):
        k, v = oracle.gen_pairs(970 + i, n, 3, 1.1, 1 << 20)
        reqs.append(("pair", k, v))
    return reqs


def _submit(srv, kind, k, v):
    import gmms

    if kind == "pair":
        return srv.submit(_pinned(k.view(np.int64)), _pinned(v.view(np.int32)), dtype=gmms.PAIR_U64_U32)
    return srv.submit(_pinned(k))


def _host(t):
    if isinstance(t, tuple):
        return t[0].numpy().view(np.uint64), t[1].numpy().view(np.uint32)
    return t.numpy()


def _prepared(gmms, kind, k, v):
    """(keys, values, dtype) as pinned host tensors, ready to submit."""
    if kind == "pair":
        return _pinned(k.view(np.int64)), _pinned(v.view(np.int32)), gmms.PAIR_U64_U32
    return _pinned(k), None, None


def test_results_same_alone_and_in_mix(gmms, oracle, dev):
    print("server: alone vs mix", flush=True)
    reqs = _requests(oracle)
    with gmms.Server(hbm_budget=1 << 30) as srv:
        alone = []
        for kind, k, v in reqs:
            t = _submit(srv, kind, k, v)
            alone.append(_host(srv.result(t)))
        assert srv.stats().failed == 0
    # all at once, queued behind a large request on a one-slot server: they
    # arrive while it runs, so same-dtype requests share a batch when it ends
    # (inputs and pinned result buffers made beforehand, so that the
    # submissions themselves take microseconds)
    import torch

    prep = [_prepared(gmms, kind, k, v) for kind, k, v in reqs]
    outs = [(torch.empty(pk.shape, dtype=pk.dtype, pin_memory=True),
             None if pv is None else torch.empty(pv.shape, dtype=pv.dtype, pin_memory=True)) for pk, pv, _ in prep]
    blocker = _pinned(oracle.gen_i32(999, 1 << 26, 32))
    with gmms.Server(hbm_budget=1 << 30, max_inflight=1) as srv:
        tb = srv.submit(blocker)
        ts = [srv.submit(pk, pv, dtype=dt, out_keys=ok, out_values=ov) for (pk, pv, dt), (ok, ov) in zip(prep, outs)]
        mixed = [_host(srv.result(t)) for t in ts]
        srv.result(tb)
        st = srv.stats()
        assert st.completed == len(reqs) + 1 and st.failed == 0
        assert st.max_batch_size >= 2
        assert len({t.batch for t in ts}) < len(ts)
    for (kind, k, v), a, m in zip(reqs, alone, mixed):
        if kind == "pair":
            wk, wv = oracle.sort_pairs(k, v)
            assert np.array_equal(a[0], wk) and np.array_equal(a[1], wv)
            assert np.array_equal(m[0], a[0]) and np.array_equal(m[1], a[1])
        else:
            w = oracle.sort(k)
            assert np.array_equal(_bits(a), _bits(w))
            assert np.array_equal(_bits(m), _bits(a))


def test_device_resident_requests(gmms, oracle, dev):
    """Requests whose buffers live on the device (copied device to device
    into the arena and out of it) give the same results as host requests,
    alone and in a mix with host requests."""
    import torch

    print("server: device-resident requests", flush=True)
    reqs = _requests(oracle)

    def to_dev(kind, k, v):
        if kind == "pair":
            return _pinned(k.view(np.int64)).to(dev), _pinned(v.view(np.int32)).to(dev)
        return _pinned(k).to(dev), None

    with gmms.Server(hbm_budget=1 << 30) as srv:
        ts = []
        for i, (kind, k, v) in enumerate(reqs):
            if i % 2:
                ts.append(_submit(srv, kind, k, v))
            else:
                dk, dv = to_dev(kind, k, v)
                ts.append(srv.submit(dk, dv, dtype=gmms.PAIR_U64_U32) if kind == "pair" else srv.submit(dk))
        outs = [srv.result(t) for t in ts]
        assert srv.stats().failed == 0
    for i, ((kind, k, v), o) in enumerate(zip(reqs, outs)):
        if kind == "pair":
            assert (o[0].is_cuda and o[1].is_cuda) == (i % 2 == 0)
            wk, wv = oracle.sort_pairs(k, v)
            assert np.array_equal(o[0].cpu().numpy().view(np.uint64), wk)
            assert np.array_equal(o[1].cpu().numpy().view(np.uint32), wv)
        else:
            assert o.is_cuda == (i % 2 == 0)
            assert np.array_equal(_bits(o.cpu().numpy()), _bits(oracle.sort(k)))


def test_caller_result_buffers(gmms, oracle, dev):
    """Results land in the output tensors the caller passes (device and
    pinned host); a mismatched output is refused before anything is queued."""
    import torch

    print("server: caller result buffers", flush=True)
    reqs = _requests(oracle)
    with gmms.Server(hbm_budget=1 << 30) as srv:
        ts, given = [], []
        for i, (kind, k, v) in enumerate(reqs):
            if kind == "pair":
                ik, iv = _pinned(k.view(np.int64)), _pinned(v.view(np.int32))
            else:
                ik, iv = _pinned(k), None
            if i % 2 == 0:
                ik = ik.to(dev)
                iv = None if iv is None else iv.to(dev)
            ok_ = torch.empty_like(ik) if ik.is_cuda else torch.empty(ik.shape, dtype=ik.dtype, pin_memory=True)
            ov_ = None
            if iv is not None:
                ov_ = torch.empty_like(iv) if iv.is_cuda else torch.empty(iv.shape, dtype=iv.dtype, pin_memory=True)
            dt = gmms.PAIR_U64_U32 if kind == "pair" else None
            ts.append(srv.submit(ik, iv, dtype=dt, out_keys=ok_, out_values=ov_))
            given.append((ok_, ov_))
        with pytest.raises(gmms.GmmsError):
            srv.submit(_pinned(reqs[0][1]), out_keys=torch.empty(3, dtype=torch.int64))
        outs = [srv.result(t) for t in ts]
        assert srv.stats().failed == 0
    for (kind, k, v), o, (gk, gv) in zip(reqs, outs, given):
        if kind == "pair":
            assert o[0] is gk and o[1] is gv
            wk, wv = oracle.sort_pairs(k, v)
            assert np.array_equal(gk.cpu().numpy().view(np.uint64), wk)
            assert np.array_equal(gv.cpu().numpy().view(np.uint32), wv)
        else:
            assert o is gk
            assert np.array_equal(_bits(gk.cpu().numpy()), _bits(oracle.sort(k)))


def test_budget_bounds_memory_and_batches_overlap(gmms, oracle, dev):
    print("server: 48 MB budget, 40 requests", flush=True)
    budget = 48 << 20
    n = 1 << 20  # 8 MiB of key buffers per request: at most 5 in flight
    with gmms.Server(hbm_budget=budget, max_batch_requests=2, max_inflight=3) as srv:
        srcs = [oracle.gen_i32(2000 + i, n, 32) for i in range(40)]
        ts = [srv.submit(_pinned(a)) for a in srcs]
        outs = [srv.result(t).numpy() for t in ts]
        st = srv.stats()
        assert st.arena_bytes == budget and st.budget == budget
        assert 0 < st.peak_reserved_bytes <= budget
        assert st.reserved_bytes == 0
        assert st.batches >= 20 and st.max_batch_size <= 2
        assert st.max_inflight_seen >= 2
        assert st.keys_sorted == 40 * n
    for a, o in zip(srcs[:5], outs[:5]):
        assert np.array_equal(o, oracle.sort(a))
    for o in outs:
        assert (np.diff(o.astype(np.int64)) >= 0).all()


def test_over_budget_request_is_refused(gmms, dev):
    import torch

    print("server: refuse over-budget", flush=True)
    with gmms.Server(hbm_budget=16 << 20) as srv:
        with pytest.raises(gmms.GmmsError) as e:
            srv.submit(torch.zeros(4 << 20, dtype=torch.int32))  # 32 MiB of key buffers
        assert e.value.status == gmms.ERROR_BUDGET
        t = srv.submit(torch.arange(1000, 0, -1, dtype=torch.int32).pin_memory())
        assert np.array_equal(srv.result(t).numpy(), np.arange(1, 1001, dtype=np.int32))
        t0 = srv.submit(torch.zeros(0, dtype=torch.int32))
        assert srv.wait(t0) and t0.status == 0
        st = srv.stats()
        assert st.rejected == 1 and st.completed == 2


def test_wait_timeout_and_drain(gmms, oracle, dev):
    print("server: wait with timeout, drain", flush=True)
    with gmms.Server(hbm_budget=256 << 20) as srv:
        ts = [srv.submit(_pinned(oracle.gen_i32(3000 + i, 1 << 18, 32))) for i in range(8)]
        srv.drain()
        for t in ts:
            assert srv.wait(t, timeout=0.0) and t.status == 0 and t.latency_ms > 0
        assert srv.stats().completed == 8


def test_device_inputs_ordered_after_current_stream(gmms, oracle, dev):
    """A device request whose input is still being written by work queued on
    torch's current stream (here the contiguous copy of a strided view,
    queued behind a few milliseconds of other kernels) is read only after
    that work: the server's copies wait for the stream."""
    import torch

    print("server: device input produced by queued work", flush=True)
    n = 1 << 21
    a = oracle.gen_i32(4000, n, 32)
    base = torch.from_numpy(a).to(dev)
    torch.cuda.synchronize()
    with gmms.Server(hbm_budget=256 << 20) as srv:
        work = torch.empty(1 << 27, device=dev)
        for _ in range(8):
            work.normal_()  # queued ahead of the producer below
        x = base.flip(0)  # written by a kernel still queued behind the chain
        t = srv.submit(x[::2])  # strided: submit copies it on the current stream
        got = srv.result(t).cpu().numpy()
    assert np.array_equal(got, oracle.sort(a[::-1][::2].copy()))


def test_dropped_tickets_keep_buffers_alive(gmms, oracle, dev):
    """Requests whose tickets the caller drops keep their buffers alive
    until they finish: tensors allocated meanwhile are never written by the
    server, and drain() reports every request done."""
    import gc

    import torch

    print("server: dropped tickets", flush=True)
    n = 1 << 20
    srcs = [oracle.gen_i32(5000 + i, n, 32) for i in range(12)]
    with gmms.Server(hbm_budget=256 << 20, max_inflight=2) as srv:
        for a in srcs:
            srv.submit(torch.from_numpy(a).to(dev)[::1])  # device input and result: ticket dropped
        gc.collect()
        fills = [torch.full((n,), 12345 + i, dtype=torch.int32, device=dev) for i in range(12)]
        srv.drain()
        torch.cuda.synchronize()
        for i, f in enumerate(fills):
            assert bool((f == 12345 + i).all())
        st = srv.stats()
        assert st.completed == len(srcs) and st.failed == 0
        assert not srv._live
