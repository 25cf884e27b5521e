#!/usr/bin/env python3
"""Benchmark: plain sort of N = 2^28 int32 keys, uniform over the full int32
range, on one MI355X through libgmms.so; prints one JSON line.

A step is one complete sort (base case + every merge round) of one resident
2^28-key array; the timed region holds exactly --steps sorts back to back,
each on its own pre-generated input (input i = input 0 XOR a per-step
constant, so no step sorts sorted data), with no allocation and no copies.

Reported:
  value      keys sorted per second over the timed region (GKeys/s)
  roofline   the dominant kernel (the K-way merge round): algorithmic bytes
             2*N*4 per launch over its average launch time, measured with HIP
             events recorded on the launch stream during the timed region;
             peak = MI355X's published 8 TB/s. traffic = calibrated HBM bytes
             per merge launch from rocprofv3 FETCH_SIZE/WRITE_SIZE
             (profiles/merge_traffic.json), if recorded for this workload.
  sort_roofline  the whole sort: the bytes of the passes it makes
             (2*N*4 per pass: base case + each merge round) over the step
             time, against 8 TB/s and against the best measured copy; the
             reference's byte formula 2*N*4*(1+ceil(log_K(N/1024))) (for its
             1024-element base runs) is reported beside it (it credits a pass
             that is not made when 4096-element base runs save a round).
  copy_probe every variant of the stream-copy probe (16-byte loads, 1-8 in
             flight per lane, plain or non-temporal stores, grid sizes) and
             torch's own device copy, 4 GiB each; the best is the copy peak.
  cpu_baseline  libstdc++ parallel sort (all host threads) of the same 2^28
             keys, median of 5 after a warm-up; single-thread std::sort of
             the same 2^28 keys (one run: the bounded sample); both sorts at
             N = 2^20 (config 1), median of 5 after a warm-up; the host's CPU
             model, nproc and hardware threads.
  configs    the other plain-sort configurations (2^20, 10^8+17, [0,1024),
             float32 total order, 2^27 pairs under four key distributions,
             uint64): HIP-event time of the sort on resident buffers (median
             of 5 after a warm-up), passes-made bytes over it against 8 TB/s,
             and a bit-exact check against the parallel CPU sort.
  batched    the batched-segment workload: int32 segments with lengths
             log-uniform in [2^10, 2^22] (mt19937_64 seed 0x6D6D7303), values
             in [0, 2^31), cut at the first prefix reaching 2^28 keys (or all
             8192 segments with --batched-literal). The autotuner picks K and
             merger sizes per size class first; then the batch is sorted as
             one HIP graph (and eagerly through gmms_seg_sort) and, for
             comparison, with one plain sort call per segment. Timed with HIP
             events on the launch stream; inputs are restored untimed.
  serving    the sort server (64 GB budget) on a mix of requests with sizes
             log-uniform in [2^10, 2^26] keys, dtypes 60% int32 / 25% float32
             (total order) / 15% (uint64 key, uint32 value) pairs: the
             saturated rate (all requests at once), then Poisson arrivals at
             25/50/75/90 % of it (p50/p99 latency, keys/s, requests per
             batch, per run), and one run with pinned host buffers.
The GPU output of input 0 is compared bit-exact with the CPU sort; the
batched output equals the per-segment output and every segment is compared
with the CPU sort of that segment; every config and served result is checked.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-mms_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

PUBLISHED_HBM_PEAK_GBPS = 8000.0
METRIC = "GKeys/s + % HBM roofline at 2^28 int32; batched-segment vs per-array launch"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1 << 28)
    ap.add_argument("--K", type=int, default=0, help="lists per merge round (0: auto)")
    ap.add_argument("--mergers", type=int, default=0)
    ap.add_argument("--prefetch", type=int, default=0)
    ap.add_argument("--kernel", type=int, default=None,
                    help="merge kernel: 0 wavefront tile (default), 1 minBlockHeap, 2 workgroup tile")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", choices=("sweep", "plain"), default="sweep",
                    help="plain: only gmms_copy (the calibration kernel of the traffic counters)")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-literal", action="store_true",
                    help="all 8192 segments (about 4e9 keys) instead of the 2^28-key prefix")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the other plain-sort configurations (dtypes, pad and duplicate cases)")
    ap.add_argument("--serve-requests", type=int, default=160)
    ap.add_argument("--serve-loads", type=str, default="0.25,0.5,0.75,0.9",
                    help="offered loads as fractions of the saturated request rate")
    ap.add_argument("--serve-arrivals", type=int, default=320, help="arrivals per offered load")
    return ap.parse_args()


def log(msg: str) -> None:
    """Progress of the run on stderr (the JSON line alone goes to stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traffic_from_profiles(kernel: int, K: int, n: int):
    """HBM bytes per merge launch of this workload, measured by rocprofv3
    FETCH_SIZE/WRITE_SIZE and calibrated with the copy kernel
    (profiles/merge_traffic.json); None when no profile matches."""
    p = os.path.join(ROOT, "profiles", "merge_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        d = json.load(f)
    if d.get("pipeline") != kernel or d.get("K") != K or d.get("n") != n:
        return None
    return d.get("hbm_bytes_per_merge_launch")


def batched(args, oracle, dev, stream, threads):
    """Batched segmented sort vs one plain sort per segment (same keys)."""
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np
    import torch

    import gmms

    seed = 0x6D6D7303
    lens = oracle.gen_seglens(seed, 8192, 10, 22)
    if not args.batched_literal:
        csum = np.cumsum(lens)
        cut = int(np.searchsorted(csum, 1 << 28)) + 1  # first prefix whose sum reaches 2^28
        lens = lens[:cut]
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    total = int(offs[-1])
    host = oracle.gen_i32(seed, total, 33)  # (int32)(u64 >> 33): [0, 2^31)
    src = torch.from_numpy(host).to(dev)
    k = torch.empty_like(src)
    ka = torch.empty_like(src)

    log(f"batched: {len(lens)} segments, {total} keys generated")
    tuned = None
    if not args.no_autotune:
        res = gmms.seg_autotune(gmms.INT32, 10, 22, batch_elems=1 << 24, reps=2, stream=stream)
        tuned = {str(c): {"K": K, "merger_elems": m, "ms_per_2^24_batch": round(ms, 4)} for c, (K, m, ms) in res.items()}
        log("batched: autotune done")
    plan = gmms.SegPlan(gmms.INT32, offs)
    info = plan.info
    ws = gmms.Workspace()
    graph = plan.graph(k, ka)
    reps = max(1, args.steps // 2)

    def timed(fn):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        k.copy_(src)
        torch.cuda.synchronize()
        e0.record(stream)
        alt = fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1), alt

    # warm-up of every path
    timed(lambda: graph.launch(stream))
    timed(lambda: plan.sort(k, ka, workspace=ws, stream=stream))
    g_ms = min(timed(lambda: graph.launch(stream))[0] for _ in range(reps))
    e_ms = min(timed(lambda: plan.sort(k, ka, workspace=ws, stream=stream))[0] for _ in range(reps))
    k.copy_(src)
    alt = graph.launch(stream)  # the batched result checked below
    torch.cuda.synchronize()
    batched_out = (ka if alt else k).clone()

    # one plain sort call per segment, on the same buffers
    per_ws = gmms.Workspace()
    wmax = max(gmms.workspace_size(gmms.INT32, int(n)) for n in lens)
    per_ws.ensure(max(wmax, 1), dev)
    per_out = torch.empty_like(src)

    # the loop calls the C entry point directly (no per-call Python wrapper work)
    import ctypes

    L = gmms.lib()
    wp, wb = per_ws.tensor.data_ptr(), per_ws.tensor.numel()
    sp = stream.cuda_stream
    kp, kap = k.data_ptr(), ka.data_ptr()
    seg_args = [(int(offs[s]) * 4, int(offs[s + 1] - offs[s])) for s in range(len(lens))]
    alt_c = ctypes.c_int(0)

    def per_array():
        for s, (byte_off, n) in enumerate(seg_args):
            st = L.gmms_sort(gmms.INT32, kp + byte_off, kap + byte_off, None, None, n, None, wp, wb, sp,
                             ctypes.byref(alt_c))
            if st:
                raise gmms.GmmsError(st, "gmms_sort")
            per_alt[s] = alt_c.value
        return None

    log(f"batched: graph {g_ms:.3f} ms, eager {e_ms:.3f} ms")
    per_alt = [0] * len(lens)
    timed(per_array)
    p_ms = min(timed(per_array)[0] for _ in range(max(1, reps // 2)))
    for s in range(len(lens)):
        a, b = int(offs[s]), int(offs[s + 1])
        per_out[a:b] = (ka if per_alt[s] else k)[a:b]
    torch.cuda.synchronize()
    same = bool(torch.equal(per_out, batched_out))

    log(f"batched: per-array {p_ms:.3f} ms")
    # every segment against the CPU sort of that segment
    got = batched_out.cpu().numpy()
    want = host.copy()

    def cpu_sort(s):
        a, b = int(offs[s]), int(offs[s + 1])
        oracle.lib.oracle_sort_i32(want[a:b].ctypes.data, b - a)

    with ThreadPoolExecutor(max_workers=max(1, min(16, threads))) as ex:
        list(ex.map(cpu_sort, range(len(lens))))
    exact = bool(np.array_equal(got, want))
    rounds_per = [plan.segment(s)[1] for s in range(len(lens))]
    per_launches = sum(1 + 2 * r for r in rounds_per)  # at most: base case + (partition + merge) per round
    formula_bytes = float(sum(2.0 * int(n) * 4 * (1 + r) for n, r in zip(lens, rounds_per)))
    graph.close()
    return {
        "workload": ("batched segments: %d int32 arrays, lengths log-uniform in [2^10, 2^22], "
                     "values in [0, 2^31)%s" % (len(lens), "" if args.batched_literal else
                                                 ", first prefix reaching 2^28 keys of the 8192-segment draw")),
        "segments": len(lens),
        "keys": total,
        "rounds": info.rounds,
        "launches_batched": info.launches,
        "launches_per_array_max": per_launches,
        "graph_ms": round(g_ms, 4),
        "eager_ms": round(e_ms, 4),
        "per_array_ms": round(p_ms, 4),
        "batched_gkeys_s": round(total / (g_ms * 1e-3) / 1e9, 4),
        "batched_eager_gkeys_s": round(total / (e_ms * 1e-3) / 1e9, 4),
        "per_array_gkeys_s": round(total / (p_ms * 1e-3) / 1e9, 4),
        "speedup_vs_per_array": round(p_ms / g_ms, 3),
        "hbm_formula_GBps": round(formula_bytes / (g_ms * 1e-3) / 1e9, 1),
        "bit_exact_vs_std_sort": exact,
        "identical_to_per_array": same,
        "autotune": tuned,
    }


def serving(args, oracle, threads):
    """The sort server on a fixed request mix: the saturated rate, then
    Poisson arrivals at fractions of it; every result checked."""
    from collections import Counter
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np
    import torch

    import gmms

    rng = np.random.default_rng(0x6D6D7305)
    nreq = args.serve_requests
    sizes = np.floor(2.0 ** rng.uniform(10, 26, nreq)).astype(np.int64)
    kinds = rng.choice(3, nreq, p=[0.60, 0.25, 0.15])  # int32, float32, pairs
    reqs = []
    for i in range(nreq):
        n = int(sizes[i])
        if kinds[i] == 0:
            k = torch.from_numpy(rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32)).pin_memory()
            reqs.append((k, None))
        elif kinds[i] == 1:
            bits = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
            k = torch.from_numpy(bits.view(np.float32)).pin_memory()  # every float32 bit pattern
            reqs.append((k, None))
        else:
            k = torch.from_numpy(rng.integers(0, 2**63, n, dtype=np.int64)).pin_memory()
            v = torch.from_numpy(np.arange(n, dtype=np.int32)).pin_memory()
            reqs.append((k, v))
    total = int(sizes.sum())
    log(f"serving: {nreq} requests generated ({total} keys)")

    # expected results (CPU sorts), kept on the device for the checks
    def expect(i):
        k, v = reqs[i]
        if v is None:
            return oracle.sort(k.numpy()), None
        wk, wv = oracle.sort_pairs(k.numpy().view(np.uint64), v.numpy().view(np.uint32))
        return wk.view(np.int64), wv.view(np.int32)

    with ThreadPoolExecutor(max_workers=max(1, min(16, threads))) as ex:
        want_host = list(ex.map(expect, range(nreq)))
    want = [(torch.from_numpy(wk).to("cuda:0"), None if wv is None else torch.from_numpy(wv).to("cuda:0"))
            for wk, wv in want_host]
    del want_host
    # the same requests resident on the device (the server copies them into
    # its arena and the results out of it, device to device)
    dreqs = [(k.to("cuda:0"), None if v is None else v.to("cuda:0")) for k, v in reqs]
    torch.cuda.synchronize()
    srv = gmms.Server()  # 64 GB budget

    def submit(k, v, outs=(None, None)):
        if v is None:
            return srv.submit(k, out_keys=outs[0])
        return srv.submit(k, v, dtype=gmms.PAIR_U64_U32, out_keys=outs[0], out_values=outs[1])

    def same(t, i):
        wk, wv = want[i % nreq]
        got_k = t.out_keys if t.out_keys.is_cuda else t.out_keys.to("cuda:0")
        if got_k.dtype == torch.float32:
            got_k, wk = got_k.view(torch.int32), wk.view(torch.int32)  # bits: NaNs compare equal to themselves
        if not torch.equal(got_k, wk):
            return False
        if wv is None:
            return True
        got_v = t.out_values if t.out_values.is_cuda else t.out_values.to("cuda:0")
        return bool(torch.equal(got_v, wv))

    def batching(ts):
        per = Counter(t.batch for t in ts)
        return {"batches": len(per), "requests_per_batch": round(len(ts) / max(1, len(per)), 3),
                "max_requests_in_a_batch": max(per.values()) if per else 0}

    # saturation: every request submitted at once (closed batch) into result
    # buffers that exist beforehand, as in a serving loop that reuses them
    outs = [(torch.empty_like(k), None if v is None else torch.empty_like(v)) for k, v in dreqs]
    # one untimed pass first: a server's first pass over this mix measured
    # 58 ms against 37-39 ms for the passes after it
    for t in [submit(k, v, o) for (k, v), o in zip(dreqs, outs)]:
        srv.wait(t)
    # the timed pass must write every result itself: fill the buffers first
    for ok_, ov_ in outs:
        ok_.view(torch.uint8).fill_(0xA5)
        if ov_ is not None:
            ov_.view(torch.uint8).fill_(0xA5)
    torch.cuda.synchronize()
    ts0 = time.perf_counter()
    sat = [submit(k, v, o) for (k, v), o in zip(dreqs, outs)]
    srv.drain()
    sat_wall = time.perf_counter() - ts0
    sat_lat = np.array([t.latency_ms for t in sat])
    ok = all(t.status == 0 for t in sat) and all(same(t, i) for i, t in enumerate(sat))
    sat_rate = nreq / sat_wall
    saturated = {"gkeys_s": round(total / sat_wall / 1e9, 4), "requests_per_s": round(sat_rate, 1),
                 "wall_s": round(sat_wall, 4), "p50_ms": round(float(np.percentile(sat_lat, 50)), 3),
                 "p99_ms": round(float(np.percentile(sat_lat, 99)), 3), **batching(sat),
                 "note": "all requests submitted at once into result buffers allocated (and overwritten) "
                         "beforehand; every result checked"}
    del outs
    log(f"serving: saturated {sat_rate:.0f} requests/s, {total / sat_wall / 1e9:.2f} GKeys/s")

    def poisson(rs, rate, arrivals, seed):
        """Poisson arrivals at `rate` of requests rs (cyclically); tickets,
        wall time from the first arrival to the last result, submit span."""
        gaps = np.random.default_rng(seed).exponential(1.0 / rate, arrivals)
        tickets = []
        t0 = time.perf_counter()
        due = t0
        for i in range(arrivals):
            due += gaps[i]
            while True:
                now = time.perf_counter()
                if now >= due:
                    break
                if due - now > 0.0005:
                    time.sleep(min(due - now - 0.0003, 0.001))
            k, v = rs[i % nreq]
            tickets.append(submit(k, v))
        t_sub = time.perf_counter()
        srv.drain()
        return tickets, time.perf_counter() - t0, t_sub - t0

    loads = []
    arrivals = args.serve_arrivals
    keys_run = int(sum(int(sizes[i % nreq]) for i in range(arrivals)))
    for j, frac in enumerate(float(x) for x in args.serve_loads.split(",") if x):
        rate = frac * sat_rate
        tickets, wall, span = poisson(dreqs, rate, arrivals, 0x5E00 + j)
        lat = np.array([t.latency_ms for t in tickets])
        good = all(t.status == 0 for t in tickets) and all(same(t, i) for i, t in enumerate(tickets))
        ok = ok and good
        loads.append({"load": frac, "offered_requests_per_s": round(rate, 1),
                      "offered_gkeys_s": round(rate * total / nreq / 1e9, 4),
                      "achieved_gkeys_s": round(keys_run / wall / 1e9, 4),
                      "p50_ms": round(float(np.percentile(lat, 50)), 3),
                      "p99_ms": round(float(np.percentile(lat, 99)), 3),
                      "mean_ms": round(float(lat.mean()), 3), "max_ms": round(float(lat.max()), 3),
                      "wall_s": round(wall, 4), "submit_span_s": round(span, 4), **batching(tickets),
                      "bit_exact": bool(good)})
        del tickets
        log(f"serving: load {frac:.2f}: p50 {loads[-1]['p50_ms']} ms, p99 {loads[-1]['p99_ms']} ms, "
            f"{loads[-1]['requests_per_batch']} requests per batch")
    # the same arrivals at half load with pinned host buffers (PCIe both ways)
    rate = 0.5 * sat_rate
    htickets, hwall, _ = poisson(reqs, rate, nreq, 0x5EFF)
    hlat = np.array([t.latency_ms for t in htickets])
    hgood = all(t.status == 0 for t in htickets) and all(same(t, i) for i, t in enumerate(htickets))
    ok = ok and hgood
    st = srv.stats()
    srv.close()
    return {
        "workload": ("sort server: %d requests, sizes log-uniform in [2^10, 2^26], dtypes 60%% int32 / "
                     "25%% float32 / 15%% (u64,u32) pairs, 64 GB budget, request buffers resident on the "
                     "device; Poisson arrivals at fractions of the saturated rate, %d arrivals each"
                     % (nreq, arrivals)),
        "requests": nreq,
        "keys": total,
        "saturated": saturated,
        "loads": loads,
        "host_buffers": {"load": 0.5, "offered_requests_per_s": round(rate, 1),
                         "p50_ms": round(float(np.percentile(hlat, 50)), 3),
                         "p99_ms": round(float(np.percentile(hlat, 99)), 3),
                         "max_ms": round(float(hlat.max()), 3),
                         "achieved_gkeys_s": round(total / hwall / 1e9, 4), **batching(htickets),
                         "note": "the same mix at half the saturated rate with pinned host buffers: the "
                                 "copies in and out cross PCIe"},
        "budget_bytes": st.budget,
        "peak_reserved_bytes": st.peak_reserved_bytes,
        "bit_exact_vs_std_sort": bool(ok),
        "all_status_ok": st.failed == 0,
    }


def other_configs(oracle, dev, stream):
    """The other plain-sort configurations: each input from the oracle's
    generators, sorted on resident buffers (HIP events on the stream around
    the sort alone; median of 5 after a warm-up, inputs restored untimed),
    every output checked bit-exact against the parallel CPU sort."""
    import numpy as np
    import torch

    import gmms

    ws = gmms.Workspace()

    def timed(keys, vals=None, dt=None):
        ka = torch.empty_like(keys)
        va = torch.empty_like(vals) if vals is not None else None
        k = torch.empty_like(keys)
        v = torch.empty_like(vals) if vals is not None else None
        ts, alt = [], 0
        for it in range(6):
            k.copy_(keys)
            if vals is not None:
                v.copy_(vals)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            alt = gmms.sort_inplace(k, ka, v, va, dtype=dt, workspace=ws, stream=stream)
            e1.record(stream)
            torch.cuda.synchronize()
            if it:
                ts.append(e0.elapsed_time(e1))
        return float(np.median(ts)), min(ts), (ka if alt else k), ((va if alt else v) if vals is not None else None)

    out = []

    def record(name, dt, n, t, tmin, ok):
        kb, vb = {gmms.INT32: (4, 0), gmms.FLOAT32: (4, 0), gmms.UINT64: (8, 0), gmms.PAIR_U64_U32: (8, 4)}[dt]
        rounds = gmms.sort_rounds(dt, n)
        pass_bytes = 2.0 * n * (kb + vb) * (1 + rounds)
        out.append({"workload": name, "n": n, "ms": round(t, 4), "ms_min": round(tmin, 4),
                    "gkeys_s": round(n / (t * 1e-3) / 1e9, 3), "passes_made": 1 + rounds,
                    "pass_GBps": round(pass_bytes / (t * 1e-3) / 1e9, 1),
                    "frac_of_8TBps": round(pass_bytes / (t * 1e-3) / 1e9 / PUBLISHED_HBM_PEAK_GBPS, 4),
                    "bit_exact_vs_std_sort": bool(ok)})
        log(f"configs: {name}: {t:.3f} ms ok={ok}")

    def int32_case(name, seed, n, shift):
        a = oracle.gen_i32(seed, n, shift)
        t, tmin, r, _ = timed(torch.from_numpy(a).to(dev))
        record(name, gmms.INT32, n, t, tmin, np.array_equal(r.cpu().numpy(), oracle.psort(a)))

    int32_case("int32 N=2^20 uniform in [0, 2^30) (config 1)", 0x6D6D7300, 1 << 20, 34)
    int32_case("int32 N=10^8+17 uniform in [0, 2^30)", 0x6D6D7302, 10**8 + 17, 34)
    int32_case("int32 N=2^28 uniform in [0, 1024)", 0x6D6D7302, 1 << 28, 54)
    n = 1 << 28
    x = oracle.gen_f32_normal(0x6D6D7304, n)
    t, tmin, r, _ = timed(torch.from_numpy(x).to(dev))
    record("float32 N=2^28 normal(0,1), total order", gmms.FLOAT32, n, t, tmin,
           np.array_equal(r.cpu().numpy().view(np.uint32), oracle.psort(x).view(np.uint32)))
    del x
    n = 1 << 27
    for dist, name in enumerate(("uniform", "sorted", "reverse sorted", "Zipf(1.1) over 2^20")):
        k, v = oracle.gen_pairs(0x6D6D7310 + dist, n, dist, 1.1, 1 << 20)
        t, tmin, rk, rv = timed(torch.from_numpy(k.view(np.int64)).to(dev), torch.from_numpy(v.view(np.int32)).to(dev),
                                gmms.PAIR_U64_U32)
        wk, wv = oracle.psort_pairs(k, v)
        ok = np.array_equal(rk.cpu().numpy().view(np.uint64), wk) and np.array_equal(rv.cpu().numpy().view(np.uint32), wv)
        record(f"(uint64 key, uint32 value=index) N=2^27, keys {name}", gmms.PAIR_U64_U32, n, t, tmin, ok)
        del k, v, wk, wv
    u = oracle.gen_u64(77, n)
    t, tmin, r, _ = timed(torch.from_numpy(u.view(np.int64)).to(dev), dt=gmms.UINT64)
    record("uint64 N=2^27 uniform", gmms.UINT64, n, t, tmin, np.array_equal(r.cpu().numpy().view(np.uint64), oracle.psort(u)))
    return out


def copy_probe(dev, stream, sweep=True):
    """Every variant of the stream-copy probe and torch's device copy over
    4 GiB (best of 5 each, HIP events on the stream)."""
    import torch

    import gmms

    cp_n = (4 << 30) // 4
    a = torch.empty(cp_n, dtype=torch.int32, device=dev)
    a.random_()
    b = torch.empty_like(a)

    def best(fn):
        t = 1e9
        for _ in range(5):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            fn()
            e1.record(stream)
            torch.cuda.synchronize()
            t = min(t, e0.elapsed_time(e1))
        return round(2.0 * cp_n * 4 / (t * 1e-3) / 1e9, 1)

    rows = [{"kernel": "gmms_copy (1 load per lane, grid 32 per CU)", "GBps": best(lambda: gmms.copy(b, a, stream=stream))},
            {"kernel": "torch copy_", "GBps": best(lambda: b.copy_(a))}]
    for unroll in (1, 2, 4, 8) if sweep else ():
        for nt in (False, True):
            for bpc in (0, 8, 32):
                rows.append({"kernel": "gmms_copy_probe", "unroll": unroll, "nontemporal": nt, "blocks_per_cu": bpc,
                             "GBps": best(lambda: gmms.copy_probe(b, a, unroll, nt, bpc, stream=stream))})
    ok = bool(torch.equal(a, b))
    del a, b
    return rows, ok


def cpu_baseline(oracle, host, threads):
    """The CPU sorts within a bounded sample: the parallel sort of the full
    input (median of 5 after a warm-up), one single-thread std::sort of the
    full input, and both at N = 2^20 (median of 5 after a warm-up)."""
    import platform

    import numpy as np

    def runs(a, thr, warm, reps):
        secs = np.zeros(reps)
        bad = oracle.lib.oracle_time_sort_i32_runs(a.ctypes.data, a.size, thr, warm, reps, secs.ctypes.data)
        if bad:
            raise RuntimeError("CPU baseline sort left the data unsorted")
        return float(np.median(secs))

    n = host.size
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    t_par = runs(host, threads, 1, 5)
    t_one = runs(host, 1, 0, 1)
    c1 = oracle.gen_i32(0x6D6D7300, 1 << 20, 34)  # config 1: 2^20 in [0, 2^30)
    c1_one, c1_par = runs(c1, 1, 1, 5), runs(c1, threads, 1, 5)
    return {
        "value": round(n / t_par / 1e9, 4),
        "unit": "GKeys/s",
        "cores": threads,
        "kind": "port",
        "what": "the oracle's CPU sorts: libstdc++ parallel sort (OpenMP) and std::sort, int32 '<'",
        "sample": f"parallel sort of the same {n} keys on {threads} threads, median of 5 after 1 warm-up "
                  f"(a fresh copy per run, not timed)",
        "single_thread": {"value": round(n / t_one / 1e9, 4), "unit": "GKeys/s", "cores": 1, "seconds": round(t_one, 3),
                          "sample": f"std::sort of the same {n} keys, one run (the bounded sample)"},
        "config1_2p20": {"single_thread_ms": round(c1_one * 1e3, 3), "parallel_ms": round(c1_par * 1e3, 3),
                         "sample": "N=2^20 int32 in [0, 2^30) (mt19937_64 seed 0x6D6D7300), median of 5 after 1 warm-up"},
        "cpu_model": model,
        "nproc": len(os.sched_getaffinity(0)),
        "hardware_threads": threads,
    }


def host_threads(oracle) -> int:
    """Host threads the CPU work may use: the hardware threads, bounded by
    this process's CPU affinity and OMP_NUM_THREADS when set."""
    t = min(oracle.lib.oracle_hw_threads(), len(os.sched_getaffinity(0)))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        t = min(t, int(env))
    return max(1, t)


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: gpu-mms-amd sorts on one GPU; runs on more than one GPU are not part of it "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import gmms
    from conftest import Oracle

    dev = torch.device("cuda:0")
    n = args.n
    oracle = Oracle()
    seed = 0x6D6D7301
    host = oracle.gen_i32(seed, n, 32)  # (int32)(u64 >> 32): the full int32 range
    base = torch.from_numpy(host).to(dev)
    total = args.warmup + args.steps
    rng = np.random.default_rng(seed)
    consts = [0] + [int(c) for c in rng.integers(1, 2**31 - 1, total - 1)]
    inputs = [base if c == 0 else torch.bitwise_xor(base, c) for c in consts]
    alt = torch.empty_like(base)
    params = gmms.make_params(K=args.K, mergers=args.mergers, prefetch=args.prefetch, pipeline=args.kernel)
    ws = gmms.Workspace()
    stream = torch.cuda.current_stream()
    resolved = gmms.resolve_params(gmms.INT32, params)
    K = resolved.kway
    kernel_name = {0: "k_wmerge (wavefront tile merge)", 1: "k_merge (pipelined minBlockHeap)",
                   2: "k_tmerge (workgroup tile merge)"}[resolved.pipeline]
    rounds = gmms.sort_rounds(gmms.INT32, n, params)  # merge rounds the sort runs
    rounds_ref = gmms.num_rounds(n, K)  # ceil(log_K(N/1024)): the reference's schedule
    # keep an untouched copy of input 0 for the bit-exact check
    check_src = base.clone()

    def step(i):
        return gmms.sort_inplace(inputs[i], alt, params=params, workspace=ws, stream=stream)

    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    gmms.stage_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.warmup, total):
        step(i)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    stages = gmms.stage_times()
    gmms.stage_timing(False)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    log(f"plain sort: {ms_per_step:.3f} ms per step")
    gkeys = n / (ms_per_step * 1e-3) / 1e9

    # per-stage averages (kind -> list of ms), merges grouped by round position
    per_kind = {}
    for kind, ms, _ in stages:
        per_kind.setdefault(gmms.STAGE_NAMES[kind], []).append(ms)
    merges = per_kind.get("merge", [])
    merge_avg = sum(merges) / len(merges) if merges else float("nan")
    merge_bytes = 2.0 * n * 4
    achieved = merge_bytes / (merge_avg * 1e-3) / 1e9 if merges else None
    per_round = []
    if merges and rounds:
        for r in range(rounds):
            vals = merges[r::rounds]
            per_round.append(round(sum(vals) / len(vals), 4))

    # measured stream-copy peak: the best of every probe variant
    copy_rows, copy_ok = copy_probe(dev, stream, args.copy_probe == "sweep")
    copy_peak = max(r["GBps"] for r in copy_rows)
    log(f"copy probe: best {copy_peak} GB/s")

    # bit-exact check of input 0 against the CPU sort
    k2 = check_src.clone()
    k2a = torch.empty_like(k2)
    res = gmms.sort_inplace(k2, k2a, params=params, workspace=ws, stream=stream)
    got = (k2a if res else k2).cpu().numpy()
    threads = host_threads(oracle)
    want = oracle.psort(host, threads)
    bit_exact = bool(np.array_equal(got, want))
    del got, want, k2, k2a

    # the bytes of the passes the sort makes (base case + each round), and
    # the reference's formula 2*N*4*(1 + ceil(log_K(N/1024))) with the K that
    # ran (with 4096-element base runs it may count a pass that is not made)
    pass_bytes = 2.0 * n * 4 * (1 + rounds)
    formula_bytes = 2.0 * n * 4 * (1 + rounds_ref)
    pass_gbps = pass_bytes / (ms_per_step * 1e-3) / 1e9
    formula_gbps = formula_bytes / (ms_per_step * 1e-3) / 1e9

    log(f"bit-exact check done: {bit_exact}")
    cpu = None if args.no_cpu_baseline else cpu_baseline(oracle, host, threads)
    if cpu:
        log("cpu baseline done")
    batch = None
    if not args.no_batched:
        batch = batched(args, oracle, dev, stream, threads)

    extra = None if args.no_configs else other_configs(oracle, dev, stream)
    serve = None
    if not args.no_serving:
        serve = serving(args, oracle, threads)

    out = {
        "metric": METRIC,
        "value": round(gkeys, 4),
        "unit": "GKeys/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": f"synthetic: mt19937_64 seed {seed:#x}, key=(int32)(u64>>32); step i XORs a constant",
        "config": {"workload": "plain sort, N=2^28 int32 uniform over the full int32 range",
                   "n": n, "K": K, "rounds": rounds, "rounds_reference_schedule": rounds_ref,
                   "base_run": resolved.base_run, "merge_kernel": resolved.pipeline,
                   "mergers": resolved.mergers},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2) if achieved else None,
                     "peak": PUBLISHED_HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / PUBLISHED_HBM_PEAK_GBPS, 4) if achieved else None,
                     "traffic": traffic_from_profiles(resolved.pipeline, K, n),
                     "kernel": kernel_name, "algorithmic_bytes_per_launch": merge_bytes,
                     "avg_launch_ms": round(merge_avg, 4)},
        "cpu_baseline": cpu,
        "bit_exact_vs_std_sort": bit_exact,
        "stream_copy_peak_GBps": round(copy_peak, 1),
        "copy_probe": {"variants": copy_rows, "copies_exact": copy_ok,
                       "note": "4 GiB D2D copies, best of 5 each; the best is the copy peak"},
        "sort_roofline": {"passes_made": 1 + rounds, "bytes": pass_bytes,
                          "achieved_GBps": round(pass_gbps, 1),
                          "frac_of_published_peak": round(pass_gbps / PUBLISHED_HBM_PEAK_GBPS, 4),
                          "frac_of_copy_peak": round(pass_gbps / copy_peak, 4),
                          "formula": "2*N*4*(1+ceil(log_K(N/1024)))", "formula_bytes": formula_bytes,
                          "formula_GBps": round(formula_gbps, 1),
                          "formula_frac_of_copy_peak": round(formula_gbps / copy_peak, 4),
                          "formula_note": ("the formula counts %d passes; the sort makes %d (4096-element "
                                           "base runs), so the formula figure credits %d pass(es) not made"
                                           % (1 + rounds_ref, 1 + rounds, rounds_ref - rounds))},
        "stage_ms_avg": {k: round(sum(v) / len(v), 4) for k, v in per_kind.items()},
        "merge_round_ms": per_round,
        "batched": batch,
        "serving": serve,
        "configs": extra,
    }
    print(json.dumps(out), flush=True)
    if not bit_exact or (batch is not None and not (batch["bit_exact_vs_std_sort"] and batch["identical_to_per_array"])):
        sys.exit(1)
    if serve is not None and not (serve["bit_exact_vs_std_sort"] and serve["all_status_ok"]):
        sys.exit(1)
    if extra is not None and not all(c["bit_exact_vs_std_sort"] for c in extra):
        sys.exit(1)


if __name__ == "__main__":
    main()
