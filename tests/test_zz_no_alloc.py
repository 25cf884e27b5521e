"""gmms_sort_noalloc: the sort with its workspace in a static device array
instead of a caller workspace. Called through the C ABI (ctypes) and compared
bit for bit with gmms_sort on the same input, twice in a row on one GPU with
the second input smaller (the scratch keeps the first call's contents), for
32-bit keys and for pairs; a size beyond the scratch is refused."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sort_ws(gmms, torch, dt, k, v, dev):
    """gmms_sort with a caller workspace: (keys, values) of the result."""
    L = gmms.lib()
    n = k.numel()
    k0, k1 = k.clone(), torch.empty_like(k)
    v0 = v.clone() if v is not None else None
    v1 = torch.empty_like(v) if v is not None else None
    nb = gmms.workspace_size(dt, n)
    ws = torch.empty(max(nb, 1), dtype=torch.uint8, device=dev)
    alt = ctypes.c_int(-1)
    s = torch.cuda.current_stream().cuda_stream
    st = L.gmms_sort(dt, k0.data_ptr(), k1.data_ptr(), None if v is None else v0.data_ptr(),
                     None if v is None else v1.data_ptr(), n, None, ws.data_ptr(), nb, s, ctypes.byref(alt))
    assert st == 0, gmms.status_string(st)
    torch.cuda.synchronize()
    return (k1, v1) if alt.value else (k0, v0)


def _sort_noalloc(gmms, torch, dt, k, v):
    L = gmms.lib()
    n = k.numel()
    k0, k1 = k.clone(), torch.empty_like(k)
    v0 = v.clone() if v is not None else None
    v1 = torch.empty_like(v) if v is not None else None
    torch.cuda.synchronize()
    alt = ctypes.c_int(-1)
    s = torch.cuda.current_stream().cuda_stream
    st = L.gmms_sort_noalloc(dt, k0.data_ptr(), k1.data_ptr(), None if v is None else v0.data_ptr(),
                             None if v is None else v1.data_ptr(), n, None, s, ctypes.byref(alt))
    assert st == 0, gmms.status_string(st) + " " + gmms.lib().gmms_last_error().decode()
    torch.cuda.synchronize()
    return (k1, v1) if alt.value else (k0, v0)


def test_scratch_size(gmms):
    b = ctypes.c_size_t(0)
    assert gmms.lib().gmms_noalloc_scratch_bytes(ctypes.byref(b)) == 0
    assert b.value == 16 << 20


@pytest.mark.parametrize("sizes", [((1 << 24) + 12345, (1 << 20) + 77)])
def test_noalloc_matches_sort_int32(gmms, oracle, dev, sizes):
    import torch

    for i, n in enumerate(sizes):
        print(f"noalloc int32 call {i} n={n}", flush=True)
        a = torch.from_numpy(oracle.gen_i32(0x6D6D7400 + i, n, 32)).to(dev)
        want, _ = _sort_ws(gmms, torch, gmms.INT32, a, None, dev)
        got, _ = _sort_noalloc(gmms, torch, gmms.INT32, a, None)
        assert torch.equal(got, want)
        if n <= (1 << 21):
            assert np.array_equal(got.cpu().numpy(), oracle.sort(a.cpu().numpy()))


def test_noalloc_matches_sort_pairs(gmms, oracle, dev):
    import torch

    for i, n in enumerate(((1 << 23) + 5, 300_001)):
        print(f"noalloc pairs call {i} n={n}", flush=True)
        k, v = oracle.gen_pairs(0x6D6D7410 + i, n, 3, 1.1, 1 << 20)
        kt = torch.from_numpy(k.view(np.int64)).to(dev)
        vt = torch.from_numpy(v.view(np.int32)).to(dev)
        wk, wv = _sort_ws(gmms, torch, gmms.PAIR_U64_U32, kt, vt, dev)
        gk, gv = _sort_noalloc(gmms, torch, gmms.PAIR_U64_U32, kt, vt)
        assert torch.equal(gk, wk) and torch.equal(gv, wv)


def test_noalloc_refuses_beyond_scratch(gmms, dev):
    import torch

    L = gmms.lib()
    n = 400_000_000  # pairs: the workspace needs more than 16 MiB
    assert gmms.workspace_size(gmms.PAIR_U64_U32, n) > (16 << 20)
    k = torch.empty(8, dtype=torch.int64, device=dev)  # never touched: refused before any launch
    v = torch.empty(8, dtype=torch.int32, device=dev)
    alt = ctypes.c_int(0)
    st = L.gmms_sort_noalloc(gmms.PAIR_U64_U32, k.data_ptr(), k.data_ptr(), v.data_ptr(), v.data_ptr(), n, None,
                             None, ctypes.byref(alt))
    assert st == gmms.ERROR_WORKSPACE
    assert b"static scratch" in L.gmms_last_error()
