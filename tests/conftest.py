import ctypes
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "gpu-mms_amd"))
sys.path.insert(0, ROOT)

ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libgmms.so on cuda:0)")


class Oracle:
    """ctypes view of oracle/liboracle.so (test infrastructure only)."""

    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            import subprocess

            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        self.lib = ctypes.CDLL(path)
        L = self.lib
        vp, i64, i32, u64, f64 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_uint64, ctypes.c_double
        L.oracle_mt64_nth.restype = u64
        L.oracle_mt64_nth.argtypes = [u64, u64]
        L.oracle_gen_u64.argtypes = [u64, i64, vp]
        L.oracle_gen_i32.argtypes = [u64, i64, i32, vp]
        L.oracle_gen_f32_normal.argtypes = [u64, i64, vp]
        L.oracle_gen_pairs.argtypes = [u64, i64, i32, f64, i64, vp, vp]
        L.oracle_gen_seglens.argtypes = [u64, i64, f64, f64, vp]
        L.oracle_glibc_rand_list.argtypes = [ctypes.c_uint32, i64, ctypes.c_int32, vp]
        L.oracle_f32_order_key.restype = ctypes.c_uint32
        L.oracle_f32_order_key.argtypes = [ctypes.c_uint32]
        L.oracle_f64_order_key.restype = u64
        L.oracle_f64_order_key.argtypes = [u64]
        for n in ("i32", "u32", "f32", "i64", "u64", "f64"):
            getattr(L, f"oracle_sort_{n}").argtypes = [vp, i64]
        L.oracle_sort_pairs.argtypes = [vp, vp, i64]
        L.oracle_psort_i32.argtypes = [vp, i64, i32]
        L.oracle_psort_u64pairs.argtypes = [vp, vp, i64, i32]
        L.oracle_psort_u64.argtypes = [vp, i64, i32]
        L.oracle_psort_f32.argtypes = [vp, i64, i32]
        L.oracle_time_sort_i32_runs.restype = i32
        L.oracle_time_sort_i32_runs.argtypes = [vp, i64, i32, i32, i32, vp]
        L.oracle_basecase_i32.argtypes = [vp, i64, i64]
        L.oracle_rounds.restype = i32
        L.oracle_rounds.argtypes = [i64, i64, i64]
        for n in ("oracle_tasks", "oracle_edge_task_size"):
            getattr(L, n).restype = i64
            getattr(L, n).argtypes = [i64, i64, i64]
        L.oracle_merger_target.restype = i64
        L.oracle_merger_target.argtypes = [i64, i64, i64]
        L.oracle_merge_round_i32.argtypes = [vp, vp, i64, i64, i64]
        L.oracle_corank_i32.argtypes = [vp, i64, vp, i32, i64, vp]
        L.oracle_time_sort_i32.restype = f64
        L.oracle_time_sort_i32.argtypes = [vp, i64, i32, i32]
        L.oracle_hw_threads.restype = i32

    @staticmethod
    def p(a):
        return a.ctypes.data_as(ctypes.c_void_p)

    # generators
    def gen_u64(self, seed, n):
        a = np.empty(n, np.uint64)
        self.lib.oracle_gen_u64(seed, n, self.p(a))
        return a

    def gen_i32(self, seed, n, shift):
        a = np.empty(n, np.int32)
        self.lib.oracle_gen_i32(seed, n, shift, self.p(a))
        return a

    def gen_f32_normal(self, seed, n):
        a = np.empty(n, np.float32)
        self.lib.oracle_gen_f32_normal(seed, n, self.p(a))
        return a

    def gen_pairs(self, seed, n, dist, s=1.1, distinct=1 << 20):
        k = np.empty(n, np.uint64)
        v = np.empty(n, np.uint32)
        self.lib.oracle_gen_pairs(seed, n, dist, s, distinct, self.p(k), self.p(v))
        return k, v

    def gen_seglens(self, seed, count, lo, hi):
        a = np.empty(count, np.int64)
        self.lib.oracle_gen_seglens(seed, count, lo, hi, self.p(a))
        return a

    def glibc_rand_list(self, seed, n, rng):
        a = np.empty(n, np.int32)
        self.lib.oracle_glibc_rand_list(seed, n, rng, self.p(a))
        return a

    # the sort oracle (returns sorted copies)
    def sort(self, a):
        b = np.ascontiguousarray(a).copy()
        name = {np.dtype(np.int32): "i32", np.dtype(np.uint32): "u32", np.dtype(np.float32): "f32",
                np.dtype(np.int64): "i64", np.dtype(np.uint64): "u64", np.dtype(np.float64): "f64"}[b.dtype]
        getattr(self.lib, f"oracle_sort_{name}")(self.p(b), b.size)
        return b

    def psort(self, a, threads=0):
        """Parallel CPU sort (all threads by default) of int32, uint64 or
        float32 (total order) keys; a sorted copy."""
        b = np.ascontiguousarray(a).copy()
        name = {np.dtype(np.int32): "i32", np.dtype(np.uint64): "u64", np.dtype(np.float32): "f32"}[b.dtype]
        getattr(self.lib, f"oracle_psort_{name}")(self.p(b), b.size, threads)
        return b

    def psort_pairs(self, k, v, threads=0):
        k2, v2 = np.ascontiguousarray(k).copy(), np.ascontiguousarray(v).copy()
        self.lib.oracle_psort_u64pairs(self.p(k2), self.p(v2), k2.size, threads)
        return k2, v2

    def sort_pairs(self, k, v):
        k2, v2 = np.ascontiguousarray(k).copy(), np.ascontiguousarray(v).copy()
        self.lib.oracle_sort_pairs(self.p(k2), self.p(v2), k2.size)
        return k2, v2

    def basecase_i32(self, a, tile=1024):
        b = a.copy()
        self.lib.oracle_basecase_i32(self.p(b), b.size, tile)
        return b

    def merge_round_i32(self, a, K, L):
        out = np.empty_like(a)
        self.lib.oracle_merge_round_i32(self.p(a), self.p(out), a.size, K, L)
        return out

    def corank_i32(self, lists_flat, stride, lens, r):
        K = len(lens)
        la = np.asarray(lens, np.int64)
        s = np.zeros(K, np.int64)
        self.lib.oracle_corank_i32(self.p(lists_flat), stride, self.p(la), K, r, self.p(s))
        return s


@pytest.fixture(scope="session")
def oracle():
    return Oracle()


def gpu_available():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gmms():
    import gmms as g

    g.lib()
    return g


@pytest.fixture(scope="session")
def dev():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    return torch.device("cuda:0")
