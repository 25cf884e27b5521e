"""Python host side of the MI355X-native multiway mergesort.

A thin ctypes layer over the C ABI of ``libgmms.so`` (declared in
``include/gpu_mms/gmms.h``). PyTorch is used only for device memory and
streams. The product path is the HIP library: importing this module on a
machine where ``libgmms.so`` is missing raises, there is no CPU fallback.

Mirrors the reference's operator interface (algoparc/GPU-MMS):

* :func:`multimergesort` -- ``multimergesort<T,cmp>(input, output, host_array,
  threadBlocks, N)`` (src/multimergesort.hxx:56-141, README.md:16-25): sorts
  ``input`` using ``output`` as the second buffer and returns whichever holds
  the result.
* :func:`sort` -- the same sort with a caller-visible workspace and stream,
  returning fresh sorted tensors (keys, or keys and values for pairs).
* :func:`squaresort`, :func:`merge_round`, :func:`find_partitions`,
  :func:`pad` -- the reference's kernels squareSort, multimergeLevel,
  findPartitions and pad as single stages.
"""
from __future__ import annotations

import ctypes
import os
import re
from dataclasses import dataclass
from typing import Optional, Tuple

_HERE = os.path.dirname(os.path.abspath(__file__))
# the in-tree build, or another build of the same library named by GMMS_LIBRARY
LIB_PATH = os.environ.get("GMMS_LIBRARY") or os.path.join(_HERE, "libgmms.so")
HEADER_PATH = os.path.normpath(os.path.join(_HERE, "..", "..", "include", "gpu_mms", "gmms.h"))

# dtype codes (gmms.h: gmms_dtype)
INT32, UINT32, FLOAT32, INT64, UINT64, FLOAT64, PAIR_U64_U32 = range(7)
DTYPE_NAMES = {
    INT32: "int32", UINT32: "uint32", FLOAT32: "float32", INT64: "int64",
    UINT64: "uint64", FLOAT64: "float64", PAIR_U64_U32: "pair_u64_u32",
}

SUCCESS = 0
ERROR_INVALID_VALUE = 1
ERROR_UNSUPPORTED = 2
ERROR_WORKSPACE = 3
ERROR_HIP = 4
ERROR_OUT_OF_MEMORY = 5
ERROR_BUDGET = 6

TILE = 1024  # base-case tile (the reference's M, src/params.h:36)


class GmmsError(RuntimeError):
    def __init__(self, status: int, where: str):
        self.status = status
        detail = ""
        if status == ERROR_HIP:
            detail = " -- " + lib().gmms_last_error().decode()
        super().__init__(f"{where}: {status_string(status)} (status {status}){detail}")


class Params(ctypes.Structure):
    """gmms_params (gmms.h); replaces src/params.h:18-38 (K, PIPELINE, BLOCKS)."""

    _fields_ = [
        ("kway", ctypes.c_int32),
        ("pipeline", ctypes.c_int32),
        ("mergers", ctypes.c_int32),
        ("prefetch", ctypes.c_int32),
        ("slack", ctypes.c_int32),
        ("base_run", ctypes.c_int32),
        ("descending", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 1),
    ]


class SegInfo(ctypes.Structure):
    """gmms_seg_info (gmms.h): the shape of a planned batch."""

    _fields_ = [
        ("n_seg", ctypes.c_int64),
        ("elements", ctypes.c_int64),
        ("extent", ctypes.c_int64),
        ("tiles", ctypes.c_int64),
        ("merge_items", ctypes.c_int64),
        ("partition_items", ctypes.c_int64),
        ("workspace_bytes", ctypes.c_size_t),
        ("rounds", ctypes.c_int32),
        ("launches", ctypes.c_int32),
    ]


SEG_CLASSES = 41  # GMMS_SEG_CLASSES


class ServerConfig(ctypes.Structure):
    """gmms_server_config (gmms.h)."""

    _fields_ = [
        ("hbm_budget", ctypes.c_size_t),
        ("max_batch_elems", ctypes.c_int64),
        ("max_batch_requests", ctypes.c_int32),
        ("max_inflight", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 4),
    ]


class ServerStats(ctypes.Structure):
    """gmms_server_stats (gmms.h)."""

    _fields_ = [
        ("submitted", ctypes.c_int64),
        ("completed", ctypes.c_int64),
        ("rejected", ctypes.c_int64),
        ("failed", ctypes.c_int64),
        ("batches", ctypes.c_int64),
        ("max_batch_size", ctypes.c_int64),
        ("max_inflight_seen", ctypes.c_int64),
        ("keys_sorted", ctypes.c_int64),
        ("budget", ctypes.c_size_t),
        ("arena_bytes", ctypes.c_size_t),
        ("reserved_bytes", ctypes.c_size_t),
        ("peak_reserved_bytes", ctypes.c_size_t),
    ]

_lib = None


def lib() -> ctypes.CDLL:
    """Load libgmms.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libgmms.so not built at {LIB_PATH}; run __graft_entry__.build()")
    # PyTorch bundles its own HIP runtime under the same soname
    # (libamdhip64.so.7). Load it first so libgmms.so binds to that one
    # instead of pulling /opt/rocm's into the same process: two HIP runtimes
    # in one process do not share devices, streams or allocations.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp, i64, i32, sz = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_size_t
    P = ctypes.POINTER(Params)
    sig = {
        "gmms_version": (ctypes.c_char_p, []),
        "gmms_status_string": (ctypes.c_char_p, [i32]),
        "gmms_last_error": (ctypes.c_char_p, []),
        "gmms_multimergesort_params": (i32, [i32, vp, vp, vp, P, i64, ctypes.POINTER(vp)]),
        "gmms_get_default_params": (i32, [P]),
        "gmms_resolve_params": (i32, [i32, P, P]),
        "gmms_element_size": (i32, [i32, ctypes.POINTER(sz), ctypes.POINTER(sz)]),
        "gmms_num_rounds": (i32, [i64, i32, ctypes.POINTER(i32)]),
        "gmms_sort_rounds": (i32, [i32, i64, P, ctypes.POINTER(i32)]),
        "gmms_workspace_size": (i32, [i32, i64, P, ctypes.POINTER(sz)]),
        "gmms_round_workspace_size": (i32, [i32, i64, i64, P, ctypes.POINTER(sz)]),
        "gmms_sort": (i32, [i32, vp, vp, vp, vp, i64, P, vp, sz, vp, ctypes.POINTER(i32)]),
        "gmms_multimergesort": (i32, [i32, vp, vp, vp, i32, i64, ctypes.POINTER(vp)]),
        "gmms_sort_noalloc": (i32, [i32, vp, vp, vp, vp, i64, P, vp, ctypes.POINTER(i32)]),
        "gmms_noalloc_scratch_bytes": (i32, [ctypes.POINTER(sz)]),
        "gmms_squaresort": (i32, [i32, vp, vp, i64, i32, i32, vp]),
        "gmms_merge_round": (i32, [i32, vp, vp, vp, vp, i64, i64, P, i32, vp, sz, vp]),
        "gmms_find_partitions": (i32, [i32, vp, vp, i64, i64, P, vp, vp, i64, ctypes.POINTER(i32), vp]),
        "gmms_pad": (i32, [i32, vp, vp, i64, vp]),
        "gmms_stage_timing": (i32, [i32]),
        "gmms_copy": (i32, [vp, vp, sz, vp]),
        "gmms_copy_probe": (i32, [vp, vp, sz, i32, i32, i32, vp]),
        "gmms_stage_times": (i32, [vp, vp, vp, i32, ctypes.POINTER(i32)]),
        "gmms_seg_plan_create": (i32, [i32, vp, i64, P, ctypes.POINTER(vp)]),
        "gmms_seg_plan_info": (i32, [vp, ctypes.POINTER(SegInfo)]),
        "gmms_seg_plan_segment": (i32, [vp, i64, ctypes.POINTER(i32), ctypes.POINTER(i32)]),
        "gmms_seg_plan_destroy": (None, [vp]),
        "gmms_seg_sort": (i32, [vp, vp, vp, vp, vp, vp, sz, vp, ctypes.POINTER(i32)]),
        "gmms_seg_graph_create": (i32, [vp, vp, vp, vp, vp, vp, sz, ctypes.POINTER(vp)]),
        "gmms_seg_graph_launch": (i32, [vp, vp, ctypes.POINTER(i32)]),
        "gmms_seg_graph_destroy": (None, [vp]),
        "gmms_segmented_sort": (i32, [i32, vp, vp, vp, vp, vp, i64, P, vp, ctypes.POINTER(i32)]),
        "gmms_seg_tuning_get": (i32, [i32, vp, vp, i32]),
        "gmms_seg_tuning_set": (i32, [i32, vp, vp, i32]),
        "gmms_seg_autotune": (i32, [i32, i32, i32, i64, i32, vp, vp, vp, vp]),
        "gmms_server_default_config": (i32, [ctypes.POINTER(ServerConfig)]),
        "gmms_server_request_bytes": (i32, [i32, i64, ctypes.POINTER(sz)]),
        "gmms_server_create": (i32, [ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]),
        "gmms_server_submit": (i32, [vp, i32, vp, vp, i64, vp, vp, ctypes.POINTER(ctypes.c_uint64)]),
        "gmms_server_submit_after": (i32, [vp, i32, vp, vp, i64, vp, vp, vp, i32,
                                           ctypes.POINTER(ctypes.c_uint64)]),
        "gmms_server_wait": (i32, [vp, ctypes.c_uint64, ctypes.c_double, ctypes.POINTER(i32), ctypes.POINTER(i32),
                                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(i64)]),
        "gmms_server_drain": (i32, [vp]),
        "gmms_server_stats_get": (i32, [vp, ctypes.POINTER(ServerStats)]),
        "gmms_server_destroy": (None, [vp]),
        "gmms_server_batch_policy": (i32, [vp, vp, i64, sz, i64, i32, vp, ctypes.POINTER(i64)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def declared_symbols(header: str = HEADER_PATH) -> list:
    """Every function name declared in include/gpu_mms/*.h (for export checks)."""
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gmms_[a-z0-9_]+)\s*\(", text)))


def status_string(status: int) -> str:
    return lib().gmms_status_string(status).decode()


def version() -> str:
    return lib().gmms_version().decode()


def _check(status: int, where: str) -> None:
    if status != SUCCESS:
        raise GmmsError(status, where)


def default_params() -> Params:
    p = Params()
    _check(lib().gmms_get_default_params(ctypes.byref(p)), "gmms_get_default_params")
    return p


def make_params(K: int = 0, mergers: int = 0, prefetch: int = 0, slack: int = 0,
                pipeline: Optional[int] = None, base_run: int = 0, descending: bool = False) -> Params:
    """pipeline=1: pipelined minBlockHeap merge; pipeline=0: tile merge.
    base_run: 1024 (the reference's tile) or 4096 sorted elements per base
    run; 0: auto."""
    p = default_params()
    if base_run:
        p.base_run = base_run
    if descending:
        p.descending = 1
    if pipeline is not None:
        p.pipeline = int(pipeline)
    if K:
        p.kway = K
    if mergers:
        p.mergers = mergers
    if prefetch:
        p.prefetch = prefetch
    if slack:
        p.slack = slack
    return p


def resolve_params(dtype: int, params: Optional[Params] = None) -> Params:
    """The K, merge kernel and merger target a sort of `dtype` runs with."""
    out = Params()
    _check(lib().gmms_resolve_params(dtype, _pp(params), ctypes.byref(out)), "gmms_resolve_params")
    return out


def num_rounds(n: int, K: int = 64) -> int:
    """Rounds of the reference's schedule (lists start at 1024)."""
    r = ctypes.c_int()
    _check(lib().gmms_num_rounds(n, K, ctypes.byref(r)), "gmms_num_rounds")
    return r.value


def sort_rounds(dtype: int, n: int, params: Optional[Params] = None) -> int:
    """Rounds gmms_sort runs for n elements with these knobs."""
    r = ctypes.c_int()
    _check(lib().gmms_sort_rounds(dtype, n, _pp(params), ctypes.byref(r)), "gmms_sort_rounds")
    return r.value


def workspace_size(dtype: int, n: int, params: Optional[Params] = None) -> int:
    b = ctypes.c_size_t()
    _check(lib().gmms_workspace_size(dtype, n, _pp(params), ctypes.byref(b)), "gmms_workspace_size")
    return b.value


def _pp(params: Optional[Params]):
    return ctypes.byref(params) if params is not None else None


# ---------------------------------------------------------------- torch side
def _torch():
    import torch  # noqa: WPS433 -- torch is only plumbing for device memory

    return torch


def dtype_of(t) -> int:
    torch = _torch()
    m = {
        torch.int32: INT32, torch.uint32: UINT32, torch.float32: FLOAT32,
        torch.int64: INT64, torch.uint64: UINT64, torch.float64: FLOAT64,
    }
    if t.dtype not in m:
        raise GmmsError(ERROR_INVALID_VALUE, f"unsupported tensor dtype {t.dtype}")
    return m[t.dtype]


def _ptr(t) -> Optional[int]:
    return None if t is None else t.data_ptr()


def _stream_ptr(stream) -> Optional[int]:
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def _check_dev(*ts) -> None:
    for t in ts:
        if t is not None and (not t.is_cuda or not t.is_contiguous()):
            raise GmmsError(ERROR_INVALID_VALUE, "tensors must be contiguous device tensors")


@dataclass
class Workspace:
    """Caller-owned device workspace (partition cuts) reused across calls."""

    tensor: object = None

    def ensure(self, nbytes: int, device, stream=None) -> Tuple[Optional[int], int]:
        """A device pointer to at least nbytes; a new block is allocated on
        `stream` (default: the current stream), the stream that uses it."""
        torch = _torch()
        if nbytes == 0:
            return None, 0
        if self.tensor is None or self.tensor.numel() < nbytes or self.tensor.device != device:
            self.tensor = None
            with torch.cuda.stream(stream if stream is not None else torch.cuda.current_stream(device)):
                self.tensor = torch.empty(nbytes, dtype=torch.uint8, device=device)
        return self.tensor.data_ptr(), self.tensor.numel()


# default workspaces, one per (device, stream): calls on different streams
# never share one (calls on one stream are ordered by the stream). Calls from
# several threads on the same stream must pass their own Workspace.
_default_ws = {}


def _workspace_for(workspace: Optional[Workspace], device, stream) -> Workspace:
    if workspace is not None:
        return workspace
    key = (str(device), _stream_ptr(stream))
    ws = _default_ws.get(key)
    if ws is None:
        ws = _default_ws[key] = Workspace()
    return ws


def _same_stream(a, b) -> bool:
    return a.cuda_stream == b.cuda_stream and a.device == b.device


def sort_inplace(keys, keys_alt, values=None, values_alt=None, *, dtype: Optional[int] = None,
                 params: Optional[Params] = None, workspace: Optional[Workspace] = None,
                 stream=None) -> int:
    """gmms_sort on preallocated buffers; returns 0 if the result is in
    keys/values, 1 if it is in keys_alt/values_alt. Asynchronous."""
    _check_dev(keys, keys_alt, values, values_alt)
    dt = dtype if dtype is not None else (PAIR_U64_U32 if values is not None else dtype_of(keys))
    n = keys.numel()
    ws = _workspace_for(workspace, keys.device, stream)
    nbytes = workspace_size(dt, n, params)
    wp, wb = ws.ensure(nbytes, keys.device, stream)
    alt = ctypes.c_int(0)
    st = lib().gmms_sort(dt, _ptr(keys), _ptr(keys_alt), _ptr(values), _ptr(values_alt), n, _pp(params),
                         wp, wb, _stream_ptr(stream), ctypes.byref(alt))
    _check(st, "gmms_sort")
    return alt.value


def sort(keys, values=None, *, params: Optional[Params] = None, workspace: Optional[Workspace] = None,
         stream=None):
    """Sort a device tensor (copy). Pairs: keys uint64/int64 viewed as uint64
    and values uint32/int32 viewed as uint32, ordered by (key, value).
    Returns the sorted keys, or (keys, values)."""
    torch = _torch()
    cur = torch.cuda.current_stream(keys.device)
    s = stream if stream is not None else cur
    if not _same_stream(s, cur):
        s.wait_stream(cur)  # the inputs come from work queued on the current stream
    # copies, buffers and the sort all on s: the results belong to s
    with torch.cuda.stream(s):
        k = keys.contiguous().clone()
        ka = torch.empty_like(k)
        if values is None:
            alt = sort_inplace(k, ka, dtype=dtype_of(k), params=params, workspace=workspace, stream=s)
            return ka if alt else k
        v = values.contiguous().clone()
        va = torch.empty_like(v)
        if k.element_size() != 8 or v.element_size() != 4 or k.numel() != v.numel():
            raise GmmsError(ERROR_INVALID_VALUE, "pairs need 64-bit keys and 32-bit values of equal length")
        alt = sort_inplace(k, ka, v, va, dtype=PAIR_U64_U32, params=params, workspace=workspace, stream=s)
        return (ka, va) if alt else (k, v)


def multimergesort(input, output, host_array=None, thread_blocks: int = 0, n: Optional[int] = None,
                   dtype: Optional[int] = None):
    """The reference's drop-in call: returns `input` or `output`, whichever
    holds the sorted data (synchronous, allocates its own workspace)."""
    _check_dev(input, output)
    dt = dtype if dtype is not None else dtype_of(input)
    n = input.numel() if n is None else n
    res = ctypes.c_void_p()
    st = lib().gmms_multimergesort(dt, _ptr(input), _ptr(output), None, thread_blocks, n, ctypes.byref(res))
    _check(st, "gmms_multimergesort")
    return output if res.value == output.data_ptr() else input


def squaresort(keys, values=None, *, dtype: Optional[int] = None, to_internal: bool = True,
               to_external: bool = True, stream=None) -> None:
    """squareSort: sort every 1024-element tile in place."""
    _check_dev(keys, values)
    dt = dtype if dtype is not None else (PAIR_U64_U32 if values is not None else dtype_of(keys))
    st = lib().gmms_squaresort(dt, _ptr(keys), _ptr(values), keys.numel(), int(to_internal), int(to_external),
                               _stream_ptr(stream))
    _check(st, "gmms_squaresort")


def merge_round(in_keys, out_keys, list_len: int, in_values=None, out_values=None, *,
                dtype: Optional[int] = None, params: Optional[Params] = None, to_external: bool = False,
                workspace: Optional[Workspace] = None, stream=None) -> None:
    """One merge round (findPartitions + multimergeLevel) on internal-order data."""
    _check_dev(in_keys, out_keys, in_values, out_values)
    dt = dtype if dtype is not None else (PAIR_U64_U32 if in_values is not None else dtype_of(in_keys))
    n = in_keys.numel()
    ws = _workspace_for(workspace, in_keys.device, stream)
    b = ctypes.c_size_t()
    _check(lib().gmms_round_workspace_size(dt, n, list_len, _pp(params), ctypes.byref(b)),
           "gmms_round_workspace_size")
    nbytes = b.value
    wp, wb = ws.ensure(nbytes, in_keys.device, stream)
    st = lib().gmms_merge_round(dt, _ptr(in_keys), _ptr(in_values), _ptr(out_keys), _ptr(out_values), n,
                                list_len, _pp(params), int(to_external), wp, wb, _stream_ptr(stream))
    _check(st, "gmms_merge_round")


def find_partitions(keys, list_len: int, values=None, *, dtype: Optional[int] = None,
                    params: Optional[Params] = None, cap_mergers: int = 1 << 16, stream=None):
    """findPartitions: returns (cuts[mergers, K], skips[mergers], m) on device."""
    torch = _torch()
    _check_dev(keys, values)
    dt = dtype if dtype is not None else (PAIR_U64_U32 if values is not None else dtype_of(keys))
    p = params if params is not None else default_params()
    K = p.kway or 64
    cuts = torch.zeros(cap_mergers * K, dtype=torch.int64, device=keys.device)
    skips = torch.zeros(cap_mergers, dtype=torch.int64, device=keys.device)
    m = ctypes.c_int(0)
    st = lib().gmms_find_partitions(dt, _ptr(keys), _ptr(values), keys.numel(), list_len, ctypes.byref(p),
                                    _ptr(cuts), _ptr(skips), cap_mergers, ctypes.byref(m), _stream_ptr(stream))
    _check(st, "gmms_find_partitions")
    return cuts.view(cap_mergers, K), skips, m.value


def pad(keys, count: int, values=None, *, dtype: Optional[int] = None, stream=None) -> None:
    """pad: write `count` maximum elements at the start of keys (and values)."""
    _check_dev(keys, values)
    dt = dtype if dtype is not None else (PAIR_U64_U32 if values is not None else dtype_of(keys))
    _check(lib().gmms_pad(dt, _ptr(keys), _ptr(values), count, _stream_ptr(stream)), "gmms_pad")


def copy(dst, src, stream=None) -> None:
    """128-bit stream copy of src into dst (the HBM peak probe)."""
    _check_dev(dst, src)
    nbytes = src.numel() * src.element_size()
    _check(lib().gmms_copy(_ptr(dst), _ptr(src), nbytes, _stream_ptr(stream)), "gmms_copy")


def copy_probe(dst, src, unroll: int = 1, nontemporal: bool = False, blocks_per_cu: int = 0, stream=None) -> None:
    """A variant of the copy probe (gmms_copy_probe)."""
    _check_dev(dst, src)
    nbytes = src.numel() * src.element_size()
    _check(lib().gmms_copy_probe(_ptr(dst), _ptr(src), nbytes, unroll, int(nontemporal), blocks_per_cu,
                                 _stream_ptr(stream)), "gmms_copy_probe")


# ------------------------------------------------------------ stage timing
STAGE_BASECASE, STAGE_PARTITION, STAGE_MERGE = 0, 1, 2
STAGE_NAMES = {STAGE_BASECASE: "basecase", STAGE_PARTITION: "partition", STAGE_MERGE: "merge"}


def stage_timing(enable: bool) -> None:
    """Record HIP events around every library kernel launch (on its stream)."""
    _check(lib().gmms_stage_timing(int(enable)), "gmms_stage_timing")


def stage_times():
    """[(stage_kind, ms, elements)] for the launches recorded since stage_timing()."""
    import numpy as np

    cnt = ctypes.c_int(0)
    _check(lib().gmms_stage_times(None, None, None, 0, ctypes.byref(cnt)), "gmms_stage_times")
    n = cnt.value
    kinds = np.zeros(max(n, 1), np.int32)
    ms = np.zeros(max(n, 1), np.float32)
    el = np.zeros(max(n, 1), np.int64)
    _check(lib().gmms_stage_times(kinds.ctypes.data, ms.ctypes.data, el.ctypes.data, n, ctypes.byref(cnt)),
           "gmms_stage_times")
    return [(int(kinds[i]), float(ms[i]), int(el[i])) for i in range(n)]


# ------------------------------------------------------- batched segments
def _offsets_array(offsets):
    import numpy as np

    a = np.ascontiguousarray(np.asarray(offsets, dtype=np.int64))
    if a.ndim != 1 or a.size < 1:
        raise GmmsError(ERROR_INVALID_VALUE, "offsets must be a 1-D sequence of n_seg + 1 entries")
    return a


class SegPlan:
    """A planned batch of segments (gmms_seg_plan): segment s is
    keys[offsets[s]:offsets[s+1]]. Planning is host work; sort() copies the
    plan's tables into a device workspace and launches on the stream."""

    def __init__(self, dtype: int, offsets, params: Optional[Params] = None):
        self.dtype = dtype
        self.offsets = _offsets_array(offsets)
        h = ctypes.c_void_p()
        st = lib().gmms_seg_plan_create(dtype, self.offsets.ctypes.data, self.offsets.size - 1, _pp(params),
                                        ctypes.byref(h))
        _check(st, "gmms_seg_plan_create")
        self._h = h
        self.info = SegInfo()
        _check(lib().gmms_seg_plan_info(self._h, ctypes.byref(self.info)), "gmms_seg_plan_info")
        self._keep = []

    def segment(self, i: int) -> Tuple[int, int]:
        """(K, merge rounds) chosen for segment i."""
        k, r = ctypes.c_int(0), ctypes.c_int(0)
        _check(lib().gmms_seg_plan_segment(self._h, i, ctypes.byref(k), ctypes.byref(r)), "gmms_seg_plan_segment")
        return k.value, r.value

    def sort(self, keys, keys_alt, values=None, values_alt=None, *, workspace: Optional[Workspace] = None,
             stream=None) -> int:
        """Sort every segment; returns 0 if the results are in keys/values,
        1 if in keys_alt/values_alt. Asynchronous; keep the plan alive until
        the stream has run it."""
        _check_dev(keys, keys_alt, values, values_alt)
        ws = _workspace_for(workspace, keys.device, stream)
        wp, wb = ws.ensure(self.info.workspace_bytes, keys.device, stream)
        alt = ctypes.c_int(0)
        st = lib().gmms_seg_sort(self._h, _ptr(keys), _ptr(keys_alt), _ptr(values), _ptr(values_alt), wp, wb,
                                 _stream_ptr(stream), ctypes.byref(alt))
        _check(st, "gmms_seg_sort")
        # the plan's pinned tables feed an asynchronous copy: hold the buffers too
        return alt.value

    def graph(self, keys, keys_alt, values=None, values_alt=None) -> "SegGraph":
        """Capture the batch's launches into a HIP graph bound to these buffers."""
        return SegGraph(self, keys, keys_alt, values, values_alt)

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().gmms_seg_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SegGraph:
    """gmms_seg_graph: the batch's launches as one HIP graph; launch() re-sorts
    whatever the bound buffers hold."""

    def __init__(self, plan: SegPlan, keys, keys_alt, values=None, values_alt=None):
        torch = _torch()
        _check_dev(keys, keys_alt, values, values_alt)
        self.plan = plan
        self.ws = torch.empty(max(plan.info.workspace_bytes, 1), dtype=torch.uint8, device=keys.device)
        self.buffers = (keys, keys_alt, values, values_alt)
        h = ctypes.c_void_p()
        st = lib().gmms_seg_graph_create(plan._h, _ptr(keys), _ptr(keys_alt), _ptr(values), _ptr(values_alt),
                                         self.ws.data_ptr(), self.ws.numel(), ctypes.byref(h))
        _check(st, "gmms_seg_graph_create")
        self._h = h

    def launch(self, stream=None) -> int:
        alt = ctypes.c_int(0)
        _check(lib().gmms_seg_graph_launch(self._h, _stream_ptr(stream), ctypes.byref(alt)), "gmms_seg_graph_launch")
        return alt.value

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().gmms_seg_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def segmented_sort(keys, offsets, values=None, *, params: Optional[Params] = None, stream=None):
    """Sort each segment keys[offsets[s]:offsets[s+1]] of a device tensor
    (copy; synchronous). Returns the sorted keys, or (keys, values) for pairs."""
    torch = _torch()
    offs = _offsets_array(offsets)
    k = keys.contiguous().clone()
    ka = torch.empty_like(k)
    v = va = None
    dt = dtype_of(k)
    if values is not None:
        v = values.contiguous().clone()
        va = torch.empty_like(v)
        dt = PAIR_U64_U32
    if offs[-1] > k.numel():
        raise GmmsError(ERROR_INVALID_VALUE, "offsets exceed the tensor")
    alt = ctypes.c_int(0)
    st = lib().gmms_segmented_sort(dt, _ptr(k), _ptr(ka), _ptr(v), _ptr(va), offs.ctypes.data, offs.size - 1,
                                   _pp(params), _stream_ptr(stream), ctypes.byref(alt))
    _check(st, "gmms_segmented_sort")
    if values is None:
        return ka if alt.value else k
    return (ka, va) if alt.value else (k, v)


def seg_tuning_get(dtype: int):
    """(K per size class, merger size per size class) of the batched planner."""
    kk = (ctypes.c_int32 * SEG_CLASSES)()
    mm = (ctypes.c_int64 * SEG_CLASSES)()
    _check(lib().gmms_seg_tuning_get(dtype, ctypes.addressof(kk), ctypes.addressof(mm), SEG_CLASSES),
           "gmms_seg_tuning_get")
    return list(kk), list(mm)


def seg_tuning_set(dtype: int, kway, merger_elems) -> None:
    kk = (ctypes.c_int32 * SEG_CLASSES)(*list(kway)[:SEG_CLASSES])
    mm = (ctypes.c_int64 * SEG_CLASSES)(*list(merger_elems)[:SEG_CLASSES])
    _check(lib().gmms_seg_tuning_set(dtype, ctypes.addressof(kk), ctypes.addressof(mm), SEG_CLASSES),
           "gmms_seg_tuning_set")


def seg_autotune(dtype: int, min_class: int, max_class: int, batch_elems: int = 1 << 24, reps: int = 2,
                 stream=None):
    """Run the batched planner's autotuner and install its choices; returns
    {class: (K, merger_elems, ms_per_batch)}."""
    kk = (ctypes.c_int32 * SEG_CLASSES)()
    mm = (ctypes.c_int64 * SEG_CLASSES)()
    tt = (ctypes.c_float * SEG_CLASSES)()
    st = lib().gmms_seg_autotune(dtype, min_class, max_class, batch_elems, reps, _stream_ptr(stream),
                                 ctypes.addressof(kk), ctypes.addressof(mm), ctypes.addressof(tt))
    _check(st, "gmms_seg_autotune")
    return {c: (kk[c], mm[c], tt[c]) for c in range(min_class, max_class + 1)}


# ------------------------------------------------------------- sort server
def server_request_bytes(dtype: int, n: int) -> int:
    b = ctypes.c_size_t(0)
    _check(lib().gmms_server_request_bytes(dtype, n, ctypes.byref(b)), "gmms_server_request_bytes")
    return b.value


def server_batch_policy(dtypes, sizes, free_bytes: int, max_batch_elems: int = 1 << 28,
                        max_batch_requests: int = 4096):
    """The server's batching policy alone (host logic): indices of the next
    batch among the pending requests (arrival order)."""
    import numpy as np

    d = np.ascontiguousarray(np.asarray(dtypes, np.int32))
    n = np.ascontiguousarray(np.asarray(sizes, np.int64))
    out = np.zeros(max(1, d.size), np.int64)
    cnt = ctypes.c_int64(0)
    st = lib().gmms_server_batch_policy(d.ctypes.data, n.ctypes.data, d.size, free_bytes, max_batch_elems,
                                        max_batch_requests, out.ctypes.data, ctypes.byref(cnt))
    _check(st, "gmms_server_batch_policy")
    return [int(i) for i in out[:cnt.value]]


class Ticket:
    """A submitted request: its id and the host buffers it reads and fills."""

    def __init__(self, tid, keys, values, out_keys, out_values):
        self.id = tid
        self.inputs = (keys, values)
        self.out_keys = out_keys
        self.out_values = out_values
        self.status = None
        self.latency_ms = None
        self.batch = None


class Server:
    """Sort server (gmms_server): requests are host tensors (pinned memory
    lets copies overlap), results come back in pinned host tensors."""

    def __init__(self, hbm_budget: int = 0, max_batch_elems: int = 0, max_batch_requests: int = 0,
                 max_inflight: int = 0):
        cfg = ServerConfig()
        _check(lib().gmms_server_default_config(ctypes.byref(cfg)), "gmms_server_default_config")
        if hbm_budget:
            cfg.hbm_budget = hbm_budget
        if max_batch_elems:
            cfg.max_batch_elems = max_batch_elems
        if max_batch_requests:
            cfg.max_batch_requests = max_batch_requests
        if max_inflight:
            cfg.max_inflight = max_inflight
        h = ctypes.c_void_p()
        _check(lib().gmms_server_create(ctypes.byref(cfg), ctypes.byref(h)), "gmms_server_create")
        self._h = h
        self.config = cfg
        # every request not yet reported done, by id: its Ticket holds the
        # buffers the server reads and writes, so they stay alive until the
        # request has finished even when the caller drops the ticket
        self._live = {}

    def submit(self, keys, values=None, *, dtype: Optional[int] = None, out_keys=None,
               out_values=None) -> Ticket:
        """Queue a sort of tensor `keys` (and `values` for pairs): host
        tensors (results in pinned host tensors) or tensors on the server's
        device (results on that device). `out_keys` / `out_values` receive
        the results when given (same shape, dtype and memory as the inputs,
        contiguous, pinned when on the host); otherwise they are allocated
        here."""
        torch = _torch()
        if values is not None and values.is_cuda != keys.is_cuda:
            raise GmmsError(ERROR_INVALID_VALUE, "keys and values must live in the same memory")
        keys = keys.contiguous()
        dt = dtype if dtype is not None else (PAIR_U64_U32 if values is not None else dtype_of(keys))

        def out_like(x, given):
            if given is not None:
                if (given.shape != x.shape or given.dtype != x.dtype or given.is_cuda != x.is_cuda
                        or not given.is_contiguous() or (not given.is_cuda and not given.is_pinned())):
                    raise GmmsError(ERROR_INVALID_VALUE, "output tensors must match their inputs "
                                    "(host outputs in pinned memory)")
                return given
            if x.is_cuda:
                return torch.empty_like(x)
            return torch.empty(x.shape, dtype=x.dtype, pin_memory=True)

        out_k = out_like(keys, out_keys)
        out_v = None
        if values is not None:
            values = values.contiguous()
            out_v = out_like(values, out_values)
        # device buffers: the server's copies wait for the work queued so far
        # on torch's current stream (which may still be writing the inputs,
        # including the copy contiguous() made, or using the result blocks)
        on_dev = keys.is_cuda or out_k.is_cuda
        stream = torch.cuda.current_stream().cuda_stream if on_dev else None
        tid = ctypes.c_uint64(0)
        st = lib().gmms_server_submit_after(self._h, dt, keys.data_ptr(), _ptr(values), keys.numel(),
                                            out_k.data_ptr(), _ptr(out_v), stream, int(on_dev), ctypes.byref(tid))
        _check(st, "gmms_server_submit_after")
        t = Ticket(tid.value, keys, values, out_k, out_v)
        self._live[t.id] = t
        return t

    def wait(self, t: Ticket, timeout: Optional[float] = None) -> bool:
        """Wait for a request; True once it is done (then t.status,
        t.latency_ms and t.batch are set)."""
        if t.status is not None:
            return True
        done, status = ctypes.c_int(0), ctypes.c_int(0)
        lat, batch = ctypes.c_double(0), ctypes.c_int64(-1)
        st = lib().gmms_server_wait(self._h, t.id, -1.0 if timeout is None else float(timeout), ctypes.byref(done),
                                    ctypes.byref(status), ctypes.byref(lat), ctypes.byref(batch))
        _check(st, "gmms_server_wait")
        if not done.value:
            return False
        t.status, t.latency_ms, t.batch = status.value, lat.value, batch.value
        self._live.pop(t.id, None)
        return True

    def result(self, t: Ticket):
        """Wait, check the status, return the sorted keys (or keys, values)."""
        self.wait(t)
        _check(t.status, "server request")
        return t.out_keys if t.out_values is None else (t.out_keys, t.out_values)

    def drain(self) -> None:
        """Wait for every request; afterwards each ticket has its status and
        the server holds no request buffers."""
        _check(lib().gmms_server_drain(self._h), "gmms_server_drain")
        for t in list(self._live.values()):
            self.wait(t, timeout=0.0)

    def stats(self) -> ServerStats:
        s = ServerStats()
        _check(lib().gmms_server_stats_get(self._h, ctypes.byref(s)), "gmms_server_stats_get")
        return s

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().gmms_server_destroy(self._h)  # finishes every request first
            self._h = None
            self._live.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
