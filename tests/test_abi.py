"""The C ABI of libgmms.so, without a GPU: the library loads, exports every
symbol include/gpu_mms/gmms.h declares, and rejects bad arguments before
touching the device."""
import ctypes
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_exports_every_declared_symbol(gmms):
    declared = gmms.declared_symbols()
    assert len(declared) >= 12
    out = subprocess.run(["nm", "-D", "--defined-only", gmms.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [s for s in declared if s not in exported]
    assert not missing, missing
    for s in declared:
        assert hasattr(gmms.lib(), s)


def test_version_and_status_strings(gmms):
    assert "gfx950" in gmms.version()
    assert gmms.status_string(0) == "success"
    assert gmms.status_string(3).startswith("workspace")
    assert gmms.status_string(99) == "unknown status"


def test_default_params(gmms):
    p = gmms.default_params()
    assert (p.kway, p.pipeline, p.mergers) == (0, 0, 0)  # K auto, wavefront merger, auto mergers
    assert 1 <= p.prefetch <= 8


def test_element_sizes(gmms):
    L = gmms.lib()
    kb, vb = ctypes.c_size_t(), ctypes.c_size_t()
    want = {gmms.INT32: (4, 0), gmms.FLOAT32: (4, 0), gmms.UINT64: (8, 0), gmms.FLOAT64: (8, 0),
            gmms.PAIR_U64_U32: (8, 4)}
    for dt, (k, v) in want.items():
        assert L.gmms_element_size(dt, ctypes.byref(kb), ctypes.byref(vb)) == 0
        assert (kb.value, vb.value) == (k, v)
    assert L.gmms_element_size(42, ctypes.byref(kb), ctypes.byref(vb)) == gmms.ERROR_INVALID_VALUE


def test_base_run_and_sort_rounds(gmms):
    assert gmms.default_params().base_run == 0
    assert gmms.resolve_params(gmms.INT32).base_run == 4096
    assert gmms.resolve_params(gmms.INT32, gmms.make_params(base_run=1024)).base_run == 1024
    with pytest.raises(gmms.GmmsError):
        gmms.resolve_params(gmms.INT32, gmms.make_params(base_run=2048))
    for n in (0, 1, 4096, 4097, 1 << 20, 1 << 28, 10**8 + 17):
        for K in (4, 16, 64):
            for run in (1024, 4096):
                r, L = 0, run
                while L < n:
                    r, L = r + 1, L * K
                assert gmms.sort_rounds(gmms.INT32, n, gmms.make_params(K=K, base_run=run)) == r
    # 2^28 with K = 16: the 4096-element runs save the fifth round
    assert gmms.num_rounds(1 << 28, 16) == 5
    assert gmms.sort_rounds(gmms.INT32, 1 << 28, gmms.make_params(K=16)) == 4


def test_descending_knob(gmms):
    assert gmms.default_params().descending == 0
    assert gmms.resolve_params(gmms.INT32, gmms.make_params(descending=True)).descending == 1
    p = gmms.make_params()
    p.descending = 2
    with pytest.raises(gmms.GmmsError):
        gmms.resolve_params(gmms.INT32, p)


def test_num_rounds_matches_reference_loop(gmms, oracle):
    for n in (0, 1, 1024, 1025, 1 << 20, 10**8 + 17, 1 << 28, 1 << 33):
        for K in (2, 8, 64):
            assert gmms.num_rounds(n, K) == oracle.lib.oracle_rounds(n, K, 1024)


def test_workspace_size(gmms):
    p = gmms.make_params(K=64, mergers=512)
    assert gmms.workspace_size(gmms.INT32, 1024, p) == 0  # no merge round, no cuts
    big = gmms.workspace_size(gmms.INT32, 1 << 28, p)
    assert 0 < big < 64 << 20


def test_argument_validation_without_gpu(gmms):
    L = gmms.lib()
    alt = ctypes.c_int()
    assert L.gmms_sort(99, None, None, None, None, 10, None, None, 0, None, ctypes.byref(alt)) == 1
    assert L.gmms_sort(0, None, None, None, None, -1, None, None, 0, None, ctypes.byref(alt)) == 1
    assert L.gmms_sort(0, None, None, None, None, 0, None, None, 0, None, ctypes.byref(alt)) == 0
    assert L.gmms_sort(0, None, None, None, None, 10, None, None, 0, None, ctypes.byref(alt)) == 1
    # pairs need both value buffers; key-only dtypes must not get them
    fake = ctypes.c_void_p(0x1000)
    assert L.gmms_sort(gmms.PAIR_U64_U32, fake, fake, None, None, 10, None, None, 0, None, None) == 1
    assert L.gmms_sort(gmms.INT32, fake, fake, fake, fake, 10, None, None, 0, None, None) == 1
    bad = gmms.make_params()
    bad.kway = 48
    assert L.gmms_sort(0, fake, fake, None, None, 10, ctypes.byref(bad), None, 0, None, None) == 1
    bad.kway = 128
    assert L.gmms_sort(0, fake, fake, None, None, 10, ctypes.byref(bad), None, 0, None, None) == 1
    # a sort that needs partitions but gets no workspace is refused up front
    p = gmms.make_params(K=64, mergers=512)
    assert L.gmms_sort(0, fake, fake, None, None, 1 << 28, ctypes.byref(p), None, 0, None, None) == 3
    res = ctypes.c_void_p()
    assert L.gmms_multimergesort(0, None, None, None, 0, -5, ctypes.byref(res)) == 1
    assert L.gmms_squaresort(9, None, None, 10, 1, 1, None) == 1
    assert L.gmms_pad(0, None, None, -1, None) == 1
    m = ctypes.c_int()
    assert L.gmms_find_partitions(0, None, None, 10, 0, None, None, None, 0, ctypes.byref(m), None) == 1


def test_header_has_reference_citations():
    text = open(os.path.join(ROOT, "include", "gpu_mms", "gmms.h")).read()
    for cite in ("src/multimergesort.hxx:56", "src/basecase/squareSort.hxx:62", "src/warp-partition.hxx:487",
                 "src/pad.hxx", "src/params.h", "src/cmp.hxx"):
        assert cite in text
