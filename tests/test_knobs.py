"""The params.h knobs of the drop-in header (the reference's src/params.h:18-38)
reach the library: tests/knobs_probe.cpp is compiled with knob macros on the
command line, as a reference caller would pass them, and prints the
parameters gmms_resolve_params makes of them. Bad knob combinations must not
compile. CPU only: the probe resolves parameters without a GPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "gpu-mms_amd", "gmms")
PROBE = os.path.join(ROOT, "tests", "knobs_probe.cpp")
TESTMM = os.path.join(ROOT, "gpu-mms_amd", "cli", "testMM.cpp")
CXX = shutil.which("g++")


def _build(tmp_path, defs):
    exe = tmp_path / "probe"
    cmd = [CXX, "-std=c++17", f"-I{INC}", *defs, PROBE, "-o", str(exe), f"-L{LIBDIR}", "-lgmms",
           f"-Wl,-rpath,{LIBDIR}", "-Wl,-rpath-link,/opt/rocm/lib"]
    return subprocess.run(cmd, capture_output=True, text=True), exe


def _run(tmp_path, defs):
    r, exe = _build(tmp_path, defs)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    return dict(kv.split("=") for kv in out.stdout.split())


@pytest.mark.skipif(CXX is None, reason="g++ missing")
@pytest.mark.parametrize("defs,want", [
    ([], {"K": "16", "PL": "4", "pipeline": "0", "blocks": "0"}),
    (["-DK=64", "-DMERGE_KERNEL=1"], {"K": "64", "PL": "6", "pipeline": "1"}),
    (["-DPIPELINE"], {"K": "16", "PL": "4", "pipeline": "1"}),
    (["-DNO_PIPELINE", "-DK=32"], {"K": "32", "PL": "5", "pipeline": "0"}),
    (["-DPIPELINE", "-DMERGE_KERNEL=0"], {"pipeline": "0"}),  # explicit MERGE_KERNEL wins
    (["-DK=8", "-DPL=3", "-DBLOCKS=482"], {"K": "8", "PL": "3", "blocks": "482"}),
    (["-DMERGE_KERNEL=2", "-DK=64"], {"K": "64", "pipeline": "2"}),
])
def test_knobs_reach_the_library(tmp_path, defs, want):
    got = _run(tmp_path, defs)
    for k, v in want.items():
        assert got[k] == v, (defs, got)


@pytest.mark.skipif(CXX is None, reason="g++ missing")
@pytest.mark.parametrize("defs", [["-DK=32", "-DPL=4"], ["-DK=48"], ["-DK=128"], ["-DPIPELINE", "-DNO_PIPELINE"],
                                  ["-DMERGE_KERNEL=3"]])
def test_bad_knobs_do_not_compile(tmp_path, defs):
    r, _ = _build(tmp_path, defs)
    assert r.returncode != 0, defs


@pytest.mark.skipif(CXX is None, reason="g++ missing")
def test_testmm_compiles_with_reference_knobs():
    """The harness CLI builds with the reference's own knob values
    (K 64, PL 6, BLOCKS 482, PIPELINE) on the command line."""
    for defs in (["-DK=64", "-DPL=6", "-DBLOCKS=482", "-DPIPELINE"], ["-DK=64", "-DMERGE_KERNEL=1"]):
        r = subprocess.run([CXX, "-std=c++17", "-fsyntax-only", f"-I{INC}", "-I/opt/rocm/include",
                            "-D__HIP_PLATFORM_AMD__", *defs, TESTMM], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


@pytest.mark.skipif(CXX is None, reason="g++ missing")
def test_custom_comparator_without_custom_header_is_refused(tmp_path):
    """A caller order through multimergesort.hxx alone (host compiler, library
    path) fails to compile with a message naming multimergesort_custom.hxx."""
    src = tmp_path / "c.cpp"
    src.write_text('#include "gpu_mms/multimergesort.hxx"\n'
                   "int my_cmp(int a, int b) { return (a & 0xFF) < (b & 0xFF); }\n"
                   "int* f(int* a, int* b, int n) { return multimergesort<int, my_cmp>(a, b, nullptr, n); }\n")
    r = subprocess.run([CXX, "-std=c++17", "-fsyntax-only", f"-I{INC}", str(src)], capture_output=True, text=True)
    assert r.returncode != 0 and "multimergesort_custom.hxx" in r.stderr
