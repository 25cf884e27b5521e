#!/usr/bin/env python3
"""Writes tests/golden/vectors.npz: seeded inputs from the oracle's
generators (mt19937_64, fixed bit operations) and their expected sorted
outputs (std::sort with the comparator contract of src/cmp.hxx:21-73 --
integers by '<', floats by IEEE total order, pairs by (key, value)), all
stored as raw bits. tests/test_golden.py pins the vectors independently
(numpy sorts, permutation and order checks) and compares the oracle and the
GPU sort against them.

Run from the repository root after building oracle/ (make -C oracle):
    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from conftest import Oracle  # noqa: E402


def main():
    o = Oracle()
    v = {}
    meta = {}

    def put(name, arr, how):
        v[name] = arr
        meta[name] = how

    a = o.gen_i32(0x6D6D7301, 4099, 32)
    put("i32_full_in", a, "gen_i32(0x6D6D7301, 4099, shift 32): full int32 range")
    put("i32_full_out", o.sort(a), "std::sort")
    a = o.gen_i32(0x6D6D7300, 5000, 34)
    put("i32_cfg1_in", a, "gen_i32(0x6D6D7300, 5000, shift 34): [0, 2^30) like the 2^20 configuration")
    put("i32_cfg1_out", o.sort(a), "std::sort")
    a = (o.gen_u64(0x6D6D7302, 6000) >> np.uint64(54)).astype(np.int32)
    put("i32_dup_in", a, "(int32)(gen_u64(0x6D6D7302, 6000) >> 54): [0, 1024), duplicate-heavy")
    put("i32_dup_out", o.sort(a), "std::sort")
    a = o.glibc_rand_list(0, 2048, 1024)
    put("i32_glibc_in", a, "glibc srand(0); rand() % 1024, 2048 values (src/buildData.h:54-65)")
    put("i32_glibc_out", o.sort(a), "std::sort")
    x = o.gen_f32_normal(0x6D6D7304, 3000)
    bits = x.view(np.uint32)
    specials = np.array([0x80000000, 0x00000000, 0x7F800000, 0xFF800000, 0x7FC00000, 0xFFC00000,
                         0x00000001, 0x80000001, 0x7FFFFFFF, 0xFFFFFFFF], np.uint32)
    bits[:200] = np.resize(specials, 200)
    put("f32_in_bits", bits.copy(), "gen_f32_normal(0x6D6D7304, 3000), first 200 = signed zeros, infinities, NaNs, denormals")
    put("f32_out_bits", o.sort(x).view(np.uint32), "std::sort, IEEE total order (-NaN < -inf < ... < -0 < +0 < ... < +NaN)")
    u = o.gen_u64(0x6D6D7306, 3000)
    put("u64_in", u, "gen_u64(0x6D6D7306, 3000)")
    put("u64_out", o.sort(u), "std::sort")
    f = o.gen_u64(0x6D6D7307, 3000).view(np.float64).copy()
    f[::29] = -0.0
    f[1::29] = 0.0
    put("f64_in_bits", f.view(np.uint64).copy(), "gen_u64(0x6D6D7307, 3000) as float64 bits, every 29th -0.0 / +0.0")
    put("f64_out_bits", o.sort(f).view(np.uint64), "std::sort, IEEE total order")
    for d in range(4):
        k, val = o.gen_pairs(0x6D6D7310 + d, 2500, d, 1.1, 1 << 20)
        wk, wv = o.sort_pairs(k, val)
        put(f"pairs{d}_k_in", k, f"gen_pairs(seed {0x6D6D7310 + d:#x}, 2500, dist {d}): "
            "0 uniform, 1 sorted, 2 reverse, 3 Zipf(1.1) over 2^20 values; value = index")
        put(f"pairs{d}_v_in", val, "value = original index")
        put(f"pairs{d}_k_out", wk, "std::sort by (key, value)")
        put(f"pairs{d}_v_out", wv, "std::sort by (key, value)")
    lens = o.gen_seglens(0x6D6D7303, 8192, 10, 22)
    put("seglens_head", lens[:64].copy(), "first 64 of gen_seglens(0x6D6D7303, 8192, 10, 22): floor(2^(10+12U))")
    put("seglens_sum", np.array([lens.sum()], np.int64), "sum of all 8192 lengths")
    slens = np.array([0, 1, 1023, 1024, 1025, 3000, 7, 2048], np.int64)
    offs = np.concatenate([[0], np.cumsum(slens)]).astype(np.int64)
    a = o.gen_i32(0x6D6D7308, int(offs[-1]), 32)
    out = a.copy()
    for s in range(len(slens)):
        out[offs[s]:offs[s + 1]] = o.sort(a[offs[s]:offs[s + 1]])
    put("seg_offsets", offs, "segment offsets (lengths 0, 1, 1023, 1024, 1025, 3000, 7, 2048)")
    put("seg_in", a, "gen_i32(0x6D6D7308, total, shift 32)")
    put("seg_out", out, "std::sort of every segment")
    np.savez_compressed(os.path.join(HERE, "vectors.npz"), **v)
    with open(os.path.join(HERE, "vectors.json"), "w") as fh:
        json.dump(meta, fh, indent=1, sort_keys=True)
    print("wrote", len(v), "arrays")


if __name__ == "__main__":
    main()
