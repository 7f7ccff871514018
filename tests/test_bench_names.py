"""CPU: the benchmark names the sliced SpMM kernels as the library launches
them, so its roofline finds the committed counter summaries.

``bench.spmm_kernel_name`` restates the library's launch rules (32-bit packed
words on 1024-thread workgroups for complex64 products of at least 8192 such
workgroups, wide palettes in LDS per 256 or 1024 threads or read through the
caches); ``bench.element_bytes`` the bytes per entry of each layout.  The
headline and Hubbard 3x4 names must be keys of the counter summaries that
``roofline.traffic`` reads.
"""

from __future__ import annotations

import json
from pathlib import Path

import bench

ROOT = Path(__file__).resolve().parent.parent


def test_element_bytes_per_layout():
    assert bench.element_bytes({"packed": 0}, 8) == 8
    assert bench.element_bytes({"packed": 1}, 16) == 0
    assert bench.element_bytes({"packed": 2}, 16) == 2


def test_workgroup_rule_names():
    packed = {"packed": 1, "n_palette": 56}
    # 48-site sector: 15.7M rows, 4 lanes per row -> 983k slices -> 61k workgroups of 16 slices
    assert bench.spmm_kernel_name(packed, "float2", 4, 4, 15_724_214) == "spmm_sell<float2,4,4,true,1024>"
    assert bench.spmm_kernel_name(packed, "double2", 10, 4, 15_724_214) == "spmm_sell<double2,10,4,true,256>"
    # Hubbard 3x4 sector: 225k rows -> 880 workgroups of 16 slices: 256 threads
    assert bench.spmm_kernel_name(packed, "float2", 6, 4, 225_440) == "spmm_sell<float2,6,4,true,256>"
    assert bench.spmm_kernel_name({"packed": 0, "n_palette": 0}, "float2", 4, 4, 15_724_214) == \
        "spmm_sell<float2,4,4,false,256>"
    wide = {"packed": 2, "n_palette": 1095}
    assert bench.spmm_kernel_name(wide, "float2", 2, 4, 5_230_170) == "spmm_sell_wide<float2,2,4,true,1024>"
    assert bench.spmm_kernel_name(wide, "double2", 5, 4, 5_230_170) == "spmm_sell_wide<double2,5,4,true,256>"
    big = {"packed": 2, "n_palette": 3000}  # complex128 palette of 48 KB
    assert bench.spmm_kernel_name(big, "double2", 5, 4, 5_230_170) == "spmm_sell_wide<double2,5,4,true,1024>"
    huge = {"packed": 2, "n_palette": 9000}  # beyond 64 KB in either precision
    assert bench.spmm_kernel_name(huge, "float2", 2, 4, 5_230_170) == "spmm_sell_wide<float2,2,4,false,256>"


def test_committed_summaries_hold_the_bench_names():
    cb = json.loads((ROOT / "profiles" / "solver_checkerboard48_traffic.json").read_text())["kernels"]
    assert bench.spmm_kernel_name({"packed": 1, "n_palette": 56}, "float2", 4, 4, 15_724_214) in cb
    hub = json.loads((ROOT / "profiles" / "spmm_hubbard3x4_traffic.json").read_text())["kernels"]
    for v in ("float2", "double2"):
        for lanes in (1, 2, 4):
            assert bench.spmm_kernel_name({"packed": 1, "n_palette": 53}, v, 6, lanes, 225_440) in hub
