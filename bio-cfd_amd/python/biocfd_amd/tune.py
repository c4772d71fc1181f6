"""Autotuner of the SOR kernel variant and planes per workgroup.

Every variant computes the same bits (the fused tile schedule is checked
against the in-place sweep, and every variant against the oracle), so the
choice is purely a speed measurement: each candidate runs a fixed number of
iterations on the block's live pc and b through `biocfd_sor_time`, which
restores pc afterwards, and the fastest is installed with `biocfd_sor_config`.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import hip as H

DEFAULT_PLANES = (8, 16, 32, 64)


def variants():
    """[(variant, tile_i, tile_j, threads, prefetch_depth)] of the compiled SOR kernels."""
    out = []
    v = 0
    a, b, t, d = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    while H.lib().biocfd_sor_variant_shape(v, C.byref(a), C.byref(b), C.byref(t), C.byref(d)) == 0:
        out.append((v, a.value, b.value, t.value, d.value))
        v += 1
    return out


def time_variant(block: int, variant: int, planes: int, iters: int, dt: float, repeats: int = 2) -> float:
    """Best of `repeats` runs, milliseconds per iteration."""
    best = float("inf")
    ms = np.zeros(1)
    for _ in range(repeats):
        H.call("biocfd_sor_time", block, variant, planes, iters, dt, H.ptr(ms))
        best = min(best, float(ms[0]) / iters)
    return best


def autotune(block: int = 1, dt: float = 1.0, iters: int = 40, planes=DEFAULT_PLANES, install: bool = True):
    """Time every (variant, planes) candidate on `block`; return
    (best_variant, best_planes, table) with table rows
    {"variant", "tile", "threads", "planes", "us_per_iteration"}."""
    table = []
    for v, ti, tj, nt, depth in variants():
        for p in (planes if v > 0 else (32,)):
            ms = time_variant(block, v, p, iters, dt)
            table.append({"variant": v, "tile": [ti, tj], "threads": nt, "depth": depth, "planes": p,
                          "us_per_iteration": round(1e3 * ms, 2)})
    best = min(table, key=lambda r: r["us_per_iteration"])
    if install:
        H.call("biocfd_sor_config", best["variant"], best["planes"])
    return best["variant"], best["planes"], table
