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


def _infos(name):
    out = []
    info = np.zeros(7, dtype=np.int32)
    v = 0
    while getattr(H.lib(), name)(v, H.ptr(info)) == 0:
        out.append((v,) + tuple(int(x) for x in info))
        v += 1
    return out


def variants():
    """[(variant, tile_i, tile_j, threads, prefetch_depth, min_wg_per_cu, iters_per_pass, family)] of
    the FP64 SOR kernels (family 0: per-phase cell lists, 1: pair-owned cells)."""
    return _infos("biocfd_sor_variant_info")


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
    for v, ti, tj, nt, depth, minb, nit, fam in variants():
        for p in (planes if v > 0 else (32,)):
            ms = time_variant(block, v, p, iters, dt)
            table.append({"variant": v, "family": fam, "tile": [ti, tj], "threads": nt, "depth": depth,
                          "min_wg_per_cu": minb, "iters_per_pass": nit, "planes": p,
                          "us_per_iteration": round(1e3 * ms, 2)})
    best = min(table, key=lambda r: r["us_per_iteration"])
    if install:
        H.call("biocfd_sor_config", best["variant"], best["planes"])
    return best["variant"], best["planes"], table


def autotune_block(block: int, dt: float = 1.0, iters: int = 40, planes=DEFAULT_PLANES, install: bool = True):
    """The FP64 SOR kernel of one block's own calls (biocfd_sor_block_config):
    every two-iteration pair-owned variant and planes count timed on the
    block's live pc and b; return (variant, planes, table)."""
    table = []
    for v, ti, tj, nt, depth, minb, nit, fam in variants():
        if fam not in (1, 2) or nit != 2:
            continue
        for p in planes:
            ms = time_variant(block, v, p, iters, dt)
            table.append({"variant": v, "tile": [ti, tj], "threads": nt, "planes": p,
                          "us_per_iteration": round(1e3 * ms, 2)})
    best = min(table, key=lambda r: r["us_per_iteration"])
    if install:
        H.call("biocfd_sor_block_config", block, best["variant"], best["planes"])
    return best["variant"], best["planes"], table


def mixed_variants():
    """[(variant, tile_i, tile_j, threads, prefetch_depth, min_wg_per_cu, sweeps_per_pass, family)] of
    the FP32 smoothing kernels."""
    return _infos("biocfd_mixed_variant_info")


def autotune_mixed(block: int = 1, dt: float = 1.0, iters: int = 40, planes=DEFAULT_PLANES, sweeps: int = 4,
                   residual_tol: float = 1e-8, install: bool = True):
    """Time every FP32 smoothing candidate on `block` (the residual of its
    current pc as rhs; pc is not changed); return (variant, planes, table)."""
    table = []
    ms = np.zeros(1)
    for v, ti, tj, nt, depth, minb, nit, fam in mixed_variants():
        for p in planes:
            best = float("inf")
            for _ in range(2):
                H.call("biocfd_mixed_time", block, v, p, iters, dt, H.ptr(ms))
                best = min(best, float(ms[0]) / iters)
            table.append({"variant": v, "family": fam, "tile": [ti, tj], "threads": nt, "depth": depth,
                          "min_wg_per_cu": minb, "sweeps_per_pass": nit, "planes": p,
                          "us_per_sweep": round(1e3 * best, 2)})
    best = min(table, key=lambda r: r["us_per_sweep"])
    if install:
        H.call("biocfd_mixed_config", best["variant"], best["planes"], sweeps, residual_tol)
    return best["variant"], best["planes"], table
