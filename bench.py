"""Throughput of the Bio-CFD time step on one MI355X, FP64 method and mixed
precision side by side, both solved to the same FP64 Poisson residual.

Workload: single block 256x256x256 on [0,8]x[0,4]x[0,4] with a rigid sphere
of diameter 1 at (2,2,2) carried by 8192 surface triangles, Re = 300,
dt = 0.0005, inflow u = 1 plus seeded 1e-3 noise, SOR omega = 1.8.  One
"step" is one full time step of the hot path (momentum + BCs + IB forcing,
pressure-correction projection, IB retagging / list rebuild / forcing), with
all fields resident in HBM.

    python bench.py [--steps K] [--warmup W] [--precision both|fp64|mixed]
    python bench.py --workload uneven [--steps K] [--warmup W]

Prints one JSON line.  `value` is the FP64 method's Mcell-steps/s (one GPU)
with every pressure solve run to max|b/dt - A pc| < 1e-8 over the fluid cells
(the FP64 residual, tested every 256 iterations; pcItaMax raised so that the
residual, not the budget, ends every solve).  `mixed` holds the mixed-precision
run of the same steps on the same inputs to the same residual (FP32 SOR
smoothing, 256 sweeps per FP64 defect correction, the residual tested after
every correction), its speed-up and its per-field error against the FP64 run.
`mixed_fp32_momentum` is the same mixed run with the dense momentum pass's
tendency evaluated in FP32 (the precision recipe's FP32 momentum stencil),
with its speed-up and per-field error against the FP64 run.
`reference_rule` is the FP64 method under the reference's own stopping rule
(update norm < 1e-8 tested every fifth iteration, pcItaMax = 500, which this
workload reaches: 501 iterations per step).  `roofline` describes the FP64
red-black SOR (the dominant kernel) of the residual run, measured with HIP
events on the runtime's stream; `cpu_baseline` times the reference's own CPU
build (oracle/_ref/ref_harness) on one step of the same workload under the
reference's rule, on one core.

`uneven` is the batched multi-block workload (a coarse block 192x128x128 and
16 fine blocks of uneven sizes, one with a heaving, pitching plate;
cases.config_uneven) under the reference's rule, FP64 with the fine blocks'
SOR calls batched: throughput, nIterPcor per block, and the roofline of its
SOR calls (algorithmic bytes over their GPU time).

--workload uneven: that workload alone, FP64 and mixed precision (equal sweep
budget), each run twice on the same inputs: the fine blocks' SOR calls
batched in one launch per pass, then one block after the other; one JSON
line with the throughputs, the speed-ups and whether the runs agree bit for
bit.  With --sor-variant auto (the default) the FP64 SOR variant is picked by
timing one batched step of the workload per candidate.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bio-cfd_amd", "python"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
ORC_RUN = os.path.join(ROOT, "oracle", "orc_run")

# algorithmic HBM bytes per interior cell of one red-black SOR iteration (FP64):
# read pc once, read b/dt once, write pc once
SOR_BYTES_PER_CELL_ITER = 24.0
# HBM bytes per fused FP64 SOR launch measured with rocprofv3 (FETCH_SIZE x 2.00
# + WRITE_SIZE x 1.00, calibrated on a 512 MiB copy); quoted only for the
# configuration measured
SOR_TRAFFIC_MEASURED = {
    # (n, family, tile_i, tile_j, threads, depth, iterations per pass, planes): bytes per launch
    # of the kernel with that many iterations per pass (profiles/sor_sphere256_hbm_traffic.md)
    (256, "cell lists", 64, 32, 512, 1, 2, 32): 5.7318e+08,
    (256, "cell lists", 64, 32, 512, 1, 1, 32): 4.2695e+08,
    (256, "pair-owned, LDS coefficients", 64, 32, 512, 1, 2, 32): 4.6074e+08,
    (256, "pair-owned, LDS coefficients", 64, 32, 512, 1, 1, 32): 4.2833e+08,
    (256, "pair-owned", 32, 32, 512, 1, 2, 64): 4.3925e+08,
    (256, "pair-owned", 32, 32, 512, 1, 1, 64): 4.2003e+08,
}


def mixed_bytes_per_cell(sweeps_per_correction: int) -> float:
    """Algorithmic bytes per cell of one correction cycle of the mixed solve:
    k FP32 sweeps (12 B each, the first reads no iterate: 8 B) and one pass
    that updates pc and takes the residual (pc 8, e 4, b/dt 8 read; pc 8,
    FP32 r 4 written)."""
    k = sweeps_per_correction
    return (12 * k - 4) + 32


def make_case(args, rule: str = "reference"):
    """The 256^3 sphere workload; rule "residual": pcItaMax raised and the
    residual tolerance as epsi, "reference": the reference's pcItaMax/epsi."""
    from biocfd_amd import cases as CS

    n = (args.n, args.n, args.n)
    if rule == "residual":
        return CS.config_sphere_256(n=n, markers=args.markers, nsteps=args.warmup + args.steps,
                                    pc_ita_max=args.resid_ita_max, epsi=args.resid_tol)
    return CS.config_sphere_256(n=n, markers=args.markers, nsteps=args.warmup + args.steps,
                                pc_ita_max=args.pc_ita_max, epsi=args.epsi)


def run_path(args, precision: str, keep_fields: bool, rule: str = "residual", momentum: str = "fp64"):
    """One run of the workload: `precision` fp64 or mixed, `rule` residual
    (FP64 residual < resid_tol) or reference (the reference's update-norm
    rule; FP64 only), `momentum` fp64 or fp32 (the dense momentum pass's
    tendency in FP32)."""
    from biocfd_amd import hip as H
    from biocfd_amd import tune
    from biocfd_amd.driver import Simulation

    case = make_case(args, rule)
    cells = case.cells()
    if args.copy_probe_mb > 0 and precision == "fp64":
        H.call("biocfd_init", 0)
        H.call("biocfd_copy_probe", int(args.copy_probe_mb) << 20, 4)
    sim = Simulation(case, precision=precision)
    H.call("biocfd_momentum_config", args.momentum_variant, 0 if args.momentum_divisions else 1)
    H.call("biocfd_momentum_precision", 1 if momentum == "fp32" else 0)
    H.call("biocfd_sor_graphs", args.sor_graphs)
    if precision == "fp64":
        if rule == "residual":
            H.call("biocfd_sor_stop_rule", 1, args.check_every)
        else:
            H.call("biocfd_sor_stop_rule", 0, 5)
    mixed_tol = args.resid_tol if rule == "residual" else args.epsi
    if precision == "fp64":
        if args.sor_variant == "auto":
            v, planes, table = tune.autotune(block=1, dt=case.dt)
            print(json.dumps({"sor_autotune": table}), file=sys.stderr)
        else:
            v, planes = int(args.sor_variant), args.sor_planes
            H.call("biocfd_sor_config", v, planes)
        shape = [t for t in tune.variants() if t[0] == v][0]
    else:
        if args.mixed_variant == "auto":
            v, planes, table = tune.autotune_mixed(block=1, dt=case.dt, sweeps=args.sweeps, residual_tol=mixed_tol)
            print(json.dumps({"mixed_autotune": table}), file=sys.stderr)
        else:
            v, planes = int(args.mixed_variant), args.mixed_planes
            H.call("biocfd_mixed_config", v, planes, args.sweeps, mixed_tol)
        shape = [t for t in tune.mixed_variants() if t[0] == v][0]
    kernel = {"variant": v, "family": ["cell lists", "pair-owned", "pair-owned, LDS coefficients"][shape[7]], "tile": [shape[1], shape[2]],
              "threads": shape[3], "depth": shape[4],
              "min_wg_per_cu": shape[5], "iters_per_pass": shape[6], "planes": planes,
              "autotuned": (args.mixed_variant if precision == "mixed" else args.sor_variant) == "auto"}
    for _ in range(args.warmup):
        sim.step()
    H.call("biocfd_synchronize")
    H.call("biocfd_timing_enable", 1)
    iters = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sim.step()
        it = np.zeros(1, dtype=np.int64)
        H.call("biocfd_sor_state", 1, H.ptr(it), H.ptr(np.zeros(1)), H.ptr(np.zeros(1)))
        iters.append(int(it[0]))
    H.call("biocfd_synchronize")
    t1 = time.perf_counter()
    sor_ms = np.zeros(1)
    sor_it = np.zeros(1, dtype=np.int64)
    sor_ci = np.zeros(1, dtype=np.int64)
    H.call("biocfd_timing_read", H.ptr(sor_ms), H.ptr(sor_it), H.ptr(sor_ci))
    H.call("biocfd_timing_enable", 0)
    out = {
        "cells": cells,
        "seconds": t1 - t0,
        "ms_per_step": 1e3 * (t1 - t0) / args.steps,
        "mcell_steps_per_s": cells * args.steps / (t1 - t0) / 1e6,
        "sor_iters_per_step": float(np.mean(iters)) if iters else 0.0,
        "sor_gpu_ms": float(sor_ms[0]),
        "sor_iterations": int(sor_it[0]),
        "sor_cell_iterations": int(sor_ci[0]),
        "kernel": kernel,
        "residual_linf": sim.residual_norm(1),
    }
    if keep_fields:
        out["fields"] = {name: sim.field(1, name) for name in ("u", "v", "w", "p")}
        res = np.zeros(4)
        kv = sim.blocks[0].kin_vec()
        H.call("biocfd_stress_calculation", 1, case.mu_f, case.rho_f, H.ptr(kv), H.ptr(res))
        out["forces"] = res.copy()
    H.call("biocfd_sor_stop_rule", 0, 5)
    H.call("biocfd_momentum_precision", 0)
    sim.close()
    return out


def tune_uneven(args):
    """The FP64 SOR variant and planes per workgroup of the uneven
    workload's batched fine blocks: one batched time step per two-iteration
    pair-owned variant and planes count (after one set-up step and the
    coarse block's own tuning), the fastest wins.  Every variant gives the
    same bits, so the steps only differ in time."""
    from biocfd_amd import cases as CS
    from biocfd_amd import hip as H
    from biocfd_amd import tune
    from biocfd_amd.driver import Simulation

    cands = [(t, p) for t in tune.variants() if t[7] in (1, 2) and t[6] == 2 for p in (16, 32, 64)]
    case = CS.config_uneven(nsteps=1 + len(cands))
    sim = Simulation(case)
    H.call("biocfd_sor_batching", 1)
    sim.step()
    if args.uneven_coarse_tune:
        tune.autotune_block(1, dt=case.dt)   # the coarse block's own kernel, as in the timed runs
    table = []
    for t, p in cands:
        H.call("biocfd_sor_config", t[0], p)
        H.call("biocfd_synchronize")
        t0 = time.perf_counter()
        sim.step()
        H.call("biocfd_synchronize")
        table.append({"variant": t[0], "tile": [t[1], t[2]], "threads": t[3], "planes": p,
                      "ms_per_step": round(1e3 * (time.perf_counter() - t0), 2)})
    sim.close()
    best = min(table, key=lambda r: r["ms_per_step"])
    return best["variant"], best["planes"], table


def run_uneven(args, batching: int, variant, precision: str = "fp64"):
    """Warm-up plus timed steps of the uneven workload (the reference's own
    rule: update norm < 1e-8 every fifth iteration, pcItaMax = 500).  In
    mixed precision every solve gets the same sweep budget (no residual
    tolerance), so the run measures the batching, not a convergence rate.
    The SOR calls are timed with HIP events in one more step after the
    timed ones (the roofline of the step's SOR calls)."""
    from biocfd_amd import cases as CS
    from biocfd_amd import hip as H
    from biocfd_amd.driver import Simulation

    case = CS.config_uneven(nsteps=args.warmup + args.steps + 1)
    sim = Simulation(case, precision=precision)
    H.call("biocfd_sor_batching", batching)
    H.call("biocfd_momentum_config", args.momentum_variant, 0 if args.momentum_divisions else 1)
    H.call("biocfd_sor_graphs", args.sor_graphs)
    H.call("biocfd_interp_batching", args.interp_batching)
    if variant is not None:
        H.call("biocfd_sor_config", variant, args.sor_planes)
    if precision == "mixed":
        H.call("biocfd_mixed_config", 26, args.uneven_mixed_planes, args.uneven_sweeps, 0.0)
    for _ in range(args.warmup):
        sim.step()
    coarse = None
    if args.uneven_coarse_tune and precision == "fp64":
        # the coarse block's own calls: its kernel timed on its live state
        from biocfd_amd import tune
        cv, cp, _ = tune.autotune_block(1, dt=case.dt)
        coarse = {"variant": cv, "planes": cp}
    H.call("biocfd_synchronize")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sim.step()
    H.call("biocfd_synchronize")
    t1 = time.perf_counter()
    # one more step with the SOR calls timed (the timing reads every call's
    # iteration count back, so it is kept out of the timed steps)
    H.call("biocfd_timing_enable", 1)
    t2 = time.perf_counter()
    sim.step()
    H.call("biocfd_synchronize")
    t3 = time.perf_counter()
    sor_ms = np.zeros(1)
    sor_it = np.zeros(1, dtype=np.int64)
    sor_ci = np.zeros(1, dtype=np.int64)
    H.call("biocfd_timing_read", H.ptr(sor_ms), H.ptr(sor_it), H.ptr(sor_ci))
    H.call("biocfd_timing_enable", 0)
    n_iter = []
    for g in range(1, len(case.blocks) + 1):
        it = np.zeros(1, dtype=np.int64)
        H.call("biocfd_sor_state", g, H.ptr(it), H.ptr(np.zeros(1)), H.ptr(np.zeros(1)))
        n_iter.append(int(it[0]))
    fields = [sim.field(g, "p") for g in range(1, len(case.blocks) + 1)]
    H.call("biocfd_sor_batching", 1)
    if precision == "mixed":
        H.call("biocfd_mixed_config", 26, 32, 128, 1e-8)
    sim.close()
    cells = case.cells()
    secs = t1 - t0
    out = {"ms_per_step": round(1e3 * secs / args.steps, 3),
           "mcell_steps_per_s": round(cells * args.steps / secs / 1e6, 3),
           "nIterPcor_last_step": n_iter,
           "coarse_sor_kernel": coarse,
           "timed_sor_step": {"ms": round(1e3 * (t3 - t2), 3), "sor_gpu_ms": round(float(sor_ms[0]), 3),
                              "sor_cell_iterations": int(sor_ci[0])}}
    bpc = SOR_BYTES_PER_CELL_ITER if precision == "fp64" else mixed_bytes_per_cell(args.uneven_sweeps) / args.uneven_sweeps
    if sor_ms[0] > 0:
        achieved = bpc * float(sor_ci[0]) / (float(sor_ms[0]) * 1e-3) / 1e9
        out["sor_roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": round(achieved / HBM_PEAK_GBS, 4),
                               "bytes_per_cell_iteration": round(bpc, 3),
                               "kernel": "every SOR call of the step (coarse block alone, fine blocks batched or "
                                         "one after the other), algorithmic bytes over the GPU time of the calls"}
    return out, fields, case


UNEVEN_TXT = ("uneven17: coarse 192x128x128 on [0,12]x[0,8]x[0,8] + 16 fine blocks (refinement ratio 2, sides {sides}), "
              "heaving/pitching plate with thin-surface cells, Re=200, dt=5e-4, the reference's rule (update norm < "
              "1e-8 every 5th iteration, pcItaMax=500)")


def uneven_summary(args):
    """The secondary line of the default run: the uneven workload, FP64 with
    the fine blocks batched, the SOR kernels autotuned for it."""
    variant, planes, _ = tune_uneven(args)
    a2 = argparse.Namespace(**vars(args))
    a2.sor_planes = planes
    b, _, case = run_uneven(a2, 1, variant)
    b["sor_kernel"] = {"variant": variant, "planes": planes}
    return {"workload": UNEVEN_TXT.format(sides=",".join(str(int(blk.n[0])) for blk in case.blocks[1:])),
            "cells": case.cells(), "steps": args.steps, "warmup": args.warmup, "dtype": "f64", **b}


def uneven_main(args):
    if args.sor_variant == "auto":
        variant, planes, table = tune_uneven(args)
        args.sor_planes = planes
        print(json.dumps({"uneven_autotune": table}), file=sys.stderr)
    else:
        variant = int(args.sor_variant)
    runs = set(args.uneven_runs.split(","))
    if runs != {"fp64-batched", "fp64-per-block", "mixed-batched", "mixed-per-block"}:
        # a subset (profiling): the runs asked for, no comparison
        out = {}
        for r in sorted(runs):
            prec, mode = r.split("-", 1)
            out[r] = run_uneven(args, 1 if mode == "batched" else 0, variant, prec)[0]
        print(json.dumps({"workload": "uneven17", "sor_variant": variant, "sor_planes": args.sor_planes, **out}))
        return
    b, fb, case = run_uneven(args, 1, variant)
    p, fp, _ = run_uneven(args, 0, variant)
    mb, fmb, _ = run_uneven(args, 1, variant, "mixed")
    mp, fmp, _ = run_uneven(args, 0, variant, "mixed")
    sizes = [int(blk.n[0]) * int(blk.n[1]) * int(blk.n[2]) for blk in case.blocks]
    print(json.dumps({
        "metric": "Mcell-steps/s of the uneven multi-block workload, fine blocks batched vs one after the other",
        "value": b["mcell_steps_per_s"], "unit": "Mcell-steps/s", "n_gpus": 1, "higher_is_better": True,
        "workload": UNEVEN_TXT.format(sides=",".join(str(int(blk.n[0])) for blk in case.blocks[1:])),
        "cells": case.cells(), "fine_cells": sum(sizes[1:]), "steps": args.steps, "warmup": args.warmup,
        "sor_variant": variant, "sor_planes": args.sor_planes, "autotuned": args.sor_variant == "auto",
        "fp64": {"batched": b, "per_block": p, "speedup": round(b["mcell_steps_per_s"] / p["mcell_steps_per_s"], 3),
                 "bit_identical": bool(all(np.array_equal(x, y) for x, y in zip(fb, fp)))},
        "mixed_equal_sweep_budget": {
            "sweeps_per_correction": args.uneven_sweeps, "batched": mb, "per_block": mp,
            "speedup": round(mb["mcell_steps_per_s"] / mp["mcell_steps_per_s"], 3),
            "bit_identical": bool(all(np.array_equal(x, y) for x, y in zip(fmb, fmp)))}}))


def adjoint_summary(args, fd: bool = True):
    """The adjoint workload (96^3 heaving, pitching ellipsoid, 40 steps, FP64,
    fixed 501 SOR iterations per solve): J = mean body x-force over the last
    20 steps, dJ/d(heave amplitude, pitch amplitude, phase, frequency) and
    dJ/d(body-frame marker nodes) of the run (the moving geometry included)
    by the adjoint time steps with checkpoints every 8 steps; wall time of
    the forward alone and of forward + backward; central differences of J
    (h = 1e-6) for the four parameters and along a random node direction."""
    from biocfd_amd import cases as CS
    from biocfd_amd import hip as H
    from biocfd_amd.adjoint import AdjointRun, THETA_NAMES

    case = CS.config_adjoint(n=args.adjoint_n, markers=args.adjoint_markers, nsteps=args.adjoint_steps,
                             pc_ita_max=args.pc_ita_max, epsi=0.0)
    H.call("biocfd_sor_graphs", args.sor_graphs)
    run = AdjointRun(case, every=8, avg_steps=max(1, args.adjoint_steps // 2), tune=True)
    try:
        run.forward()                       # warm-up (kernel loads, buffers)
        H.call("biocfd_synchronize")
        t0 = time.perf_counter()
        J = run.forward()
        H.call("biocfd_synchronize")
        t_f = time.perf_counter() - t0
        t0 = time.perf_counter()
        J2, (grad, gnodes) = run.gradient()
        H.call("biocfd_synchronize")
        t_fb = time.perf_counter() - t0
        out = {
            "workload": f"adjoint{args.adjoint_n}: single block {args.adjoint_n}^3 on [0,3]^3, ellipsoid 0.4:0.2:0.1 "
                        f"({args.adjoint_markers} markers, ID 51) at (1,1.5,1.5), heave 0.2, pitch 20 deg, phase "
                        f"90 deg, f=0.5, Re=100, dt=1e-3, {args.adjoint_steps} steps, {run.n_iter} SOR iterations "
                        f"per solve; J = mean x-force over the last {run.m} steps; gradient of the run (moving "
                        f"geometry included, discrete structure of the run)",
            "cells": case.cells(), "steps": args.adjoint_steps, "checkpoint_every": run.every, "dtype": "f64",
            "sor_kernel": run.sor,
            "J": J, "gradient": dict(zip(THETA_NAMES, (float(x) for x in grad))),
            "node_gradient_norm": float(np.linalg.norm(gnodes)), "nodes": int(run.st.nn),
            "forward_s": round(t_f, 3), "forward_backward_s": round(t_fb, 3),
            "backward_over_forward": round(t_fb / t_f, 2),
            "forward_mcell_steps_per_s": round(case.cells() * args.adjoint_steps / t_f / 1e6, 2),
            "forward_bit_identical_in_gradient_run": bool(J == J2),
        }
        if fd:
            # the run is smooth between changes of its discrete structure (tags,
            # IB cells, nearest elements): a difference is the derivative only
            # where neither perturbed run changes it, so each difference reports
            # the steps whose structure differs from the base run's
            base = []
            run.forward(states=base)

            def central(h, theta_dir=None, node_dir=None):
                sp, sm = [], []
                if node_dir is None:
                    jp = run.forward(run.theta0 + h * theta_dir, states=sp)
                    jm = run.forward(run.theta0 - h * theta_dir, states=sm)
                else:
                    n0 = run.nodes0
                    jp = run.forward(nodes=tuple(a + h * d for a, d in zip(n0, node_dir)), states=sp)
                    jm = run.forward(nodes=tuple(a - h * d for a, d in zip(n0, node_dir)), states=sm)
                changed = sum(int(a != b) for a, b in zip(sp, base)) + sum(int(a != b) for a, b in zip(sm, base))
                return (jp - jm) / (2 * h), changed

            rng = np.random.default_rng(7)
            D = [rng.standard_normal(run.st.nn) * 0.1 for _ in range(3)]
            adj_nodes = float(sum(np.dot(gnodes[:, q], D[q]) for q in range(3)))
            checks = {}
            for q, name in enumerate(THETA_NAMES):
                e = np.zeros(len(grad))
                e[q] = 1.0
                _, tl = run.tangent(e)
                row = {"adjoint": float(grad[q]), "tangent_linear": tl}
                for h in (1e-6, 1e-9):
                    v, ch = central(h, theta_dir=e)
                    row[f"central_h{h:g}"] = v
                    row[f"steps_with_changed_structure_h{h:g}"] = ch
                checks[name] = row
            _, tl = run.tangent(np.zeros(len(grad)), dnodes=D)
            row = {"adjoint": adj_nodes, "tangent_linear": tl}
            for h in (1e-6, 1e-9):
                v, ch = central(h, node_dir=D)
                row[f"central_h{h:g}"] = v
                row[f"steps_with_changed_structure_h{h:g}"] = ch
            checks["random node direction"] = row
            out["checks"] = checks
            out["max_rel_diff_adjoint_vs_tangent"] = max(
                abs(r["adjoint"] - r["tangent_linear"]) / max(abs(r["tangent_linear"]), 1e-300) for r in checks.values())
        return out
    finally:
        run.close()


def field_errors(ref, got):
    """max and L2 relative differences per field, and of the four body-force
    coefficients (viscous/pressure drag and lift)."""
    err = {}
    for k, a in ref["fields"].items():
        b = got["fields"][k]
        d = b - a
        err[k] = {"max_rel": float(np.max(np.abs(d)) / max(np.max(np.abs(a)), 1e-300)),
                  "l2_rel": float(np.linalg.norm(d) / max(np.linalg.norm(a), 1e-300))}
    fa, fb = ref["forces"], got["forces"]
    err["body_force"] = {"max_rel": float(np.max(np.abs(fb - fa)) / max(np.max(np.abs(fa)), 1e-300))}
    return err


def cpu_baseline(args):
    """The reference CPU build (serial within one block) on `cpu_steps` steps
    of the same 256^3 workload under the reference's stopping rule (501 SOR
    iterations per step; the residual rule's ~9e4 iterations per step would
    take hours on one core), setup excluded, on one OpenMP thread."""
    from biocfd_amd import cases as CS

    case = make_case(args, "reference")
    case.nsteps = args.cpu_steps
    case.dump_steps = ()
    wd = tempfile.mkdtemp(prefix="cpu_base_")
    path = os.path.join(wd, "case.bin")
    CS.write_case(case, path)
    if os.path.exists(REF_BIN):
        kind, cmd = "reference", [REF_BIN, path, wd, "--time"]
    else:
        kind, cmd = "port", [ORC_RUN, path, wd, "--time", str(args.cpu_steps)]
    env = dict(os.environ)
    # the reference parallelises over blocks only (main.F90:133-148): one block
    # runs on one thread whatever OMP_NUM_THREADS says; pin it to one
    omp_env = env.get("OMP_NUM_THREADS")
    env["OMP_NUM_THREADS"] = "1"
    r = subprocess.run(cmd, cwd=wd, capture_output=True, text=True, timeout=args.cpu_timeout, env=env)
    if r.returncode != 0:
        raise RuntimeError(f"cpu baseline failed: {r.stderr[-500:]}")
    if kind == "reference":
        txt = open(os.path.join(wd, "timing.txt")).read().split()
        steps, sec = int(txt[1]), float(txt[3])
    else:
        parts = r.stdout.split()
        steps, sec = int(parts[1]), float(parts[3])
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = None
    return {
        "value": case.cells() * steps / sec / 1e6,
        "unit": "Mcell-steps/s",
        "cores": 1,
        "kind": kind,
        "threads_used": 1,
        "omp_num_threads_run": 1,
        "omp_num_threads_env": omp_env,
        "nproc": os.cpu_count(),
        "cpus_allowed": affinity,
        "seconds": round(sec, 2),
        "sor_iterations_per_step": case.pc_ita_max + 1,
        "sample": f"{steps} time step(s) of the same 256^3 sphere workload under the reference's stopping rule "
                  f"(pcItaMax={case.pc_ita_max}, update norm < {case.epsi} every 5th iteration: "
                  f"{case.pc_ita_max + 1} SOR iterations per step), OMP_NUM_THREADS=1 (the reference is serial "
                  f"within a block), setup excluded; {sec:.1f} s; compare with reference_rule.value",
    }


def sor_roofline(args, run):
    """Roofline object of an FP64 run's red-black SOR: algorithmic bytes
    (24 B per cell and iteration, plus 16 B per cell per residual test of
    the residual rule) over the GPU time of the SOR calls (HIP events)."""
    tests = 0.0
    if run.get("rule") == "residual" and run["sor_iterations"] > 0:
        tests = run["sor_cell_iterations"] / args.check_every
    bytes_ = SOR_BYTES_PER_CELL_ITER * run["sor_cell_iterations"] + 16.0 * tests
    achieved = bytes_ / (run["sor_gpu_ms"] * 1e-3) / 1e9
    kk = run["kernel"]
    traffic = SOR_TRAFFIC_MEASURED.get((args.n, kk["family"], kk["tile"][0], kk["tile"][1], kk["threads"],
                                        kk["depth"], kk["iters_per_pass"], kk["planes"]))
    return {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "kernel": (f"red-black SOR ({'k_sor_fused' if kk['family'] == 'cell lists' else 'k_sor_pair'}<double>: "
                   f"{kk['iters_per_pass']} red+black iteration(s) per launch, tile {kk['tile'][0]}x{kk['tile'][1]})"
                   if kk["variant"] > 0 else
                   "red-black SOR iteration (2 colour-sweep launches + norm reduction launch)")
                  + ", 24 B/cell/iteration algorithmic"
                  + (f" + 16 B/cell per residual test (every {args.check_every} iterations)"
                     if run.get("rule") == "residual" else "")
                  + "; achieved over the GPU time of the whole solve; traffic = HBM bytes per launch",
        "us_per_iteration": round(1e3 * run["sor_gpu_ms"] / max(1, run["sor_iterations"]), 2),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--markers", type=int, default=8192)
    ap.add_argument("--resid-tol", type=float, default=1e-8,
                    help="max |b/dt - A pc| over the fluid cells that ends each pressure solve (both paths)")
    ap.add_argument("--resid-ita-max", type=int, default=400000,
                    help="pcItaMax of the residual-rule runs (out of reach: the residual ends every solve)")
    ap.add_argument("--check-every", type=int, default=256,
                    help="FP64 iterations between residual tests (the mixed path tests after every correction)")
    ap.add_argument("--pc-ita-max", type=int, default=500, help="pcItaMax of the reference-rule run")
    ap.add_argument("--epsi", type=float, default=1e-8, help="update-norm tolerance of the reference-rule run")
    ap.add_argument("--precision", choices=["both", "fp64", "mixed"], default="both")
    ap.add_argument("--rule", choices=["both", "residual", "reference"], default="both",
                    help="residual: the residual-rule runs (value, mixed); reference: the reference-rule run")
    ap.add_argument("--sweeps", type=int, default=256, help="FP32 sweeps per defect correction (mixed)")
    ap.add_argument("--sor-variant", default="auto",
                    help="FP64 SOR kernel variant (0 = colour-split sweeps, >=1 fused tiles) or 'auto'")
    ap.add_argument("--sor-planes", type=int, default=32, help="planes per workgroup of a fused variant")
    ap.add_argument("--mixed-variant", default="auto", help="FP32 sweep kernel variant of the mixed solve or 'auto'")
    ap.add_argument("--mixed-planes", type=int, default=16, help="planes per workgroup of the FP32 sweep kernel")
    ap.add_argument("--momentum-variant", type=int, default=4,
                    help="dense momentum kernel (include/biocfd_hip.h, biocfd_momentum_config): 0 = all three "
                         "components per thread per plane, 1 = one launch per component, 2 = as 0 with four waves "
                         "per SIMD, 3..8 = plane march of 64 x TJ column tiles (4 = TJ 4, 32 planes, the default)")
    ap.add_argument("--sor-graphs", type=int, choices=[0, 1], default=1,
                    help="1: FP64 SOR pass chunks replayed from HIP graphs (biocfd_sor_graphs); 0: pass by pass")
    ap.add_argument("--uneven-coarse-tune", type=int, choices=[0, 1], default=1,
                    help="1: the uneven workload's coarse block gets its own SOR kernel, timed on its live state")
    ap.add_argument("--interp-batching", type=int, choices=[0, 1], default=1,
                    help="1: every interface's update of a pass in one launch (biocfd_interp_batching); 0: one by one")
    ap.add_argument("--mixed-momentum", choices=["fp64", "both"], default="both",
                    help="both: a second mixed run with the FP32 momentum tendency, reported beside the first")
    ap.add_argument("--momentum-divisions", action="store_true",
                    help="keep every grid division of the momentum stencil a division (no power-of-two products)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="first run a plain device copy of this many MiB (byte-counter calibration)")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--cpu-timeout", type=float, default=600.0)
    ap.add_argument("--uneven-mixed-planes", type=int, default=8,
                    help="planes per workgroup of the uneven workload's FP32 sweeps (8 measured fastest of 4, 8, 16, 32)")
    ap.add_argument("--uneven-sweeps", type=int, default=16,
                    help="FP32 sweeps per correction of the uneven workload's mixed runs")
    ap.add_argument("--no-uneven", action="store_true", help="skip the uneven workload's secondary line")
    ap.add_argument("--uneven-runs", default="fp64-batched,fp64-per-block,mixed-batched,mixed-per-block",
                    help="--workload uneven: the runs (a subset prints them without the comparison)")
    ap.add_argument("--workload", choices=["sphere", "uneven", "adjoint"], default="sphere",
                    help="sphere: the 256^3 headline workload; uneven: batched vs per-block multi-block run; "
                         "adjoint: the adjoint workload's forward and gradient")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint workload's secondary line")
    ap.add_argument("--adjoint-n", type=int, default=96)
    ap.add_argument("--adjoint-markers", type=int, default=2048)
    ap.add_argument("--adjoint-steps", type=int, default=40)
    ap.add_argument("--adjoint-no-checks", action="store_true",
                    help="--workload adjoint: skip the tangent-linear and central-difference checks")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the block-sharded multi-GPU "
              "path is a design note, not part of this build", file=sys.stderr)
        sys.exit(2)
    if args.workload == "uneven":
        uneven_main(args)
        return
    if args.workload == "adjoint":
        print(json.dumps({"adjoint": adjoint_summary(args, fd=not args.adjoint_no_checks)}))
        return

    both = args.precision == "both"
    resid = args.rule in ("both", "residual")
    g = m = m32 = None
    if resid:
        if args.precision in ("both", "fp64"):
            g = run_path(args, "fp64", keep_fields=both, rule="residual")
            g["rule"] = "residual"
        if args.precision in ("both", "mixed"):
            m = run_path(args, "mixed", keep_fields=both, rule="residual")
            if args.mixed_momentum == "both":
                # the same mixed solve (the variant the first run chose) with the FP32 momentum tendency
                a32 = argparse.Namespace(**vars(args))
                a32.mixed_variant, a32.mixed_planes = str(m["kernel"]["variant"]), m["kernel"]["planes"]
                m32 = run_path(a32, "mixed", keep_fields=both, rule="residual", momentum="fp32")
    ref = None
    if args.rule in ("both", "reference") and args.precision in ("both", "fp64"):
        ref = run_path(args, "fp64", keep_fields=False, rule="reference")
        ref["rule"] = "reference"
    main_run = g if g is not None else (m if m is not None else ref)
    rule_txt = (f"every pressure solve to max|b/dt - A pc| < {args.resid_tol} over the fluid cells (FP64 residual, "
                f"tested every {args.check_every} FP64 iterations / after every {args.sweeps}-sweep mixed "
                f"correction; pcItaMax={args.resid_ita_max} not reached)"
                if main_run is not ref else
                f"the reference's stopping rule (update norm < {args.epsi} every 5th iteration, "
                f"pcItaMax={args.pc_ita_max})")
    out = {
        "metric": "Mcell-steps/s at 256^3 FP64 vs mixed-precision at 1e-8 residual; HBM %peak",
        "value": round(main_run["mcell_steps_per_s"], 4),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_run["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if main_run is not m else "f32+f64",
        "data": "synthetic (seeded 1e-3 noise on u=1 inflow; Fibonacci-hull sphere)",
        "config": {"workload": f"sphere{args.n}: single block {args.n}^3 on [0,8]x[0,4]x[0,4], sphere d=1, "
                               f"{args.markers} markers, Re=300, dt=5e-4, omega=1.8",
                   "stop_rule": rule_txt,
                   "value_is": ("FP64 method" if main_run is not m else "mixed precision")
                               + (" at the residual" if main_run is not ref else " under the reference's rule"),
                   "sor_iterations_per_step": main_run["sor_iters_per_step"],
                   "residual_linf_last_step": main_run["residual_linf"],
                   "sor_kernel": main_run["kernel"]},
        "roofline": None,
        "cpu_baseline": None,
    }
    if g is not None:
        out["roofline"] = sor_roofline(args, g)
    elif ref is not None:
        out["roofline"] = sor_roofline(args, ref)
    if m is not None:
        kb = mixed_bytes_per_cell(args.sweeps)
        solver_s = m["sor_gpu_ms"] * 1e-3
        mixed = {
            "value": round(m["mcell_steps_per_s"], 4),
            "ms_per_step": round(m["ms_per_step"], 3),
            "sweeps_per_step": m["sor_iters_per_step"],
            "sweeps_per_correction": args.sweeps,
            "residual_linf": m["residual_linf"],
            "solver_us_per_sweep": round(1e3 * m["sor_gpu_ms"] / max(1, m["sor_iterations"]), 2),
            "solver_achieved_gbs": round(kb / args.sweeps * m["sor_cell_iterations"] / solver_s / 1e9, 1)
            if solver_s > 0 else None,
            "solver_bytes_per_cell_sweep": kb / args.sweeps,
            "sor_kernel": m["kernel"],
        }
        if g is not None:
            mixed["speedup_vs_fp64"] = round(m["mcell_steps_per_s"] / g["mcell_steps_per_s"], 3)
            mixed["fp64_sor_iterations_per_step"] = g["sor_iters_per_step"]
            mixed["fp64_residual_linf"] = g["residual_linf"]
            mixed["rel_err_vs_fp64"] = field_errors(g, m)
        out["mixed"] = mixed
    if m32 is not None:
        f32 = {
            "momentum": "FP32 tendency in the dense pass (biocfd_momentum_precision(1)); thin-surface cells FP64",
            "value": round(m32["mcell_steps_per_s"], 4),
            "ms_per_step": round(m32["ms_per_step"], 3),
            "sweeps_per_step": m32["sor_iters_per_step"],
            "residual_linf": m32["residual_linf"],
        }
        if g is not None:
            f32["speedup_vs_fp64"] = round(m32["mcell_steps_per_s"] / g["mcell_steps_per_s"], 3)
            f32["rel_err_vs_fp64"] = field_errors(g, m32)
        out["mixed_fp32_momentum"] = f32
    if ref is not None and main_run is not ref:
        out["reference_rule"] = {
            "value": round(ref["mcell_steps_per_s"], 3),
            "ms_per_step": round(ref["ms_per_step"], 3),
            "stop_rule": f"update norm < {args.epsi} every 5th iteration, pcItaMax={args.pc_ita_max} "
                         f"(PcorVcor.F90:351-375)",
            "sor_iterations_per_step": ref["sor_iters_per_step"],
            "residual_linf_last_step": ref["residual_linf"],
            "sor_kernel": ref["kernel"],
            "roofline": sor_roofline(args, ref),
        }
    if not args.no_uneven:
        out["uneven"] = uneven_summary(args)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_summary(args, fd=False)
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(args)
        except Exception as e:  # reported, never silently replaced
            out["cpu_baseline"] = {"error": str(e)[:300]}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
