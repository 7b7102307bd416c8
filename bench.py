#!/usr/bin/env python3
"""Benchmark of the canoe dynamical-core step on one MI355X.

Workload: the 3D dry stratified convection box, 256 x 256 x 128 cells
(x1 = vertical, 128 levels), WENO5 horizontal / PLM vertical reconstruction,
HLLC fluxes, 5x5 block-tridiagonal implicit x1 solve, bottom heating,
SSPRK(3,3).  One step = one full RK3 time step of every cell; the metric is
cell-updates/s (the reference's zone-cycles/s, src/mesh_destroy.cpp:47-53).

Timed region: K steps replayed from the captured hipGraph, state resident in
HBM, bracketed by stream synchronisation.  A second, event-timed pass of the
same steps (no graph) gives the per-kernel durations used for the roofline.
The CPU baseline is the project's FP64 oracle timed on this host on a
bounded sample of the same workload (same column height, smaller footprint).

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import canoe_amd as ca  # noqa: E402

METRIC = ("cell-updates/s/GPU, mixed vs plain FP64 at ≤1e-5 rel L2; "
          "% HBM roofline")
HBM_PEAK_GBS = 8000.0       # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6     # MI355X FP64 vector


def cfl_dt(blk, th, cfg, w, cfl=0.9):
    """Time step: the signal speeds of the patched Hydro::NewBlockTimeStep
    (patches/21.new_blockdt.patch:31-45: the implicit x1 direction drops the
    sound speed), summed over directions as the unsplit RK3 update
    requires (a per-direction minimum at CFL 0.9 is unstable in 3D)."""
    sk, sj, si = blk.interior()
    ivx = 1 + blk.nmass
    rho = w[0][sk, sj, si]
    cs = np.sqrt(th.gammad * w[ivx + 3][sk, sj, si] / rho)
    inv = 0.0
    for d, dx in enumerate((blk.dx1, blk.dx2, blk.dx3)):
        if (d == 1 and blk.nx2 == 1) or (d == 2 and blk.nx3 == 1):
            continue
        v = np.abs(w[ivx + d][sk, sj, si])
        inv = inv + (v if (d == 0 and cfg.implicit_flag) else v + cs) / dx
    return float(cfl / max(float(np.max(inv)), 1e-30))


def time_gpu(ctx, blk, th, cfg, w0, dt, steps, warmup):
    run = ca.Run(ctx, blk, th, cfg)
    run.set_prim(w0)
    run.step(dt, warmup, graph=True)
    t0 = time.perf_counter()
    run.step(dt, steps, graph=True)       # returns after stream sync
    t1 = time.perf_counter()
    ms = 1e3 * (t1 - t0) / steps
    # event-timed breakdown of the same steps (no graph)
    run.set_prim(w0)
    run.set_timing(True)
    nt = max(2, min(steps, 5))
    run.step(dt, nt, graph=False)
    stats = {k: (v[0] / v[1], v[1] / nt) for k, v in run.kernel_stats().items()}
    run.set_timing(False)
    run.close()
    return ms, stats


def accuracy(ctx, blk, th, cfg, w0, dt, nsteps):
    """Relative L2 error of the FP32-flux form against the FP64 form after
    nsteps steps from the same state (the FP64 device path reproduces the
    CPU oracle bit for bit; tests/test_gpu_parity.py).  Per conserved
    variable group: density (and species), momentum, energy."""
    out = {}
    us = {}
    for prec in (ca.PREC_FP64, ca.PREC_FP32):
        cfg.precision = prec
        run = ca.Run(ctx, blk, th, cfg)
        run.set_prim(w0)
        run.step(dt, nsteps, graph=True)
        us[prec] = run.cons()
        run.close()
    sk, sj, si = blk.interior()
    a = us[ca.PREC_FP32][:, sk, sj, si]
    b = us[ca.PREC_FP64][:, sk, sj, si]
    ivx = 1 + blk.nmass
    groups = {"density": list(range(0, ivx)),
              "momentum": [ivx, ivx + 1, ivx + 2], "energy": [ivx + 3]}
    mx = 0.0
    for name, idx in groups.items():
        num = np.sqrt(np.sum((a[idx] - b[idx]) ** 2))
        den = np.sqrt(np.sum(b[idx] ** 2))
        e = float(num / den) if den > 0 else float(num)
        if not np.isfinite(e):
            e = float("inf")
        m = float(np.max(np.abs(a[idx] - b[idx])) / max(np.max(np.abs(b[idx])),
                                                         1e-300))
        out[name] = {"rel_l2": e, "rel_max": m}
        mx = e if not e <= mx else mx
    out["max_rel_l2"] = mx
    out["steps"] = nsteps
    return out


def algorithmic(blk, cfg):
    """Algorithmic HBM bytes per launch of each operation: per-unit figures
    x units per launch.  Flux divergence: read w once, write du once,
    (s_in + 8) N bytes per cell; implicit solve: read w, read du, write du,
    24 N bytes per cell."""
    N = blk.N
    cells = blk.ncells()
    s_in = 4 if cfg.precision == ca.PREC_FP32 else 8
    return {
        "flux_divergence": (s_in + 8) * N * cells,
        "implicit_solve": 24 * N * cells,
        "update": 32 * N * cells,   # read u, u1, du; write u (+ w)
    }


OPS = {
    "flux_divergence": ["k_hydrostatic_scan", "k_flux_x1", "k_flux_h<x2>",
                        "k_flux_h<x3>", "k_divergence", "k_flux32",
                        "k_mirror32"],
    "implicit_solve": ["k_implicit_fwd", "k_implicit_bwd"],
    "update": ["k_update", "k_bc_x1", "k_bc_x2", "k_bc_x3"],
}


def roofline(blk, cfg, stats):
    """Per-operation achieved HBM bandwidth from the live per-kernel
    averages; the dominant operation (largest time per step) is returned
    first."""
    alg = algorithmic(blk, cfg)
    ops = {}
    for op, names in OPS.items():
        per_launch_ms = sum(stats[n][0] for n in names if n in stats)
        per_step_ms = sum(stats[n][0] * stats[n][1] for n in names
                          if n in stats)
        if per_launch_ms <= 0:
            continue
        ach = alg[op] / (per_launch_ms * 1e-3) / 1e9
        ops[op] = {
            "kernels": [n for n in names if n in stats],
            "avg_launch_ms": round(per_launch_ms, 4),
            "ms_per_step": round(per_step_ms, 3),
            "algorithmic_bytes": alg[op],
            "achieved_GBs": round(ach, 1),
            "frac": round(ach / HBM_PEAK_GBS, 4),
        }
    dom = max(ops, key=lambda o: ops[o]["ms_per_step"])
    return dom, ops


def traffic_from_profiles(op):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE,
    WRITE_SIZE, calibrated against a copy kernel), if a summary is
    committed under profiles/."""
    p = os.path.join(ROOT, "profiles", "traffic_box256_fp64.json")
    if not os.path.exists(p):
        return None
    try:
        d = json.load(open(p))
        return d.get(op)
    except Exception:
        return None


def cpu_baseline(seconds, threads):
    """Oracle step on a bounded sample of the workload: 128 levels x 32 x 32
    columns (same physics, same column height)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orl
    blk, th, cfg = ca.pgen_setup(2, 128, 32, 32)
    w = ca.pgen_init(2, 1, blk, th, cfg)
    u = orl.prim_to_cons(blk, th, w)
    dt = cfl_dt(blk, th, cfg, w)
    cells = blk.ncells()
    n = 0
    t0 = time.perf_counter()
    while True:
        u, w = orl.step(blk, th, cfg, dt, 1, u, w, threads=threads)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds or n >= 200:
            break
    return {"value": round(cells * n / el, 1), "unit": "cell-updates/s",
            "cores": threads, "kind": "port",
            "sample": "%d steps of 128x32x32 cells (config-2 physics, %d "
                      "columns of 128 levels) in %.1f s" % (n, 32 * 32, el)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--acc-steps", type=int, default=100)
    ap.add_argument("--nx", type=int, nargs=3, default=[128, 256, 256],
                    help="nx1 (vertical) nx2 nx3")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(columns would shard by horizontal MeshBlocks; see "
              "DESIGN.md).  Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    blk, th, cfg = ca.pgen_setup(2, *args.nx)
    w0 = ca.pgen_init(2, 1, blk, th, cfg)
    dt = cfl_dt(blk, th, cfg, w0)
    cells = blk.ncells()
    print("# config-2 box %dx%dx%d, dt=%.4f s" % (blk.nx1, blk.nx2, blk.nx3,
                                                  dt), file=sys.stderr)
    ctx = ca.Context(0)
    forms = {}
    for name, prec in (("plain_fp64", ca.PREC_FP64),
                       ("mixed", ca.PREC_FP32)):
        cfg.precision = prec
        ms, stats = time_gpu(ctx, blk, th, cfg, w0, dt, args.steps,
                             args.warmup)
        dom, ops = roofline(blk, cfg, stats)
        forms[name] = {"ms_per_step": round(ms, 4),
                       "cell_updates_per_s": cells / (ms * 1e-3),
                       "dominant": dom, "ops": ops,
                       "kernels_avg_ms": {k: round(v[0], 4)
                                          for k, v in stats.items()}}
        print("# %s: %.3f ms/step, %.3e cell-updates/s" %
              (name, ms, cells / (ms * 1e-3)), file=sys.stderr)
    acc = accuracy(ctx, blk, th, cfg, w0, dt, args.acc_steps)
    forms["mixed"]["error_vs_fp64"] = acc
    forms["mixed"]["speedup_vs_plain"] = round(
        forms["plain_fp64"]["ms_per_step"] / forms["mixed"]["ms_per_step"], 3)
    print("# mixed vs FP64 after %d steps: %s" % (args.acc_steps, acc),
          file=sys.stderr)
    ctx.close()

    ok = acc["max_rel_l2"] <= 1e-5
    chosen = forms["mixed"] if ok else forms["plain_fp64"]
    dom = chosen["dominant"]
    op = chosen["ops"][dom]
    rf = {"bound": "hbm", "op": dom, "kernels": op["kernels"],
          "achieved": op["achieved_GBs"], "peak": HBM_PEAK_GBS,
          "unit": "GB/s", "frac": op["frac"],
          "traffic": traffic_from_profiles(dom)}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_seconds, 1)
        thr = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or \
            min(16, os.cpu_count() or 1)
        cpu_all = cpu_baseline(max(3.0, args.cpu_seconds / 3), thr)
        cpu["all_cores"] = {"value": cpu_all["value"], "cores": thr}

    out = {
        "metric": METRIC,
        "value": round(chosen["cell_updates_per_s"], 1),
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": chosen["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if chosen is forms["plain_fp64"] else "f32/f64",
        "data": "synthetic",
        "config": {"workload": "3D dry stratified box %dx%dx%d (x1 vertical), "
                               "WENO5-h/PLM-v, HLLC, 5x5 implicit x1, "
                               "SSPRK(3,3), bottom heating"
                               % (blk.nx2, blk.nx3, blk.nx1),
                   "cells": cells, "dt": dt,
                   "form": "plain FP64" if chosen is forms["plain_fp64"]
                   else "FP32 fluxes + FP64 state"},
        "roofline": rf,
        "cpu_baseline": cpu,
        "forms": forms,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
