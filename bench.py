#!/usr/bin/env python3
"""Benchmark of the canoe dynamical-core step on one MI355X.

Headline workload (`value`): the 3D dry stratified convection box,
256 x 256 x 128 cells (x1 = vertical, 128 levels), WENO5 horizontal / PLM
vertical reconstruction, HLLC fluxes, 5x5 block-tridiagonal implicit x1
solve, bottom heating, SSPRK(3,3), in the form chosen by accuracy: FP32
reconstruction + Riemann fluxes with FP64 state and implicit solve when its
100-step relative L2 error against the plain FP64 form is <= 1e-5, else the
plain FP64 form.  One step = one full RK3 time step of every cell; the
metric is cell-updates/s (the reference's zone-cycles/s,
src/mesh_destroy.cpp:47-53).

Also reported (not `value`): the moist box (the dry box + vapour, cloud,
precipitation, saturation adjustment) as plain FP64 with Newton from the
all-vapour state vs FP32 fluxes with the network-guided Newton start, with
the network trained on this GPU from pairs of the plain adjustment; and the
uneven-column batch (32 MeshBlocks, heights 40..256, footprints 32..128)
with the autotuned launch shape.

Timed region: K steps, state resident in HBM, bracketed by stream
synchronisation.  A second, event-timed pass of the same steps gives the
per-kernel durations used for the roofline.  The CPU baseline is the
project's FP64 oracle timed on this host on a bounded sample of the
headline workload (same column height, smaller footprint).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--quick]
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
from canoe_amd import mlp as cm  # noqa: E402

METRIC = ("cell-updates/s/GPU, mixed vs plain FP64 at ≤1e-5 rel L2; "
          "% HBM roofline")
HBM_PEAK_GBS = 8000.0       # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6     # MI355X FP64 vector
BF16_PEAK_TFLOPS = 2500.0   # MI355X dense bf16 MFMA
MLP_FLOP = 2 * (4 * 64 + 64 * 64 + 64 * 64 + 64 * 3)  # per sample forward
USE_GRAPH = True


def log(msg):
    print("# " + msg, file=sys.stderr, flush=True)


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


def make_run(ctx, blocks, th, cfg, w0s, net=None, tune=None):
    run = ca.Run(ctx, None, th, cfg, blocks=blocks)
    if net is not None:
        run.set_mlp(net)
    if tune is not None:
        run.set_tune(tune.implicit_wpb, tune.implicit_grid, tune.flux_block,
                     tune.face_variant, tune.march_seg, tune.implicit_fused)
    for b, w in enumerate(w0s):
        run.set_prim(w, b)
    return run


def time_form(ctx, blocks, th, cfg, w0s, dt, steps, warmup, net=None,
              tune=None, autotune=False):
    """ms per step (graph replay, or direct launches when graphs are off),
    per-kernel event-timed averages, the launch shape used, and the median
    of 20 single synchronised steps."""
    run = make_run(ctx, blocks, th, cfg, w0s, net, tune)
    try:
        if autotune:
            run.autotune(dt)
            for b, w in enumerate(w0s):
                run.set_prim(w, b)
        t = run.tune()
        run.step(dt, warmup, graph=USE_GRAPH)
        t0 = time.perf_counter()
        run.step(dt, steps, graph=USE_GRAPH)       # returns after stream sync
        t1 = time.perf_counter()
        ms = 1e3 * (t1 - t0) / steps
        # the median of single synchronised steps (each a graph replay)
        single = []
        for _ in range(20):
            s0 = time.perf_counter()
            run.step(dt, 1, graph=USE_GRAPH)
            single.append(1e3 * (time.perf_counter() - s0))
        median = float(np.median(single))
        for b, w in enumerate(w0s):
            run.set_prim(w, b)
        run.set_timing(True)
        nt = max(2, min(steps, 5))
        run.step(dt, nt, graph=False)
        stats = {k: (v[0] / v[1], v[1] / nt)
                 for k, v in run.kernel_stats().items()}
        run.set_timing(False)
        return ms, stats, t.as_dict(), median
    finally:
        run.close()


def main_loop(ctx, blk, th, cfg, w0, cfl, cycles, warmup, tune=None):
    """The device main loop (cg_run_advance): per cycle one step, the
    NewBlockTimeStep reduction over the resident state and Mesh::NewTimeStep
    on the host from one read-back double.  ms per cycle, the reduction
    kernel's average duration, and the steps taken."""
    run = make_run(ctx, [blk], th, cfg, [w0], tune=tune)
    try:
        t, dt, n = run.advance(cfl, nlim=warmup, graph=USE_GRAPH)
        t0 = time.perf_counter()
        t, dt, n = run.advance(cfl, nlim=warmup + cycles, time=t, dt=dt,
                               ncycle=n, graph=USE_GRAPH)
        t1 = time.perf_counter()
        run.set_timing(True)
        run.new_dt(cfl)
        ms_dt = run.kernel_stats().get("k_new_dt", (0.0, 1))
        run.set_timing(False)
        return {"cfl": cfl, "cycles": cycles,
                "ms_per_cycle": round(1e3 * (t1 - t0) / cycles, 4),
                "k_new_dt_ms": round(ms_dt[0] / max(ms_dt[1], 1), 4),
                "k_new_dt_GBs": round(blk.N * 8 * blk.ncells()
                                      / (ms_dt[0] / max(ms_dt[1], 1) * 1e-3)
                                      / 1e9, 1) if ms_dt[0] > 0 else None,
                "last_dt": dt, "time": t}
    finally:
        run.close()


def final_state(ctx, blocks, th, cfg, w0s, dt, nsteps, net=None, tune=None):
    run = make_run(ctx, blocks, th, cfg, w0s, net, tune)
    try:
        run.step(dt, nsteps, graph=USE_GRAPH)
        return [run.cons(b) for b in range(len(blocks))]
    finally:
        run.close()


def errors(blocks, ua, ub):
    """Relative L2 and max errors of ua against ub over the interior cells of
    all blocks, per conserved group: density, each tracer, momentum,
    energy."""
    nm = blocks[0].nmass
    ivx = 1 + nm
    groups = {"density": [0]}
    for n in range(1, ivx):
        groups["tracer%d" % n] = [n]
    groups["momentum"] = [ivx, ivx + 1, ivx + 2]
    groups["energy"] = [ivx + 3]
    out = {}
    mx = 0.0
    for name, idx in groups.items():
        num = den = 0.0
        dmax = bmax = 0.0
        for blk, a, b in zip(blocks, ua, ub):
            sk, sj, si = blk.interior()
            x = a[idx][:, sk, sj, si]
            y = b[idx][:, sk, sj, si]
            num += float(np.sum((x - y) ** 2))
            den += float(np.sum(y ** 2))
            dmax = max(dmax, float(np.max(np.abs(x - y))))
            bmax = max(bmax, float(np.max(np.abs(y))))
        e = float(np.sqrt(num / den)) if den > 0 else float(np.sqrt(num))
        if not np.isfinite(e):
            e = float("inf")
        m = dmax / bmax if bmax > 0 else dmax
        out[name] = {"rel_l2": e, "rel_max": m}
        if name in ("density", "momentum", "energy"):
            mx = e if not e <= mx else mx
    out["max_rel_l2"] = mx
    return out


def algorithmic(blk, cfg):
    """Algorithmic HBM bytes per launch of each operation: per-unit figures
    x units per launch.  Flux divergence: read w once, write du once,
    (s_in + 8) N bytes per cell, s_in = 8 in both forms (the FP32-flux
    kernels read the FP64 state w; no FP32 copy of it is stored).  Column
    solve of a run: the launch solves the x1 columns and, in its downward
    sweep, applies the stage (u, u1 and du combined, ConservedToPrimitive)
    and writes the next stage's hydrostatic pressures, so one launch reads
    w, du and u (and u1 at stages 2 and 3) and writes u, w, psf and psv:
    (8 (5 + 2/3) N + 16) bytes per cell, averaged over the three stages."""
    N = blk.N
    cells = blk.ncells()
    s_in = 8
    return {
        "flux_divergence": (s_in + 8) * N * cells,
        "implicit_solve": int(round((8 * (5 + 2.0 / 3.0) * N + 16) * cells)),
        "update": 32 * N * cells,   # read u, u1, du; write u (+ w)
    }


OPS = {
    "flux_divergence": ["k_hydrostatic_scan", "k_flux_x1", "k_flux_h<x2>",
                        "k_flux_h<x3>", "k_divergence", "k_flux32",
                        "k_mix_fused"],
    "implicit_solve": ["k_implicit_fwd", "k_implicit_bwd", "k_implicit_fused"],
    "update": ["k_update", "k_bc_x1", "k_bc_x2", "k_bc_x3",
               "k_satadj_run_mlp"],
}


def roofline(blk, cfg, stats):
    """Per-operation achieved HBM bandwidth from the live per-kernel
    averages; the dominant operation is the largest time per step."""
    alg = algorithmic(blk, cfg)
    ops = {}
    for op, names in OPS.items():
        # implicit runs update the state inside the downward sweep: the
        # ghost kernels left in "update" move no algorithmic bytes
        if op == "update" and "k_update" not in stats:
            continue
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


def traffic_from_profiles(op, tune):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE,
    WRITE_SIZE, calibrated against a copy kernel), if a summary is
    committed under profiles/; the column solve's figure for the sweeps
    fused per group or in two launches, as the autotuner chose."""
    p = os.path.join(ROOT, "profiles", "traffic_box256.json")
    if not os.path.exists(p):
        return None
    if op == "implicit_solve" and not tune.get("implicit_fused", 1):
        op = "implicit_solve_two_launches"
    try:
        d = json.load(open(p))
        return d.get(op)
    except Exception:
        return None


def fp64_from_profiles(avg_ms, cells):
    """FP64 rate of the column solve: the FP64 lane operations per cell of
    each column kernel, counted with rocprofv3 (SQ_INSTS_VALU_*_F64, a
    committed summary under profiles/), over the kernels' live event-timed
    durations."""
    p = os.path.join(ROOT, "profiles", "fp64_ops_box256.json")
    if not os.path.exists(p):
        return None
    try:
        per = json.load(open(p))["flop_per_cell"]
    except Exception:
        return None
    names = {"k_implicit_fused": "k_implicit_fused_batch",
             "k_implicit_fwd": "k_implicit_fwd_batch",
             "k_implicit_bwd": "k_implicit_bwd_update_batch"}
    flop = ms = 0.0
    for k, key in names.items():
        if k in avg_ms and key in per:
            flop += per[key] * cells
            ms += avg_ms[k]
    if ms <= 0:
        return None
    tf = flop / (ms * 1e-3) / 1e12
    return {"flop_per_cell": round(flop / cells, 1), "achieved": round(tf, 2),
            "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(tf / FP64_PEAK_TFLOPS, 4)}


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
            "sample": "%d steps of 128x32x32 cells (dry-box physics, %d "
                      "columns of 128 levels) in %.1f s" % (n, 32 * 32, el)}


# ---------------------------------------------------------------------------
def bench_box(ctx, nx, steps, warmup, acc_steps):
    blk, th, cfg = ca.pgen_setup(2, *nx)
    w0 = ca.pgen_init(2, 1, blk, th, cfg)
    dt = cfl_dt(blk, th, cfg, w0)
    cells = blk.ncells()
    log("dry box %dx%dx%d, dt=%.4f s" % (blk.nx1, blk.nx2, blk.nx3, dt))
    forms = {}
    tunes = {}
    for name, prec in (("plain_fp64", ca.PREC_FP64), ("mixed", ca.PREC_FP32)):
        cfg.precision = prec
        ms, stats, tune, med = time_form(ctx, [blk], th, cfg, [w0], dt, steps,
                                    warmup, autotune=True)
        dom, ops = roofline(blk, cfg, stats)
        forms[name] = {"ms_per_step": round(ms, 4),
                       "ms_per_step_median_of_20": round(med, 4),
                       "cell_updates_per_s": cells / (ms * 1e-3),
                       "dominant": dom, "ops": ops, "tune": tune,
                       "kernels_avg_ms": {k: round(v[0], 4)
                                          for k, v in stats.items()}}
        tunes[name] = tune
        log("%s: %.3f ms/step, %.3e cell-updates/s, tune %s"
            % (name, ms, cells / (ms * 1e-3), tune))
    us = {}
    for name, prec in (("plain_fp64", ca.PREC_FP64), ("mixed", ca.PREC_FP32)):
        cfg.precision = prec
        us[prec] = final_state(ctx, [blk], th, cfg, [w0], dt, acc_steps,
                               tune=ca.Tune(**tunes[name]))
    acc = errors([blk], us[ca.PREC_FP32], us[ca.PREC_FP64])
    acc["steps"] = acc_steps
    forms["mixed"]["error_vs_fp64"] = acc
    forms["mixed"]["speedup_vs_plain"] = round(
        forms["plain_fp64"]["ms_per_step"] / forms["mixed"]["ms_per_step"], 3)
    log("mixed vs FP64 after %d steps: max rel L2 %.3e"
        % (acc_steps, acc["max_rel_l2"]))
    # the same form driven by the device main loop: per-direction CFL 0.45
    # gives the step of the summed-direction bound at 0.9 on this box
    cfg.precision = ca.PREC_FP32
    forms["mixed"]["main_loop"] = main_loop(
        ctx, blk, th, cfg, w0, 0.45, steps, warmup,
        tune=ca.Tune(**tunes["mixed"]))
    log("main loop: %s" % forms["mixed"]["main_loop"])
    return blk, cfg, dt, cells, forms, acc


def moist_states(blk, th, w):
    """(rho, ie, q) of the interior cells of a moist primitive state: ie is
    the sensible internal energy density p / (gamma(q) - 1) = rho c_v T."""
    sk, sj, si = blk.interior()
    ivx = 1 + blk.nmass
    rho = np.ascontiguousarray(w[0][sk, sj, si].ravel())
    q = np.ascontiguousarray(np.stack([w[n][sk, sj, si].ravel()
                                       for n in (1, 2, 3)]))
    p = w[ivx + 3][sk, sj, si].ravel()
    cvd = th.rd / (th.gammad - 1)
    rv = th.rd * (1 + th.inv_mu_ratio_m1[0])
    qd = 1 - q.sum(0)
    R = th.rd * qd + rv * q[0]
    cv = cvd * qd + sum(cvd * (1 + th.cv_ratio_m1[n]) * q[n] for n in range(3))
    return rho, np.ascontiguousarray(cv * p / R), q


def supersaturated(th, rho, ie, q, n, seed):
    """n states drawn from the given cells, water scaled by 1.1..2.2 into
    vapour and temperature shifted by -2..2 K (most become saturated)."""
    rng = np.random.default_rng(seed)
    b = rng.integers(0, rho.size, n)
    cvd = th.rd / (th.gammad - 1)
    cvv = cvd * (1 + th.cv_ratio_m1[0])
    cvp = cvd * (1 + th.cv_ratio_m1[2])
    qb = q[:, b].copy()
    qd = 1 - qb.sum(0)
    cvb = cvd * qd + cvv * qb[0] + cvd * (1 + th.cv_ratio_m1[1]) * qb[1] \
        + cvp * qb[2]
    T = ie[b] / (rho[b] * cvb) + rng.uniform(-2, 2, n)
    qn = qb.copy()
    qn[0] = (qb[0] + qb[1]) * rng.uniform(1.1, 2.2, n)
    qn[1] = 0.0
    cvn = cvd * (1 - qn.sum(0)) + cvv * qn[0] + cvp * qn[2]
    r = rho[b].copy()
    return r, r * cvn * T, np.ascontiguousarray(qn)


def adjust_event_ms(ctx, th, rho, ie, q, net, reps=5):
    """cg_saturation_adjust on device-resident arrays, plain and guided:
    median over reps of the event-timed call (the inputs are restored from a
    device copy between calls, outside the timed region)."""
    L = cm._lib()
    n = rho.size
    d = {k: ctx.array(v) for k, v in (("rho", rho), ("ie", ie), ("q", q))}
    keep = {k: ctx.array(v) for k, v in (("ie", ie), ("q", q))}
    it = ctx.array(shape=(n,), dtype=np.int32)
    out = []
    try:
        for nh in (None, net):
            ts = []
            for _ in range(reps + 1):
                for k in ("ie", "q"):
                    ctx._check(ca.lib().cg_memcpy_d2d(
                        ctx.h, d[k].ptr, keep[k].ptr, d[k].nbytes), "d2d")
                ctx.timer_start()
                ctx._check(L.cg_saturation_adjust(
                    ctx.h, ca._ref(th), n, d["rho"].ptr, d["ie"].ptr,
                    d["q"].ptr, nh.h if nh is not None else None, it.ptr),
                    "cg_saturation_adjust")
                ts.append(ctx.timer_stop())
            out.append(float(np.median(ts[1:])))
    finally:
        for a in list(d.values()) + list(keep.values()) + [it]:
            a.free()
    return out


def bench_moist(ctx, nx, steps, warmup, acc_steps, npairs, train_steps):
    blk, th, cfg = ca.pgen_setup(4, *nx)
    w0 = ca.pgen_init(4, 1, blk, th, cfg)
    dt = cfl_dt(blk, th, cfg, w0)
    cells = blk.ncells()
    log("moist box %dx%dx%d, dt=%.4f s" % (blk.nx1, blk.nx2, blk.nx3, dt))
    out = {"workload": "moist box %dx%dx%d: dry box + vapour/cloud/precip, "
                       "saturation adjustment (Newton to 1e-10 K)"
                       % (blk.nx2, blk.nx3, blk.nx1), "cells": cells, "dt": dt}
    # training pairs from the state of a moist run (the plain adjustment
    # labels them); the network is trained on this GPU
    cfg.precision = ca.PREC_FP64
    cfg.satadj = ca.SATADJ_NEWTON
    run = make_run(ctx, [blk], th, cfg, [w0])
    run.step(dt, 10, graph=USE_GRAPH)
    wrun = run.prim()
    run.close()
    rho, ie, q = moist_states(blk, th, wrun)
    t0 = time.perf_counter()
    dx, dtt, _ = cm.make_pairs(ctx, th, rho, ie, q, npairs, seed=5,
                               device=True)
    t1 = time.perf_counter()
    net = cm.Mlp(ctx, seed=7)
    losses = net.train(dx, dtt, steps=train_steps, batch=8192, lr=3e-3,
                       lr_final=1e-5, seed=9)
    t2 = time.perf_counter()
    xh, th_, _ = cm.make_pairs(ctx, th, rho, ie, q, 1 << 18, seed=77)
    yh = net.forward(xh)
    sat = th_[:, 2] > 0
    rms_k = float(np.sqrt(np.mean((yh[sat, 0] - th_[sat, 0]) ** 2)) * 10)
    # inference throughput on the matrix cores: forward of 8M samples
    nbig = 1 << 23
    xb = ctx.array(np.tile(xh, (nbig // xh.shape[0], 1)), dtype=np.float32)
    yb = ctx.array(shape=(nbig, 3), dtype=np.float32)
    L = cm._lib()
    ctx._check(L.cg_mlp_forward(net.h, nbig, xb.ptr, yb.ptr), "forward")
    reps = 5
    t3 = time.perf_counter()
    for _ in range(reps):
        ctx._check(L.cg_mlp_forward(net.h, nbig, xb.ptr, yb.ptr), "forward")
    t4 = time.perf_counter()
    xb.free()
    yb.free()
    tf = MLP_FLOP * nbig * reps / (t4 - t3) / 1e12
    # guided vs plain adjustment on held-out supersaturated states
    rr, en, qn = supersaturated(th, rho, ie, q, 1 << 20, seed=3)
    ie_p, y_p, it_p = cm.saturation_adjust(ctx, th, rr, en, qn)
    ie_g, y_g, it_g = cm.saturation_adjust(ctx, th, rr, en, qn, net)
    adj_ms = adjust_event_ms(ctx, th, rr, en, qn, net)
    s = it_p > 0
    out["mlp"] = {
        "net": "4-64-64-64-3 tanh-GELU, bf16 MFMA (fp32 accumulate)",
        "pairs": npairs, "pair_gen_s": round(t1 - t0, 3),
        "train_steps": train_steps, "batch": 8192,
        "train_s": round(t2 - t1, 3),
        "loss_first": float(np.mean(losses[:20])),
        "loss_last": float(np.mean(losses[-100:])),
        "heldout_T_rms_K": round(rms_k, 4),
        "newton_iters_plain": round(float(it_p[s].mean()), 3),
        "newton_iters_guided": round(float(it_g[s].mean()), 3),
        "saturated_cells": int(s.sum()),
        "adjust_states": int(rr.size),
        "adjust_ms_plain": round(adj_ms[0], 4),
        "adjust_ms_guided": round(adj_ms[1], 4),
        "max_rel_diff_ie": float(np.max(np.abs(ie_g - ie_p) / np.abs(ie_p))),
        "forward_tflops": round(tf, 3),
        "forward_mfma_frac": round(tf / BF16_PEAK_TFLOPS, 5),
    }
    log("mlp: %s" % out["mlp"])
    forms = {}
    variants = (("plain_fp64_newton", ca.PREC_FP64, ca.SATADJ_NEWTON, None),
                ("mixed_newton", ca.PREC_FP32, ca.SATADJ_NEWTON, None),
                ("mixed_guided", ca.PREC_FP32, ca.SATADJ_MLP, net))
    for name, prec, mode, n in variants:
        cfg.precision = prec
        cfg.satadj = mode
        ms, stats, tune, med = time_form(ctx, [blk], th, cfg, [w0], dt, steps,
                                    warmup, net=n, autotune=True)
        dom, ops = roofline(blk, cfg, stats)
        forms[name] = {"ms_per_step": round(ms, 4),
                       "ms_per_step_median_of_20": round(med, 4),
                       "cell_updates_per_s": cells / (ms * 1e-3),
                       "dominant": dom, "ops": ops, "tune": tune,
                       "kernels_avg_ms": {k: round(v[0], 4)
                                          for k, v in stats.items()}}
        log("moist %s: %.3f ms/step, tune %s" % (name, ms, tune))
    us = {}
    for name, prec, mode, n in variants:
        cfg.precision = prec
        cfg.satadj = mode
        us[name] = final_state(ctx, [blk], th, cfg, [w0], dt, acc_steps, net=n)
    plain = forms["plain_fp64_newton"]["ms_per_step"]
    for name in ("mixed_newton", "mixed_guided"):
        acc = errors([blk], us[name], us["plain_fp64_newton"])
        acc["steps"] = acc_steps
        forms[name]["error_vs_fp64"] = acc
        forms[name]["speedup_vs_plain"] = round(
            plain / forms[name]["ms_per_step"], 3)
    # the chosen moist form: the faster mixed form within 1e-5 of FP64
    ok = [n for n in ("mixed_newton", "mixed_guided")
          if forms[n]["error_vs_fp64"]["max_rel_l2"] <= 1e-5]
    out["chosen"] = (min(ok, key=lambda n: forms[n]["ms_per_step"]) if ok
                     else "plain_fp64_newton")
    out["forms"] = forms
    out["kessler"] = bench_kessler(ctx, blk, th, cfg, w0, dt, steps, warmup,
                                   forms["mixed_newton"]["tune"])
    net.close()
    dx.free()
    dtt.free()
    cfg.satadj = ca.SATADJ_NEWTON
    return out


def bench_lines(ctx, nx, steps, warmup):
    """SolveImplicit3D in every direction on the dry box, plain FP64:
    implicit_flag 15 = x3 lines, then x2 lines (both with periodic ends,
    PeriodicForwardSweep's cyclic recurrences), then the x1 columns, in each
    of the three stages.  Algorithmic bytes of one direction's solve: read w,
    read du, write du, 24 N bytes per cell; the line kernels' time per
    direction and stage is their time per step over 6."""
    blk, th, cfg = ca.pgen_setup(2, *nx)
    w0 = ca.pgen_init(2, 1, blk, th, cfg)
    dt = cfl_dt(blk, th, cfg, w0)
    cfg.implicit_flag = 15
    cfg.precision = ca.PREC_FP64
    cells = blk.ncells()
    ms, stats, tune, med = time_form(ctx, [blk], th, cfg, [w0], dt, steps, warmup)
    lk = stats.get("k_line_*", (0.0, 0))
    line_ms = lk[0] * lk[1]
    per_dir = line_ms / 6.0
    alg = 24 * blk.N * cells
    ach = alg / (per_dir * 1e-3) / 1e9 if per_dir > 0 else None
    log("lines (flag 15): %.3f ms/step, line kernels %.3f ms/step"
        % (ms, line_ms))
    return {"workload": "dry box %dx%dx%d, implicit_flag 15 (x3, x2 periodic "
                        "lines + x1 columns), plain FP64"
                        % (blk.nx2, blk.nx3, blk.nx1),
            "ms_per_step": round(ms, 4),
            "line_kernels_ms_per_step": round(line_ms, 4),
            "line_ms_per_direction_stage": round(per_dir, 4),
            "algorithmic_bytes_per_direction_stage": alg,
            "achieved_GBs": round(ach, 1) if ach else None,
            "frac": round(ach / HBM_PEAK_GBS, 4) if ach else None,
            "kernels_avg_ms": {k: round(v[0], 4) for k, v in stats.items()}}


def bench_host_solve(ctx, nx, reps=3):
    """cg_implicit_solve_host on the dry box (implicit_flag 9, plain FP64):
    du and w stay in pageable host memory, as the reference's AthenaArrays
    do; the call copies both to the device, solves the x1 columns and copies
    du back, so its time includes PCIe transfers of 3 N cells_with_ghosts x 8
    bytes.  The rate of the host-buffer boundary, beside (not instead of) the
    resident step."""
    blk, th, cfg = ca.pgen_setup(2, *nx)
    w = np.ascontiguousarray(ca.pgen_init(2, 1, blk, th, cfg), np.float64)
    dt = cfl_dt(blk, th, cfg, w)
    du = np.zeros(blk.shape())
    sk, sj, si = blk.interior()
    du[:, sk, sj, si] = np.random.default_rng(3).normal(
        size=du[:, sk, sj, si].shape) * np.array(
            [1e-3, 1e-2, 1e-2, 1e-2, 3e2])[:, None, None, None]
    times = []
    for r in range(reps + 1):
        d = du.copy()
        t0 = time.perf_counter()
        rc = ca.lib().cg_implicit_solve_host(ctx.h, ca._ref(blk), ca._ref(th),
                                             ca._ptr(d), ca._ptr(w), dt, 9)
        t1 = time.perf_counter()
        if rc != ca.OK:
            raise ca.CanoeGpuError(rc, "cg_implicit_solve_host")
        if r:
            times.append(t1 - t0)
    ms = float(np.median(times)) * 1e3
    cells = blk.ncells()
    pcie = 3 * du.nbytes
    # the parts of a call: the same transfers between pageable host memory
    # and device buffers allocated once (du and w in, du out), and the
    # device solve on those buffers (event-timed); the rest of the call is
    # its staging and scratch memory (taken from the device's pool, which the
    # context keeps between calls) and host overhead
    L = ca.lib()
    dd, dw = ctx.array(du), ctx.array(w)
    d = du.copy()
    tx, ts = [], []
    try:
        for r in range(reps + 1):
            t0 = time.perf_counter()
            ctx._check(L.cg_memcpy_h2d(ctx.h, dd.ptr, ca._ptr(d), du.nbytes), "h2d")
            ctx._check(L.cg_memcpy_h2d(ctx.h, dw.ptr, ca._ptr(w), w.nbytes), "h2d")
            ctx._check(L.cg_memcpy_d2h(ctx.h, ca._ptr(d), dd.ptr, du.nbytes), "d2h")
            t1 = time.perf_counter()
            dd.upload(du)
            ctx.timer_start()
            ctx._check(L.cg_implicit_solve(ctx.h, ca._ref(blk), ca._ref(th),
                                           dd.ptr, dw.ptr, dt, 9), "solve")
            t2 = ctx.timer_stop()
            if r:
                tx.append(1e3 * (t1 - t0))
                ts.append(t2)
    finally:
        dd.free()
        dw.free()
    ms_x = float(np.median(tx))
    ms_s = float(np.median(ts))
    log("implicit solve on host arrays: %.2f ms per call (transfers %.2f, "
        "solve %.2f)" % (ms, ms_x, ms_s))
    return {"workload": "cg_implicit_solve_host, dry box %dx%dx%d, flag 9, "
                        "pageable host du and w" % (blk.nx2, blk.nx3, blk.nx1),
            "ms_per_call": round(ms, 3),
            "pcie_bytes": pcie,
            "pcie_inclusive_GBs": round(pcie / (ms * 1e-3) / 1e9, 2),
            "cell_solves_per_s": cells / (ms * 1e-3),
            "transfers_ms": round(ms_x, 3),
            "transfers_GBs": round(pcie / (ms_x * 1e-3) / 1e9, 2),
            "device_solve_ms": round(ms_s, 3),
            "allocation_and_other_ms": round(ms - ms_x - ms_s, 3)}


def mlp_run(nx, train_steps):
    """A short, fixed run of the saturation-guess network for rocprofv3:
    pairs from the moist box's initial state (2^20), train_steps Adam steps
    of batch 8192, then 3 forwards of 2^23 samples; prints the measured
    forward rate against the bf16 MFMA peak."""
    ctx = ca.Context(0)
    blk, th, cfg = ca.pgen_setup(4, *nx)
    w0 = ca.pgen_init(4, 1, blk, th, cfg)
    rho, ie, q = moist_states(blk, th, w0)
    dx, dtt, _ = cm.make_pairs(ctx, th, rho, ie, q, 1 << 20, seed=5,
                               device=True)
    net = cm.Mlp(ctx, seed=7)
    t0 = time.perf_counter()
    net.train(dx, dtt, steps=train_steps, batch=8192, lr=3e-3, lr_final=1e-5,
              seed=9)
    t1 = time.perf_counter()
    nbig = 1 << 23
    xb = ctx.array(shape=(nbig, 4), dtype=np.float32)
    xb.upload(np.tile(dx.download()[:1 << 20], (nbig >> 20, 1)))
    yb = ctx.array(shape=(nbig, 3), dtype=np.float32)
    L = cm._lib()
    reps = 3
    t2 = time.perf_counter()
    for _ in range(reps):
        ctx._check(L.cg_mlp_forward(net.h, nbig, xb.ptr, yb.ptr), "forward")
    t3 = time.perf_counter()
    tf = MLP_FLOP * nbig * reps / (t3 - t2) / 1e12
    for a in (xb, yb, dx, dtt):
        a.free()
    net.close()
    ctx.close()
    print(json.dumps({"profile": "saturation-guess network 4-64-64-64-3",
                      "train_steps": train_steps, "batch": 8192,
                      "train_ms_per_step": round(1e3 * (t1 - t0) /
                                                 max(train_steps, 1), 4),
                      "forward_samples": nbig, "forward_tflops": round(tf, 2),
                      "forward_mfma_frac": round(tf / BF16_PEAK_TFLOPS, 5)}))


def bench_batch(ctx, seed, nblocks, steps, warmup, acc_steps):
    bl = ca.batch_blocks(seed, nblocks)
    blocks = [b[0] for b in bl]
    th, cfg = bl[0][1], bl[0][2]
    w0s = [b[3] for b in bl]
    dt = min(cfl_dt(b[0], b[1], b[2], b[3]) for b in bl)
    cells = sum(b.ncells() for b in blocks)
    cols = sum(b.nx2 * b.nx3 for b in blocks)
    log("batch of %d blocks, %d cells, %d columns, dt=%.4f"
        % (nblocks, cells, cols, dt))
    out = {"workload": "uneven-column batch: %d blocks, heights %s, "
                       "footprints %s, seed %d"
                       % (nblocks, sorted(set(b.nx1 for b in blocks)),
                          sorted(set(b.nx2 for b in blocks)), seed),
           "cells": cells, "columns": cols, "dt": dt}
    forms = {}
    for name, prec in (("plain_fp64", ca.PREC_FP64), ("mixed", ca.PREC_FP32)):
        cfg.precision = prec
        ms, stats, tune, med = time_form(ctx, blocks, th, cfg, w0s, dt, steps,
                                    warmup, autotune=True)
        forms[name] = {"ms_per_step": round(ms, 4),
                       "ms_per_step_median_of_20": round(med, 4),
                       "cell_updates_per_s": cells / (ms * 1e-3),
                       "tune": tune,
                       "kernels_avg_ms": {k: round(v[0], 4)
                                          for k, v in stats.items()}}
        log("batch %s: %.3f ms/step, tune %s" % (name, ms, tune))
    us = {}
    for prec in (ca.PREC_FP64, ca.PREC_FP32):
        cfg.precision = prec
        us[prec] = final_state(ctx, blocks, th, cfg, w0s, dt, acc_steps)
    acc = errors(blocks, us[ca.PREC_FP32], us[ca.PREC_FP64])
    acc["steps"] = acc_steps
    forms["mixed"]["error_vs_fp64"] = acc
    out["forms"] = forms
    return out


# Kessler rates of examples/2019-Li-snap/earth-precip.yaml:44-53 and
# water-kessler.yaml:4-18 (s^-1), precipitation fall speed (m/s)
KESSLER_RATES = dict(k_auto=0.01, k_accr=0.02, k_evap=0.1, v_fall=8.0)


def bench_kessler(ctx, blk, th, cfg, w0, dt, steps, warmup, tune):
    """The moist box's chosen form (FP32 fluxes, plain Newton) with Kessler
    warm rain and precipitation sedimentation after the last stage.  With
    the box's reflecting x1 walls the last stage's downward column sweep
    applies the microphysics to each cell as its new state appears, so no
    kernel of its own runs (k_kessler / k_kessler_fill, one read and one
    write of u and w per cell plus psf / psv, 32 N + 16 algorithmic bytes,
    where a run still needs one: periodic x1 sides, the guided adjustment,
    explicit runs).  ms per step, the sweep's event-timed average, and the
    separate kernel's average and HBM rate when it runs."""
    cfg.precision = ca.PREC_FP32
    cfg.satadj = ca.SATADJ_NEWTON
    for k, v in KESSLER_RATES.items():
        setattr(cfg, k, v)
    t = ca.Tune(*[tune[k] for k, _ in ca.Tune._fields_]) if tune else None
    try:
        ms, stats, _, med = time_form(ctx, [blk], th, cfg, [w0], dt, steps,
                                      warmup, tune=t)
    finally:
        for k in KESSLER_RATES:
            setattr(cfg, k, 0.0)
    cells = blk.ncells()
    kk = stats.get("k_kessler", (0.0, 0))
    sw = stats.get("k_implicit_fused", (0.0, 0))
    alg = (32 * blk.N + 16) * cells
    ach = alg / (kk[0] * 1e-3) / 1e9 if kk[0] > 0 else None
    log("kessler: %.3f ms/step, sweep %.4f ms, k_kessler %.4f ms" % (ms, sw[0], kk[0]))
    return {"workload": "moist box %dx%dx%d, FP32 fluxes + plain Newton, "
                        "Kessler rates %s" % (blk.nx1, blk.nx2, blk.nx3,
                                              KESSLER_RATES),
            "ms_per_step": round(ms, 4),
            "ms_per_step_median_of_20": round(med, 4),
            "microphysics": ("in the last stage's downward sweep" if kk[1] == 0
                             else "separate column kernel"),
            "implicit_sweep_ms": round(sw[0], 4),
            "k_kessler_ms": round(kk[0], 4) if kk[1] else None,
            "k_kessler_launches_per_step": round(kk[1], 2),
            "algorithmic_bytes": alg if kk[1] else None,
            "achieved_GBs": round(ach, 1) if ach else None,
            "frac": round(ach / HBM_PEAK_GBS, 4) if ach else None,
            "kernels_avg_ms": {k: round(v[0], 4) for k, v in stats.items()}}


def bench_sphere(ctx, steps, warmup, dt=60.0, nx=(64, 96, 96), acc_steps=20):
    """Config 3: the cubed sphere of six 96 x 96 panels of 64 levels with the
    Held-Suarez forcing and rotation, 3 x 3 implicit x1 columns (hs94.inp),
    fixed step dt, in the plain FP64 form and the FP32-flux form (face
    states and Riemann fluxes of all three directions in FP32, the x1
    decomposition and face pressures in FP64).  ms per step (graph replay),
    per-kernel event averages, the state after the timed steps, and the FP32
    form's error against FP64 after acc_steps steps."""
    blk, th, cfg = ca.pgen_setup(3, *nx)
    sp = ca.pgen_sphere()
    w = ca.pgen_sphere_init(2024, blk, th)
    cells = 6 * blk.ncells()
    sk, sj, si = blk.interior()
    I = (slice(None), slice(None), sk, sj, si)
    log("cubed sphere 6x%dx%dx%d, dt=%.1f s" % (blk.nx2, blk.nx3, blk.nx1, dt))
    forms, finals = {}, {}
    for name, prec in (("plain_fp64", ca.PREC_FP64), ("fp32_faces", ca.PREC_FP32)):
        cfg.precision = prec
        run = ca.SphereRun(ctx, blk, th, cfg, sp)
        try:
            run.set_all(w)
            run.step(dt, warmup, graph=USE_GRAPH)
            t0 = time.perf_counter()
            run.step(dt, steps, graph=USE_GRAPH)
            t1 = time.perf_counter()
            ms = 1e3 * (t1 - t0) / steps
            u1 = run.cons_all()
            run.set_all(w)
            run.step(dt, acc_steps, graph=USE_GRAPH)
            finals[name] = run.cons_all()
            run.set_timing(True)
            nt = max(2, min(steps, 5))
            run.step(dt, nt, graph=False)
            stats = {k: (v[0] / v[1], v[1] / nt) for k, v in run.kernel_stats().items()}
            run.set_timing(False)
        finally:
            run.close()
        vmax = float(np.abs(u1[:, 1:4][:, :, sk, sj, si] / u1[:, 0:1][:, :, sk, sj, si]).max())
        forms[name] = {"ms_per_step": round(ms, 4),
                       "cell_updates_per_s": cells / (ms * 1e-3),
                       "finite": bool(np.isfinite(u1[I]).all()),
                       "max_momentum_over_density": round(vmax, 4),
                       "kernels_avg_ms": {k: round(v[0], 4) for k, v in stats.items()},
                       "kernel_launches_per_step": {k: round(v[1], 2)
                                                    for k, v in stats.items()}}
        log("sphere %s: %.3f ms/step, %.3e cell-updates/s"
            % (name, ms, forms[name]["cell_updates_per_s"]))
    a, b = finals["fp32_faces"][I], finals["plain_fp64"][I]
    names = ["density", "momentum1", "momentum2", "momentum3", "energy"]
    err = {names[n]: float(np.linalg.norm(a[:, n] - b[:, n]) / np.linalg.norm(b[:, n]))
           for n in range(5)}
    err["steps"] = acc_steps
    forms["fp32_faces"]["rel_l2_vs_fp64"] = err
    forms["fp32_faces"]["speedup_vs_plain"] = round(
        forms["plain_fp64"]["ms_per_step"] / forms["fp32_faces"]["ms_per_step"], 3)
    return {"workload": "cubed sphere 6 x %d x %d x %d (R 6.371e6 m, 25 km deep), "
                        "WENO5 + HLLC with panel frames, 3x3 implicit x1, "
                        "Held-Suarez forcing, Coriolis, SSPRK(3,3)"
                        % (blk.nx2, blk.nx3, blk.nx1),
            "cells": cells, "dt": dt, "steps": steps, "forms": forms}


# the launch shapes the autotuner picks between on one MI355X: column solve
# with 4 waves per workgroup on one workgroup per compute unit, sweeps fused
# per group or in two launches; 32 x 8 face tiles marching 16 levels
PROFILE_TUNES = (ca.Tune(4, 256, 256, 3, 16, 1, 0.0),
                 ca.Tune(4, 256, 256, 3, 16, 0, 0.0))


def profile_run(nx, steps):
    """A short, fixed run for rocprofv3: a streaming copy of 1 GiB (the
    calibration of the HBM traffic counters: it reads and writes exactly
    2^30 bytes) three times, then, for each launch shape of PROFILE_TUNES,
    `steps` graph-replayed steps of the dry box in the FP32-flux form, each
    followed by the NewBlockTimeStep reduction."""
    ctx = ca.Context(0)
    nbytes = 1 << 30
    a = ctx.array(shape=(nbytes // 8,), dtype=np.float64)
    b = ctx.array(shape=(nbytes // 8,), dtype=np.float64)
    L = ca.lib()
    for _ in range(3):
        ctx._check(L.cg_memcpy_d2d(ctx.h, b.ptr, a.ptr, nbytes), "d2d")
    a.free()
    b.free()
    blk, th, cfg = ca.pgen_setup(2, *nx)
    w0 = ca.pgen_init(2, 1, blk, th, cfg)
    dt = cfl_dt(blk, th, cfg, w0)
    cfg.precision = ca.PREC_FP32
    ms = []
    for tune in PROFILE_TUNES:
        run = make_run(ctx, [blk], th, cfg, [w0], tune=tune)
        t0 = time.perf_counter()
        for _ in range(steps):
            run.step(dt, 1, graph=USE_GRAPH)
            run.new_dt(0.45)
        t1 = time.perf_counter()
        run.close()
        ms.append(round(1e3 * (t1 - t0) / steps, 4))
    ctx.close()
    print(json.dumps({"profile": "dry box %dx%dx%d FP32 fluxes" % tuple(nx),
                      "steps": steps, "copy_bytes": nbytes,
                      "tunes": [t.as_dict() for t in PROFILE_TUNES],
                      "ms_per_step": ms}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true",
                    help="headline workload only (no moist box, no batch)")
    ap.add_argument("--acc-steps", type=int, default=100)
    ap.add_argument("--pairs", type=int, default=10_000_000)
    ap.add_argument("--train-steps", type=int, default=20000)
    ap.add_argument("--nx", type=int, nargs=3, default=[128, 256, 256],
                    help="nx1 (vertical) nx2 nx3")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--profile", action="store_true",
                    help="short fixed run for rocprofv3 (copy calibration + "
                         "a few steps of the dry box, FP32 fluxes)")
    ap.add_argument("--mlp", action="store_true",
                    help="short fixed run of the network for rocprofv3 "
                         "(pairs, a few training steps, 3 forwards of 8M)")
    ap.add_argument("--sphere", action="store_true",
                    help="only the cubed sphere (config 3)")
    ap.add_argument("--kessler", action="store_true",
                    help="only the moist box with Kessler microphysics")
    ap.add_argument("--lines", action="store_true",
                    help="only the dry box with implicit x2/x3 lines "
                         "(implicit_flag 15)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path here "
              "(the step would shard by horizontal MeshBlock columns, each "
              "GPU owning whole columns, with a 3-cell halo exchange per "
              "stage).  Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    global USE_GRAPH
    USE_GRAPH = not args.no_graph
    if args.profile:
        profile_run(args.nx, max(1, min(args.steps, 5)))
        return
    if args.mlp:
        mlp_run(args.nx, 200)
        return
    if args.sphere:
        ctx = ca.Context(0)
        print(json.dumps(bench_sphere(ctx, max(1, args.steps), args.warmup)))
        ctx.close()
        return
    if args.kessler:
        ctx = ca.Context(0)
        blk, th, cfg = ca.pgen_setup(4, *args.nx)
        w0 = ca.pgen_init(4, 1, blk, th, cfg)
        print(json.dumps(bench_kessler(ctx, blk, th, cfg, w0,
                                       cfl_dt(blk, th, cfg, w0),
                                       max(1, args.steps), args.warmup, None)))
        ctx.close()
        return
    if args.lines:
        ctx = ca.Context(0)
        print(json.dumps(bench_lines(ctx, args.nx, max(1, args.steps),
                                     args.warmup)))
        ctx.close()
        return

    ctx = ca.Context(0)
    blk, cfg, dt, cells, forms, acc = bench_box(ctx, args.nx, args.steps,
                                                args.warmup, args.acc_steps)
    moist = batch = lines = host = sphere = None
    if not args.quick:
        sphere = bench_sphere(ctx, max(2, args.steps // 2), 1)
        lines = bench_lines(ctx, args.nx, max(2, args.steps // 5), 1)
        host = bench_host_solve(ctx, args.nx)
        moist = bench_moist(ctx, args.nx, args.steps, args.warmup,
                            args.acc_steps, args.pairs, args.train_steps)
        batch = bench_batch(ctx, 2024, 32, max(3, args.steps // 4),
                            args.warmup, max(10, args.acc_steps // 5))
    ctx.close()

    ok = acc["max_rel_l2"] <= 1e-5
    chosen = forms["mixed"] if ok else forms["plain_fp64"]
    dom = chosen["dominant"]
    op = chosen["ops"][dom]
    rf = {"bound": "hbm", "op": dom, "kernels": op["kernels"],
          "achieved": op["achieved_GBs"], "peak": HBM_PEAK_GBS,
          "unit": "GB/s", "frac": op["frac"],
          "traffic": traffic_from_profiles(dom, chosen["tune"])}
    # every operation of the chosen form: algorithmic rate, and the counter
    # traffic of one launch against its algorithmic bytes
    rf["ops"] = {}
    for name, o in chosen["ops"].items():
        tr = traffic_from_profiles(name, chosen["tune"])
        rf["ops"][name] = {
            "kernels": o["kernels"], "avg_launch_ms": o["avg_launch_ms"],
            "algorithmic_bytes": o["algorithmic_bytes"],
            "achieved": o["achieved_GBs"], "frac": o["frac"],
            "traffic": tr,
            "traffic_over_algorithmic": (round(tr / o["algorithmic_bytes"], 3)
                                         if tr else None)}
    # the FP64 side of the column solve, and the whole step against the
    # step's algorithmic bytes B_step = 88 N per cell-update (3 stages x
    # 24 N for U_s, U_n, U_{s+1} + 16 N for the adjustment pass)
    if dom == "implicit_solve":
        rf["fp64"] = fp64_from_profiles(chosen["kernels_avg_ms"], cells)
    bstep = 88 * blk.N
    sgb = bstep * chosen["cell_updates_per_s"] / 1e9
    rf["step"] = {"bytes_per_cell_update": bstep, "achieved": round(sgb, 1),
                  "unit": "GB/s", "frac": round(sgb / HBM_PEAK_GBS, 4)}

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
                   "cells": cells, "dt": dt, "graph": USE_GRAPH,
                   "form": "plain FP64" if chosen is forms["plain_fp64"]
                   else "FP32 fluxes + FP64 state"},
        "roofline": rf,
        "cpu_baseline": cpu,
        "forms": forms,
        "moist_box": moist,
        "batch": batch,
        "implicit_lines": lines,
        "implicit_solve_host": host,
        "cubed_sphere": sphere,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
