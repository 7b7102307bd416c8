"""bench.py's reporting helpers on the CPU: the algorithmic byte model, the
per-operation roofline from kernel averages, the committed traffic and
FP64 summaries it reads, and the one-GPU guard (no GPU is used)."""
import json
import os
import subprocess
import sys

import pytest

import bench
import canoe_amd as ca

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def dry_box():
    blk, th, cfg = ca.pgen_setup(2, 0, 0, 0)
    return blk, cfg


def test_algorithmic_bytes_of_the_dry_box():
    blk, cfg = dry_box()
    cells = 128 * 256 * 256
    assert blk.N == 5 and blk.ncells() == cells
    alg = bench.algorithmic(blk, cfg)
    # read w once, write du once: (8 + 8) N bytes per cell
    assert alg["flux_divergence"] == 80 * cells
    # w, du, u in and u, w out, u1 at two stages of three, psf and psv
    per_cell = 8 * (5 + 2.0 / 3.0) * 5 + 16
    assert alg["implicit_solve"] == int(round(per_cell * cells))
    assert alg["update"] == 160 * cells


def test_roofline_fraction_and_dominant_operation():
    blk, cfg = dry_box()
    alg = bench.algorithmic(blk, cfg)
    # (average ms per launch, launches per step)
    stats = {"k_mix_fused": (0.5, 3.0), "k_implicit_fused": (1.25, 3.0),
             "k_bc_x2": (0.02, 3.0)}
    dom, ops = bench.roofline(blk, cfg, stats)
    assert dom == "implicit_solve"
    assert "update" not in ops  # no k_update: the sweep applies the stage
    f = ops["flux_divergence"]
    assert f["kernels"] == ["k_mix_fused"]
    assert f["ms_per_step"] == pytest.approx(1.5)
    gbs = alg["flux_divergence"] / 0.5e-3 / 1e9
    assert f["achieved_GBs"] == pytest.approx(gbs, abs=0.05)
    assert f["frac"] == pytest.approx(gbs / bench.HBM_PEAK_GBS, abs=1e-4)


def test_committed_traffic_summary_matches_the_profile_tables():
    d = json.load(open(os.path.join(ROOT, "profiles", "traffic_box256.json")))
    pk = d["per_kernel_bytes"]
    # the copy kernel of 2^30 bytes calibrates both counters
    assert pk["cg::k_copy16"] == pytest.approx(2 ** 31, rel=1e-3)
    assert d["implicit_solve"] == pk["cg::k_implicit_fused_batch<0, 5, 4>"]
    assert d["implicit_solve_two_launches"] == (
        pk["cg::k_implicit_fwd_batch<0, 5, 4>"]
        + pk["cg::k_implicit_bwd_update_batch<0, 5, 4>"])
    assert d["flux_divergence"] == pk["cg::k_mix_fused_batch<0, 5, 2, 32, 8>"]
    assert bench.traffic_from_profiles("implicit_solve", {"implicit_fused": 1}) \
        == d["implicit_solve"]
    assert bench.traffic_from_profiles("implicit_solve", {"implicit_fused": 0}) \
        == d["implicit_solve_two_launches"]
    # the archive round trip (2 x 30 doubles per cell) is most of the solve's
    # traffic: at least the archive plus one read and one write of w
    cells = 128 * 256 * 256
    assert d["implicit_solve"] > (2 * 30 + 2 * 5) * 8 * cells


def test_fp64_rate_from_the_committed_operation_counts():
    cells = 128 * 256 * 256
    r = bench.fp64_from_profiles({"k_implicit_fused": 1.25}, cells)
    assert r is not None and r["unit"] == "TFLOP/s"
    assert r["achieved"] == pytest.approx(r["flop_per_cell"] * cells / 1.25e-3 / 1e12,
                                          rel=1e-3)


def test_more_than_one_gpu_is_refused_with_the_sharding_note():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode != 0
    assert "one GPU" in (p.stdout + p.stderr)
