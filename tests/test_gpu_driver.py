"""The C++ host driver (canoe_amd/lib/canoe_step) runs the reference's main
loop over the C-ABI: it must finish, conserve mass to round-off with
reflecting / periodic walls and report zone-cycles."""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "canoe_amd", "lib", "canoe_step")


@pytest.mark.parametrize("args", [["2", "--nx", "32", "16", "16"],
                                  ["1", "--nx", "16", "64", "1"],
                                  ["2", "--nx", "32", "16", "16", "--fp32"],
                                  ["2", "--nx", "32", "16", "16",
                                   "--implicit-flag", "11"]])
def test_cpp_driver(args):
    print("canoe_step", " ".join(args), flush=True)
    r = subprocess.run([EXE] + args + ["--nlim", "6", "--out", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = r.stdout
    nx = [int(a) for a in args[2:5]]
    m = re.search(r"zone-cycles = (\d+)", out)
    assert m and int(m.group(1)) == nx[0] * nx[1] * nx[2] * 6
    v0 = re.search(r"var 0: (\S+) -> (\S+)", out)
    a, b = float(v0.group(1)), float(v0.group(2))
    assert abs(b - a) <= 1e-13 * abs(a)
    assert "zone-cycles/wall_second" in out


def test_cpp_driver_stops_at_tlim():
    """The device main loop cuts the last step to end at tlim."""
    r = subprocess.run([EXE, "2", "--nx", "32", "16", "16", "--tlim", "3.0",
                        "--cfl", "0.45", "--nlim", "1000", "--out", "5"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    last = re.findall(r"cycle=(\d+) time=(\S+) dt=(\S+)", r.stdout)[-1]
    assert abs(float(last[1]) - 3.0) <= 1e-12
    assert int(last[0]) > 1


def test_cpp_driver_history_and_debug_files(tmp_path):
    """--hst / --dbg write the reference's text layouts: one history row per
    diagnostics point (mass conserved to round-off) and N rows of extremes."""
    base = str(tmp_path / "box")
    r = subprocess.run([EXE, "2", "--nx", "32", "16", "16", "--nlim", "6",
                        "--out", "2", "--cfl", "0.45", "--hst", "--dbg",
                        "--basename", base], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    hst = open(base + ".hst").read().splitlines()
    assert hst[0] == "# Athena++ history data" and "[10]=tot-E" in hst[1]
    rows = [list(map(float, ln.split())) for ln in hst[2:]]
    assert len(rows) == 4  # cycles 0, 2, 4 and the final state
    assert [round(rw[0] > 0) for rw in rows] == [0, 1, 1, 1]
    mass = [rw[2] for rw in rows]
    assert max(abs(m - mass[0]) for m in mass) <= 1e-13 * mass[0]
    dbg = open(base + ".dbg").read().splitlines()
    assert dbg[0] == "# Athena++ debug data"
    assert len(dbg) == 2 + 4 * 5 and " --   4  -- " in dbg[6]


def test_cpp_driver_netcdf_dumps(tmp_path):
    """--nc writes the primitive variables and (temp, theta) in the layout of
    the reference's netcdf output; the Straka run to 900 s read back the way
    examples/2019-Li-snap/test_straka.py reads it (theta[-1, :, :, 0])."""
    from scipy.io import netcdf_file
    base = str(tmp_path / "straka")
    r = subprocess.run([EXE, "1", "--tlim", "900", "--cfl", "0.45", "--nlim",
                        "100000", "--out", "2500", "--nc", "--basename", base],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    with netcdf_file(base + ".uov.nc", "r", mmap=False) as nc:
        t = nc.variables["time"][:]
        # cycles 0, 2500, 5000 and the last (t = 900 s)
        assert t[0] == 0.0 and t[-1] == 900.0 and len(t) == 4
        theta = nc.variables["theta"][-1, :, :, 0].astype(np.float64)
        assert theta.shape == (64, 512)
        tp = theta - 300.0
        assert -10.0 <= tp.min() <= -8.3 and tp.max() < 0.5
        # the initial record: the bubble is -15 K of temperature at 3 km,
        # i.e. theta' = -15 (p0 / p)^(R/cp) = -15 x 300 / 270.7 = -16.6 K
        th0 = nc.variables["theta"][0, :, :, 0].astype(np.float64)
        assert -16.8 <= th0.min() - 300.0 <= -16.4
    with netcdf_file(base + ".prim.nc", "r", mmap=False) as nc:
        assert [v for v in ("rho", "press", "vel1", "vel2", "vel3")
                if v not in nc.variables] == []
        rho = nc.variables["rho"][-1, :, :, 0]
        assert rho.shape == (64, 512) and np.all(rho > 0)
        assert nc.variables["x2"][:].shape == (512,)


def test_cpp_driver_cubed_sphere():
    """Config 3 from the C++ driver: six panels advanced by fixed steps."""
    r = subprocess.run([EXE, "3", "--nx", "16", "8", "8", "--nlim", "4", "--out", "2",
                        "--dt", "100"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    m = re.search(r"zone-cycles = (\d+)", r.stdout)
    assert m and int(m.group(1)) == 6 * 16 * 8 * 8 * 4
    s = re.search(r"panel density sum \(initial -> final\): (\S+) -> (\S+)", r.stdout)
    a, b = float(s.group(1)), float(s.group(2))
    assert np.isfinite(b) and abs(b - a) <= 1e-3 * abs(a)


def test_cpp_driver_kessler():
    """--kessler on the moist box: the run finishes and conserves the total
    mass (dry air plus the three water species; the implicit write-back
    moves mass between dry air and water, forward_backward.hpp:101-124)."""
    r = subprocess.run([EXE, "4", "--nx", "20", "12", "10", "--kessler", "--nlim", "4",
                        "--out", "2", "--cfl", "0.45"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    a = b = 0.0
    for n in range(4):
        v = re.search(r"var %d: (\S+) -> (\S+)" % n, r.stdout)
        a += float(v.group(1))
        b += float(v.group(2))
    assert abs(b - a) <= 1e-13 * abs(a)
