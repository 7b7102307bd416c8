"""The NetCDF writer (cg_netcdf_write, host code) against an independent
reader: scipy's NetCDF classic reader must see the reference's layout
(src/outputs/netcdf.cpp:100-420): dimensions time (unlimited), x1, x1f, x2,
x2f, x3, x3f, coordinate values at cell centres and faces, float fields
(time, x1, x2, x3) equal to the interior cells of the block, units /
long_name attributes, and one record per call."""
import numpy as np
import pytest
from scipy.io import netcdf_file

import canoe_amd as ca


def _fields(blk, rng, names=("rho", "press", "theta")):
    shp = blk.shape()[1:]
    return [(n, "u%d" % i, "long %s" % n, rng.normal(size=shp))
            for i, n in enumerate(names)]


def _interior_ijk(blk, a):
    sk, sj, si = blk.interior()
    return np.transpose(a[sk, sj, si], (2, 1, 0))  # (k, j, i) -> (i, j, k)


def test_layout_values_and_records(tmp_path):
    blk, th, cfg = ca.pgen_setup(2, 8, 6, 5)
    blk.x1min, blk.x2min, blk.x3min = 10.0, -300.0, 50.0
    rng = np.random.default_rng(3)
    f0, f1 = _fields(blk, rng), _fields(blk, rng)
    path = str(tmp_path / "box.nc")
    ca.netcdf_write(path, blk, 0.0, f0)
    ca.netcdf_write(path, blk, 1.5, f1, append=True)
    with netcdf_file(path, "r", mmap=False) as nc:
        assert nc.version_byte == 2
        dims = dict(nc.dimensions)
        assert dims == {"time": None, "x1": 8, "x1f": 9, "x2": 6, "x2f": 7,
                        "x3": 5, "x3f": 6}
        assert list(nc.variables["time"][:]) == [0.0, 1.5]
        for d, (n, x0, dx) in enumerate([(8, 10.0, blk.dx1), (6, -300.0, blk.dx2),
                                         (5, 50.0, blk.dx3)]):
            v = nc.variables["x%d" % (d + 1)]
            np.testing.assert_array_equal(
                v[:], (x0 + (np.arange(n) + 0.5) * dx).astype(np.float32))
            vf = nc.variables["x%df" % (d + 1)]
            np.testing.assert_array_equal(
                vf[:], (x0 + np.arange(n + 1) * dx).astype(np.float32))
            assert v.axis == ["Z", "Y", "X"][d].encode()
            assert list(v.domain_decomposition) == [1, n, 1, n]
        assert nc.NumFilesInSet == 1
        for rec, fs in enumerate((f0, f1)):
            for name, units, long_name, a in fs:
                v = nc.variables[name]
                assert v.dimensions == ("time", "x1", "x2", "x3")
                assert v.units == units.encode()
                assert v.long_name == long_name.encode()
                np.testing.assert_array_equal(
                    v[rec], _interior_ijk(blk, a).astype(np.float32))


def test_two_dimensional_block_slices_like_the_reference(tmp_path):
    # test_straka.py reads theta[-1, :, :, 0]: (x1, x2) of the last record
    blk, th, cfg = ca.pgen_setup(1, 12, 20, 1)
    rng = np.random.default_rng(4)
    fs = _fields(blk, rng, names=("temp", "theta"))
    path = str(tmp_path / "straka.nc")
    ca.netcdf_write(path, blk, 900.0, fs)
    with netcdf_file(path, "r", mmap=False) as nc:
        assert "x3f" not in nc.dimensions and nc.dimensions["x3"] == 1
        th_nc = nc.variables["theta"][-1, :, :, 0]
        assert th_nc.shape == (12, 20)
        sk, sj, si = blk.interior()
        np.testing.assert_array_equal(
            th_nc, fs[1][3][0, sj, si].T.astype(np.float32))


def test_append_needs_the_same_layout_and_bad_paths_fail(tmp_path):
    blk, th, cfg = ca.pgen_setup(2, 8, 6, 5)
    rng = np.random.default_rng(5)
    path = str(tmp_path / "a.nc")
    ca.netcdf_write(path, blk, 0.0, _fields(blk, rng))
    with pytest.raises(ca.CanoeGpuError) as e:
        ca.netcdf_write(path, blk, 1.0, _fields(blk, rng, names=("rho",)),
                        append=True)
    assert e.value.code == ca.ERR_ARG
    with pytest.raises(ca.CanoeGpuError) as e:
        ca.netcdf_write(str(tmp_path / "missing" / "b.nc"), blk, 0.0,
                        _fields(blk, rng))
    assert e.value.code == ca.ERR_IO
    with pytest.raises(ca.CanoeGpuError):
        ca.netcdf_write(str(tmp_path / "new.nc"), blk, 0.0, _fields(blk, rng),
                        append=True)
