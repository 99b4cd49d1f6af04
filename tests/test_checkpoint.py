"""grim's HDF5 primitive dumps and restart files (grim_amd/libgrim_io.so,
include/grim_io.h; grid::dump / grid::load, src/grid/grid.cpp:367-381,
482-499).

CPU: dump -> load round trips bit for bit on 1D, 2D and 3D grids (ghost zones
untouched by load); the file holds one dataset in PETSc's DMDA natural order,
checked with an independent reader written here against the HDF5 C library
(dims (N3,) (N2,) N1, dof; the variable fastest); shape mismatches and
missing files are errors with a message; the restart and output file names
follow torus/problem.cpp.
GPU: a run stopped by the wall-clock limit (restart files written) and
continued from them matches the uninterrupted run.
PETSc's own reader cannot run here, so agreement with files that PETSc
writes rests on this layout (parity unpinned beyond it).
"""
import ctypes as C
import os

import numpy as np
import pytest

from grim_amd import capi
from grim_amd import checkpoint as ck
from grim_amd.timestepper import grid_config


def hdf5():
    for cand in (os.environ.get("GRIM_HDF5_LIB"), "libhdf5.so", "/opt/conda/lib/libhdf5.so"):
        if not cand:
            continue
        try:
            return C.CDLL(cand)
        except OSError:
            continue
    pytest.skip("no HDF5 C library on this machine")


def read_dataset(path, name):
    """Shape and values of a FP64 dataset, through the HDF5 C API."""
    H = hdf5()
    H.H5open()
    hid = C.c_int64
    H.H5Fopen.restype = H.H5Dopen2.restype = H.H5Dget_space.restype = hid
    H.H5Fopen.argtypes = [C.c_char_p, C.c_uint, hid]
    H.H5Dopen2.argtypes = [hid, C.c_char_p, hid]
    H.H5Dget_space.argtypes = [hid]
    H.H5Sget_simple_extent_ndims.argtypes = [hid]
    H.H5Sget_simple_extent_dims.argtypes = [hid, C.c_void_p, C.c_void_p]
    H.H5Dread.argtypes = [hid, hid, hid, hid, hid, C.c_void_p]
    for f in ("H5Sclose", "H5Dclose", "H5Fclose"):
        getattr(H, f).argtypes = [hid]
    f = H.H5Fopen(os.fsencode(path), 0, 0)
    assert f >= 0
    d = H.H5Dopen2(f, name.encode(), 0)
    assert d >= 0
    s = H.H5Dget_space(d)
    nd = H.H5Sget_simple_extent_ndims(s)
    dims = (C.c_ulonglong * nd)()
    H.H5Sget_simple_extent_dims(s, dims, None)
    shape = tuple(dims)
    out = np.empty(shape)
    native_double = hid.in_dll(H, "H5T_NATIVE_DOUBLE_g").value
    assert H.H5Dread(d, native_double, 0, 0, 0, out.ctypes.data) >= 0
    H.H5Sclose(s)
    H.H5Dclose(d)
    H.H5Fclose(f)
    return out


GRIDS = {
    "1d": dict(N1=12, N2=1, N3=1, dim=1),
    "2d": dict(N1=10, N2=7, N3=1, dim=2),
    "3d": dict(N1=6, N2=5, N3=4, dim=3),
}


@pytest.mark.parametrize("case", list(GRIDS))
def test_dump_load_round_trip_and_layout(tmp_path, case):
    hdf5()
    cfg = grid_config(**GRIDS[case])
    n = capi.lib().grim_grid_zones(C.byref(cfg))
    dof = 10
    rng = np.random.default_rng(7)
    prim = rng.standard_normal((dof, n))
    path = str(tmp_path / "primVarsT000000.h5")
    ck.dump(cfg, prim, path)
    # independent reader: PETSc's DMDA natural order, variable fastest
    data = read_dataset(path, "primitives")
    g = cfg.numGhost
    N = [cfg.N1, cfg.N2, cfg.N3][:cfg.dim]
    assert data.shape == tuple(reversed(N)) + (dof,)
    shape = [cfg.N3 + (2 * g if cfg.dim > 2 else 0), cfg.N2 + (2 * g if cfg.dim > 1 else 0), cfg.N1 + 2 * g]
    planes = prim.reshape(dof, *shape)
    bulk = planes[:, g:g + cfg.N3, :, :] if cfg.dim > 2 else planes
    bulk = bulk[:, :, g:g + cfg.N2, :] if cfg.dim > 1 else bulk
    bulk = bulk[:, :, :, g:g + cfg.N1]
    want = np.moveaxis(bulk, 0, -1).reshape(data.shape)
    assert np.array_equal(data, want)
    # load into a NaN-filled grid: the bulk returns bit for bit, ghosts stay NaN
    back = np.full_like(prim, np.nan)
    ck.load(cfg, back, path)
    ghost = np.isnan(back)
    assert np.array_equal(back[~ghost], prim[~ghost])
    assert ghost.sum() == dof * (n - cfg.N1 * cfg.N2 * cfg.N3)


def test_load_rejects_other_shapes_and_missing_files(tmp_path):
    hdf5()
    cfg = grid_config(N1=8, N2=8)
    n = capi.lib().grim_grid_zones(C.byref(cfg))
    path = str(tmp_path / "a.h5")
    ck.dump(cfg, np.zeros((8, n)), path)
    other = grid_config(N1=8, N2=6)
    with pytest.raises(RuntimeError, match="does not have the grid's shape"):
        ck.load(other, np.zeros((8, capi.lib().grim_grid_zones(C.byref(other)))), path)
    with pytest.raises(RuntimeError, match="does not have the grid's shape"):
        ck.load(cfg, np.zeros((10, n)), path)
    with pytest.raises(RuntimeError, match="no dataset"):
        ck.load(cfg, np.zeros((8, n)), path, dataset="nothing")
    with pytest.raises(RuntimeError, match="cannot open"):
        ck.load(cfg, np.zeros((8, n)), str(tmp_path / "missing.h5"))


def test_output_file_names():
    # torus/problem.cpp:781-806 (WriteIdx = floor(t / WriteDataEveryDt), 6 digits)
    # and :889-896 (floor(100 t), 8 digits)
    assert ck.prim_vars_name(4.1, 2.) == "primVarsT000002.h5"
    assert ck.restart_vars_name(12.345) == "restartVarsT00001234.h5"


@pytest.mark.gpu
def test_timestepper_dump_and_load(tmp_path):
    """TimeStepper.dump / load: a torus state written after one step and read
    into a fresh context returns bit for bit in the bulk; load refills the
    periodic (X3) ghost zones from the bulk, as grid::load does."""
    import bench
    from grim_amd.timestepper import TimeStepper

    hdf5()
    cfg = bench.torus_config(48, 24, 4)
    p = bench.torus_params_emhd()
    a = TimeStepper(p, cfg)
    a.initTorus(capi.torus_params())
    a.set_time(0., 0.005)
    a.timeStep()
    path = str(tmp_path / "primVarsT000000.h5")
    a.dump(path)
    want = a.primOld.reshape(p.dof, *cfg.shape)
    a.close()
    b = TimeStepper(p, cfg)
    b.load(path)
    got = b.primOld.reshape(p.dof, *cfg.shape)
    b.close()
    g = cfg.numGhost
    bulk = (slice(None), slice(g, -g), slice(g, -g), slice(g, -g))
    assert np.array_equal(got[bulk], want[bulk])
    # X3 is periodic: its ghost planes are the periodic images of the bulk
    assert np.array_equal(got[:, :g, g:-g, g:-g], got[:, -2 * g:-g, g:-g, g:-g])
