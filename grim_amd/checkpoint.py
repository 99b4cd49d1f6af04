"""ctypes binding of grim_amd/libgrim_io.so (include/grim_io.h): grim's HDF5
primitive dumps (grid::dump / grid::load, src/grid/grid.cpp:367-381,
482-499) and the torus's restart protocol (torus/problem.cpp:880-906,
src/timestepper/timestepper.cpp:387-441) on host arrays."""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

from . import capi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgrim_io.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        G = C.POINTER(capi.GridConfig)
        L.grim_io_dump.argtypes = [G, C.c_int, C.c_void_p, C.c_char_p, C.c_char_p]
        L.grim_io_dump.restype = C.c_int
        L.grim_io_load.argtypes = [G, C.c_int, C.c_void_p, C.c_char_p, C.c_char_p]
        L.grim_io_load.restype = C.c_int
        L.grim_io_last_error.argtypes = []
        L.grim_io_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _check(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"{what}: {lib().grim_io_last_error().decode()}")


def _planes(cfg, prim):
    n = capi.lib().grim_grid_zones(C.byref(cfg))
    assert prim.dtype == np.float64 and prim.flags.c_contiguous and prim.size % n == 0
    return prim.size // n


def dump(cfg: capi.GridConfig, prim: np.ndarray, file: str, dataset: str = "primitives"):
    """grid::dump: the bulk of the ghosted planes `prim` (dof x zones) as `dataset`."""
    prim = np.ascontiguousarray(prim, dtype=np.float64)
    _check(lib().grim_io_dump(C.byref(cfg), _planes(cfg, prim), prim.ctypes.data, dataset.encode(),
                              os.fsencode(file)), "grim_io_dump")


def load(cfg: capi.GridConfig, prim: np.ndarray, file: str, dataset: str = "primitives") -> np.ndarray:
    """grid::load into the bulk of `prim` (modified in place and returned)."""
    assert prim.dtype == np.float64 and prim.flags.c_contiguous
    _check(lib().grim_io_load(C.byref(cfg), _planes(cfg, prim), prim.ctypes.data, dataset.encode(),
                              os.fsencode(file)), "grim_io_load")
    return prim


def restart_vars_name(time: float) -> str:
    """restartVarsT%08d.h5 with index floor(100 t) (torus/problem.cpp:889-896)."""
    return f"restartVarsT{math.floor(time * 100):08d}.h5"


def prim_vars_name(time: float, every: float) -> str:
    """primVarsT%06d.h5 with index floor(t / WriteDataEveryDt) (torus/problem.cpp:781-806)."""
    return f"primVarsT{math.floor(time / every):06d}.h5"
