"""ctypes binding of grim_amd/libgrim_io.so (include/grim_io.h): grim's HDF5
primitive dumps (grid::dump / grid::load, src/grid/grid.cpp:367-381,
482-499) and the torus's restart protocol (torus/problem.cpp:880-906,
src/timestepper/timestepper.cpp:387-441) on host arrays, and the VTS dumps
of the dump variables (grid::dumpVTS, grid.cpp:383-480)."""
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
        L.grim_io_dump_vts.argtypes = [G, C.c_int, C.POINTER(C.c_char_p), C.c_void_p, C.c_char_p]
        L.grim_io_dump_vts.restype = C.c_int
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


def dump_vts_name(time: float, every: float) -> str:
    """dump%06d.vts with index floor(t / WriteDataEveryDt) (torus/problem.cpp:781-804)."""
    return f"dump{math.floor(time / every):06d}.vts"


def dump_vts(cfg: capi.GridConfig, names, data: np.ndarray, file: str):
    """grid::dumpVTS: data is (len(names) + 3, N1*N2*N3), the named planes
    then x1, x2, x3 of the points (TimeStepper.dump_vars' layout)."""
    data = np.ascontiguousarray(data, dtype=np.float64)
    nb = cfg.N1 * cfg.N2 * cfg.N3
    if data.shape != (len(names) + 3, nb):
        raise ValueError(f"dump_vts: data shape {data.shape}, expected {(len(names) + 3, nb)}")
    arr = (C.c_char_p * len(names))(*[n.encode() for n in names])
    _check(lib().grim_io_dump_vts(C.byref(cfg), len(names), arr, data.ctypes.data, os.fsencode(file)),
           "grim_io_dump_vts")


def read_vts(file: str):
    """Reads a StructuredGrid written by dump_vts: (extent, points (N, 3),
    {name: array (N,)}), N = N1*N2*N3 with X1 fastest."""
    import re

    with open(file, "rb") as f:
        raw = f.read()
    head, _, body = raw.partition(b'<AppendedData encoding="raw">')
    start = body.index(b"_") + 1
    text = head.decode()
    if 'header_type="UInt64"' not in text or 'byte_order="LittleEndian"' not in text:
        raise ValueError(f"{file}: not a little-endian UInt64-header VTS")
    extent = [int(v) for v in re.search(r'WholeExtent="([^"]*)"', text).group(1).split()]

    def block(offset):
        o = start + offset
        size = int(np.frombuffer(body, np.uint64, 1, o)[0])
        return np.frombuffer(body, np.float64, size // 8, o + 8)

    pts = re.search(r'<Points>\s*<DataArray[^>]*offset="(\d+)"', text)
    points = block(int(pts.group(1))).reshape(-1, 3)
    arrays = {m.group(1): block(int(m.group(2)))
              for m in re.finditer(r'<DataArray type="Float64" Name="([^"]+)"[^>]*offset="(\d+)"', text)}
    return extent, points, arrays
