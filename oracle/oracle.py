"""ctypes wrapper of the grim oracle (TEST INFRASTRUCTURE ONLY).

The oracle is the CPU FP64 restatement of grim's per-zone implicit solve
(`oracle/grim_oracle.hpp`).  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module; the product
(`grim_amd`) never does.

All arrays are structure-of-arrays: shape ``(planes, n)`` float64, C order.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgrim_oracle.so")

CLOSURE_LINEAR_MODES = 0
CLOSURE_PER_ZONE = 1
CLOSURE_TORUS = 2
NEWTON_PER_ZONE = 0
NEWTON_GLOBAL = 1
MAX_VARS = 10


class Params(C.Structure):
    """Mirror of ``grim_oracle_params`` (grim's ``namespace params``)."""

    _fields_ = [
        ("conduction", C.c_int),
        ("viscosity", C.c_int),
        ("highOrderTermsConduction", C.c_int),
        ("highOrderTermsViscosity", C.c_int),
        ("adiabaticIndex", C.c_double),
        ("rhoFloorInFluidElement", C.c_double),
        ("uFloorInFluidElement", C.c_double),
        ("bSqrFloorInFluidElement", C.c_double),
        ("temperatureFloorInFluidElement", C.c_double),
        ("ConductionAlpha", C.c_double),
        ("ViscosityAlpha", C.c_double),
        ("closure", C.c_int),
        ("tau0", C.c_double),
        ("ConductionClosureFactor", C.c_double),
        ("ViscosityClosureFactor", C.c_double),
        ("maxNonLinearIter", C.c_int),
        ("maxLineSearchIters", C.c_int),
        ("nonlinearsolve_atol", C.c_double),
        ("JacobianAssembleEpsilon", C.c_double),
        ("linesearchfloor", C.c_double),
    ]

    @property
    def nf(self) -> int:
        return 5 + self.conduction + self.viscosity

    @property
    def dof(self) -> int:
        return 8 + self.conduction + self.viscosity


class _Batch(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("prim", C.c_void_p),
        ("geom", C.c_void_p),
        ("rhs", C.c_void_p),
        ("primOld", C.c_void_p),
        ("primCoef", C.c_void_p),
        ("closure", C.c_void_p),
        ("radius", C.c_void_p),
        ("dtPrime", C.c_double),
        ("residualMask", C.c_void_p),
    ]


def build() -> str:
    """Compile the oracle with its Makefile (g++, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, i64, d = C.c_void_p, C.c_int64, C.c_double
        P = C.POINTER(Params)
        B = C.POINTER(_Batch)
        L.grim_oracle_default_params.argtypes = [P, C.c_int]
        L.grim_oracle_mks_geometry.argtypes = [i64, vp, vp, d, d, vp, vp, vp, vp, vp]
        L.grim_oracle_pack_geometry.argtypes = [i64, vp, vp, vp, vp, vp]
        L.grim_oracle_conserved.argtypes = [P, i64, vp, vp, vp, vp, vp]
        L.grim_oracle_residual.argtypes = [P, B, vp]
        L.grim_oracle_known_root_rhs.argtypes = [P, B, vp]
        L.grim_oracle_element.argtypes = [P, i64, vp, vp, vp, vp, vp]
        L.grim_oracle_residual_pieces.argtypes = [P, B, vp, vp, vp]
        L.grim_oracle_fd_jacobian.argtypes = [P, B, vp]
        L.grim_oracle_dual_jacobian.argtypes = [P, B, vp]
        L.grim_oracle_newton.argtypes = [P, B, C.c_int, vp, vp, vp, C.c_int]
        L.grim_oracle_newton.restype = i64
        L.grim_oracle_lu_solve.argtypes = [C.c_int, vp, vp]
        L.grim_oracle_lu_solve.restype = C.c_int
        L.grim_oracle_count_ops.argtypes = [P, B, vp, vp]
        _lib = L
    return _lib


def _ptr(a):
    if a is None:
        return None
    assert a.dtype in (np.float64, np.int32, np.int64) and a.flags.c_contiguous
    return a.ctypes.data


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def default_params(emhd: bool = False) -> Params:
    p = Params()
    lib().grim_oracle_default_params(C.byref(p), 1 if emhd else 0)
    return p


def mks_geometry(X1, X2, a, h):
    """grim MKS metric (no cylindrification) at points (X1, X2).

    Returns dict with gCov (16, n), gCon (16, n), g (n), alpha (n) and
    gamma (64, n) = Gamma^mu_{nu lambda} by grim's central differences.
    """
    X1 = _f64(X1)
    X2 = _f64(X2)
    n = X1.size
    out = {
        "gCov": np.empty((16, n)),
        "gCon": np.empty((16, n)),
        "g": np.empty(n),
        "alpha": np.empty(n),
        "gamma": np.empty((64, n)),
    }
    lib().grim_oracle_mks_geometry(n, _ptr(X1), _ptr(X2), a, h, *(_ptr(out[k]) for k in ("gCov", "gCon", "g", "alpha", "gamma")))
    return out


def pack_geometry(gCov, gCon, g, alpha):
    """16-plane batch geometry (gCov upper triangle, gCon row 0, alpha, g)."""
    gCov, gCon, g, alpha = map(_f64, (gCov, gCon, g, alpha))
    n = g.size
    out = np.empty((16, n))
    lib().grim_oracle_pack_geometry(n, _ptr(gCov), _ptr(gCon), _ptr(g), _ptr(alpha), _ptr(out))
    return out


def conserved(p: Params, prim, geom, closure=None, radius=None):
    prim, geom, closure, radius = map(_f64, (prim, geom, closure, radius))
    n = prim.shape[1]
    out = np.empty((p.dof, n))
    lib().grim_oracle_conserved(C.byref(p), n, _ptr(prim), _ptr(geom), _ptr(closure), _ptr(radius), _ptr(out))
    return out


class Batch:
    """One batch of independent zones in the oracle's layout."""

    def __init__(self, prim, geom, rhs, primOld=None, primCoef=None, closure=None, radius=None, dtPrime=1.0,
                 residualMask=None):
        self.prim = np.array(prim, dtype=np.float64, order="C", copy=True)
        self.geom = _f64(geom)
        self.rhs = _f64(rhs)
        self.primOld = _f64(primOld)
        self.primCoef = _f64(primCoef)
        self.closure = _f64(closure)
        self.radius = _f64(radius)
        self.dtPrime = float(dtPrime)
        self.residualMask = _f64(residualMask)
        self.n = self.prim.shape[1]

    def _c(self):
        return _Batch(self.n, _ptr(self.prim), _ptr(self.geom), _ptr(self.rhs), _ptr(self.primOld),
                      _ptr(self.primCoef), _ptr(self.closure), _ptr(self.radius), self.dtPrime,
                      _ptr(self.residualMask))

    def copy(self):
        return Batch(self.prim.copy(), self.geom, self.rhs, self.primOld, self.primCoef, self.closure, self.radius,
                     self.dtPrime, self.residualMask)


ELEMENT_FIELDS = (["uCon%d" % i for i in range(4)] + ["uCov%d" % i for i in range(4)]
                  + ["bCon%d" % i for i in range(4)] + ["bCov%d" % i for i in range(4)]
                  + ["bSqr", "q", "deltaP", "pressure", "rho", "u", "tau", "chi", "nu"]
                  + ["T%d%d" % (m, k) for m in range(4) for k in range(4)])


def element(p: Params, prim, geom, closure=None, radius=None):
    """fluidElement::set internals, dict name -> (n,) array (see ELEMENT_FIELDS)."""
    prim, geom, closure, radius = map(_f64, (prim, geom, closure, radius))
    n = prim.shape[1]
    out = np.empty((len(ELEMENT_FIELDS), n))
    lib().grim_oracle_element(C.byref(p), n, _ptr(prim), _ptr(geom), _ptr(closure), _ptr(radius), _ptr(out))
    return dict(zip(ELEMENT_FIELDS, out))


def residual_pieces(p: Params, b: Batch):
    """(U, Si, Std): conserved variables, implicit and time-derivative sources."""
    U, Si, Std = (np.empty((p.nf, b.n)) for _ in range(3))
    cb = b._c()
    lib().grim_oracle_residual_pieces(C.byref(p), C.byref(cb), _ptr(U), _ptr(Si), _ptr(Std))
    return U, Si, Std


def known_root_rhs(p: Params, b: Batch):
    """rhs making b.prim an exact root of the batched residual."""
    out = np.empty((p.nf, b.n))
    tmp = Batch(b.prim, b.geom, np.zeros((p.nf, b.n)), b.primOld, b.primCoef, b.closure, b.radius, b.dtPrime,
                b.residualMask)
    cb = tmp._c()
    lib().grim_oracle_known_root_rhs(C.byref(p), C.byref(cb), _ptr(out))
    return out


def residual(p: Params, b: Batch):
    out = np.empty((p.nf, b.n))
    cb = b._c()
    lib().grim_oracle_residual(C.byref(p), C.byref(cb), _ptr(out))
    return out


def fd_jacobian(p: Params, b: Batch):
    out = np.empty((p.nf * p.nf, b.n))
    cb = b._c()
    lib().grim_oracle_fd_jacobian(C.byref(p), C.byref(cb), _ptr(out))
    return out


def dual_jacobian(p: Params, b: Batch):
    out = np.empty((p.nf * p.nf, b.n))
    cb = b._c()
    lib().grim_oracle_dual_jacobian(C.byref(p), C.byref(cb), _ptr(out))
    return out


def newton(p: Params, b: Batch, mode: int = NEWTON_PER_ZONE, nthreads: int = 0):
    """Solve every zone of ``b`` in place; returns (unconverged, iters, finalL2, resEvals)."""
    iters = np.zeros(b.n, dtype=np.int32)
    l2 = np.zeros(b.n)
    ev = np.zeros(b.n, dtype=np.int32)
    cb = b._c()
    bad = lib().grim_oracle_newton(C.byref(p), C.byref(cb), mode, _ptr(iters), _ptr(l2), _ptr(ev), nthreads)
    return int(bad), iters, l2, ev


def lu_solve(A, b):
    A = np.array(A, dtype=np.float64, order="C", copy=True)
    x = np.array(b, dtype=np.float64, order="C", copy=True)
    info = lib().grim_oracle_lu_solve(A.shape[0], _ptr(A), _ptr(x))
    return x, info


def count_ops(p: Params, b: Batch):
    """(residual_ops, lu_ops) as arrays [add, mul, div, sqrt, other]."""
    r = np.zeros(5, dtype=np.int64)
    l = np.zeros(5, dtype=np.int64)
    cb = b._c()
    lib().grim_oracle_count_ops(C.byref(p), C.byref(cb), _ptr(r), _ptr(l))
    return r, l
