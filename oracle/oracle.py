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
CLOSURE_SHOCK_TESTS = 3
CLOSURE_BUOYANCY = 4
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


METRIC_MINKOWSKI = 0
METRIC_MKS = 1
BC_PERIODIC = 0
BC_OUTFLOW = 1
BC_MIRROR = 2
RECON_MINMOD = 0
RECON_WENO5 = 1
RECON_PPM = 2
RIEMANN_HLL = 0
RIEMANN_LLF = 1
LOC_PLANES = 24


class Grid(C.Structure):
    """Mirror of ``grim_oracle_grid`` (grid, geometry and scheme switches)."""

    _fields_ = [
        ("N1", C.c_int), ("N2", C.c_int), ("N3", C.c_int), ("dim", C.c_int), ("numGhost", C.c_int),
        ("X1Start", C.c_double), ("X1End", C.c_double), ("X2Start", C.c_double),
        ("X2End", C.c_double), ("X3Start", C.c_double), ("X3End", C.c_double),
        ("metric", C.c_int), ("blackHoleSpin", C.c_double), ("hSlope", C.c_double),
        ("DerefineThetaHorizon", C.c_int), ("DoCylindrify", C.c_int),
        ("X1cyl", C.c_double), ("X2cyl", C.c_double),
        ("boundaryLeft", C.c_int), ("boundaryRight", C.c_int), ("boundaryBottom", C.c_int),
        ("boundaryTop", C.c_int), ("boundaryBack", C.c_int), ("boundaryFront", C.c_int),
        ("reconstruction", C.c_int), ("riemannSolver", C.c_int), ("slopeLimTheta", C.c_double),
        ("CourantFactor", C.c_double), ("maxDtIncrement", C.c_double),
        ("residualScaling", C.c_int),
    ]

    @property
    def shape(self):
        """(N3T, N2T, N1T): numpy shape of one plane (X1 fastest)."""
        g = self.numGhost
        return (self.N3 + (2 * g if self.dim > 2 else 0), self.N2 + (2 * g if self.dim > 1 else 0),
                self.N1 + 2 * g)

    @property
    def n(self) -> int:
        a, b, c = self.shape
        return a * b * c

    def bulk(self):
        """Index expression selecting the bulk of a (…, N3T, N2T, N1T) array."""
        g = self.numGhost
        sl = [slice(g, g + self.N1)]
        sl.insert(0, slice(g, g + self.N2) if self.dim > 1 else slice(None))
        sl.insert(0, slice(g, g + self.N3) if self.dim > 2 else slice(None))
        return (Ellipsis, *sl)


class _Geometry(C.Structure):
    _fields_ = [("center", C.c_void_p), ("faceX1", C.c_void_p), ("faceX2", C.c_void_p),
                ("gamma", C.c_void_p)]


class StepStats(C.Structure):
    _fields_ = [("dt", C.c_double), ("unconverged", C.c_int64 * 2), ("maxIterations", C.c_int32 * 2),
                ("nonfinite", C.c_int64 * 2)]


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
        L.grim_oracle_fluxes.argtypes = [P, i64, vp, vp, vp, vp, C.c_int, vp]
        L.grim_oracle_residual_pieces.argtypes = [P, B, vp, vp, vp]
        L.grim_oracle_fd_jacobian.argtypes = [P, B, vp]
        L.grim_oracle_dual_jacobian.argtypes = [P, B, vp]
        L.grim_oracle_newton.argtypes = [P, B, C.c_int, vp, vp, vp, C.c_int]
        L.grim_oracle_newton.restype = i64
        L.grim_oracle_lu_solve.argtypes = [C.c_int, vp, vp]
        L.grim_oracle_lu_solve.restype = C.c_int
        L.grim_oracle_count_ops.argtypes = [P, B, vp, vp]
        L.grim_oracle_count_live_ops.argtypes = [P, B, vp]
        L.grim_oracle_set_threads.argtypes = [C.c_int]
        Gp = C.POINTER(Grid)
        Gm = C.POINTER(_Geometry)
        L.grim_oracle_grid_zones.argtypes = [Gp]
        L.grim_oracle_grid_zones.restype = i64
        L.grim_oracle_grid_xcoords.argtypes = [Gp, vp]
        L.grim_oracle_build_geometry.argtypes = [Gp, Gm]
        L.grim_oracle_apply_boundaries.argtypes = [Gp, C.c_int, vp]
        L.grim_oracle_reconstruct.argtypes = [Gp, C.c_int, vp, C.c_int, vp, vp]
        L.grim_oracle_riemann_flux.argtypes = [Gp, P, Gm, vp, C.c_int, vp]
        L.grim_oracle_div_fluxes.argtypes = [Gp, P, Gm, vp, vp, vp, C.c_int]
        L.grim_oracle_uniform.argtypes = [C.c_uint64, i64, vp]
        L.grim_oracle_explicit_sources.argtypes = [Gp, P, Gm, vp, vp]
        L.grim_oracle_compute_dt.argtypes = [Gp, P, Gm, vp, d]
        L.grim_oracle_compute_dt.restype = d
        Tp = C.POINTER(Torus)
        L.grim_oracle_stage.argtypes = [Gp, P, Gm, Tp, C.c_int, d, vp, vp, vp, vp, C.c_int, C.POINTER(StepStats)]
        L.grim_oracle_time_step.argtypes = [Gp, P, Gm, Tp, vp, vp, vp, vp, C.c_int, C.POINTER(StepStats)]
        L.grim_oracle_torus_ic.argtypes = [Gp, P, Tp, Gm, vp]
        L.grim_oracle_apply_floor.argtypes = [Gp, P, Tp, Gm, vp]
        L.grim_oracle_torus_bcs.argtypes = [Gp, P, Tp, Gm, vp]
        L.grim_oracle_full_step_diagnostics.argtypes = [Gp, P, Tp, Gm, vp]
        L.grim_oracle_torus_observers.argtypes = [Gp, P, Gm, vp, vp]
        L.grim_oracle_dump_vars.argtypes = [Gp, P, Gm, vp, vp]
        L.grim_oracle_set_ghost_state.argtypes = [vp]
        L.grim_oracle_set_x1_side_state.argtypes = [vp]
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


def fluxes(p: Params, prim, geom, d: int, closure=None, radius=None):
    """computeFluxes(d) per zone (d = 0: conserved variables)."""
    prim, geom = _f64(prim), _f64(geom)
    n = prim.shape[-1]
    out = np.zeros((p.dof, n))
    lib().grim_oracle_fluxes(C.byref(p), n, _ptr(prim), _ptr(geom), _ptr(_f64(closure)), _ptr(_f64(radius)), d,
                             _ptr(out))
    return out


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


def set_threads(n: int) -> None:
    """OpenMP threads of the oracle's later parallel loops."""
    lib().grim_oracle_set_threads(int(n))


def count_live_ops(p: Params, b: Batch):
    """Live operations of one residual evaluation of zone 0 (the residual
    depends on them and they depend on the guess) as [add, mul, div, sqrt,
    other, select]."""
    r = np.zeros(6, dtype=np.int64)
    cb = b._c()
    lib().grim_oracle_count_live_ops(C.byref(p), C.byref(cb), _ptr(r))
    return r


# ---------------------------------------------------------------------------
# Full time step
# ---------------------------------------------------------------------------
def grid(**kw) -> Grid:
    """A Grid with grim's defaults (Minkowski, periodic, MINMOD+HLL) overridden by kw."""
    g = Grid(N1=64, N2=64, N3=1, dim=2, numGhost=3, X1Start=0., X1End=1., X2Start=0., X2End=1.,
             X3Start=0., X3End=1., metric=METRIC_MINKOWSKI, blackHoleSpin=0.9375, hSlope=0.3,
             DerefineThetaHorizon=0, DoCylindrify=0, X1cyl=0., X2cyl=0.,
             boundaryLeft=BC_PERIODIC, boundaryRight=BC_PERIODIC, boundaryBottom=BC_PERIODIC,
             boundaryTop=BC_PERIODIC, boundaryBack=BC_PERIODIC, boundaryFront=BC_PERIODIC,
             reconstruction=RECON_MINMOD, riemannSolver=RIEMANN_HLL, slopeLimTheta=2.,
             CourantFactor=0.9, maxDtIncrement=1.3)
    for k, v in kw.items():
        setattr(g, k, v)
    return g


class Geometry:
    """Geometry arrays of a grid: centre/faceX1/faceX2 (23 planes), gamma (64)."""

    def __init__(self, g: Grid, center=None, faceX1=None, faceX2=None, gamma=None, build=True):
        n = g.n
        self.center = _f64(center) if center is not None else np.zeros((LOC_PLANES, n))
        self.faceX1 = _f64(faceX1) if faceX1 is not None else np.zeros((LOC_PLANES, n))
        self.faceX2 = _f64(faceX2) if faceX2 is not None else np.zeros((LOC_PLANES, n))
        self.gamma = _f64(gamma) if gamma is not None else np.zeros((64, n))
        if build and center is None:
            lib().grim_oracle_build_geometry(C.byref(g), C.byref(self._c()))

    def _c(self):
        return _Geometry(_ptr(self.center), _ptr(self.faceX1), _ptr(self.faceX2), _ptr(self.gamma))


def xcoords(g: Grid):
    out = np.zeros((3, g.n))
    lib().grim_oracle_grid_xcoords(C.byref(g), _ptr(out))
    return out


def apply_boundaries(g: Grid, prim):
    prim = _f64(prim).copy()
    lib().grim_oracle_apply_boundaries(C.byref(g), prim.shape[0], _ptr(prim))
    return prim


def reconstruct(g: Grid, prim, d: int):
    prim = _f64(prim)
    left = np.full_like(prim, np.nan)
    right = np.full_like(prim, np.nan)
    lib().grim_oracle_reconstruct(C.byref(g), prim.shape[0], _ptr(prim), d, _ptr(left), _ptr(right))
    return left, right


def riemann_flux(g: Grid, p: Params, gm: Geometry, prim, d: int):
    prim = _f64(prim)
    out = np.full((p.dof, g.n), np.nan)
    lib().grim_oracle_riemann_flux(C.byref(g), C.byref(p), C.byref(gm._c()), _ptr(prim), d, _ptr(out))
    return out


def div_fluxes(g: Grid, p: Params, gm: Geometry, prim, want_fluxes=False, flux_filter=False):
    """computeDivOfFluxes; flux_filter applies the torus's flux filter."""
    prim = _f64(prim)
    div = np.full((p.dof, g.n), np.nan)
    fl = np.zeros((g.dim * p.dof, g.n)) if want_fluxes else None
    lib().grim_oracle_div_fluxes(C.byref(g), C.byref(p), C.byref(gm._c()), _ptr(prim), _ptr(div), _ptr(fl),
                                 1 if flux_filter else 0)
    return (div, fl) if want_fluxes else div


def uniform(seed: int, n: int) -> np.ndarray:
    """The torus noise stream: uniform in [0, 1) of zones 0..n-1."""
    out = np.empty(n)
    lib().grim_oracle_uniform(C.c_uint64(seed), n, _ptr(out))
    return out


def explicit_sources(g: Grid, p: Params, gm: Geometry, prim):
    prim = _f64(prim)
    out = np.full((p.dof, g.n), np.nan)
    lib().grim_oracle_explicit_sources(C.byref(g), C.byref(p), C.byref(gm._c()), _ptr(prim), _ptr(out))
    return out


def compute_dt(g: Grid, p: Params, gm: Geometry, prim, dt: float) -> float:
    return float(lib().grim_oracle_compute_dt(C.byref(g), C.byref(p), C.byref(gm._c()), _ptr(_f64(prim)), dt))


def stage(g: Grid, p: Params, gm: Geometry, st: int, dt: float, prim_old, prim_half=None, cons_old=None,
          semantics=NEWTON_PER_ZONE, torus: "Torus | None" = None):
    """One stage; returns (prim_new, cons_old, stats); prim_old/prim_half get their boundaries applied."""
    cons_old = np.zeros((p.dof, g.n)) if cons_old is None else cons_old
    new = np.zeros((p.dof, g.n))
    st_ = StepStats()
    lib().grim_oracle_stage(C.byref(g), C.byref(p), C.byref(gm._c()), C.byref(torus) if torus else None, st, dt,
                            _ptr(prim_old), _ptr(prim_half), _ptr(cons_old), _ptr(new), semantics, C.byref(st_))
    return new, cons_old, st_


class Simulation:
    """timeStepper state: primOld, the state elemOld was set from, dt, time
    (and the torus problem's hooks when torus is given)."""

    def __init__(self, g: Grid, p: Params, prim, dt: float, gm: Geometry | None = None, time: float = 0.,
                 torus: "Torus | None" = None, ghost_state=None, side_state=None):
        self.g, self.p = g, p
        # bondi_inflow's hook: ghost zones of primOld from this state
        self.ghost_state = None if ghost_state is None else _f64(ghost_state).copy()
        # buoyancy_instabilities' hook: rho, u, u~^1 of the X1 sides (3 planes)
        self.side_state = None if side_state is None else _f64(side_state).copy()
        self.gm = gm if gm is not None else Geometry(g)
        self.prim = _f64(prim).copy()
        self.prim_dt = self.prim.copy()
        self.dt = C.c_double(dt)
        self.time = C.c_double(time)
        self.torus = torus

    def step(self, semantics=NEWTON_PER_ZONE) -> StepStats:
        st = StepStats()
        lib().grim_oracle_set_ghost_state(None if self.ghost_state is None else _ptr(self.ghost_state))
        lib().grim_oracle_set_x1_side_state(None if self.side_state is None else _ptr(self.side_state))
        try:
            lib().grim_oracle_time_step(C.byref(self.g), C.byref(self.p), C.byref(self.gm._c()),
                                        C.byref(self.torus) if self.torus else None, _ptr(self.prim),
                                        _ptr(self.prim_dt), C.byref(self.dt), C.byref(self.time), semantics,
                                        C.byref(st))
        finally:
            lib().grim_oracle_set_ghost_state(None)
            lib().grim_oracle_set_x1_side_state(None)
        return st


# ---------------------------------------------------------------------------
# Torus problem
# ---------------------------------------------------------------------------
class Torus(C.Structure):
    """Mirror of grim_oracle_torus (src/problem/torus/params.cpp)."""

    _fields_ = [
        ("InnerEdgeRadius", C.c_double), ("PressureMaxRadius", C.c_double), ("Adiabat", C.c_double),
        ("MinPlasmaBeta", C.c_double), ("MagneticLoops", C.c_int), ("UseMADdisk", C.c_int),
        ("InitialPerturbationAmplitude", C.c_double), ("MaxLorentzFactor", C.c_double),
        ("RhoFloorAmpl", C.c_double), ("UFloorAmpl", C.c_double), ("RhoFloorSlope", C.c_double),
        ("UFloorSlope", C.c_double), ("BsqrOverRhoMax", C.c_double), ("BsqrOverUMax", C.c_double),
        ("seed", C.c_uint64), ("minRho", C.c_void_p), ("minU", C.c_void_p),
    ]


def torus_params(**kw) -> Torus:
    """The SANE single-loop torus: r_in = 6, r_max = 12, kappa = 1e-3, beta_min = 100,
    4% noise, grim's floors (torus/params.cpp:38-80), seed 20261202."""
    t = Torus(InnerEdgeRadius=6., PressureMaxRadius=12., Adiabat=1e-3, MinPlasmaBeta=100., MagneticLoops=1,
              UseMADdisk=0, InitialPerturbationAmplitude=4e-2, MaxLorentzFactor=10., RhoFloorAmpl=1e-3,
              UFloorAmpl=1e-5, RhoFloorSlope=-2.5, UFloorSlope=-2.5, BsqrOverRhoMax=10., BsqrOverUMax=500.,
              seed=20261202)
    for k, v in kw.items():
        setattr(t, k, v)
    return t


def torus_ic(g: Grid, p: Params, t: Torus, gm: Geometry):
    out = np.zeros((p.dof, g.n))
    lib().grim_oracle_torus_ic(C.byref(g), C.byref(p), C.byref(t), C.byref(gm._c()), _ptr(out))
    return out


def apply_floor(g: Grid, p: Params, t: Torus, gm: Geometry, prim):
    prim = _f64(prim).copy()
    lib().grim_oracle_apply_floor(C.byref(g), C.byref(p), C.byref(t), C.byref(gm._c()), _ptr(prim))
    return prim


def full_step_diagnostics(g: Grid, p: Params, t: Torus, gm: Geometry, prim):
    """fullStepDiagnostics' NaN scrub and applyFloor (torus/problem.cpp:719-746)."""
    prim = _f64(prim).copy()
    lib().grim_oracle_full_step_diagnostics(C.byref(g), C.byref(p), C.byref(t), C.byref(gm._c()), _ptr(prim))
    return prim


def torus_bcs(g: Grid, p: Params, t: Torus, gm: Geometry, prim):
    prim = _f64(prim).copy()
    lib().grim_oracle_torus_bcs(C.byref(g), C.byref(p), C.byref(t), C.byref(gm._c()), _ptr(prim))
    return prim


OBSERVABLES = ("baryon_mass", "magnetic_energy", "thermal_energy", "mdot_in", "mdot_out", "mdot_out_unbound",
               "mdot_out_relativistic", "rho_max", "beta_min")


def torus_observers(g: Grid, p: Params, gm: Geometry, prim) -> dict:
    """ComputeEnergyIntegrals, ComputeMinMaxVariables, ComputeBoundaryFluxes
    (ObserversTorus.hpp) of the element set at prim, as a dict."""
    out = np.zeros(9)
    lib().grim_oracle_torus_observers(C.byref(g), C.byref(p), C.byref(gm._c()), _ptr(_f64(prim)), _ptr(out))
    return dict(zip(OBSERVABLES, out.tolist()))


def dump_var_names(p: Params) -> list:
    """dumpVars names in plane order (torus/problem.cpp:808-825,
    params.cpp:136-147), then the VTS points x1, x2, x3."""
    names = ["rho", "u", "u1", "u2", "u3"]
    if p.conduction:
        names.append("q")
    if p.viscosity:
        names.append("dP")
    return names + ["B1", "B2", "B3", "bSqr", "gamma", "x1", "x2", "x3"]


def dump_vars(g: Grid, p: Params, gm: Geometry, prim) -> np.ndarray:
    """fullStepDiagnostics' dump variables of the element set at prim, as
    (planes, N1*N2*N3) with X1 fastest (names: dump_var_names)."""
    out = np.zeros((len(dump_var_names(p)), g.N1 * g.N2 * g.N3))
    lib().grim_oracle_dump_vars(C.byref(g), C.byref(p), C.byref(gm._c()), _ptr(_f64(prim)), _ptr(out))
    return out
