"""ctypes binding of the C ABI in ``include/grim_hip.h``.

The library ``grim_amd/libgrim_hip.so`` is the only implementation: when it
is missing, or no gfx950 device is visible, calls raise ``GrimError``; there
is no CPU fallback.  Device buffers are torch tensors (``torch.float64`` on
``cuda``), passed to the library by ``data_ptr()``; torch is used for
allocation and streams only.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgrim_hip.so")

GRIM_OK = 0
GRIM_ERR_HIP = -1
GRIM_ERR_BAD_ARG = -2
GRIM_ERR_NO_DEVICE = -3
GRIM_ERR_OOM = -4

CLOSURE_LINEAR_MODES = 0
CLOSURE_PER_ZONE = 1
CLOSURE_TORUS = 2
CLOSURE_SHOCK_TESTS = 3
CLOSURE_BUOYANCY = 4
JAC_FD = 0
JAC_ANALYTIC = 1
LU_FP64 = 0
LU_FP32 = 1
LU_FP32_FP64 = 2
NEWTON_PER_ZONE = 0
NEWTON_GLOBAL = 1
GEOM_PLANES = 16
MAX_VARS = 10


class GrimError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what} failed with status {code}: {last_error()}")
        self.code = code


class Params(C.Structure):
    """grim_params: grim's ``namespace params`` subset (src/params.hpp:96-210)."""

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


class ZoneBatch(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("prim", C.c_void_p),
        ("geom", C.c_void_p),
        ("rhs", C.c_void_p),
        ("prim_old", C.c_void_p),
        ("prim_coef", C.c_void_p),
        ("closure", C.c_void_p),
        ("radius", C.c_void_p),
        ("residual_mask", C.c_void_p),
        ("dt_prime", C.c_double),
        ("dt_prime_device", C.c_void_p),
        ("guess", C.c_void_p),
        ("geom_columns", C.c_int64),
    ]


class NewtonOpts(C.Structure):
    _fields_ = [
        ("jacobian", C.c_int),
        ("lu_precision", C.c_int),
        ("semantics", C.c_int),
        ("compaction", C.c_int),
        ("block_size", C.c_int),
        ("fused", C.c_int),
        ("phase1_iters", C.c_int),
        ("phase2_group", C.c_int),
        ("workspace", C.c_void_p),
        ("workspace_bytes", C.c_int64),
    ]


class NewtonStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_void_p),
        ("residual_evals", C.c_void_p),
        ("final_l2", C.c_void_p),
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


class GridConfig(C.Structure):
    """grim_grid_config: grid, geometry, boundary and scheme params."""

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
        ("residualScaling", C.c_int), ("explicitFast", C.c_int),
    ]

    @property
    def shape(self):
        """(N3T, N2T, N1T): shape of one plane, X1 fastest."""
        g = self.numGhost
        return (self.N3 + (2 * g if self.dim > 2 else 0), self.N2 + (2 * g if self.dim > 1 else 0),
                self.N1 + 2 * g)

    @property
    def n(self) -> int:
        a, b, c = self.shape
        return a * b * c


class StepStats(C.Structure):
    _fields_ = [("dt", C.c_double), ("unconverged", C.c_int64 * 2), ("max_iterations", C.c_int32 * 2),
                ("nonfinite", C.c_int64 * 2)]


ELEMENT_FIELDS = ("rho", "u", "pressure", "temperature", "gammaLorentzFactor", "soundSpeed", "bSqr", "bNorm", "q",
                  "deltaP", "tau", "chi_emhd", "nu_emhd", "uCon0", "uCon1", "uCon2", "uCon3", "uCov0", "uCov1",
                  "uCov2", "uCov3", "bCon0", "bCon1", "bCon2", "bCon3", "bCov0", "bCov1", "bCov2", "bCov3")


class StageArrays(C.Structure):
    """grim_stage_arrays: device arrays of a context under grim's timeStepper names."""

    _fields_ = [(name, C.c_void_p) for name in (
        "primOld", "primHalfStep", "prim", "consOld", "sourcesExplicit", "sourcesImplicitOld", "rhs",
        "residualMask", "fluxesX1", "fluxesX2", "fluxesX3")]


class TorusObservables(C.Structure):
    """grim_torus_observables (ObserversTorus.hpp:6-251)."""

    _fields_ = [(name, C.c_double) for name in (
        "baryon_mass", "magnetic_energy", "thermal_energy", "mdot_in", "mdot_out", "mdot_out_unbound",
        "mdot_out_relativistic", "rho_max", "beta_min")]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class TorusParams(C.Structure):
    """grim_torus_params (src/problem/torus/params.cpp:38-80)."""

    _fields_ = [
        ("InnerEdgeRadius", C.c_double), ("PressureMaxRadius", C.c_double), ("Adiabat", C.c_double),
        ("MinPlasmaBeta", C.c_double), ("MagneticLoops", C.c_int), ("UseMADdisk", C.c_int),
        ("InitialPerturbationAmplitude", C.c_double), ("MaxLorentzFactor", C.c_double),
        ("RhoFloorAmpl", C.c_double), ("UFloorAmpl", C.c_double), ("RhoFloorSlope", C.c_double),
        ("UFloorSlope", C.c_double), ("BsqrOverRhoMax", C.c_double), ("BsqrOverUMax", C.c_double),
        ("seed", C.c_uint64),
    ]


def torus_params(**kw) -> TorusParams:
    """SANE single-loop torus: r_in = 6, r_max = 12, kappa = 1e-3, beta_min = 100,
    4% internal-energy noise, grim's floors; noise seed 20261202."""
    t = TorusParams(InnerEdgeRadius=6., PressureMaxRadius=12., Adiabat=1e-3, MinPlasmaBeta=100., MagneticLoops=1,
                    UseMADdisk=0, InitialPerturbationAmplitude=4e-2, MaxLorentzFactor=10., RhoFloorAmpl=1e-3,
                    UFloorAmpl=1e-5, RhoFloorSlope=-2.5, UFloorSlope=-2.5, BsqrOverRhoMax=10., BsqrOverUMax=500.,
                    seed=20261202)
    for k, v in kw.items():
        setattr(t, k, v)
    return t


# Every symbol include/grim_hip.h declares, with its ctypes signature.
_vp, _i64, _d, _i = C.c_void_p, C.c_int64, C.c_double, C.c_int
_P = C.POINTER(Params)
_B = C.POINTER(ZoneBatch)
_G = C.POINTER(GridConfig)
_pvp = C.POINTER(C.c_void_p)
SIGNATURES = {
    "grim_default_params": (None, [_P, _i]),
    "grim_last_error": (C.c_char_p, []),
    "grim_device_info": (_i, [C.c_char_p, _i]),
    "grim_conserved": (_i, [_P, _i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "grim_residual": (_i, [_P, _B, _vp, _vp]),
    "grim_jacobian": (_i, [_P, _B, _i, _vp, _vp]),
    "grim_batch_linear_solve": (_i, [_i64, _i, _vp, _vp, _vp, _i, _vp]),
    "grim_residual_tangent": (_i, [_P, _B, _vp, _vp, _i, _vp]),
    "grim_residual_adjoint": (_i, [_P, _B, _vp, _vp, _i, _vp]),
    "grim_batched_newton": (_i, [_P, _B, C.POINTER(NewtonOpts), C.POINTER(NewtonStats), _vp]),
    "grim_batched_newton_noalloc": (_i, [_P, _B, C.POINTER(NewtonOpts), C.POINTER(NewtonStats), _vp]),
    "grim_count_above": (_i, [_i64, _vp, _d, C.POINTER(C.c_int64), _vp]),
    "grim_generate_kerr_schild_batch": (_i, [_P, _i64, C.c_uint64, _d, _d, _vp, _vp, _vp, _vp, _vp, _vp]),
    "grim_known_root_rhs": (_i, [_P, _B, _vp, _vp, _vp]),
    "grim_newton_workspace_bytes": (C.c_int64, [_i64]),
    "grim_autotune_newton": (_i, [_P, _B, C.POINTER(NewtonOpts), _vp]),
    "grim_grid_zones": (_i64, [_G]),
    "grim_create": (_i, [_P, _G, C.POINTER(NewtonOpts), _vp, _pvp]),
    "grim_destroy": (_i, [_vp]),
    "grim_prim": (_i, [_vp, _pvp]),
    "grim_geometry": (_i, [_vp, _i, _pvp]),
    "grim_geometry_columns": (C.c_int64, [_vp]),
    "grim_set_prim": (_i, [_vp, _vp, _i]),
    "grim_get_prim": (_i, [_vp, _vp, _i]),
    "grim_get_geometry": (_i, [_vp, _i, _vp, _i]),
    "grim_set_time": (_i, [_vp, _d, _d]),
    "grim_get_time": (_i, [_vp, C.POINTER(_d), C.POINTER(_d)]),
    "grim_time_step": (_i, [_vp, C.POINTER(StepStats)]),
    "grim_apply_boundaries": (_i, [_vp, _vp, _i]),
    "grim_communicate": (_i, [_vp, _vp, _i]),
    "grim_reconstruct": (_i, [_vp, _vp, _i, _i, _vp, _vp]),
    "grim_riemann_flux": (_i, [_vp, _vp, _i, _vp]),
    "grim_div_fluxes": (_i, [_vp, _vp, _vp, _vp]),
    "grim_explicit_sources": (_i, [_vp, _vp, _vp]),
    "grim_compute_dt": (_i, [_vp, _vp, _d, C.POINTER(_d)]),
    "grim_init_linear_mode": (_i, [_vp, _i, _d]),
    "grim_init_torus": (_i, [_vp, C.POINTER(TorusParams)]),
    "grim_apply_floor": (_i, [_vp, _vp]),
    "grim_full_step_diagnostics": (_i, [_vp, _vp]),
    "grim_torus_bcs": (_i, [_vp, _vp]),
    "grim_torus_observers": (_i, [_vp, _vp, _vp]),
    "grim_dump_vars": (_i, [_vp, _vp, _vp, _i]),
    "grim_set_ghost_state": (_i, [_vp, _vp, _i]),
    "grim_set_x1_side_state": (_i, [_vp, _vp, _i]),
    "grim_riemann_solve": (_i, [_vp, _vp, _vp, _i, _vp]),
    "grim_element_fluxes": (_i, [_vp, _vp, _i, _i, _vp]),
    "grim_element_char_speeds": (_i, [_vp, _vp, _i, _i, _vp, _vp]),
    "grim_element_fields": (_i, [_vp, _vp, _i, _vp]),
    "grim_element_implicit_sources": (_i, [_vp, _vp, _vp, _vp]),
    "grim_element_time_deriv_sources": (_i, [_vp, _vp, _vp, _vp, _d, _vp]),
    "grim_step_begin": (_i, [_vp]),
    "grim_fold_rhs": (_i, [_vp, _i]),
    "grim_solve": (_i, [_vp, _i, _vp, C.POINTER(StepStats)]),
    "grim_stage_residual": (_i, [_vp, _i, _vp, _vp]),
    "grim_stage_finish": (_i, [_vp, _i]),
    "grim_step_end": (_i, [_vp]),
    "grim_get_stage_arrays": (_i, [_vp, C.POINTER(StageArrays)]),
    "grim_last_solve_stats": (_i, [_vp, _vp, _vp, _i]),
}

_lib = None


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GrimError(GRIM_ERR_NO_DEVICE, f"loading {LIB_PATH} (not built: run __graft_entry__.build())")
        # torch ships its own HIP runtime under the same sonames
        # (libamdhip64.so.7, libhsa-runtime64.so.1): whichever loads first
        # serves the whole process.  Load torch's first, so that torch and
        # this library share one runtime whatever the import order.
        import torch  # noqa: F401

        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    try:
        msg = lib().grim_last_error()
    except Exception:  # noqa: BLE001 - only used while formatting an error
        return "(library not loaded)"
    return msg.decode(errors="replace") if msg else ""


def check(rc: int, what: str) -> None:
    if rc != GRIM_OK:
        raise GrimError(rc, what)


def default_params(emhd: bool = False) -> Params:
    p = Params()
    lib().grim_default_params(C.byref(p), 1 if emhd else 0)
    return p


def _stream(stream):
    if stream is None:
        import torch

        return torch.cuda.current_stream().cuda_stream
    return getattr(stream, "cuda_stream", stream)


def _ptr(t, dtype="float64", numel: int = 0):
    """Device address of a contiguous GPU tensor for the library.

    dtype names the element type the entry point reads ("float64", "int32";
    None for raw workspace bytes) and numel the number of elements it reads
    at most.  A tensor of another type or a shorter one raises ValueError
    here, before anything is enqueued: the library only sees the address and
    would read past the end of the allocation."""
    if t is None:
        return None
    import torch

    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.is_contiguous()):
        raise ValueError("expected a contiguous device tensor")
    if dtype is not None and t.dtype != getattr(torch, dtype):
        raise ValueError(f"expected a torch.{dtype} tensor, got {t.dtype}")
    if t.numel() < numel:
        raise ValueError(f"tensor holds {t.numel()} elements, the call reads {numel}")
    return t.data_ptr()


def device_info() -> str:
    buf = C.create_string_buffer(64)
    check(lib().grim_device_info(buf, 64), "grim_device_info")
    return buf.value.decode()


def make_batch(prim, geom, rhs, prim_old=None, prim_coef=None, closure=None, radius=None,
               residual_mask=None, dt_prime=1.0, guess=None) -> ZoneBatch:
    """guess: optional starting fluid rows (numFluidVars x n) that the solve
    reads instead of prim's; the result still goes to prim."""
    n = prim.shape[-1]
    return ZoneBatch(n, _ptr(prim, numel=8 * n), _ptr(geom, numel=GEOM_PLANES * n), _ptr(rhs, numel=5 * n),
                     _ptr(prim_old, numel=8 * n), _ptr(prim_coef, numel=8 * n), _ptr(closure, numel=2 * n),
                     _ptr(radius, numel=n), _ptr(residual_mask, numel=n), float(dt_prime), None,
                     _ptr(guess, numel=5 * n))


def conserved(p: Params, prim, geom, closure=None, radius=None, out=None, stream=None):
    import torch

    n = prim.shape[-1]
    if out is None:
        out = torch.empty((p.dof, n), dtype=torch.float64, device=prim.device)
    check(lib().grim_conserved(C.byref(p), n, _ptr(prim, numel=p.dof * n), _ptr(geom, numel=GEOM_PLANES * n),
                               _ptr(closure, numel=2 * n), _ptr(radius, numel=n), _ptr(out, numel=p.dof * n),
                               _stream(stream)), "grim_conserved")
    return out


def residual(p: Params, batch: ZoneBatch, out, stream=None):
    check(lib().grim_residual(C.byref(p), C.byref(batch), _ptr(out, numel=p.nf * batch.n), _stream(stream)),
          "grim_residual")
    return out


def jacobian(p: Params, batch: ZoneBatch, kind: int, out, stream=None):
    check(lib().grim_jacobian(C.byref(p), C.byref(batch), kind, _ptr(out, numel=p.nf * p.nf * batch.n),
                              _stream(stream)), "grim_jacobian")
    return out


def residual_tangent(p: Params, batch: ZoneBatch, v, out, precision=LU_FP64, stream=None):
    """out = J v per zone (analytic tangent of the residual at batch.prim)."""
    n = batch.n
    check(lib().grim_residual_tangent(C.byref(p), C.byref(batch), _ptr(v, numel=p.nf * n), _ptr(out, numel=p.nf * n),
                                      precision, _stream(stream)), "grim_residual_tangent")
    return out


def residual_adjoint(p: Params, batch: ZoneBatch, w, out, precision=LU_FP64, stream=None):
    """out = J^T w per zone (hand-written reverse sweep of the residual at batch.prim)."""
    n = batch.n
    check(lib().grim_residual_adjoint(C.byref(p), C.byref(batch), _ptr(w, numel=p.nf * n), _ptr(out, numel=p.nf * n),
                                      precision, _stream(stream)), "grim_residual_adjoint")
    return out


def batch_linear_solve(A, b, x, lu_precision=LU_FP64, stream=None):
    nvar = b.shape[0]
    n = b.shape[-1]
    check(lib().grim_batch_linear_solve(n, nvar, _ptr(A, numel=nvar * nvar * n), _ptr(b, numel=nvar * n),
                                        _ptr(x, numel=nvar * n), lu_precision, _stream(stream)),
          "grim_batch_linear_solve")
    return x


def newton_opts(jacobian=JAC_FD, lu_precision=LU_FP64, semantics=NEWTON_PER_ZONE, compaction=0,
                block_size=0, fused=0, phase1_iters=0, phase2_group=0, workspace=None) -> NewtonOpts:
    return NewtonOpts(jacobian, lu_precision, semantics, compaction, block_size, fused, phase1_iters,
                      phase2_group, _ptr(workspace, dtype=None), 0 if workspace is None else workspace.numel() *
                      workspace.element_size())


def workspace_bytes(n: int) -> int:
    return int(lib().grim_newton_workspace_bytes(n))


def autotune_newton(p: Params, batch: ZoneBatch, opts: NewtonOpts, stream=None) -> NewtonOpts:
    """Pick the fastest launch configuration for this batch (in place on a copy)."""
    o = NewtonOpts.from_buffer_copy(opts)
    check(lib().grim_autotune_newton(C.byref(p), C.byref(batch), C.byref(o), _stream(stream)),
          "grim_autotune_newton")
    return o


def newton_stats(n: int, iterations=None, residual_evals=None, final_l2=None) -> NewtonStats:
    """grim_newton_stats of optional per-zone device arrays (int32, int32, float64)."""
    return NewtonStats(_ptr(iterations, "int32", n), _ptr(residual_evals, "int32", n), _ptr(final_l2, numel=n))


def batched_newton(p: Params, batch: ZoneBatch, jacobian=JAC_FD, lu_precision=LU_FP64,
                   semantics=NEWTON_PER_ZONE, compaction=0, block_size=0, iterations=None,
                   residual_evals=None, final_l2=None, stream=None, fused=0, opts=None):
    if opts is None:
        opts = newton_opts(jacobian, lu_precision, semantics, compaction, block_size, fused)
    st = newton_stats(batch.n, iterations, residual_evals, final_l2)
    check(lib().grim_batched_newton(C.byref(p), C.byref(batch), C.byref(opts), C.byref(st), _stream(stream)),
          "grim_batched_newton")


NOALLOC_MAX_ZONES = 8388608


def batched_newton_noalloc(p: Params, batch: ZoneBatch, opts: NewtonOpts, iterations=None, residual_evals=None,
                           final_l2=None, stream=None):
    """grim_batched_newton_noalloc: the per-zone solve with the compaction work
    list in the library's own device array (no allocation in the call)."""
    st = newton_stats(batch.n, iterations, residual_evals, final_l2)
    check(lib().grim_batched_newton_noalloc(C.byref(p), C.byref(batch), C.byref(opts), C.byref(st),
                                            _stream(stream)), "grim_batched_newton_noalloc")


def count_above(final_l2, tol, stream=None) -> int:
    c = C.c_int64(0)
    check(lib().grim_count_above(final_l2.shape[-1], _ptr(final_l2, numel=final_l2.shape[-1]), float(tol),
                                 C.byref(c), _stream(stream)),
          "grim_count_above")
    return int(c.value)


def generate_kerr_schild_batch(p: Params, n: int, seed: int, a: float = 0.9375, h: float = 0.3,
                               device="cuda", stream=None):
    """Synthetic isolated-zone batch; returns dict of device tensors."""
    import torch

    f = dict(dtype=torch.float64, device=device)
    out = {
        "geom": torch.empty((GEOM_PLANES, n), **f),
        "prim_star": torch.empty((p.dof, n), **f),
        "guess": torch.empty((p.dof, n), **f),
        "prim_old": torch.empty((p.dof, n), **f) if (p.conduction or p.viscosity) else None,
        "closure": torch.empty((2, n), **f) if (p.conduction or p.viscosity) else None,
    }
    check(lib().grim_generate_kerr_schild_batch(C.byref(p), n, seed, a, h, _ptr(out["geom"]),
                                                 _ptr(out["prim_star"]), _ptr(out["guess"]),
                                                 _ptr(out["prim_old"]), _ptr(out["closure"]),
                                                 _stream(stream)), "grim_generate_kerr_schild_batch")
    return out


def known_root_rhs(p: Params, batch: ZoneBatch, rhs, mask=None, stream=None):
    check(lib().grim_known_root_rhs(C.byref(p), C.byref(batch), _ptr(rhs, numel=p.nf * batch.n),
                                    _ptr(mask, numel=batch.n), _stream(stream)),
          "grim_known_root_rhs")
    return rhs
