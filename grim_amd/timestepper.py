"""Host-side mirror of grim's ``timeStepper`` (src/timestepper/timestepper.hpp:19-132)
over the C ABI: one object owns one ``grim_ctx`` (all grid arrays live on the
device), ``timeStep()`` advances primOld by one predictor/corrector step.

Names follow the reference: ``timeStep``, ``computeDt``, ``primOld``,
``time``, ``dt``; the stage-level pieces (``applyBoundaryConditions``,
``reconstruct``, ``riemannFlux``, ``computeDivOfFluxes``,
``computeExplicitSources``) take and return torch device tensors of shape
(vars, N3T, N2T, N1T).  There is no CPU path: every call goes to
``libgrim_hip.so`` and raises ``GrimError`` when it fails.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi


def grid_config(**kw) -> capi.GridConfig:
    """A grid config with grim's linear-modes defaults (2D Minkowski, periodic,
    MINMOD + HLL, Courant 0.9; src/problem/linear_modes/params.cpp:8-64)
    overridden by keyword."""
    g = capi.GridConfig(N1=64, N2=64, N3=1, dim=2, numGhost=3, X1Start=0., X1End=1., X2Start=0., X2End=1.,
                        X3Start=0., X3End=1., metric=capi.METRIC_MINKOWSKI, blackHoleSpin=0.9375, hSlope=0.3,
                        DerefineThetaHorizon=0, DoCylindrify=0, X1cyl=0., X2cyl=0.,
                        boundaryLeft=capi.BC_PERIODIC, boundaryRight=capi.BC_PERIODIC,
                        boundaryBottom=capi.BC_PERIODIC, boundaryTop=capi.BC_PERIODIC,
                        boundaryBack=capi.BC_PERIODIC, boundaryFront=capi.BC_PERIODIC,
                        reconstruction=capi.RECON_MINMOD, riemannSolver=capi.RIEMANN_HLL, slopeLimTheta=2.,
                        CourantFactor=0.9, maxDtIncrement=1.3)
    for k, v in kw.items():
        if not hasattr(g, k):
            raise AttributeError(f"grim_grid_config has no field {k}")
        setattr(g, k, v)
    return g


class TimeStepper:
    def __init__(self, params: capi.Params, grid: capi.GridConfig, opts: capi.NewtonOpts | None = None,
                 stream=None):
        """Allocate the grid and build its geometry.  Set primOld (or call
        initLinearMode), then set_time(time, dt) before the first timeStep."""
        import torch

        self.params, self.grid = params, grid
        self.dof = params.dof
        self.shape = grid.shape
        self.n = grid.n
        self.device = torch.device("cuda", torch.cuda.current_device())
        self._stream = stream
        ctx = C.c_void_p()
        o = C.byref(opts) if opts is not None else None
        capi.check(capi.lib().grim_create(C.byref(params), C.byref(grid), o, capi._stream(stream), C.byref(ctx)),
                   "grim_create")
        self._ctx = ctx

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if getattr(self, "_ctx", None) is not None and self._ctx.value:
            capi.lib().grim_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass

    # -- state ------------------------------------------------------------
    @property
    def primOld(self) -> np.ndarray:
        out = np.empty((self.dof, self.n))
        capi.check(capi.lib().grim_get_prim(self._ctx, out.ctypes.data, 0), "grim_get_prim")
        return out

    @primOld.setter
    def primOld(self, prim):
        prim = np.ascontiguousarray(prim, dtype=np.float64).reshape(self.dof, self.n)
        capi.check(capi.lib().grim_set_prim(self._ctx, prim.ctypes.data, 0), "grim_set_prim")

    def geometry(self, which: int = 0) -> np.ndarray:
        planes = 64 if which == 3 else capi.LOC_PLANES
        out = np.empty((planes, self.n))
        capi.check(capi.lib().grim_get_geometry(self._ctx, which, out.ctypes.data, 0), "grim_get_geometry")
        return out

    def set_time(self, time: float, dt: float):
        """time and the previous dt (timeStepper(..., Time, InitialDt, ...));
        also marks the current primOld as the state computeDt reads."""
        capi.check(capi.lib().grim_set_time(self._ctx, time, dt), "grim_set_time")

    @property
    def time(self) -> float:
        t, d = C.c_double(), C.c_double()
        capi.check(capi.lib().grim_get_time(self._ctx, C.byref(t), C.byref(d)), "grim_get_time")
        return t.value

    @property
    def dt(self) -> float:
        t, d = C.c_double(), C.c_double()
        capi.check(capi.lib().grim_get_time(self._ctx, C.byref(t), C.byref(d)), "grim_get_time")
        return d.value

    def initLinearMode(self, mode: int, amplitude: float):
        capi.check(capi.lib().grim_init_linear_mode(self._ctx, mode, amplitude), "grim_init_linear_mode")

    def initTorus(self, torus: capi.TorusParams):
        """Torus initial conditions into primOld; enables the torus's per-stage hooks."""
        capi.check(capi.lib().grim_init_torus(self._ctx, C.byref(torus)), "grim_init_torus")

    def _grid_ptr(self, t, planes: int):
        """Address of a float64 device tensor holding at least `planes` planes
        of the grid's zones (raises ValueError before any call otherwise)."""
        if t is None:
            return None
        if t.dim() < 1 or t.shape[0] < planes:
            raise ValueError(f"expected {planes} planes of {self.n} zones, got shape {tuple(t.shape)}")
        return capi._ptr(t, numel=planes * self.n)

    def applyFloor(self, prim):
        capi.check(capi.lib().grim_apply_floor(self._ctx, self._grid_ptr(prim, self.dof)), "grim_apply_floor")
        return prim

    def fullStepDiagnostics(self, prim):
        """The NaN scrub and applyFloor of the torus's fullStepDiagnostics
        (torus/problem.cpp:719-746) on device primitives, in place."""
        capi.check(capi.lib().grim_full_step_diagnostics(self._ctx, self._grid_ptr(prim, self.dof)),
                   "grim_full_step_diagnostics")
        return prim

    def setProblemSpecificBCs(self, prim):
        capi.check(capi.lib().grim_torus_bcs(self._ctx, self._grid_ptr(prim, self.dof)), "grim_torus_bcs")
        return prim

    def dump(self, file: str, dataset: str = "primitives"):
        """grid::dump of primOld (grim's HDF5 layout, grim_amd/checkpoint.py)."""
        from . import checkpoint

        checkpoint.dump(self.grid, self.primOld, file, dataset)

    def load(self, file: str, dataset: str = "primitives"):
        """grid::load into primOld's bulk, then communicate (periodic ghosts)."""
        import torch

        from . import checkpoint

        prim = checkpoint.load(self.grid, self.primOld, file, dataset)
        dev = torch.from_numpy(prim).to(self.device)
        self.communicate(dev)
        torch.cuda.synchronize()
        self.primOld = dev.cpu().numpy()

    def observers(self, prim=None) -> dict:
        """ComputeEnergyIntegrals, ComputeMinMaxVariables and
        ComputeBoundaryFluxes (ObserversTorus.hpp:6-251) of prim (a device
        tensor of dof planes; None: primOld), as a dict."""
        out = capi.TorusObservables()
        capi.check(capi.lib().grim_torus_observers(self._ctx, self._grid_ptr(prim, self.dof), C.byref(out)),
                   "grim_torus_observers")
        return out.as_dict()

    def set_ghost_state(self, prim_ic):
        """setProblemSpecificBCs of bondi_inflow (problem.cpp:157-183): every
        ghost zone of primOld is reset to prim_ic (dof x zones, host array or
        device tensor) in each half stage; None switches it off."""
        if prim_ic is None:
            capi.check(capi.lib().grim_set_ghost_state(self._ctx, None, 0), "grim_set_ghost_state")
            return
        on_device = hasattr(prim_ic, "is_cuda") and prim_ic.is_cuda
        if on_device:
            if prim_ic.numel() != self.dof * self.n:
                raise ValueError("set_ghost_state: expected dof x zones values")
            ptr = capi._ptr(prim_ic, numel=self.dof * self.n)
        else:
            prim_ic = np.ascontiguousarray(prim_ic, dtype=np.float64)
            if prim_ic.size != self.dof * self.n:
                raise ValueError("set_ghost_state: expected dof x zones values")
            ptr = prim_ic.ctypes.data
        capi.check(capi.lib().grim_set_ghost_state(self._ctx, ptr, int(on_device)), "grim_set_ghost_state")

    def set_x1_side_state(self, prim_ic):
        """setProblemSpecificBCs of buoyancy_instabilities (problem.cpp:
        221-293): in both stages the X1 sides of the stage state (X1 index
        0..numGhost, the first bulk zone included, and the right ghosts) take
        rho, u, u~^1 of prim_ic (3 x zones, host array or device tensor),
        zero u~^2, u~^3, B and q~; None switches it off."""
        if prim_ic is None:
            capi.check(capi.lib().grim_set_x1_side_state(self._ctx, None, 0), "grim_set_x1_side_state")
            return
        on_device = hasattr(prim_ic, "is_cuda") and prim_ic.is_cuda
        if on_device:
            if prim_ic.numel() != 3 * self.n:
                raise ValueError("set_x1_side_state: expected 3 x zones values")
            ptr = capi._ptr(prim_ic, numel=3 * self.n)
        else:
            prim_ic = np.ascontiguousarray(prim_ic, dtype=np.float64)
            if prim_ic.size != 3 * self.n:
                raise ValueError("set_x1_side_state: expected 3 x zones values")
            ptr = prim_ic.ctypes.data
        capi.check(capi.lib().grim_set_x1_side_state(self._ctx, ptr, int(on_device)), "grim_set_x1_side_state")

    def dump_vars(self, prim=None):
        """fullStepDiagnostics' dump variables (torus/problem.cpp:808-875) of
        prim (None: primOld) as a device tensor (dof + 5, N1*N2*N3), X1
        fastest: rho, u, u1..3, [q], [dP], B1..3, bSqr, gamma, x1..3."""
        import torch

        nb = self.grid.N1 * self.grid.N2 * self.grid.N3
        out = torch.empty((self.params.dof + 5, nb), dtype=torch.float64, device=self.device)
        capi.check(capi.lib().grim_dump_vars(self._ctx, self._grid_ptr(prim, self.dof),
                                             capi._ptr(out, numel=out.numel()), 1), "grim_dump_vars")
        return out

    # -- the step -----------------------------------------------------------
    def timeStep(self, stats: bool = True) -> capi.StepStats | None:
        """One step.  stats=False enqueues it without a host synchronisation
        (dt and time advance on the device), e.g. under HIP graph capture."""
        st = capi.StepStats() if stats else None
        capi.check(capi.lib().grim_time_step(self._ctx, C.byref(st) if stats else None), "grim_time_step")
        return st

    # -- stage-level pieces on torch device tensors ------------------------
    def _empty(self, planes):
        import torch

        return torch.full((planes, self.n), float("nan"), dtype=torch.float64, device=self.device)

    def applyBoundaryConditions(self, prim):
        capi.check(capi.lib().grim_apply_boundaries(self._ctx, self._grid_ptr(prim, prim.shape[0]), prim.shape[0]),
                   "grim_apply_boundaries")
        return prim

    def communicate(self, prim):
        capi.check(capi.lib().grim_communicate(self._ctx, self._grid_ptr(prim, prim.shape[0]), prim.shape[0]),
                   "grim_communicate")
        return prim

    def reconstruct(self, prim, dir: int):
        left, right = self._empty(prim.shape[0]), self._empty(prim.shape[0])
        nv = prim.shape[0]
        capi.check(capi.lib().grim_reconstruct(self._ctx, self._grid_ptr(prim, nv), nv, dir,
                                               self._grid_ptr(left, nv), self._grid_ptr(right, nv)),
                   "grim_reconstruct")
        return left, right

    def riemannFlux(self, prim, dir: int):
        flux = self._empty(self.dof)
        capi.check(capi.lib().grim_riemann_flux(self._ctx, self._grid_ptr(prim, self.dof), dir,
                                                self._grid_ptr(flux, self.dof)),
                   "grim_riemann_flux")
        return flux

    def computeDivOfFluxes(self, prim, want_fluxes: bool = False):
        div = self._empty(self.dof)
        fl = self._empty(self.grid.dim * self.dof) if want_fluxes else None
        capi.check(capi.lib().grim_div_fluxes(self._ctx, self._grid_ptr(prim, self.dof), self._grid_ptr(div, self.dof),
                                              self._grid_ptr(fl, self.grid.dim * self.dof)),
                   "grim_div_fluxes")
        return (div, fl) if want_fluxes else div

    def computeExplicitSources(self, prim):
        src = self._empty(self.params.nf)
        capi.check(capi.lib().grim_explicit_sources(self._ctx, self._grid_ptr(prim, self.dof),
                                                    self._grid_ptr(src, self.params.nf)),
                   "grim_explicit_sources")
        return src

    def riemannSolve(self, primLeft, primRight, dir: int):
        """riemannSolver::solve from reconstructed states (reconstruct's output)."""
        flux = self._empty(self.dof)
        capi.check(capi.lib().grim_riemann_solve(self._ctx, self._grid_ptr(primLeft, self.dof),
                                                 self._grid_ptr(primRight, self.dof), dir,
                                                 self._grid_ptr(flux, self.dof)), "grim_riemann_solve")
        return flux

    # -- fluidElement on the grid (src/physics/physics.hpp:14-108) ----------
    def elementFluxes(self, prim, dir: int, location: int = 0):
        """computeFluxes(dir) of the element set from prim at a location
        (0 centre, 1 X1 faces, 2 X2 faces); dir 0 = conserved variables."""
        out = self._empty(self.dof)
        capi.check(capi.lib().grim_element_fluxes(self._ctx, self._grid_ptr(prim, self.dof), location, dir,
                                                  self._grid_ptr(out, self.dof)), "grim_element_fluxes")
        return out

    def elementCharSpeeds(self, prim, dir: int, location: int = 0):
        mn, mx = self._empty(1), self._empty(1)
        capi.check(capi.lib().grim_element_char_speeds(self._ctx, self._grid_ptr(prim, self.dof), location, dir,
                                                       self._grid_ptr(mn, 1), self._grid_ptr(mx, 1)),
                   "grim_element_char_speeds")
        return mn[0], mx[0]

    def elementFields(self, prim, location: int = 0) -> dict:
        out = self._empty(len(capi.ELEMENT_FIELDS))
        capi.check(capi.lib().grim_element_fields(self._ctx, self._grid_ptr(prim, self.dof), location,
                                                  self._grid_ptr(out, len(capi.ELEMENT_FIELDS))),
                   "grim_element_fields")
        return {name: out[q] for q, name in enumerate(capi.ELEMENT_FIELDS)}

    def elementImplicitSources(self, prim, tauDamp):
        out = self._empty(self.dof)
        capi.check(capi.lib().grim_element_implicit_sources(self._ctx, self._grid_ptr(prim, self.dof),
                                                            self._grid_ptr(tauDamp.reshape(1, -1), 1),
                                                            self._grid_ptr(out, self.dof)),
                   "grim_element_implicit_sources")
        return out

    def elementTimeDerivSources(self, prim, primOld, primNew, dt: float):
        out = self._empty(self.dof)
        capi.check(capi.lib().grim_element_time_deriv_sources(self._ctx, self._grid_ptr(prim, self.dof),
                                                              self._grid_ptr(primOld, self.dof),
                                                              self._grid_ptr(primNew, self.dof), float(dt),
                                                              self._grid_ptr(out, self.dof)),
                   "grim_element_time_deriv_sources")
        return out

    # -- the step in its pieces (timestep.cpp:3-334) -----------------------
    def stepBegin(self):
        capi.check(capi.lib().grim_step_begin(self._ctx), "grim_step_begin")

    def foldRHS(self, stage: int):
        capi.check(capi.lib().grim_fold_rhs(self._ctx, stage), "grim_fold_rhs")

    def solve(self, stage: int, prim=None, stats: bool = False):
        st = capi.StepStats() if stats else None
        capi.check(capi.lib().grim_solve(self._ctx, stage, self._grid_ptr(prim, self.dof),
                                         C.byref(st) if stats else None), "grim_solve")
        return st

    def computeResidual(self, stage: int, prim):
        out = self._empty(self.params.nf)
        capi.check(capi.lib().grim_stage_residual(self._ctx, stage, self._grid_ptr(prim, self.dof),
                                                  self._grid_ptr(out, self.params.nf)), "grim_stage_residual")
        return out

    def stageFinish(self, stage: int):
        capi.check(capi.lib().grim_stage_finish(self._ctx, stage), "grim_stage_finish")

    def stepEnd(self):
        capi.check(capi.lib().grim_step_end(self._ctx), "grim_step_end")

    def stageArrays(self) -> dict:
        """Device addresses of the context's arrays under grim's names."""
        a = capi.StageArrays()
        capi.check(capi.lib().grim_get_stage_arrays(self._ctx, C.byref(a)), "grim_get_stage_arrays")
        return {name: getattr(a, name) for name, _ in a._fields_}

    def lastSolveStats(self):
        """(iterations, final sum R^2) per zone of the last solve, host arrays."""
        it = np.empty(self.n, dtype=np.int32)
        l2 = np.empty(self.n)
        capi.check(capi.lib().grim_last_solve_stats(self._ctx, it.ctypes.data, l2.ctypes.data, 0),
                   "grim_last_solve_stats")
        return it, l2

    def computeDt(self, prim, dt: float) -> float:
        out = C.c_double()
        capi.check(capi.lib().grim_compute_dt(self._ctx, self._grid_ptr(prim, self.dof), dt, C.byref(out)),
                   "grim_compute_dt")
        return out.value
