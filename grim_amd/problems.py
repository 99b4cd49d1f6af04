"""Initial states of further grim problems on the same time step, host side:
the 1D EMHD shock tests (src/problem/shock_tests/), the Orszag-Tang vortex
(src/problem/orzag_tang/) and Bondi inflow (src/problem/bondi_inflow/, whose
ghost-zone hook is TimeStepper.set_ghost_state).  Each gives the grid, the parameters and
the primitive state, set into a TimeStepper with ``primOld = ...``.

Both run on grim's time step unchanged.  The shock tests supply a closure
(GRIM_CLOSURE_SHOCK_TESTS: tau = 0.1, chi = ConductionAlpha c_s^2 tau,
nu = ViscosityAlpha c_s^2 tau, shock_tests/problem.cpp:13-18) and the
initial state; stationary_shock_BVP_input reads solution files that the
reference does not hold, so it is not offered.  The vortex is ideal MHD.
"""
from __future__ import annotations

import numpy as np

from . import capi

# (rho, pressure, u1, u2, u3, B1, B2, B3) left and right, finalTime
# (shock_tests/problem.cpp:31-146).  The stationary shock's pressures are
# multiples of (adiabaticIndex - 1): see shock_tube_prim.
SHOCK_TESTS = {
    "fast_shock": ((1., 1., 25., 0., 0., 20., 25.02, 0.), (25.48, 367.5, 1.091, 0.3923, 0., 20., 49., 0.), 2.5),
    "slow_shock": ((1., 10., 1.53, 0., 0., 10., 18.28, 0.), (3.323, 55.36, 0.9571, -0.6822, 0., 10., 14.49, 0.),
                   2.),
    "switch_on_slow": ((1.78e-3, .1, -.765, -1.386, 0., 1., 1.022, 0.), (0.01, 1., 0., 0., 0., 1., 0., 0.), 2.),
    "switch_off_fast": ((0.1, 1., -2., 0., 0., 2., 0., 0.), (0.562, 10., -0.212, -0.590, 0., 2., 4.71, 0.), 1.),
    "alfven_wave": ((1., 1., 0., 0., 0., 3., 3., 0.), (1., 1., 3.7, 5.76, 0., 3., -6.857, 0.), 2.),
    "shock_tube_1": ((1., 1000., 0., 0., 0., 1., 0., 0.), (.1, 1., 0., 0., 0., 1., 0., 0.), 1.),
    "shock_tube_2": ((1., 30., 0., 0., 0., 0., 20., 0.), (.1, 1., 0., 0., 0., 0., 0., 0.), 1.),
    "collision": ((1., 1., 5., 0., 0., 10., 10., 0.), (1., 1., -5., 0., 0., 10., -10., 0.), 1.22),
    "stationary_shock": ((1., 1., 1., 0., 0., 0.00001, 0., 0.), (3.08312999, 4.94577705, 0.32434571, 0., 0.,
                                                               0.00001, 0., 0.), None),
}


def shock_grid(N1: int = 512) -> capi.GridConfig:
    """src/problem/shock_tests/params.cpp: 1D Minkowski X1 in [-0.5, 1.5],
    outflow in X1, MINMOD (theta 2) + LLF, Courant 0.25."""
    from .timestepper import grid_config

    return grid_config(N1=N1, N2=1, N3=1, dim=1, X1Start=-.5, X1End=1.5, boundaryLeft=capi.BC_OUTFLOW,
                       boundaryRight=capi.BC_OUTFLOW, reconstruction=capi.RECON_MINMOD,
                       riemannSolver=capi.RIEMANN_LLF, slopeLimTheta=2., CourantFactor=0.25, maxDtIncrement=1.3)


def shock_params(test: str = "shock_tube_2") -> capi.Params:
    """grim's shock-test parameters (params.cpp:8-94): EMHD with high-order
    terms, ConductionAlpha 5, ViscosityAlpha 3, Gamma 4/3, floors 1e-20,
    3 Newton iterations of at most 3 line-search steps."""
    p = capi.default_params(True)
    p.conduction = p.viscosity = 1
    p.highOrderTermsConduction = p.highOrderTermsViscosity = 1
    p.ConductionAlpha, p.ViscosityAlpha = 5., 3.
    p.adiabaticIndex = 4. / 3
    p.rhoFloorInFluidElement = p.uFloorInFluidElement = 1e-20
    p.bSqrFloorInFluidElement = p.temperatureFloorInFluidElement = 1e-20
    p.closure = capi.CLOSURE_SHOCK_TESTS
    p.maxNonLinearIter, p.maxLineSearchIters = 3, 3
    p.nonlinearsolve_atol, p.JacobianAssembleEpsilon, p.linesearchfloor = 1e-20, 4e-8, 1e-24
    return p


def shock_tube_prim(grid: capi.GridConfig, p: capi.Params, test: str) -> np.ndarray:
    """timeStepper::initialConditions (problem.cpp:20-163): the left state on
    array indices [0, N1/2) of every X1 line (ghost zones included, the
    inclusive af::seq(0, N1/2) overwritten at N1/2 by the right half), the
    right state on [N1/2, N1T); u = pressure / (Gamma - 1); q~ = dP~ = 0."""
    left, right, _ = SHOCK_TESTS[test]
    gam = p.adiabaticIndex
    prim = np.zeros((p.dof, *grid.shape))
    b1 = p.dof - 3
    for sl, st in ((slice(0, grid.N1 // 2), left), (slice(grid.N1 // 2, None), right)):
        rho, pres, u1, u2, u3, B1, B2, B3 = st
        if test == "stationary_shock":
            pres = pres * (gam - 1.)  # pressureLeft = 1*(Gamma-1), pressureRight = 4.94...*(Gamma-1)
        prim[0][..., sl] = rho
        prim[1][..., sl] = pres / (gam - 1.)
        prim[2][..., sl], prim[3][..., sl], prim[4][..., sl] = u1, u2, u3
        prim[b1][..., sl], prim[b1 + 1][..., sl], prim[b1 + 2][..., sl] = B1, B2, B3
    return prim.reshape(p.dof, -1)


# ---------------------------------------------------------------------------
# Orszag-Tang vortex (src/problem/orzag_tang/)
# ---------------------------------------------------------------------------
def orszag_tang_grid(N: int = 256) -> capi.GridConfig:
    """orzag_tang/params.cpp: 2D Minkowski [0, 1]^2, periodic, MINMOD (theta
    2) + LLF, Courant 0.9."""
    from .timestepper import grid_config

    return grid_config(N1=N, N2=N, reconstruction=capi.RECON_MINMOD, riemannSolver=capi.RIEMANN_LLF,
                       slopeLimTheta=2., CourantFactor=0.9, maxDtIncrement=1.3)


def orszag_tang_params() -> capi.Params:
    """orzag_tang/params.cpp: ideal MHD, Gamma 5/3, floors 1e-20, 5 Newton
    iterations of at most 3 line-search steps to sum R^2 <= 1e-12."""
    p = capi.default_params(False)
    p.adiabaticIndex = 5. / 3
    p.rhoFloorInFluidElement = p.uFloorInFluidElement = 1e-20
    p.bSqrFloorInFluidElement = p.temperatureFloorInFluidElement = 1e-20
    p.maxNonLinearIter, p.maxLineSearchIters = 5, 3
    p.nonlinearsolve_atol, p.JacobianAssembleEpsilon, p.linesearchfloor = 1e-12, 4e-8, 1e-24
    return p


def orszag_tang_prim(grid: capi.GridConfig, p: capi.Params) -> np.ndarray:
    """timeStepper::initialConditions (orzag_tang/problem.cpp:3-34) on every
    zone, ghosts included: rho = 25/(36 pi), u = 5/(12 pi (Gamma-1)),
    v = 0.5 (-sin 2 pi X2, sin 2 pi X1), u~ = gamma v,
    B = (-sin 2 pi X2, sin 4 pi X1) / sqrt(4 pi), at the zone centres
    X = Start + (i + 1/2) dX of the bulk index i (grid.cpp:200-203)."""
    n3, n2, n1 = grid.shape
    g = grid.numGhost
    X1 = grid.X1Start + (np.arange(n1) - g + 0.5) * ((grid.X1End - grid.X1Start) / grid.N1)
    X2 = grid.X2Start + (np.arange(n2) - (g if grid.dim > 1 else 0) + 0.5) * ((grid.X2End - grid.X2Start) / grid.N2)
    X1, X2 = np.broadcast_to(X1[None, None, :], grid.shape), np.broadcast_to(X2[None, :, None], grid.shape)
    prim = np.zeros((p.dof, *grid.shape))
    prim[0] = 25. / (36. * np.pi)
    prim[1] = 5. / (12. * np.pi * (p.adiabaticIndex - 1.))
    v1 = -0.5 * np.sin(2 * np.pi * X2)
    v2 = 0.5 * np.sin(2 * np.pi * X1)
    gamma = 1. / np.sqrt(1. - v1 * v1 - v2 * v2 - 0.)
    prim[2], prim[3] = gamma * v1, gamma * v2
    b1 = p.dof - 3
    prim[b1] = -1. / np.sqrt(4 * np.pi) * np.sin(2 * np.pi * X2)
    prim[b1 + 1] = 1. / np.sqrt(4 * np.pi) * np.sin(4 * np.pi * X1)
    return prim.reshape(p.dof, -1)


# ---------------------------------------------------------------------------
# Bondi inflow onto a Schwarzschild hole (src/problem/bondi_inflow/)
# ---------------------------------------------------------------------------
# The reference directory holds the problem but no params.cpp: sonicRadius,
# mDot and bMag (params.hpp:201-203) get the values below, the usual
# hydrodynamic Bondi test with a weak radial monopole field.
BONDI = dict(sonicRadius=8., mDot=1., bMag=1e-2)


def bondi_grid(N1: int = 256, Rin: float = 3., Rout: float = 40.) -> capi.GridConfig:
    """1D modified Kerr-Schild grid, spin 0, X1 = ln r on [ln Rin, ln Rout],
    the equator (X2 = 1/2 with hSlope 1), WENO5 + LLF, Courant 0.5."""
    from .timestepper import grid_config

    return grid_config(N1=N1, N2=1, N3=1, dim=1, metric=capi.METRIC_MKS, blackHoleSpin=0., hSlope=1.,
                       X1Start=np.log(Rin), X1End=np.log(Rout), X2Start=0., X2End=1., X3Start=0.,
                       X3End=2 * np.pi, boundaryLeft=capi.BC_OUTFLOW, boundaryRight=capi.BC_OUTFLOW,
                       reconstruction=capi.RECON_WENO5, riemannSolver=capi.RIEMANN_LLF, CourantFactor=0.5,
                       maxDtIncrement=1.3)


def bondi_params() -> capi.Params:
    """Ideal MHD, Gamma 4/3; the problem's closure is the linear-modes form
    (tau = 1, chi = nu = c_s^2 tau, bondi_inflow/problem.cpp:3-8)."""
    p = capi.default_params(False)
    p.adiabaticIndex = 4. / 3
    p.closure = capi.CLOSURE_LINEAR_MODES
    return p


def _bondi_T(R, C1, C2, n):
    """SolveT (bondi_inflow/problem.cpp:20-55): regula falsi on FuncT for the
    temperature at radius R, same bracket and stopping rule."""
    def f(T):
        return (1. + (1. + n) * T) ** 2 * (1. - 2. / R + (C1 / R / R / T ** n) ** 2) - C2

    Tmin = 0.6 * (np.sqrt(C2) - 1.) / (n + 1.)
    Tmax = (C1 * np.sqrt(2. / R ** 3)) ** (1. / n)
    T0, T1 = 0.6 * Tmin, Tmax
    f0, f1 = f(T0), f(T1)
    if f0 * f1 > 0.:
        raise ValueError(f"Bondi: no temperature bracket at R = {R}")
    Th = (f1 * T0 - f0 * T1) / (f1 - f0)
    fh = f(Th)
    eps = 1e-12 * (Tmin + Tmax)
    while abs(Th - T0) > eps and abs(Th - T1) > eps and abs(fh) > 1e-14:
        if fh * f0 < 0.:
            T0, f0 = Th, fh
        else:
            T1, f1 = Th, fh
        Th = (f1 * T0 - f0 * T1) / (f1 - f0)
        fh = f(Th)
    return Th


def bondi_constants(p: capi.Params, sonicRadius=BONDI["sonicRadius"], mDot=BONDI["mDot"]):
    """(n, C1, C2, Kp) of the Bondi solution through the sonic point
    (bondi_inflow/problem.cpp:59-67)."""
    n = 1. / (p.adiabaticIndex - 1.)
    Rc = sonicRadius
    uc = np.sqrt(0.5 / Rc)
    vc2 = n * uc * uc / (1. - 3. * uc * uc)
    Tc = vc2 / (1. - vc2) / (n + 1.)
    C1 = Tc ** n * uc * Rc * Rc
    C2 = (1. + (n + 1.) * Tc) ** 2 * (1. - 2. / Rc + uc * uc)
    Kp = (mDot / 4. / np.pi / C1) ** (-1. / n)
    return n, C1, C2, Kp


def bondi_prim(grid: capi.GridConfig, p: capi.Params, **kw) -> np.ndarray:
    """timeStepper::initialConditions (bondi_inflow/problem.cpp:57-145) on
    every zone, ghosts included: per radius r = e^X1 the temperature from
    SolveT, rho = (T/Kp)^n, u = n rho T, u^r = -C1 / (r^2 T^n), u^t from
    u.u = -1 in Kerr-Schild, u~^1 = (u^r + u^t 2/(r+2)) / r, B^1 = bMag / r^3
    (rho floor, u = 0, u~ = 0 inside r = 2)."""
    bMag = kw.pop("bMag", BONDI["bMag"])
    n, C1, C2, Kp = bondi_constants(p, **kw)
    n3, n2, n1 = grid.shape
    X1 = grid.X1Start + (np.arange(n1) - grid.numGhost + 0.5) * ((grid.X1End - grid.X1Start) / grid.N1)
    prim = np.zeros((p.dof, n3, n2, n1))
    b1 = p.dof - 3
    for i, R in enumerate(np.exp(X1)):
        if R > 2. + 1e-8:
            T = _bondi_T(R, C1, C2, n)
            rho = (T / Kp) ** n
            ur = -C1 / R / R / T ** n
            ut = (2. * ur / R + np.sqrt(ur * ur + 1. - 2. / R)) / (1. - 2. / R)
        else:
            T, rho, ur, ut = 1., p.rhoFloorInFluidElement, 0., 1.
        prim[0][..., i] = rho
        prim[1][..., i] = n * rho * T
        prim[2][..., i] = (ur + ut * 2. / (R + 2.)) / R
        prim[b1][..., i] = bMag / R ** 3
    return prim.reshape(p.dof, -1)


# ---------------------------------------------------------------------------
# Buoyancy instabilities of a conducting atmosphere (src/problem/
# buoyancy_instabilities/)
# ---------------------------------------------------------------------------
# params.cpp:8-96.  grim's 2D grid has one X3 zone of width (X3End -
# X3Start) / N3 with N3 = 128 (the DMDA of a 2D run ignores N3, the
# coordinates do not): X3 = pi / 256 at the zone centre, which only the
# dumped coordinates see.
BUOYANCY = dict(Rin=200., Rout=300., InitialPerturbationAmplitude=4e-2, N3=128, X3End=np.pi)
BUOYANCY_FILES = ("atmosphere_soln_rho.txt", "atmosphere_soln_u.txt", "atmosphere_soln_phi.txt",
                  "atmosphere_soln_rCoords.txt")


def buoyancy_grid(N1: int = 64, N2: int = 64) -> capi.GridConfig:
    """params.cpp: 2D modified Kerr-Schild, spin 0, hSlope 1, r in [Rin, Rout]
    with Rin = 100 (1 + sqrt(1 - a^2)) = 200 and Rout = 300, X2 in [1e-8,
    1 - 1e-8], theta derefined and cylindrified (X1cyl = ln 4 Rin, X2cyl =
    3 / N2), outflow in X1, mirror in X2, MINMOD (theta 2) + LLF, Courant
    0.5."""
    from .timestepper import grid_config

    Rin = BUOYANCY["Rin"]
    return grid_config(N1=N1, N2=N2, N3=1, dim=2, metric=capi.METRIC_MKS, blackHoleSpin=0., hSlope=1.,
                       X1Start=np.log(Rin), X1End=np.log(BUOYANCY["Rout"]), X2Start=1e-8, X2End=1. - 1e-8,
                       X3Start=0., X3End=BUOYANCY["X3End"] / BUOYANCY["N3"], DerefineThetaHorizon=1,
                       DoCylindrify=1, X1cyl=np.log(4. * Rin), X2cyl=3. / N2, boundaryLeft=capi.BC_OUTFLOW,
                       boundaryRight=capi.BC_OUTFLOW, boundaryBottom=capi.BC_MIRROR, boundaryTop=capi.BC_MIRROR,
                       boundaryBack=capi.BC_PERIODIC, boundaryFront=capi.BC_PERIODIC,
                       reconstruction=capi.RECON_MINMOD, riemannSolver=capi.RIEMANN_LLF, slopeLimTheta=2.,
                       CourantFactor=0.5, maxDtIncrement=1.3)


def buoyancy_params() -> capi.Params:
    """params.cpp: conduction with the high-order rescaling, no viscosity,
    ConductionAlpha 1, Gamma 5/3, fluid-element floors 1e-20, 3 Newton
    iterations of at most 3 line-search steps to sum R^2 <= 1e-6, and the
    problem's closure chi = sqrt(r), tau = 1.2 chi / c_s^2
    (GRIM_CLOSURE_BUOYANCY, problem.cpp:11-21)."""
    p = capi.default_params(True)
    p.conduction, p.viscosity = 1, 0
    p.highOrderTermsConduction = p.highOrderTermsViscosity = 1
    p.ConductionAlpha = p.ViscosityAlpha = 1.
    p.ConductionClosureFactor = p.ViscosityClosureFactor = 1.
    p.adiabaticIndex = 5. / 3
    p.rhoFloorInFluidElement = p.uFloorInFluidElement = 1e-20
    p.bSqrFloorInFluidElement = p.temperatureFloorInFluidElement = 1e-20
    p.closure = capi.CLOSURE_BUOYANCY
    p.maxNonLinearIter, p.maxLineSearchIters = 3, 3
    p.nonlinearsolve_atol, p.JacobianAssembleEpsilon, p.linesearchfloor = 1e-6, 4e-8, 1e-24
    return p


def zone_radii(grid: capi.GridConfig) -> np.ndarray:
    """r = e^X1 of the N1 + 2 numGhost zone centres of an X1 line (MKS)."""
    X1 = grid.X1Start + (np.arange(grid.N1 + 2 * grid.numGhost) - grid.numGhost + 0.5) * (
        (grid.X1End - grid.X1Start) / grid.N1)
    return np.exp(X1)


def hydrostatic_atmosphere(grid: capi.GridConfig, p: capi.Params, rho0: float = 1., T0: float = 5e-3):
    """The four input profiles of the problem (rho, u, q~, r per X1 zone,
    ghosts included), whose generator the reference does not hold: a static
    atmosphere in hydrostatic equilibrium around a Schwarzschild hole (M =
    1) with T = p / rho = T0 r0 / r, r0 = the first profile radius.  For a
    static perfect fluid dp/dr = -(rho + u + p) / (r^2 (1 - 2/r)) with u =
    p / (Gamma - 1), which integrates in closed form:

        p(r) = p0 ((r - 2) / (r0 - 2))^(-1 / (T0 r0))
                  (((r - 2) / r) / ((r0 - 2) / r0))^(-Gamma / (2 (Gamma - 1)))

    with p0 = rho0 T0; q~ = 0 (no heat flux along the toroidal field)."""
    r = zone_radii(grid)
    r0 = r[0]
    gam = p.adiabaticIndex
    pres = rho0 * T0 * ((r - 2.) / (r0 - 2.)) ** (-1. / (T0 * r0)) * (
        ((r - 2.) / r) / ((r0 - 2.) / r0)) ** (-gam / (2. * (gam - 1.)))
    T = T0 * r0 / r
    return pres / T, pres / (gam - 1.), np.zeros_like(r), r


def write_atmosphere_files(directory: str, rho, u, q, rCoords) -> None:
    """The problem's input files (problem.cpp:3-9): one value per line, X1
    zones -numGhost .. N1 + numGhost - 1, printed to 17 significant digits so
    that atof reads the same doubles back."""
    import os

    for name, a in zip(BUOYANCY_FILES, (rho, u, q, rCoords)):
        with open(os.path.join(directory, name), "w") as f:
            f.writelines(f"{v:.17e}\n" for v in np.asarray(a, dtype=np.float64))


def read_atmosphere_files(grid: capi.GridConfig, directory: str = "."):
    """The file reading of initialConditions (problem.cpp:24-114): N1 + 2
    numGhost lines of each file; a missing file or a short one is an error,
    with grim's messages."""
    import os

    count = grid.N1 + 2 * grid.numGhost
    out = []
    for name in BUOYANCY_FILES:
        path = os.path.join(directory, name)
        if not os.path.exists(path):
            raise FileNotFoundError("Input data files not found!")
        with open(path) as f:
            lines = f.readlines()
        if len(lines) < count:
            raise ValueError("Found the input data files but error in reading them! "
                             "Check number of grid zones in python script")
        out.append(np.array([float(v) for v in lines[:count]]))
    return tuple(out)


def noise_stream(seed: int, n: int) -> np.ndarray:
    """Uniforms in [0, 1) of zones 0..n-1 from the library's counter-based
    stream (splitmix64 of seed + G (32 z + 1), G = 0x9E3779B97F4A7C15), the
    stream the torus perturbation draws; it replaces PETSc's RAND48."""
    z = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = np.uint64(seed) + np.uint64(0x9E3779B97F4A7C15) * (np.uint64(32) * z + np.uint64(1))
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        x = x ^ (x >> np.uint64(31))
    return (x >> np.uint64(11)).astype(np.float64) * (1. / 9007199254740992.)


def buoyancy_prim(grid: capi.GridConfig, p: capi.Params, geom_center: np.ndarray, rho1D, u1D, q1D, rCoords1D,
                  amplitude: float = BUOYANCY["InitialPerturbationAmplitude"], seed: int = 20261203):
    """timeStepper::initialConditions (problem.cpp:22-216) on every zone:

    * rho, u and (conduction) q~ of X1 zone i from the profiles, after
      checking |r(i) - rCoords1D[i]| <= 1e-10 on the first X2 / X3 line;
    * u~^1 = (c u^r / r - sqrt(-a b^4 - a b^2 c (u^r)^2 / r^2 - b^2 c)) /
      (a b^2 + c) with a = g_11, b = g^01, c = g^00 and u^r = 0 (the
      static Boyer-Lindquist observer, in grim's expression order), times
      1 + amplitude (uniform - 1/2) on the bulk zones (the noise stream in
      place of PETSc's RAND48); u~^2 = u~^3 = 0;
    * B^1 = B^2 = 0, B^3 = 1e-5 / sqrt(-g).

    geom_center: the zone-centre geometry planes (TimeStepper.geometry(0)).
    Returns (prim, side_state): side_state holds rho, u and u~^1 of primIC,
    the planes setProblemSpecificBCs imposes on the X1 sides
    (TimeStepper.set_x1_side_state)."""
    n3, n2, n1 = grid.shape
    n = n1 * n2 * n3
    geom = np.asarray(geom_center).reshape(-1, n)
    r = geom[22].reshape(n3, n2, n1)
    rC = np.asarray(rCoords1D, dtype=np.float64)
    for i in range(n1):
        if abs(r[0, 0, i] - rC[i]) > 1e-10:
            raise ValueError(f"Mismatch in rCoords! Check r coords in python script. r = {r[0, 0, i]:.18f}, "
                             f"rCoords = {rC[i]:.18f}, i = {i}")
    prim = np.zeros((p.dof, n3, n2, n1))
    prim[0] = np.asarray(rho1D, dtype=np.float64)[None, None, :]
    prim[1] = np.asarray(u1D, dtype=np.float64)[None, None, :]
    if p.conduction:
        prim[5] = np.asarray(q1D, dtype=np.float64)[None, None, :]
    a = geom[4].reshape(n3, n2, n1)   # gCov[1][1]
    b = geom[11].reshape(n3, n2, n1)  # gCon[0][1]
    c = geom[10].reshape(n3, n2, n1)  # gCon[0][0]
    uCon1 = 0. * prim[2]
    vUpr = (c * uCon1 / r - np.sqrt(-a * b * b * b * b - a * b * b * c * uCon1 * uCon1 / (r * r) - b * b * c)) / (
        a * b * b + c)
    u1 = vUpr.copy()
    g = grid.numGhost
    bulk = (slice(0, 1) if grid.dim < 3 else slice(g, n3 - g), slice(g, n2 - g) if grid.dim > 1 else slice(0, 1),
            slice(g, n1 - g))
    rand = noise_stream(seed, n).reshape(n3, n2, n1)
    u1[bulk] = vUpr[bulk] * (1. + amplitude * (rand[bulk] - 0.5))
    prim[2] = u1
    b1 = p.dof - 3
    prim[b1 + 2] = 1e-5 / geom[15].reshape(n3, n2, n1)
    side = np.stack([prim[0], prim[1], prim[2]]).reshape(3, n)
    return prim.reshape(p.dof, -1), side
