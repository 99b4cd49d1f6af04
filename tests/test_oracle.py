"""Pins for the oracle (CPU only).

The oracle (oracle/grim_oracle.hpp) restates grim's per-zone algorithm; these
tests check it against what fixes each part independently of the
restatement, chosen so that a dropped term, wrong sign or swapped index in
the oracle fails at least one of them.

Pinning status per oracle part:
  mks geometry (setgCovInXCoords, setgDetAndgConFromgCov): closed-form MKS
      metric (McKinney & Gammie 2004, as restated in the reference test
      src/geometry/test_geometry.py:97-150), exact inverse and determinant.
  connection coefficients (computeConnectionCoeffs): complex-step
      derivatives of the closed-form metric (machine precision) and spot
      closed forms of src/geometry/test_geometry.py:178-346.
  fluidElement::set / computeFluxes(0): u.u = -1, b.u = 0; the contraction
      identities of the EMHD stress tensor (trace, u T b, b T b); Minkowski
      reduction to the textbook special-relativistic MHD conserved variables.
  computeImplicitSources: closed form g qTilde / tau.
  computeTimeDerivSources: vanishes for identical states, scales as 1/dt;
      its detailed form is parity unpinned here (pinned later by the EMHD
      linear-mode solution of the full step).
  computeResidual: known-root construction, mask linearity.
  FD Jacobian: agrees with the exact dual-number Jacobian to FD accuracy;
      dual Jacobian agrees with central differences.
  LU with partial pivoting: LAPACK (numpy.linalg.solve) on random systems,
      a system that needs every row exchange.
  Newton + line search: constructed roots recovered; lock-step and per-zone
      readings agree; iteration cap leaves zones unconverged.
"""
import numpy as np
import pytest

import batchgen

SPINS = [(0.9375, 0.3), (0.5, 0.7)]


# ----------------------------------------------------------------- geometry
def closed_form_mks(X1, X2, a, h, xp=np):
    """Closed-form MKS g_mu_nu (r = e^X1, theta = pi X2 + (1-h)/2 sin 2 pi X2)."""
    r = xp.exp(X1)
    th = np.pi * X2 + 0.5 * (1.0 - h) * xp.sin(2.0 * np.pi * X2)
    dr = r
    dth = np.pi * (1.0 + (1.0 - h) * xp.cos(2.0 * np.pi * X2))
    sig = r * r + (a * xp.cos(th)) ** 2
    s2 = xp.sin(th) ** 2
    z = 0.0 * r
    g = [[z] * 4 for _ in range(4)]
    g[0][0] = -(1.0 - 2.0 * r / sig)
    g[0][1] = g[1][0] = (2.0 * r / sig) * dr
    g[0][3] = g[3][0] = -2.0 * a * r * s2 / sig
    g[1][1] = (1.0 + 2.0 * r / sig) * dr * dr
    g[1][3] = g[3][1] = -a * (1.0 + 2.0 * r / sig) * s2 * dr
    g[2][2] = sig * dth * dth
    g[3][3] = s2 * (sig + a * a * (1.0 + 2.0 * r / sig) * s2)
    return xp.stack([xp.stack(row) for row in g])  # (4, 4, n)


def random_X(n, seed=7):
    rng = np.random.default_rng(seed)
    X1 = rng.uniform(np.log(1.5), np.log(40.0), n)
    X2 = rng.uniform(0.06, 0.94, n)
    return X1, X2


@pytest.mark.parametrize("a,h", SPINS)
def test_mks_metric_closed_form(oracle, a, h):
    X1, X2 = random_X(400)
    geo = oracle.mks_geometry(X1, X2, a, h)
    G = closed_form_mks(X1, X2, a, h)
    np.testing.assert_allclose(geo["gCov"].reshape(4, 4, -1), G, rtol=1e-13, atol=1e-13)
    Gz = np.moveaxis(G, -1, 0)
    inv = np.linalg.inv(Gz)
    np.testing.assert_allclose(np.moveaxis(geo["gCon"].reshape(4, 4, -1), -1, 0), inv, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(geo["g"], np.sqrt(-np.linalg.det(Gz)), rtol=1e-12)
    np.testing.assert_allclose(geo["alpha"], 1.0 / np.sqrt(-inv[:, 0, 0]), rtol=1e-12)


def complex_step_gamma(X1, X2, a, h):
    """Gamma^mu_{nu lam} from complex-step derivatives of the closed form."""
    hstep = 1e-30
    G = closed_form_mks(X1, X2, a, h)
    dG = np.zeros((4, 4, 4, X1.size))  # dG[k] = d g / d X^k
    dG[1] = closed_form_mks(X1 + 1j * hstep, X2 + 0j, a, h).imag / hstep
    dG[2] = closed_form_mks(X1 + 0j, X2 + 1j * hstep, a, h).imag / hstep
    down = np.zeros((4, 4, 4, X1.size))
    for e in range(4):
        for m in range(4):
            for l in range(4):
                down[e, m, l] = 0.5 * (dG[l][e, m] + dG[m][e, l] - dG[e][m, l])
    Ginv = np.moveaxis(np.linalg.inv(np.moveaxis(G, -1, 0)), 0, -1)
    return np.einsum("mez,enlz->mnlz", Ginv, down)


@pytest.mark.parametrize("a,h", SPINS)
def test_connection_vs_complex_step(oracle, a, h):
    X1, X2 = random_X(300, seed=3)
    geo = oracle.mks_geometry(X1, X2, a, h)
    gam = geo["gamma"].reshape(4, 4, 4, -1)
    ref = complex_step_gamma(X1, X2, a, h)
    scale = np.abs(ref).max(axis=(0, 1, 2))
    err = np.abs(gam - ref).max(axis=(0, 1, 2)) / scale
    assert err.max() < 2e-8  # grim's 2nd-order central difference, eps = 1e-5


def test_connection_spot_closed_forms(oracle):
    # Entries of src/geometry/test_geometry.py:178-346 at a=0.9375, h=0.3
    a, h = 0.9375, 0.3
    X1, X2 = random_X(200, seed=11)
    gam = oracle.mks_geometry(X1, X2, a, h)["gamma"].reshape(4, 4, 4, -1)
    r = np.exp(X1)
    th = np.pi * X2 + 0.5 * (1 - h) * np.sin(2 * np.pi * X2)
    dth = np.pi * (1 + (1 - h) * np.cos(2 * np.pi * X2))
    sig = r ** 2 + (a * np.cos(th)) ** 2
    sigm = r ** 2 - (a * np.cos(th)) ** 2
    checks = {
        (0, 0, 0): 2 * r * sigm / sig ** 3,
        (0, 2, 2): -2 * r ** 2 * dth ** 2 / sig,
        (1, 2, 2): -(a ** 2 + r * (r - 2)) * dth ** 2 / sig,
        (2, 1, 2): r ** 2 / sig,
        (3, 0, 0): a * sigm / sig ** 3,
        (3, 2, 2): -a * r * dth ** 2 / sig,
    }
    for (m, n_, l), val in checks.items():
        np.testing.assert_allclose(gam[m, n_, l], val, rtol=2e-7, atol=1e-9)


# ----------------------------------------------------------- fluid element
def _minkowski_geom(n):
    geom = np.zeros((16, n))
    geom[0] = -1.0
    geom[4] = geom[7] = geom[9] = 1.0  # g11, g22, g33
    geom[10] = -1.0                    # g^00
    geom[14] = 1.0                     # alpha
    geom[15] = 1.0                     # sqrt(-g)
    return geom


def _random_prims(n, emhd, seed=5):
    rng = np.random.default_rng(seed)
    nf = 7 if emhd else 5
    P = np.zeros((nf + 3, n))
    P[0] = 10 ** rng.uniform(-3, 0, n)
    P[1] = P[0] * 10 ** rng.uniform(-2, 1, n)
    P[2:5] = rng.uniform(-2, 2, (3, n))
    if emhd:
        P[5:7] = rng.uniform(-0.3, 0.3, (2, n))
    P[nf:nf + 3] = rng.uniform(-1, 1, (3, n))
    return P


def test_minkowski_textbook_srmhd(oracle):
    """U of grim's computeFluxes(0) vs textbook SR-MHD conserved variables."""
    n = 500
    p = oracle.default_params(False)
    P = _random_prims(n, False)
    U = oracle.conserved(p, P, _minkowski_geom(n))
    rho, u = P[0], P[1]
    press = (p.adiabaticIndex - 1) * u
    w = rho + u + press
    ut = P[2:5]
    gam = np.sqrt(1 + (ut * ut).sum(0))
    v = ut / gam
    B = P[5:8]
    vB = (v * B).sum(0)
    b0 = gam * vB
    bi = B / gam + b0 * v
    bsq = (B * B).sum(0) / gam ** 2 + vB ** 2
    D = rho * gam
    E = (w + bsq) * gam ** 2 - (press + 0.5 * bsq) - b0 ** 2
    S = (w + bsq) * gam ** 2 * v - b0 * bi
    np.testing.assert_allclose(U[0], D, rtol=1e-13)
    np.testing.assert_allclose(U[1], D - E, rtol=1e-10, atol=1e-12 * np.abs(E).max())
    np.testing.assert_allclose(U[2:5], S, rtol=1e-11, atol=1e-12)
    # induction rows: B^i
    np.testing.assert_allclose(U[5:8], B, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("emhd", [False, True])
def test_element_contractions(oracle, emhd):
    """Contraction identities of the (E)MHD stress tensor on MKS zones."""
    n = 2000
    p = oracle.default_params(emhd)
    b = batchgen.kerr_schild_batch(oracle, n, emhd=emhd)
    P = b["prim_star"]
    if emhd:
        p.closure = oracle.CLOSURE_PER_ZONE
    e = oracle.element(p, P, b["geom"], b["closure"])
    uC = np.stack([e["uCon%d" % i] for i in range(4)])
    uL = np.stack([e["uCov%d" % i] for i in range(4)])
    bC = np.stack([e["bCon%d" % i] for i in range(4)])
    bL = np.stack([e["bCov%d" % i] for i in range(4)])
    T = np.stack([np.stack([e["T%d%d" % (m, k)] for k in range(4)]) for m in range(4)])
    np.testing.assert_allclose((uC * uL).sum(0), -1.0, atol=1e-11 * (uC[0] ** 2).max())
    bu = (bC * uL).sum(0)
    assert np.all(np.abs(bu) <= 1e-10 * np.sqrt(e["bSqr"]) * np.abs(uC).max(0) * np.abs(uL).max(0) + 1e-300)
    np.testing.assert_allclose((bC * bL).sum(0) + p.bSqrFloorInFluidElement, e["bSqr"], rtol=1e-9)
    scale = (e["rho"] + e["u"] + e["bSqr"]) * (uC[0] * np.abs(uL).max(0))
    # trace: T^mu_mu = 3p - rho - u (q and deltaP terms are traceless)
    tr = np.einsum("mmz->z", T)
    np.testing.assert_allclose(tr, 3 * e["pressure"] - e["rho"] - e["u"], atol=1e-9 * scale.max())
    bn = np.sqrt(e["bSqr"])
    # u_mu T^mu_nu bhat^nu = -q
    uTb = np.einsum("mz,mnz,nz->z", uL, T, bC) / bn
    np.testing.assert_allclose(uTb, -e["q"], atol=1e-8 * (scale * uC[0]).max())
    # bhat_mu T^mu_nu bhat^nu = p - b^2/2 - 2/3 deltaP
    bTb = np.einsum("mz,mnz,nz->z", bL, T, bC) / e["bSqr"]
    np.testing.assert_allclose(bTb, e["pressure"] - 0.5 * e["bSqr"] - 2.0 / 3.0 * e["deltaP"],
                               atol=1e-8 * (scale * uC[0]).max())


def _emhd_batch(oracle, n, dt=1.0):
    p = oracle.default_params(True)
    p.closure = oracle.CLOSURE_PER_ZONE
    b = batchgen.kerr_schild_batch(oracle, n, emhd=True)
    B = oracle.Batch(b["prim_star"], b["geom"], np.zeros((p.nf, n)), b["prim_old"], None, b["closure"],
                     None, dt)
    return p, b, B


def test_implicit_sources_closed_form(oracle):
    n = 500
    p, b, B = _emhd_batch(oracle, n)
    U, Si, Std = oracle.residual_pieces(p, B)
    tau = p.tau0
    g = b["geom"][15]
    np.testing.assert_allclose(Si[5], g * b["prim_star"][5] / tau, rtol=1e-15)
    np.testing.assert_allclose(Si[6], g * b["prim_star"][6] / tau, rtol=1e-15)
    assert np.all(Si[:5] == 0) and np.all(Std[:5] == 0)


def test_time_derivative_sources_structure(oracle):
    n = 500
    p, b, B = _emhd_batch(oracle, n)
    # identical old / new states: every time derivative vanishes, leaving the
    # high-order -1/2 g divu x terms, which vanish too
    same = oracle.Batch(b["prim_old"], b["geom"], np.zeros((p.nf, n)), b["prim_old"], None, b["closure"])
    _, _, Std0 = oracle.residual_pieces(p, same)
    assert np.all(Std0 == 0)
    # Std is proportional to 1/dt at fixed states
    _, _, S1 = oracle.residual_pieces(p, B)
    B2 = oracle.Batch(B.prim, B.geom, B.rhs, B.primOld, None, B.closure, None, 2.0)
    _, _, S2 = oracle.residual_pieces(p, B2)
    np.testing.assert_allclose(S2[5:], 0.5 * S1[5:], rtol=1e-12, atol=1e-300)


# --------------------------------------------------------------- residual
@pytest.mark.parametrize("emhd", [False, True])
def test_known_root_and_mask(oracle, emhd):
    n = 1000
    p = oracle.default_params(emhd)
    if emhd:
        p.closure = oracle.CLOSURE_PER_ZONE
    b = batchgen.kerr_schild_batch(oracle, n, emhd=emhd)
    B = oracle.Batch(b["prim_star"], b["geom"], np.zeros((p.nf, n)), b["prim_old"], None, b["closure"])
    rhs = oracle.known_root_rhs(p, B)
    mask = 1.0 / np.abs(rhs).max(axis=0)
    R = oracle.residual(p, oracle.Batch(b["prim_star"], b["geom"], rhs, b["prim_old"], None, b["closure"],
                                        None, 1.0, mask))
    assert np.abs(R).max() < 1e-13
    Rg = oracle.residual(p, oracle.Batch(b["guess"], b["geom"], rhs, b["prim_old"], None, b["closure"]))
    Rm = oracle.residual(p, oracle.Batch(b["guess"], b["geom"], rhs, b["prim_old"], None, b["closure"],
                                         None, 1.0, mask))
    np.testing.assert_allclose(Rm, Rg * mask, rtol=1e-15)
    # ideal rows of the folded form: R = (U - rhs)/dt exactly
    U, _, _ = oracle.residual_pieces(p, oracle.Batch(b["guess"], b["geom"], rhs, b["prim_old"], None,
                                                     b["closure"]))
    np.testing.assert_array_equal(Rg[:5], U[:5] - rhs[:5])


# ------------------------------------------------------------ Jacobians
@pytest.mark.parametrize("emhd", [False, True])
def test_fd_jacobian_vs_dual(oracle, emhd):
    n = 2000
    p = oracle.default_params(emhd)
    if emhd:
        p.closure = oracle.CLOSURE_PER_ZONE
    b = batchgen.kerr_schild_batch(oracle, n, emhd=emhd)
    B = oracle.Batch(b["prim_star"], b["geom"], np.zeros((p.nf, n)), b["prim_old"], None, b["closure"])
    rhs = oracle.known_root_rhs(p, B)
    G = oracle.Batch(b["guess"], b["geom"], rhs, b["prim_old"], None, b["closure"])
    Jfd = oracle.fd_jacobian(p, G).reshape(p.nf, p.nf, n)
    Jd = oracle.dual_jacobian(p, G).reshape(p.nf, p.nf, n)
    # Error relative to each column's largest entry.  The forward difference
    # loses digits to cancellation where |U| >> |P dU/dP| (magnetically
    # dominated zones with u/rho ~ 1e-3: ~1e-4 relative), so the tail is
    # loose; a transposed or mis-signed entry would be O(1).
    # The EMHD batch's near-floor zones (rho ~ 1e-8, u ~ 1e-11) are excluded:
    # there grim's step eps*u ~ 1e-18 leaves the forward difference wrong by
    # up to 100x while the dual Jacobian still matches central differences
    # (test_dual_jacobian_vs_central_differences).
    if emhd:
        keep = batchgen.uniform(batchgen.SEED, np.arange(n, dtype=np.uint64), 21) >= 0.1
        Jfd, Jd = Jfd[:, :, keep], Jd[:, :, keep]
    colscale = np.abs(Jd).max(axis=0, keepdims=True) + 1e-300
    err = np.abs(Jfd - Jd) / colscale
    worst = err.max(axis=(0, 1))
    f = 10.0 if emhd else 1.0
    assert np.median(worst) < 1e-6 * f
    assert np.percentile(worst, 90) < 1e-4 * f
    if not emhd:
        assert worst.max() < 1e-2


@pytest.mark.parametrize("emhd", [False, True])
def test_fd_jacobian_assembly_rule(oracle, emhd):
    """grim's rule, solve.cpp:82-109, rebuilt column by column from residuals."""
    n = 300
    p = oracle.default_params(emhd)
    if emhd:
        p.closure = oracle.CLOSURE_PER_ZONE
    b = batchgen.kerr_schild_batch(oracle, n, emhd=emhd)
    B = oracle.Batch(b["prim_star"], b["geom"], np.zeros((p.nf, n)), b["prim_old"], None, b["closure"])
    rhs = oracle.known_root_rhs(p, B)
    P0 = b["guess"].copy()
    P0[2, :50] = 1e-9  # |P| < eps/2: absolute step eps
    G = oracle.Batch(P0, b["geom"], rhs, b["prim_old"], None, b["closure"])
    J = oracle.fd_jacobian(p, G).reshape(p.nf, p.nf, n)
    R0 = oracle.residual(p, G)
    eps = p.JacobianAssembleEpsilon
    for j in range(p.nf):
        P = P0.copy()
        small = np.abs(P0[j]) < 0.5 * eps
        P[j] = np.where(small, P0[j] + eps, P0[j] + eps * P0[j])
        Rp = oracle.residual(p, oracle.Batch(P, b["geom"], rhs, b["prim_old"], None, b["closure"]))
        col = (Rp - R0) / (P[j] - P0[j])
        np.testing.assert_array_equal(J[:, j, :], col)  # row i = residual, column j = primitive


@pytest.mark.parametrize("emhd", [False, True])
def test_dual_jacobian_vs_central_differences(oracle, emhd):
    n = 300
    p = oracle.default_params(emhd)
    if emhd:
        p.closure = oracle.CLOSURE_PER_ZONE
    b = batchgen.kerr_schild_batch(oracle, n, emhd=emhd)
    B = oracle.Batch(b["prim_star"], b["geom"], np.zeros((p.nf, n)), b["prim_old"], None, b["closure"])
    rhs = oracle.known_root_rhs(p, B)
    P0 = b["guess"]
    Jd = oracle.dual_jacobian(p, oracle.Batch(P0, b["geom"], rhs, b["prim_old"], None, b["closure"]))
    Jd = Jd.reshape(p.nf, p.nf, n)
    for j in range(p.nf):
        h = 1e-5 * np.abs(P0[j]) + 1e-12

        def R_at(s):
            P = P0.copy()
            P[j] = P0[j] + s * h
            return oracle.residual(p, oracle.Batch(P, b["geom"], rhs, b["prim_old"], None, b["closure"]))
        # 4th-order central difference
        col = (-R_at(2) + 8 * R_at(1) - 8 * R_at(-1) + R_at(-2)) / (12 * h)
        scale = np.abs(Jd[:, j, :]).max(axis=0) + 1e-300
        err = np.abs(col - Jd[:, j, :]).max(axis=0) / scale
        assert np.median(err) < 1e-8, (j, np.median(err))
        assert np.percentile(err, 99) < 1e-4, (j, np.percentile(err, 99))


# ------------------------------------------------------------------- LU
def test_lu_vs_lapack(oracle):
    rng = np.random.default_rng(1)
    for nvar in (1, 2, 5, 7, 10):
        for _ in range(20):
            A = rng.standard_normal((nvar, nvar)) * 10 ** rng.uniform(-3, 3, (nvar, 1))
            bvec = rng.standard_normal(nvar)
            x, info = oracle.lu_solve(A, bvec)
            assert info == 0
            ref = np.linalg.solve(A, bvec)
            np.testing.assert_allclose(x, ref, rtol=1e-9 * np.linalg.cond(A) / 1e3 + 1e-12, atol=1e-12)


def test_lu_needs_every_row_exchange(oracle):
    # anti-diagonal dominant: partial pivoting must swap at every step
    A = np.array([[1e-12, 0, 0, 1.0], [0, 1e-12, 2.0, 0], [0, 3.0, 1e-12, 0], [4.0, 0, 0, 1e-12]])
    bvec = np.array([1.0, 2.0, 3.0, 4.0])
    x, info = oracle.lu_solve(A, bvec)
    np.testing.assert_allclose(A @ x, bvec, rtol=1e-14)
    np.testing.assert_allclose(x, np.linalg.solve(A, bvec), rtol=1e-13)
    _, info0 = oracle.lu_solve(np.zeros((3, 3)), np.ones(3))
    assert info0 == 1


# --------------------------------------------------------------- Newton
@pytest.mark.parametrize("emhd", [False, True])
def test_newton_recovers_constructed_root(oracle, emhd):
    n = 4000
    p = oracle.default_params(emhd)
    p.maxNonLinearIter = 30
    if emhd:
        p.closure = oracle.CLOSURE_PER_ZONE
    b = batchgen.kerr_schild_batch(oracle, n, emhd=emhd)
    B = oracle.Batch(b["prim_star"], b["geom"], np.zeros((p.nf, n)), b["prim_old"], None, b["closure"])
    rhs = oracle.known_root_rhs(p, B)
    mask = 1.0 / np.abs(rhs).max(axis=0)
    G = oracle.Batch(b["guess"], b["geom"], rhs, b["prim_old"], None, b["closure"], None, 1.0, mask)
    bad, it, l2, ev = oracle.newton(p, G)
    ok = l2 <= p.nonlinearsolve_atol
    if not emhd:
        assert bad == 0 and ok.all()
        assert it.max() <= 4
    else:
        assert ok.mean() > 0.95
    err = np.abs(G.prim[:p.nf] - b["prim_star"][:p.nf]) / np.maximum(np.abs(b["prim_star"][:p.nf]), 1e-30)
    assert np.median(err[:, ok].max(axis=0)) < 1e-9
    assert np.all(G.prim[p.nf:] == b["guess"][p.nf:])  # B held fixed


def test_newton_global_vs_per_zone(oracle):
    n = 2000
    p = oracle.default_params(False)
    p.maxNonLinearIter = 10
    b = batchgen.kerr_schild_batch(oracle, n)
    B = oracle.Batch(b["prim_star"], b["geom"], np.zeros((p.nf, n)))
    rhs = oracle.known_root_rhs(p, B)
    mask = 1.0 / np.abs(rhs).max(axis=0)
    A = oracle.Batch(b["guess"], b["geom"], rhs, residualMask=mask)
    Gl = A.copy()
    _, itA, l2A, _ = oracle.newton(p, A, oracle.NEWTON_PER_ZONE)
    _, itG, l2G, _ = oracle.newton(p, Gl, oracle.NEWTON_GLOBAL)
    assert np.all(itG == itA.max())            # lock-step: every zone updated each iteration
    assert np.all(l2G <= p.nonlinearsolve_atol)
    same = itA == itA.max()
    np.testing.assert_array_equal(A.prim[:, same], Gl.prim[:, same])
    # Extra updates of converged zones move them only inside the tolerance
    # ball; in ill-conditioned zones (u/rho ~ 1e-3, b^2/rho ~ 50) that is
    # ~1e-5 relative in u, elsewhere far less.
    rel = np.abs(A.prim[:5] - Gl.prim[:5]) / np.abs(Gl.prim[:5])
    assert np.median(rel.max(axis=0)) < 1e-9


def test_newton_iteration_cap_leaves_unconverged(oracle):
    n = 500
    p = oracle.default_params(False)
    p.maxNonLinearIter = 1
    b = batchgen.kerr_schild_batch(oracle, n)
    B = oracle.Batch(b["prim_star"], b["geom"], np.zeros((p.nf, n)))
    rhs = oracle.known_root_rhs(p, B)
    mask = 1.0 / np.abs(rhs).max(axis=0)
    G = oracle.Batch(b["guess"], b["geom"], rhs, residualMask=mask)
    bad, it, l2, _ = oracle.newton(p, G)
    assert bad > 0 and np.all(it == 1)
    assert bad == int((l2 > p.nonlinearsolve_atol).sum())


# -------------------------------------------------------- operation counts
def test_bench_operation_tallies_match_oracle_counts(oracle):
    """The per-zone operation tallies bench.py prices its VALU fraction with:
    the exact residual's entry equals the oracle's live count (operations the
    residual rows depend on and that depend on the guess), and the LU's
    entry equals the counting scalar's tally of one 5x5 solve.  The live
    count is a subset of every operation grim's formulation performs."""
    import bench

    ref = batchgen.kerr_schild_batch(oracle, 16, emhd=False, seed=20261202)
    p = oracle.default_params(False)
    B = oracle.Batch(ref["guess"], ref["geom"], np.ones((p.nf, 16)))
    total_res, lu = oracle.count_ops(p, B)
    live = oracle.count_live_ops(p, B)
    assert int(live.sum()) == bench.OPS["res_exact"]
    assert int(lu[:4].sum()) == bench.OPS["lu"] and lu[3] == 0 and lu[4] == 0
    # live operations never exceed all operations of the same kind
    assert (live[:5] <= total_res).all() and live[:5].sum() < total_res.sum()
    # one square root (the Lorentz factor) per ideal residual, no transcendental
    assert live[3] == 1 and live[4] == 0


def test_live_count_is_a_property_of_the_formulation(oracle):
    """The same count at every zone of a batch (no data-dependent branches
    outside the floors), and more work for EMHD, whose rows add the closure,
    the heat-flux and viscous stresses and the time-derivative sources."""
    counts = {}
    for emhd in (False, True):
        ref = batchgen.kerr_schild_batch(oracle, 8, emhd=emhd, seed=20261202)
        p = oracle.default_params(emhd)
        if emhd:
            p.closure = 1
        per_zone = []
        for z in range(8):
            sl = slice(z, z + 1)
            B = oracle.Batch(ref["guess"][:, sl], ref["geom"][:, sl], np.ones((p.nf, 1)),
                             None if not emhd else ref["prim_old"][:, sl], None,
                             None if not emhd else ref["closure"][:, sl])
            per_zone.append(tuple(oracle.count_live_ops(p, B)))
        assert len(set(per_zone)) == 1, per_zone
        counts[emhd] = np.array(per_zone[0])
    assert counts[True].sum() > counts[False].sum()
