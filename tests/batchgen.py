"""Seeded synthetic zone batches for the oracle tests (numpy, CPU).

Same sampling rule and counter-based RNG as the HIP generator in
``grim_amd/csrc/batchgen.hpp`` (the GPU tests check the two agree):

* RNG: uniform ``u_k(z)`` = top 53 bits of ``splitmix64(seed + G*(32*z + k + 1))``
  with ``G = 0x9E3779B97F4A7C15``.
* Kerr-Schild batch (isolated per-zone solves): ``r`` log-uniform in
  [1.5, 40] (X1 = ln r), ``theta`` uniform in (0.05 pi, 0.95 pi) with X2
  from inverting grim's ``theta(X2) = pi X2 + (1-h)/2 sin(2 pi X2)``,
  ``rho = 10^U[-5,0]``, ``u = rho 10^U[-3,1]``, Lorentz factor uniform in
  [1,10] along a direction uniform on S^2 in the orthonormal frame of
  ``g_ij``, magnetic field direction uniform on S^2 scaled so that
  ``b^2/rho = 10^U[-3, log10 50]``.  The guess is the root perturbed by
  ``1e-3 xi`` (relative, ``xi ~ U[-1,1]``) in every fluid variable.
* EMHD batch: additionally ``qTilde, dPTilde ~ 0.1 U[-1,1]``, per-zone
  conduction/viscosity alphas ``10^U[-2,1]``, an old state perturbed by
  ``1e-2 xi'`` and 10% of zones with ``rho = 1e-8 (1 + U[0,1])``.
"""
from __future__ import annotations

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
SEED = 20261202
R_MIN, R_MAX = 1.5, 40.0
SPIN, HSLOPE = 0.9375, 0.3


def splitmix64(x):
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, z, k):
    """u_k(z) in [0, 1): top 53 bits of splitmix64(seed + G*(32 z + k + 1))."""
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = np.uint64(seed) + GOLDEN * (np.uint64(32) * z + np.uint64(k + 1))
    return (splitmix64(x) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def theta_to_X2(theta, h):
    """Invert grim's theta(X2) = pi X2 + (1-h)/2 sin(2 pi X2) (12 Newton steps)."""
    X2 = theta / np.pi
    for _ in range(12):
        f = np.pi * X2 + 0.5 * (1.0 - h) * np.sin(2.0 * np.pi * X2) - theta
        df = np.pi + np.pi * (1.0 - h) * np.cos(2.0 * np.pi * X2)
        X2 = X2 - f / df
    return X2


def _spatial_frame(gCov):
    """Rows of L^{-T} for the Cholesky factor g_ij = L L^T (i, j = 1..3)."""
    n = gCov.shape[1]
    g = np.empty((n, 3, 3))
    for i in range(3):
        for j in range(3):
            g[:, i, j] = gCov[(i + 1) * 4 + (j + 1)]
    L = np.linalg.cholesky(g)
    return np.linalg.inv(np.transpose(L, (0, 2, 1)))  # L^{-T}


def _unit_vectors(u_cos, u_phi):
    ct = 2.0 * u_cos - 1.0
    st = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
    ph = 2.0 * np.pi * u_phi
    return np.stack([st * np.cos(ph), st * np.sin(ph), ct], axis=1)


def kerr_schild_batch(oracle, n, emhd=False, seed=SEED, a=SPIN, h=HSLOPE):
    """Returns dict(prim_star, guess, geom, closure, prim_old, X1, X2)."""
    z = np.arange(n, dtype=np.uint64)
    u = [uniform(seed, z, k) for k in range(32)]
    X1 = np.log(R_MIN) + u[0] * (np.log(R_MAX) - np.log(R_MIN))
    theta = np.pi * (0.05 + 0.9 * u[1])
    X2 = theta_to_X2(theta, h)
    geo = oracle.mks_geometry(X1, X2, a, h)
    geom = oracle.pack_geometry(geo["gCov"], geo["gCon"], geo["g"], geo["alpha"])

    nf = 7 if emhd else 5
    dof = nf + 3
    rho = 10.0 ** (-5.0 + 5.0 * u[2])
    if emhd:
        near_floor = u[21] < 0.1
        rho = np.where(near_floor, 1e-8 * (1.0 + u[2]), rho)
    uu = rho * 10.0 ** (-3.0 + 4.0 * u[3])
    lorentz = 1.0 + 9.0 * u[4]
    LinvT = _spatial_frame(geo["gCov"])
    nvec = _unit_vectors(u[5], u[6])
    e = np.einsum("nij,nj->ni", LinvT, nvec)
    utilde = np.sqrt(lorentz * lorentz - 1.0)[:, None] * e

    # b^2 for a unit-norm field direction: b^2 = (B.B + (B^i u_i)^2)/(u^0)^2
    alpha = geo["alpha"]
    gCov = geo["gCov"].reshape(4, 4, n)
    gCon0 = geo["gCon"][0:4]
    lor = np.sqrt(1.0 + np.einsum("ijn,ni,nj->n", gCov[1:, 1:], utilde, utilde))
    uCon = np.empty((4, n))
    uCon[0] = lor / alpha
    for i in range(3):
        uCon[i + 1] = utilde[:, i] - lor * gCon0[i + 1] * alpha
    uCov = np.einsum("mkn,kn->mn", gCov, uCon)
    mvec = _unit_vectors(u[7], u[8])
    eB = np.einsum("nij,nj->ni", LinvT, mvec)
    Bdotu = np.einsum("ni,in->n", eB, uCov[1:])
    bsq_unit = (1.0 + Bdotu * Bdotu) / (uCon[0] * uCon[0])
    sigma = 10.0 ** (-3.0 + (np.log10(50.0) + 3.0) * u[9])
    s = np.sqrt(sigma * rho / bsq_unit)
    B = s[:, None] * eB

    prim = np.zeros((dof, n))
    prim[0] = rho
    prim[1] = uu
    prim[2:5] = utilde.T
    if emhd:
        prim[5] = 0.1 * (2.0 * u[17] - 1.0)
        prim[6] = 0.1 * (2.0 * u[18] - 1.0)
    prim[nf:nf + 3] = B.T
    guess = prim.copy()
    for v in range(nf):
        guess[v] = prim[v] * (1.0 + 1e-3 * (2.0 * u[10 + v] - 1.0))
    out = {"prim_star": prim, "guess": guess, "geom": geom, "X1": X1, "X2": X2,
           "closure": None, "prim_old": None}
    if emhd:
        out["closure"] = np.stack([10.0 ** (-2.0 + 3.0 * u[19]), 10.0 ** (-2.0 + 3.0 * u[20])])
        old = prim.copy()
        for v in range(nf):
            old[v] = prim[v] * (1.0 + 1e-2 * (2.0 * u[22 + v] - 1.0))
        out["prim_old"] = old
    return out
