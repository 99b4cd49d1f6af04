"""The integrands of the torus observers (src/problem/torus/ObserversTorus.hpp)
evaluated with vectorised numpy from the primitive planes and the metric
planes of a grid, independently of the oracle's fluidElement: the checker of
the oracle's grim_oracle_torus_observers and the scale of the GPU tolerance.

Metric planes (centre location): gCov 00 01 02 03 11 12 13 22 23 33, gCon
row 0, alpha, sqrt(-g).
"""
import numpy as np

COV10 = [(0, 0), (0, 1), (0, 2), (0, 3), (1, 1), (1, 2), (1, 3), (2, 2), (2, 3), (3, 3)]
NAMES = ("baryon_mass", "magnetic_energy", "thermal_energy", "mdot_in", "mdot_out", "mdot_out_unbound",
         "mdot_out_relativistic", "rho_max", "beta_min")


def observers(g, p, center, prim):
    """(values, scales): each observable, and for the sums the sum of the
    absolute integrands (the rounding scale of a sum in another order)."""
    bulk = lambda a: a.reshape(a.shape[0], *g.shape)[g.bulk()]  # noqa: E731
    P = bulk(np.asarray(prim))  # (dof, N3, N2, N1): X1 last
    G = bulk(np.asarray(center)[:16])
    cov = np.empty((4, 4) + G.shape[1:])
    for q, (a, b) in enumerate(COV10):
        cov[a, b] = cov[b, a] = G[q]
    gcon0, alpha, sqrtg = G[10:14], G[14], G[15]
    rho, u, ut, B = P[0], P[1], P[2:5], P[p.nf:p.nf + 3]
    uf = np.maximum(u, p.uFloorInFluidElement)
    # u^0 = gamma / alpha, u^i = ut^i - gamma alpha g^{0i}  (physics.cpp:74-94)
    gamma = np.sqrt(1. + np.einsum("ij...,i...,j...->...", cov[1:, 1:], ut, ut))
    ucon = np.concatenate([(gamma / alpha)[None], ut - gamma * alpha * gcon0[1:]])
    ucov = np.einsum("mn...,n...->m...", cov, ucon)
    # b^0 = B^i u_i, b^i = (B^i + b^0 u^i) / u^0  (physics.cpp:96-112)
    b0 = np.einsum("i...,i...->...", B, ucov[1:])
    bcon = np.concatenate([b0[None], (B + b0 * ucon[1:]) / ucon[0]])
    bsq = np.einsum("m...,mn...,n...->...", bcon, cov, bcon) + p.bSqrFloorInFluidElement
    pressure = (p.adiabaticIndex - 1.) * uf
    dX = [(g.X1End - g.X1Start) / g.N1, (g.X2End - g.X2Start) / g.N2, (g.X3End - g.X3Start) / g.N3]
    vol = float(np.prod(dX[:g.dim]))
    w = vol * sqrtg * gamma / alpha
    mdot = rho * vol * sqrtg * ucon[1]
    first, last = mdot[..., 0], mdot[..., -1]
    bern = (1. + u / rho * p.adiabaticIndex) * ucov[0]
    unbound = last * (bern[..., -1] < -1.)
    rel = unbound * (gamma[..., -1] > 2.)
    terms = dict(baryon_mass=rho * w, magnetic_energy=bsq * w, thermal_energy=u * w, mdot_in=first,
                 mdot_out=last, mdot_out_unbound=unbound, mdot_out_relativistic=rel)
    values = {k: float(v.sum()) for k, v in terms.items()}
    scales = {k: float(np.abs(v).sum()) for k, v in terms.items()}
    values["rho_max"] = float(np.nanmax(rho))
    values["beta_min"] = float(np.nanmin(2. * (pressure + 1e-13) / (bsq + 1e-18)))
    return values, scales
