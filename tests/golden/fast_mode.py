"""Derive the ideal SR-MHD fast magnetosonic eigenmode used by the 2D
Minkowski linear-wave workload, and write it to fast_mode.json.

Background: rho = 1, u = 2, utilde = 0, B = (0.1, 0.3, 0), Gamma = 4/3,
k = (2 pi, 2 pi).  The flux Jacobians A_mu = dF^mu/dP come from the oracle's
computeFluxes by fourth-order central differences; the fast mode is the
eigenvector of (k1 A1 + k2 A2) r = omega A0 r with the largest omega.  The
perturbation P = P0 + amp r cos(k.x - omega t) then solves the linearised
equations exactly.

Run: python tests/golden/fast_mode.py   (needs oracle/libgrim_oracle.so)
"""
import json
import os
import sys

import numpy as np
import scipy.linalg

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402

K = (2 * np.pi, 2 * np.pi)
P0 = np.array([1.0, 2.0, 0.0, 0.0, 0.0, 0.1, 0.3, 0.0])
MINKOWSKI_GEOM16 = np.array([-1, 0, 0, 0, 1, 0, 0, 1, 0, 1, -1, 0, 0, 0, 1, 1], dtype=float)


def jacobians():
    p = O.default_params(False)
    geom = MINKOWSKI_GEOM16[:, None].repeat(1, axis=1)
    A = np.zeros((3, 8, 8))
    for j in range(8):
        h = 1e-3 * max(abs(P0[j]), 1e-1)
        cols = []
        for s in (-2, -1, 1, 2):
            P = P0.copy()
            P[j] += s * h
            cols.append(np.stack([O.fluxes(p, P[:, None], geom, d)[:, 0] for d in range(3)]))
        fm2, fm1, fp1, fp2 = cols
        A[:, :, j] = (fm2 - 8 * fm1 + 8 * fp1 - fp2) / (12 * h)
    return A


def fast_mode():
    A = jacobians()
    M = K[0] * A[1] + K[1] * A[2]
    w, V = scipy.linalg.eig(M, A[0])
    w = w.real
    i = int(np.argmax(w))
    r = V[:, i].real
    r = r / np.linalg.norm(r)
    if r[0] < 0:
        r = -r
    resid = np.linalg.norm(M @ r - w[i] * (A[0] @ r))
    return float(w[i]), r, float(resid), sorted(w.tolist())


if __name__ == "__main__":
    omega, r, resid, spectrum = fast_mode()
    out = {"k": list(K), "background": P0.tolist(), "adiabaticIndex": 4.0 / 3.0, "omega": omega,
           "eigenvector": r.tolist(), "eigen_residual": resid, "spectrum": spectrum}
    with open(os.path.join(HERE, "fast_mode.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))
