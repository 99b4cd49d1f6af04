"""Benchmark: batched per-zone implicit solve of grim's stage on one MI355X.

Workload (the configuration the headline metric is quoted on): 2^22
independent zones of ideal MHD (8 primitives, 5 Newton unknowns) on
modified Kerr-Schild geometry, spin 0.9375, r log-uniform in [1.5, 40],
theta uniform in (0.05 pi, 0.95 pi), Lorentz factor uniform in [1, 10],
b^2/rho log-uniform in [1e-3, 50]; guess = known root perturbed by 1e-3
relative; Newton to sum R^2 <= 1e-20 on the residual normalised per zone.

One step = restore the guess into the working primitives (grim's "Set a
guess for prim", src/timestepper/timestep.cpp:72-76) + one batched Newton
solve of all zones.  Inputs stay resident in HBM; nothing crosses PCIe in
the timed region.  Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "zone-updates/s, 1 MI355X: analytic-Jac+FP32 LU vs FD-Jac FP64 @ resid≤1e-10"
N_ZONES = 1 << 22
SEED = 20261202
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--zones", type=int, default=N_ZONES)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-reps", type=int, default=5)
    return ap.parse_args()


def flops_per_zone_solve(iters, res_evals, nf=5):
    """Algorithmic FP64 operations of the FD-Jacobian solve from per-zone counts.

    Tallied by hand from grim_amd/csrc/zone_physics.hpp and newton.hip
    (add/sub/mul/div/sqrt = 1 each, dead code excluded):
      residual evaluation (ideal MHD)   151  (kinematics 110, t-row + U 26,
                                              (U - rhs)/dt 10, mask 5)
      FD quotients per Jacobian          65  (5 columns x (3 + 2 x 5))
      LU solve 5x5, partial pivoting    120  (elimination 95, substitution 25)
      update bookkeeping                 12  (f0, f'0, P += lambda d)
      per line-search / test evaluation  24  (trial, 0.5 sum R^2, Armijo test)
    """
    return (res_evals * RESIDUAL_OPS[nf] + (res_evals - nf * iters) * 24
            + iters * (FD_QUOTIENT_OPS[nf] + LU_OPS[nf] + 12))


RESIDUAL_OPS = {5: 151}
FD_QUOTIENT_OPS = {5: 65}
LU_OPS = {5: 120}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the isolated-zone batch shards by zone range with no exchange; see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from grim_amd import capi

    assert torch.cuda.is_available(), "no GPU"
    dev = capi.device_info()
    n = args.zones
    p = capi.default_params(False)
    p.maxNonLinearIter = 30
    print(f"# device {dev}, {n} zones, generating batch", file=sys.stderr, flush=True)
    b = capi.generate_kerr_schild_batch(p, n, SEED)
    rhs = torch.empty((p.nf, n), dtype=torch.float64, device="cuda")
    mask = torch.empty(n, dtype=torch.float64, device="cuda")
    capi.known_root_rhs(p, capi.make_batch(b["prim_star"], b["geom"], rhs), rhs, mask)
    prim = b["guess"].clone()
    it = torch.empty(n, dtype=torch.int32, device="cuda")
    ev = torch.empty(n, dtype=torch.int32, device="cuda")
    l2 = torch.empty(n, dtype=torch.float64, device="cuda")
    batch = capi.make_batch(prim, b["geom"], rhs, residual_mask=mask)
    stream = torch.cuda.current_stream()
    torch.cuda.synchronize()

    def step(ev_mid=None):
        prim[: p.nf].copy_(b["guess"][: p.nf])
        if ev_mid is not None:
            ev_mid.record(stream)
        capi.batched_newton(p, batch, capi.JAC_FD, capi.LU_FP64, capi.NEWTON_PER_ZONE,
                            iterations=it, residual_evals=ev, final_l2=l2, stream=stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    print("# timing", file=sys.stderr, flush=True)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True),
            torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e0, e1, e2 in evs:
        e0.record(stream)
        step(e1)
        e2.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_ms = [e0.elapsed_time(e2) for e0, _, e2 in evs]
    kern_ms = [e1.elapsed_time(e2) for _, e1, e2 in evs]
    ms_per_step = wall * 1e3 / args.steps
    value = n / (ms_per_step * 1e-3)

    unconverged = capi.count_above(l2, p.nonlinearsolve_atol)
    iters = it.to(torch.int64)
    evals = ev.to(torch.int64)
    tot_flops = flops_per_zone_solve(float(iters.sum()), float(evals.sum()))
    kern_s = float(np.mean(kern_ms)) * 1e-3
    achieved_tflops = tot_flops / kern_s / 1e12
    hist = torch.bincount(it.to(torch.int64)).tolist()
    rel = ((prim[: p.nf] - b["prim_star"][: p.nf]).abs() / b["prim_star"][: p.nf].abs()).amax(dim=0)
    # algorithmic HBM bytes of one solve: read 8 prim + 16 geom + 5 rhs + mask,
    # write 5 prim + iterations/evals (int32) + final l2
    bytes_per_zone = (8 + 16 + 5 + 1) * 8 + 5 * 8 + 4 + 4 + 8

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Kerr-Schild zone batch, known roots)",
        "config": {
            "workload": "isolated batched per-zone Newton, ideal MHD, 2^22 zones, MKS a=0.9375, "
                        "FD Jacobian + FP64 LU, resid<=1e-10 (per-zone normalised)",
            "zones": n, "newton_unknowns": p.nf, "primitives": p.dof, "path": "fd_fp64",
            "semantics": "per_zone",
        },
        "roofline": {
            "bound": "valu",
            "pipe": "VALU FP64 (no MFMA: 5x5 per-zone systems)",
            "achieved": achieved_tflops,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / FP64_PEAK_TFLOPS,
            "traffic": None,
            "kernel": "newtonPerZoneFD<Layout<false,false>>",
            "kernel_ms": float(np.mean(kern_ms)),
            "flops_per_launch": tot_flops,
            "hbm_gbs_algorithmic": bytes_per_zone * n / kern_s / 1e9,
            "hbm_frac": bytes_per_zone * n / kern_s / 1e9 / HBM_PEAK_GBS,
        },
        "newton": {
            "unconverged": unconverged,
            "iterations_hist": hist,
            "mean_residual_evals": float(evals.float().mean()),
            "max_rel_err_vs_root_median": float(rel.median()),
            "step_ms_min": float(min(step_ms)),
            "kernel_ms_min": float(min(kern_ms)),
        },
    }

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, b, rhs, mask, n, args.cpu_reps)
    print(json.dumps(out), flush=True)


def cpu_baseline(p, b, rhs, mask, n, reps):
    """The oracle (FP64 FD Jacobian, grim's algorithm) timed on the host
    cores over the same batch: median of `reps` full passes after one
    warm-up pass, OpenMP over zones on every available core."""
    import numpy as np

    from oracle import oracle as O

    O.lib()
    cores = len(os.sched_getaffinity(0))
    q = O.Params()
    for name, _ in O.Params._fields_:
        setattr(q, name, getattr(p, name))
    g = b["geom"].cpu().numpy()
    guess = b["guess"].cpu().numpy()
    rh = rhs.cpu().numpy()
    mk = mask.cpu().numpy()
    times = []
    for r in range(reps + 1):
        B = O.Batch(guess, g, rh, residualMask=mk)
        t = time.perf_counter()
        O.newton(q, B, O.NEWTON_PER_ZONE, nthreads=cores)
        dt = time.perf_counter() - t
        print(f"# cpu pass {r}: {dt:.2f} s", file=sys.stderr, flush=True)
        if r > 0:
            times.append(dt)
    med = float(np.median(times))
    return {"value": n / med, "unit": "zone-updates/s", "cores": cores, "kind": "port",
            "sample": f"full {n}-zone batch, median of {reps} passes after 1 warm-up, "
                      f"OpenMP over zones, {cores} threads"}


if __name__ == "__main__":
    main()
