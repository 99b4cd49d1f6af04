"""Benchmark: batched per-zone implicit solve of grim's stage on one MI355X.

Workload (the configuration the headline metric is quoted on): 2^22
independent zones of ideal MHD (8 primitives, 5 Newton unknowns) on
modified Kerr-Schild geometry, spin 0.9375, r log-uniform in [1.5, 40],
theta uniform in (0.05 pi, 0.95 pi), Lorentz factor uniform in [1, 10],
b^2/rho log-uniform in [1e-3, 50]; guess = known root perturbed by 1e-3
relative; Newton to sum R^2 <= 1e-20 on the residual normalised per zone.

One step = one batched Newton solve of all zones from the guess planes
into the working primitives: grim's "Set a guess for prim" copy
(src/timestepper/timestep.cpp:72-76) is the solve's first read of the guess
(grim_zone_batch.guess), so each step starts from the same guess without a
separate copy.  Inputs stay resident in HBM; nothing crosses PCIe in the
timed region.  Prints ONE JSON line.
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
    ap.add_argument("--phase1", type=int, default=2)
    ap.add_argument("--phase2-group", type=int, default=4)
    ap.add_argument("--no-full-step", action="store_true", help="skip the full time-step workloads")
    ap.add_argument("--no-emhd-batch", action="store_true", help="skip the EMHD uneven-convergence batch")
    ap.add_argument("--torus", type=int, nargs=3, default=[192, 192, 128], metavar=("N1", "N2", "N3"))
    ap.add_argument("--torus-steps", type=int, default=20)
    ap.add_argument("--torus-paths", default="fd_fp64_lockstep,fd_fp64,mixed,fd_fp64_fast,mixed_fast",
                    help="comma-separated subset of the torus Newton paths (profiling)")
    ap.add_argument("--no-wave", action="store_true", help="skip the 64x64 linear-wave workload")
    ap.add_argument("--only-torus", action="store_true",
                    help="run only the torus paths of --torus-paths and print their JSON (profiling)")
    ap.add_argument("--no-torus-error", action="store_true",
                    help="skip the one-step mixed vs FD-FP64 comparison of the torus")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="run one elementwise negation of this many MiB first: reference bytes that calibrate "
                         "the FETCH_SIZE / WRITE_SIZE counters of a profiled run")
    ap.add_argument("--paths", default=",".join(["fd_fp64", "fd_fp64_split", "fd_fp64_fused", "analytic_fp64", "mixed",
                                                 "mixed_compacted"]),
                    help="comma-separated subset of the Newton paths to time (profiling)")
    return ap.parse_args()


# Operation tallies per zone (add/sub/mul/div/sqrt/max = 1 each; a fused
# multiply-add counts as its two operations).  "Algorithmic" operations are
# those a compiled evaluation must repeat: the exact residual's 177 are the
# operations of grim's residual (src/timestepper/residual.cpp:3-180 over
# src/physics/physics.cpp:53-217) that the residual rows depend on and that
# depend on the guess, counted by the oracle's tracing scalar (72 add/sub,
# 93 mul, 9 div, 1 sqrt, 2 max for ideal MHD; a CPU test keeps this table
# equal to that count, and the LU's 120 equal to the counting scalar's).
# grim's set() also forms quantities the residual never reads (333 operations
# in all), which are not counted.  The other rows are tallied by hand from
# grim_amd/csrc/zone_physics.hpp, analytic.hpp and newton_kernels.hpp:
#   exact residual (grim order, ideal MHD)          177 FP64
#   fast residual (identity form, fused paths)      115 FP64
#   per test / line-search evaluation bookkeeping    24 FP64 (trial, 0.5 sum R^2, Armijo)
#   FD quotients per Jacobian, grim order            65 FP64 (5 x (3 + 2 x 5))
#   FD quotients per Jacobian, reciprocal form       75 FP64 (5 x (5 + 2 x 5))
#   LU 5x5 with partial pivoting                    120 (FP64 or FP32)
#   analytic Jacobian (primal 85, rho/u columns 21, velocity columns 5 + 3 x 57)
#                                                   282 (FP64 or FP32)
#   row/column equilibration around the FP32 LU      90 FP32
#   update bookkeeping (f0, f'0, P += lambda d)      12 FP64
OPS = {
    "res_exact": 177, "res_fast": 115, "eval_bookkeeping": 24, "fd_quotients": 65,
    "fd_quotients_fused": 75, "lu": 120, "analytic_jac": 282, "equilibration": 90, "update": 12,
}


def path_flops(path, iters, res_evals, nf=5):
    """(FP64 ops, FP32 ops) of one solve of the batch from per-zone counts."""
    ls = res_evals - nf * iters  # tests + line-search evaluations (FD paths)
    if path in ("fd_fp64", "fd_fp64_split"):
        f64 = (res_evals * OPS["res_exact"] + ls * OPS["eval_bookkeeping"]
               + iters * (OPS["fd_quotients"] + OPS["lu"] + OPS["update"]))
        return f64, 0.0
    if path == "fd_fp64_fused":
        f64 = (res_evals * OPS["res_fast"] + ls * OPS["eval_bookkeeping"]
               + iters * (OPS["fd_quotients_fused"] + OPS["lu"] + OPS["update"]))
        return f64, 0.0
    if path.startswith("mixed"):
        f64 = res_evals * (OPS["res_fast"] + OPS["eval_bookkeeping"]) + iters * OPS["update"]
        f32 = iters * (OPS["analytic_jac"] + OPS["equilibration"] + OPS["lu"])
        return f64, f32
    if path == "analytic_fp64":
        f64 = (res_evals * (OPS["res_exact"] + OPS["eval_bookkeeping"])
               + iters * (OPS["analytic_jac"] + OPS["lu"] + OPS["update"]))
        return f64, 0.0
    f64 = res_evals * (OPS["res_fast"] + OPS["eval_bookkeeping"]) + iters * OPS["update"]
    f32 = iters * (OPS["analytic_jac"] + OPS["equilibration"] + OPS["lu"])
    return f64, f32


PATHS = {
    "fd_fp64": ("JAC_FD", "LU_FP64", 0, "newtonPerZone<Layout<false,false>, MODE_FD>"),
    "fd_fp64_split": ("JAC_FD", "LU_FP64", 0, "newtonPerZoneSplit<Layout<false,false>, MODE_FD>"),
    "fd_fp64_fused": ("JAC_FD", "LU_FP64", 1, "newtonPerZone<Layout<false,false>, MODE_FD_FAST>"),
    "analytic_fp64": ("JAC_ANALYTIC", "LU_FP64", 0, "newtonPerZone<Layout<false,false>, MODE_ANALYTIC64>"),
    "mixed": ("JAC_ANALYTIC", "LU_FP32", 1, "newtonPerZone<Layout<false,false>, MODE_MIXED>"),
    "mixed_compacted": ("JAC_ANALYTIC", "LU_FP32", 1, "newtonPhase1+newtonPhase2<Layout<false,false>, MODE_MIXED>"),
}
HEADLINE = "mixed"
# MI355X's published FP32 vector peak.  It counts packed FP32 issue, which
# this build leaves off, so the FP32 share is priced at twice the rate the
# kernels can reach (a conservative fraction).
FP32_PEAK_TFLOPS = 157.3


TRAFFIC_FILE = "profiles/newton_mixed_ks2e22.md"


def measured_traffic():
    """HBM bytes per launch of the headline kernel from the calibrated
    FETCH_SIZE + WRITE_SIZE counter passes recorded in TRAFFIC_FILE (counters
    cannot be read from inside a timed run); None when the summary is absent."""
    import re

    try:
        text = open(os.path.join(ROOT, TRAFFIC_FILE)).read()
    except OSError:
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*(\d+)\s*$", text, re.M)
    return int(m.group(1)) if m else None


def measured_issue():
    """The headline kernel's VALU issue counters recorded in TRAFFIC_FILE
    (SQ_INSTS_VALU per wavefront and the fraction of the VALU issue slots of
    the kernel's busy cycles it used); None when absent."""
    import re

    try:
        text = open(os.path.join(ROOT, TRAFFIC_FILE)).read()
    except OSError:
        return None
    f = re.search(r"^valu_issue_frac:\s*([\d.]+)\s*$", text, re.M)
    w = re.search(r"^valu_instr_per_wave:\s*([\d.]+)\s*$", text, re.M)
    if not f:
        return None
    return {"frac": float(f.group(1)), "valu_instr_per_wave": float(w.group(1)) if w else None,
            "source": TRAFFIC_FILE,
            "definition": "SQ_INSTS_VALU x 4 cycles / (1024 SIMDs x GRBM_GUI_ACTIVE / 8 XCDs): the share of "
                          "the wave64 VALU issue slots the kernel used"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the isolated-zone batch shards by zone range with no exchange; the grid step by slabs with "
              "one halo exchange per stage)",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from grim_amd import capi

    assert torch.cuda.is_available(), "no GPU"
    dev = capi.device_info()
    if args.copy_probe_mib:
        src = torch.ones(args.copy_probe_mib << 17, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.neg(src, out=dst)  # reads and writes copy_probe_mib MiB once each
        torch.cuda.synchronize()
        del src, dst
    if args.only_torus:
        args.no_wave = True
        print(json.dumps({"full_step": full_step_bench(args)}), flush=True)
        return
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
    # every step solves from the same guess planes into prim (grim's copy
    # prim = primOld before the solve, folded into the solve's first read)
    batch = capi.make_batch(prim, b["geom"], rhs, residual_mask=mask, guess=b["guess"][: p.nf])
    ws = torch.empty(capi.workspace_bytes(n), dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()
    torch.cuda.synchronize()
    results = {}
    finals = {}
    converged = {}
    selected = [q for q in PATHS if q in args.paths.split(",")]
    for path in selected:
        jac, lu, fused, kname = PATHS[path]
        jac, lu = getattr(capi, jac), getattr(capi, lu)

        opts = capi.newton_opts(jac, lu, fused=fused)
        if path == "fd_fp64_split":
            # grim's FD solve, stragglers gathered inside each workgroup
            opts = capi.newton_opts(jac, lu, fused=fused, compaction=2, phase1_iters=args.phase1)
        if path == "mixed_compacted":
            opts = capi.newton_opts(jac, lu, fused=fused, compaction=1, phase1_iters=args.phase1,
                                    phase2_group=args.phase2_group, workspace=ws)

        def step(ev_mid=None):
            if ev_mid is not None:
                ev_mid.record(stream)
            capi.batched_newton(p, batch, iterations=it, residual_evals=ev, final_l2=l2, stream=stream,
                                opts=opts)

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        print(f"# timing {path}", file=sys.stderr, flush=True)
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
        iters = float(it.to(torch.int64).sum())
        evals = float(ev.to(torch.int64).sum())
        f64, f32 = path_flops(path, iters, evals)
        kern_s = float(np.mean(kern_ms)) * 1e-3
        t_peak = f64 / (FP64_PEAK_TFLOPS * 1e12) + f32 / (FP32_PEAK_TFLOPS * 1e12)
        rel = ((prim[: p.nf] - b["prim_star"][: p.nf]).abs() / b["prim_star"][: p.nf].abs()).amax(dim=0)
        finals[path] = prim[: p.nf].clone()
        converged[path] = l2 <= p.nonlinearsolve_atol
        results[path] = {
            "zone_updates_per_s": n / (ms_per_step * 1e-3),
            "ms_per_step": ms_per_step,
            "kernel": kname,
            "kernel_ms": float(np.mean(kern_ms)),
            "kernel_ms_min": float(min(kern_ms)),
            "step_ms_min": float(min(step_ms)),
            "unconverged": capi.count_above(l2, p.nonlinearsolve_atol),
            "iterations_hist": torch.bincount(it.to(torch.int64)).tolist(),
            "mean_residual_evals": evals / n,
            "fp64_ops": f64,
            "fp32_ops": f32,
            "achieved_tflops": (f64 + f32) / kern_s / 1e12,
            "valu_frac": t_peak / kern_s,
            "err_vs_root_p99": float(torch.quantile(rel.float(), 0.99)),
            "err_vs_root_max": float(rel.max()),
        }
    for ref in ("fd_fp64", "analytic_fp64"):
        for path in selected:
            if path == ref or ref not in finals:
                continue
            d = ((finals[path] - finals[ref]).abs() / finals[ref].abs()).amax(dim=0)
            results[path][f"err_vs_{ref}_p99"] = float(torch.quantile(d.float(), 0.99))
            results[path][f"err_vs_{ref}_max"] = float(d.max())
            # split by convergence: zones that met the tolerance in both paths
            both = converged[path] & converged[ref]
            if bool(both.any()):
                results[path][f"err_vs_{ref}_converged_max"] = float(d[both].max())
                results[path][f"err_vs_{ref}_converged_p99"] = float(torch.quantile(d[both].float(), 0.99))
            results[path][f"zones_unconverged_in_either_vs_{ref}"] = int((~both).sum())
            if bool((~both).any()):
                results[path][f"err_vs_{ref}_unconverged_max"] = float(d[~both].max())

    head = results[HEADLINE] if HEADLINE in results else results[selected[-1]]
    fd = results.get("fd_fp64", head)
    kern_s = head["kernel_ms"] * 1e-3
    bytes_per_zone = (8 + 16 + 5 + 1) * 8 + 5 * 8 + 4 + 4 + 8
    mixed_peak = (head["fp64_ops"] + head["fp32_ops"]) / (
        head["fp64_ops"] / FP64_PEAK_TFLOPS + head["fp32_ops"] / FP32_PEAK_TFLOPS)
    out = {
        "metric": METRIC,
        "value": head["zone_updates_per_s"],
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 residual/update + f32 Jacobian/LU",
        "data": "synthetic (seeded Kerr-Schild zone batch, known roots)",
        "config": {
            "workload": "isolated batched per-zone Newton, ideal MHD, 2^22 zones, MKS a=0.9375, "
                        "resid<=1e-10 (per-zone normalised); headline = analytic Jacobian + FP32 LU",
            "zones": n, "newton_unknowns": p.nf, "primitives": p.dof, "semantics": "per_zone",
        },
        "speedup_vs_fd_fp64": head["zone_updates_per_s"] / fd["zone_updates_per_s"],
        # against the faster of grim's FD solve's two schedules (one pass, or
        # stragglers gathered inside each workgroup)
        "speedup_vs_fd_fp64_best": head["zone_updates_per_s"] / max(
            fd["zone_updates_per_s"], results.get("fd_fp64_split", fd)["zone_updates_per_s"]),
        "speedup_vs_fd_fp64_fused": head["zone_updates_per_s"] / results.get("fd_fp64_fused", head)["zone_updates_per_s"],
        # Arithmetic intensity (counted ops / algorithmic bytes) is below
        # the FP64 ridge point (78.6 TFLOP/s / 8 TB/s = 9.8 op/B), so the
        # roofline bound is HBM; the VALU fraction is reported beside it.
        "roofline": {
            "bound": "hbm",
            "achieved": bytes_per_zone * n / kern_s / 1e9,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": bytes_per_zone * n / kern_s / 1e9 / HBM_PEAK_GBS,
            "traffic": measured_traffic(),
            "traffic_source": TRAFFIC_FILE,
            "kernel": head["kernel"],
            "kernel_ms": head["kernel_ms"],
            "bytes_per_zone": bytes_per_zone,
            "arithmetic_intensity": (head["fp64_ops"] + head["fp32_ops"]) / (bytes_per_zone * n),
            "valu": {
                "pipe": "VALU FP64 + FP32 (no MFMA: 5x5 per-zone systems)",
                "achieved": head["achieved_tflops"],
                "peak": mixed_peak,
                "unit": "TFLOP/s",
                "frac": head["valu_frac"],
            },
            # what bounds the kernel: counted VALU issue (profile counters)
            "valu_issue": measured_issue(),
        },
        "paths": results,
    }

    del b, rhs, mask, prim, batch, ws, finals, converged
    torch.cuda.empty_cache()
    if not args.no_emhd_batch:
        out["emhd_batch"] = emhd_batch_bench(args)
        torch.cuda.empty_cache()
    if not args.no_full_step:
        out["full_step"] = full_step_bench(args)
    if not args.no_cpu:
        p = capi.default_params(False)
        p.maxNonLinearIter = 30
        b = capi.generate_kerr_schild_batch(p, n, SEED)
        rhs = torch.empty((p.nf, n), dtype=torch.float64, device="cuda")
        mask = torch.empty(n, dtype=torch.float64, device="cuda")
        capi.known_root_rhs(p, capi.make_batch(b["prim_star"], b["geom"], rhs), rhs, mask)
        out["cpu_baseline"] = cpu_baseline(p, b, rhs, mask, n, args.cpu_reps)
    print(json.dumps(out), flush=True)


def torus_config(N1, N2, N3, explicit_fast=0):
    """Grid of the 3D Kerr torus workload (src/problem/torus/params.cpp:51-105 with
    a = 0.9375, WENO5 + LLF)."""
    import numpy as np

    from grim_amd import capi
    from grim_amd.timestepper import grid_config

    a = 0.9375
    Rin = 0.85 * (1 + np.sqrt(1 - a * a))
    return grid_config(N1=N1, N2=N2, N3=N3, dim=3, metric=capi.METRIC_MKS, blackHoleSpin=a, X1Start=np.log(Rin),
                       X1End=np.log(3000.), X2Start=1e-8, X2End=1 - 1e-8, X3Start=0., X3End=2 * np.pi,
                       DerefineThetaHorizon=1, DoCylindrify=1, X1cyl=np.log(4 * Rin), X2cyl=1. / N2,
                       boundaryLeft=capi.BC_OUTFLOW, boundaryRight=capi.BC_OUTFLOW, boundaryBottom=capi.BC_MIRROR,
                       boundaryTop=capi.BC_MIRROR, boundaryBack=capi.BC_PERIODIC, boundaryFront=capi.BC_PERIODIC,
                       reconstruction=capi.RECON_WENO5, riemannSolver=capi.RIEMANN_LLF, residualScaling=1,
                       explicitFast=explicit_fast)


def time_steps(ts, steps, warmup, graph=False):
    """Wall time per step of `steps` grim_time_step calls after `warmup`.  The
    timed steps run without statistics (no host synchronisation: dt and time
    advance on the device); with graph=True one step is captured into a HIP
    graph and replayed.  One more step with statistics follows the timed
    region and is returned."""
    import torch

    for _ in range(warmup):
        ts.timeStep()
    torch.cuda.synchronize()
    if graph:
        s = ts._stream
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            ts.timeStep(stats=False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            g.replay()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        del g
    else:
        t0 = time.perf_counter()
        for _ in range(steps):
            ts.timeStep(stats=False)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
    return ms, ts.timeStep()


def torus_params_emhd():
    from grim_amd import capi

    p = capi.default_params(True)
    p.adiabaticIndex = 5. / 3
    p.closure = capi.CLOSURE_TORUS
    p.maxNonLinearIter = 10
    return p


def torus_one_step_error(cfg, N1, N2, N3):
    """Primitive error of one mixed-precision step against one FD-FP64 step
    taken from the same initial state (per-zone semantics for both)."""
    import numpy as np
    import torch

    from grim_amd import capi
    from grim_amd.timestepper import TimeStepper

    out = {}
    g = N1 + 6, N2 + 6, N3 + 6
    bulk = lambda a, k: a.reshape(k, g[2], g[1], g[0])[:, 3:-3, 3:-3, 3:-3].reshape(k, -1)  # noqa: E731
    for name, opts in (("fd_fp64", capi.newton_opts()),
                       ("mixed", capi.newton_opts(jacobian=capi.JAC_ANALYTIC, lu_precision=capi.LU_FP32_FP64))):
        p = torus_params_emhd()
        ts = TimeStepper(p, cfg, opts=opts)
        ts.initTorus(capi.torus_params())
        ts.set_time(0., 0.005)
        st = ts.timeStep()
        _, l2 = ts.lastSolveStats()  # the full stage's final sum R^2 per zone
        out[name] = (bulk(ts.primOld, p.dof), [int(st.unconverged[0]), int(st.unconverged[1])],
                     bulk(l2[None, :], 1)[0] <= p.nonlinearsolve_atol)
        ts.close()
        del ts
        torch.cuda.empty_cache()
    ref, mine = out["fd_fp64"][0], out["mixed"][0]
    both = out["fd_fp64"][2] & out["mixed"][2]
    diff = np.abs(mine - ref)
    rel = diff[:2] / np.abs(ref[:2])
    scaled = diff / np.abs(ref).max(axis=1, keepdims=True)
    relz = rel.max(axis=0)
    scaledz = scaled.max(axis=0)
    res = {"rho_u_rel_p99": float(np.percentile(rel, 99)), "rho_u_rel_median": float(np.median(rel)),
           "rho_u_rel_max": float(rel.max()), "scaled_p99": float(np.percentile(scaled, 99)),
           "scaled_max": float(scaled.max()), "unconverged_fd_fp64": out["fd_fp64"][1],
           "unconverged_mixed": out["mixed"][1],
           "full_stage_converged_in_both": int(both.sum()), "full_stage_unconverged_in_either": int((~both).sum()),
           "converged": {"rho_u_rel_max": float(relz[both].max()), "rho_u_rel_p99": float(np.percentile(relz[both], 99)),
                         "scaled_max": float(scaledz[both].max()),
                         "scaled_p99": float(np.percentile(scaledz[both], 99))},
           "note": "rho, u relative to themselves; every variable relative to its largest value; 'converged' = "
                   "bulk zones whose full-stage residual met the tolerance in both paths"}
    if (~both).any():
        res["unconverged"] = {"rho_u_rel_max": float(relz[~both].max()), "scaled_max": float(scaledz[~both].max())}
        res["max_in"] = "unconverged zones" if relz[~both].max() >= relz[both].max() else "converged zones"
    return res


OBS_BYTES_PER_ZONE = (8 + 16) * 8  # rho, u, u~^i, B^i + gCov (10), gCon row 0, alpha, g
# Algorithmic bytes of one zone-update of the 3D EMHD torus step (stored
# geometry, interior zones; nP = 10 primitives, nf = 7 unknowns, G = 22
# metric doubles, 40 connection coefficients): per stage 3 (2 nP + G) for the
# face sweeps + (3 nP + nP + G + 40 + nf + 3) for divergence, CT, sources and
# right-hand side + (2 nP + nf + G) for the Newton solve + 2 nP for the
# floors = 307 doubles; two stages plus (nP + G) doubles of computeDt.
# The count keeps the geometry per zone, as grim stores it; this build keeps
# one geometry value per (X1, X2) column (the metric does not depend on X3),
# so the 2 (3 G + G + 40 + G) + G = 322 geometry doubles of it are read
# from the caches, not HBM: the measured traffic (profiles) is well below it.
TORUS_BYTES_PER_ZONE_STEP = (2 * 307 + 10 + 22) * 8
TORUS_PROFILE_FILE = "profiles/torus_step_192x192x128.md"


def torus_profile():
    """Per-kernel durations and calibrated HBM traffic of the torus step's
    kernels from the committed rocprofv3 summary (lines 'kernel: name us GB'),
    or None when it is absent."""
    import re

    try:
        text = open(os.path.join(ROOT, TORUS_PROFILE_FILE)).read()
    except OSError:
        return None
    out = {}
    for m in re.finditer(r"^kernel_summary:\s*(\S+)\s+([\d.]+)\s+([\d.]+)\s+([\d.]+)\s*$", text, re.M):
        name, us, gb, valu = m.group(1), float(m.group(2)), float(m.group(3)), float(m.group(4))
        out[name] = {"us": us, "hbm_GB": gb, "GBps": gb / (us * 1e-6), "frac_of_hbm_peak": gb / (us * 1e-6) /
                     HBM_PEAK_GBS, "valu_issue_frac": valu}
    return {"source": TORUS_PROFILE_FILE, "per_launch": out} if out else None


TORUS_PROFILE = torus_profile()


def observers_bench(ts, zones, calls=20):
    """ComputeEnergyIntegrals + ComputeMinMaxVariables + ComputeBoundaryFluxes
    (grim_torus_observers) on primOld: HIP events on the context's stream
    around each call (both reduction kernels and the 72-byte readback)."""
    import torch

    s = ts._stream
    for _ in range(2):
        ts.observers()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(calls)]
    t0 = time.perf_counter()
    for e0, e1 in ev:
        e0.record(s)
        obs = ts.observers()
        e1.record(s)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / calls
    ms = sum(e0.elapsed_time(e1) for e0, e1 in ev) / calls
    return {"ms_per_call": ms, "wall_ms_per_call": wall, "bytes_per_zone": OBS_BYTES_PER_ZONE,
            "achieved_GBps": OBS_BYTES_PER_ZONE * zones / (ms * 1e-3) / 1e9,
            "frac_of_hbm_peak": OBS_BYTES_PER_ZONE * zones / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "values_after_steps": obs}


def full_step_bench(args):
    """The whole time step (timeStepper::timeStep) on two workloads:
    the 3D EMHD Kerr torus at 192x192x128 with both Newton paths, and the 2D
    64x64 Minkowski linear fast wave (latency-bound)."""
    import torch

    from grim_amd import capi
    from grim_amd.timestepper import TimeStepper, grid_config

    res = {}
    N1, N2, N3 = args.torus
    cfg = torus_config(N1, N2, N3)
    zones = N1 * N2 * N3
    torus = {}
    if "mixed" in args.torus_paths.split(",") and not args.no_torus_error:
        torus["one_step_error_mixed_vs_fd_fp64"] = torus_one_step_error(cfg, N1, N2, N3)
    states = {}
    torus_state = {}
    mixed_opts = lambda: capi.newton_opts(jacobian=capi.JAC_ANALYTIC, lu_precision=capi.LU_FP32_FP64)  # noqa: E731
    for name, opts in (("fd_fp64_lockstep", capi.newton_opts(semantics=capi.NEWTON_GLOBAL)),
                       ("fd_fp64", capi.newton_opts()),
                       ("mixed", mixed_opts()),
                       ("fd_fp64_fast", capi.newton_opts()),
                       ("mixed_fast", mixed_opts())):
        if name not in args.torus_paths.split(","):
            continue
        cfg = torus_config(N1, N2, N3, explicit_fast=1 if name.endswith("_fast") else 0)
        p = capi.default_params(True)
        p.adiabaticIndex = 5. / 3
        p.closure = capi.CLOSURE_TORUS
        p.maxNonLinearIter = 10
        print(f"# full step: torus {N1}x{N2}x{N3} EMHD, {name}", file=sys.stderr, flush=True)
        ts = TimeStepper(p, cfg, opts=opts)
        ts.initTorus(capi.torus_params())
        ts.set_time(0., 0.005)
        if not args.no_cpu and "cpu_state" not in torus_state:
            # the initial state and geometry the CPU baseline starts from
            torus_state["cpu_state"] = (p, ts.primOld, [ts.geometry(w) for w in range(4)], ts.geometry(4))
        ms, last = time_steps(ts, args.torus_steps, 2)
        torus[name] = {
            "ms_per_step": ms,
            "zone_updates_per_s": zones / (ms * 1e-3),
            "unconverged_next_step": [int(last.unconverged[0]), int(last.unconverged[1])],
            "max_iterations_next_step": [int(last.max_iterations[0]), int(last.max_iterations[1])],
            "dt_next_step": last.dt,
        }
        states[name] = ts.primOld
        if name == "mixed":
            torus["observers"] = observers_bench(ts, zones)
        ts.close()
        del ts
        torch.cuda.empty_cache()
    import numpy as np

    g = N1 + 6, N2 + 6, N3 + 6
    bulk = lambda a: a.reshape(p.dof, g[2], g[1], g[0])[:, 3:-3, 3:-3, 3:-3].reshape(p.dof, -1)  # noqa: E731
    ref = bulk(states["fd_fp64"]) if "fd_fp64" in states else None
    for name in states:
        mine = bulk(states[name])
        torus[name]["nonfinite_zones"] = int((~np.isfinite(mine).all(axis=0)).sum())
        if ref is None or name == "fd_fp64":
            continue
        # rho and u (positive) relative to themselves; every variable also
        # relative to its largest value (velocities, field components, q and
        # dP pass through zero, where a pointwise relative error means nothing)
        ok = np.isfinite(mine).all(axis=0) & np.isfinite(ref).all(axis=0)
        scale = np.abs(ref[:, ok]).max(axis=1, keepdims=True)
        diff = np.abs(mine[:, ok] - ref[:, ok])
        rel = diff[:2] / np.abs(ref[:2][:, ok])
        scaled = diff / scale
        torus[name]["vs_fd_fp64_rho_u_rel_p99"] = float(np.percentile(rel, 99))
        torus[name]["vs_fd_fp64_rho_u_rel_median"] = float(np.median(rel))
        torus[name]["vs_fd_fp64_scaled_p99"] = float(np.percentile(scaled, 99))
        torus[name]["vs_fd_fp64_scaled_max"] = float(scaled.max())
    for head, base, key in (("mixed", "fd_fp64_lockstep", "speedup_mixed_vs_fd_fp64_lockstep"),
                            ("mixed", "fd_fp64", "speedup_mixed_vs_fd_fp64_per_zone"),
                            ("mixed_fast", "fd_fp64_lockstep", "speedup_mixed_fast_vs_fd_fp64_lockstep"),
                            ("mixed_fast", "fd_fp64", "speedup_mixed_fast_vs_fd_fp64_per_zone"),
                            ("mixed_fast", "fd_fp64_fast", "speedup_mixed_fast_vs_fd_fp64_fast")):
        if head in torus and base in torus:
            torus[key] = torus[head]["zone_updates_per_s"] / torus[base]["zone_updates_per_s"]
    # HBM roofline of the whole step: algorithmic bytes per zone-step (stored
    # geometry, interior zones) over the measured time per step
    for name in ("mixed", "fd_fp64", "fd_fp64_lockstep", "fd_fp64_fast", "mixed_fast"):
        if name in torus:
            gbps = TORUS_BYTES_PER_ZONE_STEP * zones / (torus[name]["ms_per_step"] * 1e-3) / 1e9
            torus[name]["roofline"] = {"bound": "hbm", "achieved": gbps, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                       "frac": gbps / HBM_PEAK_GBS,
                                       "bytes_per_zone_step": TORUS_BYTES_PER_ZONE_STEP,
                                       "kernels": TORUS_PROFILE}
    if "cpu_state" in torus_state:
        del states
        torch.cuda.empty_cache()
        pc, prim0, geoms, floors = torus_state.pop("cpu_state")
        print("# cpu baseline: torus step", file=sys.stderr, flush=True)
        cb = cpu_baseline_step(cfg, pc, prim0, geoms, floors, steps=1, dt0=0.005)
        cb["sample"] = (f"1 step (of the GPU run's {args.torus_steps}) of the oracle's full time step, FD-FP64 per-zone "
                        f"Newton, from the same initial state and geometry, {cb['cores']} threads")
        torus["cpu_baseline"] = cb
        del prim0, geoms, floors
        states = {}
    torus["config"] = (f"3D MKS torus {N1}x{N2}x{N3} (a=0.9375, r_in=6, r_max=12, SANE beta_min=100), EMHD "
                       f"(torus closure), WENO5+LLF, {args.torus_steps} steps after 2 warm-up, Newton to "
                       "relative residual 1e-10 (maxNonLinearIter 10); fd_fp64_lockstep = grim's solve loop "
                       "(every zone updated while any zone is unconverged, solve.cpp:75-78), fd_fp64 = the same "
                       "FD-FP64 algorithm stopping each zone at its own convergence, mixed = analytic Jacobian "
                       "with FP32 assembly and LU for the first two updates of each zone and FP64 for the "
                       "compacted stragglers (FP64 residual and update throughout); the *_fast paths run the "
                       "same Newton with the explicit stage in its fast form (shared reciprocals, fused "
                       "multiply-adds, 40-plane connection; equal to grim's order to rounding), the others with "
                       "the explicit stage in grim's expression order")
    res["torus_emhd"] = torus
    if args.no_wave:
        return res

    p = capi.default_params(False)
    p.maxNonLinearIter = 10
    wave = {"config": "2D Minkowski fast magnetosonic wave 64x64, WENO5+LLF, FD-FP64 Newton, 10 steps after 2 "
                      "warm-up; eager = steps enqueued one by one, graph = one step captured in a HIP graph "
                      "and replayed"}
    for mode in ("eager", "graph"):
        ts = TimeStepper(p, grid_config(N1=64, N2=64, reconstruction=capi.RECON_WENO5,
                                        riemannSolver=capi.RIEMANN_LLF), stream=torch.cuda.Stream())
        ts.initLinearMode(0, 1e-4)
        ts.set_time(0., 0.5 / 64)
        if mode == "eager" and not args.no_cpu:
            wave_ic = (ts.primOld, [ts.geometry(w) for w in range(4)])
        ms, _ = time_steps(ts, 10, 2, graph=(mode == "graph"))
        wave[mode] = {"ms_per_step": ms, "zone_updates_per_s": 64 * 64 / (ms * 1e-3)}
        ts.close()
    if not args.no_cpu:
        wcfg = grid_config(N1=64, N2=64, reconstruction=capi.RECON_WENO5, riemannSolver=capi.RIEMANN_LLF)
        cb = cpu_baseline_step(wcfg, p, wave_ic[0], wave_ic[1], None, steps=10, dt0=0.5 / 64, torus=False)
        cb["sample"] = f"the same 10 steps in the oracle (FD-FP64 per-zone Newton), {cb['cores']} threads"
        from oracle import oracle as O

        O.set_threads(1)
        one = cpu_baseline_step(wcfg, p, wave_ic[0], wave_ic[1], None, steps=10, dt0=0.5 / 64, torus=False)
        O.set_threads(host_cores())
        cb["one_core"] = {"value": one["value"], "cores": 1, "sample": "the same 10 steps on one core"}
        wave["cpu_baseline"] = cb
    res["linear_wave_64x64"] = wave

    # grim's 1D EMHD shock tube (src/problem/shock_tests/, N1 = 512, MINMOD +
    # LLF, conduction + viscosity, grim's three Newton iterations)
    from grim_amd import problems

    cfg = problems.shock_grid(512)
    p = problems.shock_params()
    shock = {"config": "1D Minkowski EMHD shock_tube_2, N1=512 (src/problem/shock_tests/params.cpp), FD-FP64 "
                       "Newton, 20 steps after 2 warm-up from dt0 = 1e-5"}
    for mode in ("eager", "graph"):
        ts = TimeStepper(p, cfg, stream=torch.cuda.Stream())
        ts.primOld = problems.shock_tube_prim(cfg, p, "shock_tube_2")
        ts.set_time(0., 1e-5)
        ms, _ = time_steps(ts, 20, 2, graph=(mode == "graph"))
        shock[mode] = {"ms_per_step": ms, "zone_updates_per_s": 512 / (ms * 1e-3)}
        ts.close()
    res["shock_tube_1d_512"] = shock

    # grim's Orszag-Tang vortex (src/problem/orzag_tang/params.cpp: 256^2,
    # MINMOD + LLF, ideal MHD, 5 Newton iterations)
    cfg = problems.orszag_tang_grid(256)
    p = problems.orszag_tang_params()
    ot = {"config": "2D Minkowski Orszag-Tang vortex 256x256 (src/problem/orzag_tang/params.cpp), FD-FP64 Newton, "
                    "20 steps after 2 warm-up from dt0 = 0.002"}
    ts = TimeStepper(p, cfg, stream=torch.cuda.Stream())
    ts.primOld = problems.orszag_tang_prim(cfg, p)
    ts.set_time(0., 0.002)
    ms, _ = time_steps(ts, 20, 2)
    ot["eager"] = {"ms_per_step": ms, "zone_updates_per_s": 256 * 256 / (ms * 1e-3)}
    ts.close()
    res["orszag_tang_256"] = ot

    # Bondi inflow (src/problem/bondi_inflow/): 1D Kerr-Schild r in [3, 40],
    # 256 zones, WENO5 + LLF, ghost zones held at the initial state
    cfg = problems.bondi_grid(256)
    p = problems.bondi_params()
    bondi = {"config": "1D Kerr-Schild Bondi inflow, 256 zones, r in [3, 40], sonic radius 8, WENO5+LLF, FD-FP64 "
                       "Newton, ghost zones held at the initial state, 20 steps after 2 warm-up from dt0 = 1e-3"}
    ts = TimeStepper(p, cfg, stream=torch.cuda.Stream())
    prim = problems.bondi_prim(cfg, p)
    ts.primOld = prim
    ts.set_ghost_state(prim)
    ts.set_time(0., 1e-3)
    ms, _ = time_steps(ts, 20, 2)
    bondi["eager"] = {"ms_per_step": ms, "zone_updates_per_s": 256 / (ms * 1e-3)}
    ts.close()
    res["bondi_1d_256"] = bondi
    return res


def emhd_batch_bench(args):
    """Isolated EMHD batch with uneven convergence: 2^22 zones, 7 Newton
    unknowns, per-zone closure coefficients over three decades, 10 % of the
    zones near the density floor.  Paths: grim's FD-FP64 solve, the one-pass
    mixed solve, and the mixed solve in the configuration the autotuner picks
    (workgroup size, compaction, phase budgets), replayed from a HIP graph."""
    import numpy as np
    import torch

    from grim_amd import capi

    n = args.zones
    p = capi.default_params(True)
    p.closure = capi.CLOSURE_PER_ZONE
    p.maxNonLinearIter = 20
    print(f"# EMHD batch: {n} zones, generating", file=sys.stderr, flush=True)
    b = capi.generate_kerr_schild_batch(p, n, SEED)
    rhs = torch.empty((p.nf, n), dtype=torch.float64, device="cuda")
    mask = torch.empty(n, dtype=torch.float64, device="cuda")
    capi.known_root_rhs(p, capi.make_batch(b["prim_star"], b["geom"], rhs, b["prim_old"], None, b["closure"]),
                        rhs, mask)
    prim = b["guess"].clone()
    it = torch.zeros(n, dtype=torch.int32, device="cuda")
    l2 = torch.zeros(n, dtype=torch.float64, device="cuda")
    ws = torch.empty(capi.workspace_bytes(n), dtype=torch.uint8, device="cuda")
    batch = capi.make_batch(prim, b["geom"], rhs, b["prim_old"], None, b["closure"], residual_mask=mask,
                            guess=b["guess"][: p.nf])
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    tuned = capi.autotune_newton(p, batch, capi.newton_opts(capi.JAC_ANALYTIC, capi.LU_FP32, workspace=ws),
                                 stream=stream)
    torch.cuda.synchronize()
    paths = {
        "fd_fp64": (capi.newton_opts(), False),
        "fd_fp64_split": (capi.newton_opts(compaction=2, phase1_iters=2), False),
        "mixed": (capi.newton_opts(capi.JAC_ANALYTIC, capi.LU_FP32), False),
        "mixed_tuned_graph": (tuned, True),
    }
    res = {"autotuned": {"block_size": tuned.block_size, "compaction": tuned.compaction,
                         "phase1_iters": tuned.phase1_iters, "phase2_group": tuned.phase2_group}}
    finals = {}
    for name, (opts, graph) in paths.items():
        def step():
            capi.batched_newton(p, batch, iterations=it, final_l2=l2, opts=opts,
                                stream=torch.cuda.current_stream())

        with torch.cuda.stream(stream):
            for _ in range(args.warmup):
                step()
        torch.cuda.synchronize()
        g = None
        if graph:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=stream):
                step()
            torch.cuda.synchronize()
        print(f"# EMHD batch: timing {name}", file=sys.stderr, flush=True)
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(args.steps)]
        t0 = time.perf_counter()
        with torch.cuda.stream(stream):
            for e0, e1 in evs:
                e0.record(stream)
                if g is not None:
                    g.replay()
                else:
                    step()
                e1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ms = [e0.elapsed_time(e1) for e0, e1 in evs]
        finals[name] = prim[: p.nf].clone()
        res[name] = {
            "zone_updates_per_s": n / (wall / args.steps),
            "ms_per_step": wall * 1e3 / args.steps,
            "event_ms_mean": float(np.mean(ms)),
            "unconverged": capi.count_above(l2, p.nonlinearsolve_atol),
            "iterations_hist": torch.bincount(it.to(torch.int64)).tolist(),
        }
        del g
    star = b["prim_star"][: p.nf]
    scale = star.abs().amax(dim=1, keepdim=True)
    for name in finals:
        # against the known root: rho, u, u^i relative; q and dP (drawn in
        # [-0.1, 0.1]) relative to their largest value
        d = (finals[name] - star).abs()
        d = torch.cat([d[:5] / star[:5].abs(), d[5:] / scale[5:]]).amax(dim=0)
        res[name]["err_vs_root_p99"] = float(torch.quantile(d.float(), 0.99))
        res[name]["err_vs_root_median"] = float(d.median())
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline_emhd(p, b, rhs, mask, n)
    res["speedup_tuned_vs_fd_fp64"] = res["mixed_tuned_graph"]["zone_updates_per_s"] / res["fd_fp64"]["zone_updates_per_s"]
    res["speedup_tuned_vs_fd_fp64_best"] = res["mixed_tuned_graph"]["zone_updates_per_s"] / max(
        res["fd_fp64"]["zone_updates_per_s"], res["fd_fp64_split"]["zone_updates_per_s"])
    res["speedup_tuned_vs_mixed_one_pass"] = (res["mixed_tuned_graph"]["zone_updates_per_s"]
                                              / res["mixed"]["zone_updates_per_s"])
    res["config"] = (f"isolated EMHD batch, {n} zones, 7 unknowns, per-zone closure alphas 10^[-2,1], 10% near-floor "
                     "zones, relative residual 1e-10, maxNonLinearIter 20")
    return res


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores() -> int:
    return len(os.sched_getaffinity(0))


def oracle_params(O, p):
    q = O.Params()
    for name, _ in O.Params._fields_:
        setattr(q, name, getattr(p, name))
    return q


def cpu_newton_rate(O, q, guess, geom, rhs, mask, prim_old=None, closure=None, reps=3, threads=0):
    """zone-updates/s of the oracle's FD-FP64 per-zone solve on these zones:
    median of `reps` passes after one warm-up pass."""
    import numpy as np

    n = guess.shape[-1]
    times = []
    for r in range(reps + 1):
        B = O.Batch(guess, geom, rhs, prim_old, None, closure, residualMask=mask)
        t = time.perf_counter()
        O.newton(q, B, O.NEWTON_PER_ZONE, nthreads=threads)
        dt = time.perf_counter() - t
        if r > 0:
            times.append(dt)
    return n / float(np.median(times))


def cpu_baseline(p, b, rhs, mask, n, reps):
    """The oracle (FP64 FD Jacobian, grim's algorithm) timed on the host
    cores over the same batch: median of `reps` full passes after one
    warm-up pass, OpenMP over zones on every available core; then the same
    on one core over the first 2^16 zones."""
    from oracle import oracle as O

    O.lib()
    cores = host_cores()
    q = oracle_params(O, p)
    g = b["geom"].cpu().numpy()
    guess = b["guess"].cpu().numpy()
    rh = rhs.cpu().numpy()
    mk = mask.cpu().numpy()
    print("# cpu baseline: headline batch", file=sys.stderr, flush=True)
    value = cpu_newton_rate(O, q, guess, g, rh, mk, reps=reps, threads=cores)
    m = min(n, 1 << 16)
    one = cpu_newton_rate(O, q, guess[:, :m].copy(), g[:, :m].copy(), rh[:, :m].copy(), mk[:m].copy(), reps=3,
                          threads=1)
    O.set_threads(cores)
    return {"value": value, "unit": "zone-updates/s", "cores": cores, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"full {n}-zone batch, median of {reps} passes after 1 warm-up, OpenMP over zones, "
                      f"{cores} threads",
            "one_core": {"value": one, "cores": 1, "sample": f"first {m} zones, median of 3 passes"}}


def cpu_baseline_emhd(p, b, rhs, mask, n):
    """The oracle's FD-FP64 solve of the EMHD batch on a 2^18-zone sample
    (every host core) and a 2^14-zone sample (one core)."""
    from oracle import oracle as O

    O.lib()
    cores = host_cores()
    q = oracle_params(O, p)
    m = min(n, 1 << 18)
    host = lambda t, k: t[..., :k].cpu().numpy().copy()  # noqa: E731
    print("# cpu baseline: EMHD batch", file=sys.stderr, flush=True)
    value = cpu_newton_rate(O, q, host(b["guess"], m), host(b["geom"], m), host(rhs, m), host(mask, m),
                            host(b["prim_old"], m), host(b["closure"], m), reps=3, threads=cores)
    k = min(n, 1 << 14)
    one = cpu_newton_rate(O, q, host(b["guess"], k), host(b["geom"], k), host(rhs, k), host(mask, k),
                          host(b["prim_old"], k), host(b["closure"], k), reps=3, threads=1)
    O.set_threads(cores)
    return {"value": value, "unit": "zone-updates/s", "cores": cores, "kind": "port", "cpu_model": cpu_model(),
            "sample": f"first {m} zones of the batch, median of 3 passes after 1 warm-up, {cores} threads",
            "one_core": {"value": one, "cores": 1, "sample": f"first {k} zones, median of 3 passes"}}


def cpu_baseline_step(cfg, p, prim, geoms, floors, steps, dt0, torus=True):
    """The oracle's full time step (both stages, FD-FP64 per-zone Newton, the
    torus hooks when torus) on the same grid from the same state and
    geometry as the GPU run, timed over `steps` steps on every host core."""
    from oracle import oracle as O

    O.lib()
    O.set_threads(host_cores())
    g = O.grid(**{f: getattr(cfg, f) for f, _ in cfg._fields_})
    gm = O.Geometry(g, center=geoms[0], faceX1=geoms[1], faceX2=geoms[2], gamma=geoms[3])
    t = O.torus_params(minRho=floors[0].ctypes.data, minU=floors[1].ctypes.data) if torus else None
    sim = O.Simulation(g, oracle_params(O, p), prim, dt=dt0, gm=gm, torus=t)
    zones = cfg.N1 * cfg.N2 * cfg.N3
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step()
    sec = (time.perf_counter() - t0) / steps
    return {"value": zones / sec, "unit": "zone-updates/s", "cores": host_cores(), "kind": "port",
            "cpu_model": cpu_model(), "seconds_per_step": sec}


if __name__ == "__main__":
    main()
