"""The C++ interface (grim_amd/host/timestepper.hpp: grim's timeStepper
names over the C ABI) and the grim-style driver built on it
(grim_amd/grim_step, grim_amd/host/grim_main.cpp).

CPU: the driver is built, and without a GPU it fails loudly (exit status 1
and the library's message) instead of falling back to anything.
GPU: the driver's steps equal, bit for bit, the same steps taken through the
Python mirror of the interface (same library, same parameters), for the
64x64 fast wave with grim's FD-FP64 solve and for a small EMHD torus with the
escalating mixed solve; the torus observers it prints after every step
(--observe-every) match the Python interface's.
"""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(os.path.dirname(HERE), "grim_amd", "grim_step")


def test_driver_is_built_and_fails_loudly_without_a_gpu():
    assert os.path.exists(DRIVER), "grim_amd/grim_step not built (run __graft_entry__.build())"
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    r = subprocess.run([DRIVER, "--steps", "1"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1
    assert "grim_create" in r.stderr


def test_driver_rejects_bad_arguments():
    r = subprocess.run([DRIVER, "--solver", "nonsense"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "unknown solver" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["fastmode_fd", "torus_mixed"])
def test_driver_matches_python_interface(tmp_path, case):
    import torch  # noqa: F401

    from grim_amd import capi
    from grim_amd.timestepper import TimeStepper, grid_config

    dump = str(tmp_path / "prim.bin")
    if case == "fastmode_fd":
        args = ["--problem", "fastmode", "--steps", "5", "--solver", "fd"]
        p = capi.default_params(False)
        p.maxNonLinearIter = 10
        cfg = grid_config(N1=64, N2=64, reconstruction=capi.RECON_WENO5, riemannSolver=capi.RIEMANN_LLF)
        opts, steps = capi.newton_opts(), 5
    else:
        args = ["--problem", "torus", "--N", "96", "48", "4", "--steps", "2", "--solver", "mixed",
                "--observe-every", "1e-9"]
        import bench

        p = bench.torus_params_emhd()
        cfg = bench.torus_config(96, 48, 4)
        opts, steps = capi.newton_opts(capi.JAC_ANALYTIC, capi.LU_FP32_FP64), 2
    print("driver:", " ".join(args), flush=True)
    r = subprocess.run([DRIVER, *args, "--dump", dump], capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    got = np.fromfile(dump, dtype=np.float64)
    ts = TimeStepper(p, cfg, opts=opts)
    if case == "fastmode_fd":
        ts.initLinearMode(0, 1e-4)
        ts.set_time(0., 0.5 / 64)
    else:
        ts.initTorus(capi.torus_params())
        ts.set_time(0., 0.005)
    for _ in range(steps):
        ts.timeStep()
    ref = ts.primOld.ravel()
    ts.close()
    assert got.shape == ref.shape
    assert np.array_equal(got, ref, equal_nan=True)
    if case == "torus_mixed":
        # the observers printed after the last step (grim's format, %e)
        ts = TimeStepper(p, cfg, opts=opts)
        ts.primOld = ref
        obs = ts.observers()
        ts.close()
        line = [ln for ln in r.stdout.splitlines() if ln.startswith("Baryon Mass = ")][-1]
        mass = float(line.split("=")[1].split(";")[0])
        assert r.stdout.count("Global quantities at t = ") == steps
        assert mass == pytest.approx(obs["baryon_mass"], rel=1e-5)


@pytest.mark.gpu
def test_driver_buoyancy_reads_grim_atmosphere_files(tmp_path):
    """--problem buoyancy (buoyancy_instabilities/problem.cpp:22-216): the
    driver reads the four atmosphere_soln_*.txt files written by
    problems.write_atmosphere_files, sets the state through the C++
    interface and its X1-side hook, and after 4 steps holds the Python
    path's state bitwise (same profiles, noise seed and amplitude); without
    the files it stops with grim's message."""
    import torch  # noqa: F401

    from grim_amd import problems as PR
    from grim_amd.timestepper import TimeStepper

    r = subprocess.run([DRIVER, "--problem", "buoyancy", "--steps", "1", "--atmosphere", str(tmp_path)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 1 and "Input data files not found!" in r.stderr
    cfg, p = PR.buoyancy_grid(64, 64), PR.buoyancy_params()
    prof = PR.hydrostatic_atmosphere(cfg, p)
    PR.write_atmosphere_files(str(tmp_path), *prof)
    assert np.array_equal(np.array(PR.read_atmosphere_files(cfg, str(tmp_path))), np.array(prof))
    dump = str(tmp_path / "prim.bin")
    r = subprocess.run([DRIVER, "--problem", "buoyancy", "--steps", "4", "--atmosphere", str(tmp_path),
                        "--dump", dump], capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(dump, dtype=np.float64)
    ts = TimeStepper(p, cfg)
    prim, side = PR.buoyancy_prim(cfg, p, ts.geometry(0), *PR.read_atmosphere_files(cfg, str(tmp_path)),
                                  amplitude=4e-2, seed=20261203)
    ts.primOld = prim
    ts.set_x1_side_state(side)
    ts.set_time(0., 0.01)
    for _ in range(4):
        ts.timeStep()
    ref = ts.primOld.ravel()
    ts.close()
    assert got.shape == ref.shape and np.array_equal(got, ref)


@pytest.mark.gpu
def test_driver_restart_continues_the_run(tmp_path):
    """CheckWallClockTermination + the constructor's restart branch: a run
    stopped after its first step (restartVarsT*.h5, restartFileName.txt,
    restartFileTime.txt written) and started again in the same directory
    continues to the state of the uninterrupted run.  grim's restart keeps no
    dt (the next computeDt is bounded by maxDtIncrement * InitialDt), so the
    runs start from an InitialDt above the Courant limit: every step then
    takes the Courant dt.  The restart file holds the bulk only, so ghost
    zones (and through them computeDt) may differ in the last bits: the
    comparison is to 1e-9 of each variable's scale."""
    whole, part = tmp_path / "whole", tmp_path / "part"
    whole.mkdir()
    part.mkdir()
    base = [DRIVER, "--problem", "fastmode", "--solver", "fd", "--dt0", "0.05"]
    r = subprocess.run(base + ["--steps", "4", "--dump", "prim.bin"], cwd=whole, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    r = subprocess.run(base + ["--steps", "4", "--max-wall-time", "0"], cwd=part, capture_output=True, text=True,
                       timeout=300)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0 and "Termination reason: WallClock" in r.stdout
    name = (part / "restartFileName.txt").read_text().split()[0]
    assert name.startswith("restartVarsT") and (part / name).exists()
    assert float((part / "restartFileTime.txt").read_text()) > 0.
    r = subprocess.run(base + ["--steps", "3", "--dump", "prim.bin"], cwd=part, capture_output=True, text=True,
                       timeout=300)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0 and "Restarting from file" in r.stdout
    g = 3
    a = np.fromfile(whole / "prim.bin").reshape(8, 64 + 2 * g, 64 + 2 * g)[:, g:-g, g:-g]
    b = np.fromfile(part / "prim.bin").reshape(8, 64 + 2 * g, 64 + 2 * g)[:, g:-g, g:-g]
    scale = np.abs(a).reshape(8, -1).max(axis=1)[:, None, None] + 1e-300
    diff = (np.abs(a - b) / scale).max()
    print(f"restart vs uninterrupted: max scaled difference {diff:.2e}", flush=True)
    assert diff < 1e-9


@pytest.mark.gpu
def test_driver_write_every_writes_primitives_and_vts(tmp_path):
    """WriteData (torus/problem.cpp:777-876): at the start (the initial
    conditions' fullStepDiagnostics, :539, with the metric files) and every
    time the run crosses a multiple of --write-every it writes
    primVarsT%06d.h5 and dump%06d.vts;
    after the last step they hold the final state (rho and u of the VTS equal
    the primitives, the field within the 1e-10 that dx/dX by central
    differences leaves, the points the zone centres)."""
    from grim_amd import checkpoint
    from grim_amd.timestepper import grid_config

    args = [DRIVER, "--problem", "fastmode", "--solver", "fd", "--dt0", "0.05", "--steps", "2",
            "--write-every", "1e-9", "--dump", "prim.bin"]
    r = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0, r.stderr
    vts = sorted(tmp_path.glob("dump*.vts"))
    h5 = sorted(tmp_path.glob("primVarsT*.h5"))
    # the initial conditions' WriteIdx 0 and one write per step
    assert len(vts) == len(h5) == 3
    assert h5[0].name == "primVarsT000000.h5" and vts[0].name == "dump000000.vts"
    # the metric files of the first write (torus/problem.cpp:780-792):
    # Minkowski, so gCov = gCon = diag(-1, 1, 1, 1), sqrt(-g) = 1, and the
    # zone centres' x = X (X3 = 1/2 on the 2D grid)
    from test_checkpoint import read_dataset

    eta = np.diag([-1., 1., 1., 1.]).ravel()
    for name in ("gCov", "gCon"):
        data = read_dataset(str(tmp_path / f"{name}.h5"), name)
        assert data.shape[-1] == 16 and int(np.prod(data.shape[:-1])) == 64 * 64
        np.testing.assert_array_equal(data.reshape(-1, 16), np.tile(eta, (64 * 64, 1)))
    data = read_dataset(str(tmp_path / "sqrtDetg.h5"), "sqrtDetg")
    np.testing.assert_array_equal(data.ravel(), 1.)
    data = read_dataset(str(tmp_path / "xCoords.h5"), "xCoords")
    xc = data.reshape(64 * 64, 3)
    Xc = (np.arange(64) + 0.5) / 64
    np.testing.assert_array_equal(xc[:, 0], np.tile(Xc, 64))
    np.testing.assert_array_equal(xc[:, 1], np.repeat(Xc, 64))
    np.testing.assert_array_equal(xc[:, 2], 0.5)
    g = 3
    prim = np.fromfile(tmp_path / "prim.bin").reshape(8, 64 + 2 * g, 64 + 2 * g)[:, g:-g, g:-g].reshape(8, -1)
    extent, pts, arrays = checkpoint.read_vts(str(vts[-1]))
    assert extent == [0, 63, 0, 63, 0, 0]
    assert list(arrays) == ["rho", "u", "u1", "u2", "u3", "B1", "B2", "B3", "bSqr", "gamma"]
    np.testing.assert_array_equal(arrays["rho"], prim[0])
    np.testing.assert_array_equal(arrays["u"], prim[1])
    for i in range(3):
        np.testing.assert_allclose(arrays[f"B{i + 1}"], prim[5 + i], rtol=1e-10, atol=1e-14)
    X = (np.arange(64) + 0.5) / 64
    np.testing.assert_array_equal(pts[:, 0], np.tile(X, 64))
    np.testing.assert_array_equal(pts[:, 1], np.repeat(X, 64))
    np.testing.assert_array_equal(pts[:, 2], 0.)  # a 2D DMDA has (x, y) coordinates only
    cfg = grid_config(N1=64, N2=64)
    full = np.zeros((8, cfg.N1 + 2 * g, cfg.N2 + 2 * g))
    loaded = checkpoint.load(cfg, full.reshape(8, -1).copy(), str(h5[-1]))
    np.testing.assert_array_equal(loaded.reshape(8, 70, 70)[:, g:-g, g:-g].reshape(8, -1), prim)


FACADE = os.path.join(os.path.dirname(HERE), "grim_amd", "facade_check")


def read_facade(outdir):
    out = {}
    for line in open(os.path.join(outdir, "manifest.txt")):
        name, count = line.split()
        out[name] = np.fromfile(os.path.join(outdir, name + ".f64"), dtype=np.float64)
        assert out[name].size == int(count)
    return out


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return bool(((a == b) | (np.isnan(a) & np.isnan(b))).all())


@pytest.mark.gpu
def test_cpp_interfaces_match_python_path(tmp_path):
    """grim's C++ interfaces of grim_amd/host (grid, geometry, fluidElement,
    riemannSolver, reconstruction::reconstruct, timeStepper::solve /
    computeResidual / computeDivOfFluxes / batchLinearSolve and the step
    written out stage by stage), run by the g++-built grim_amd/facade_check
    on a 48x32x8 EMHD torus, against the same operations taken through the
    Python bindings (and, where one exists, through a second kernel path)."""
    import torch

    import bench
    from grim_amd import capi
    from grim_amd.timestepper import TimeStepper

    assert os.path.exists(FACADE), "grim_amd/facade_check not built (run __graft_entry__.build())"
    print("facade_check on a 48x32x8 EMHD torus", flush=True)
    r = subprocess.run([FACADE, str(tmp_path)], capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr
    F = read_facade(tmp_path)
    p = bench.torus_params_emhd()
    cfg = bench.torus_config(48, 32, 8)
    ts = TimeStepper(p, cfg)
    ts.initTorus(capi.torus_params())
    ts.set_time(0., 0.005)
    dof, n = p.dof, cfg.n
    prim0 = ts.primOld
    assert same(F["primOld"], prim0.ravel())
    P = torch.from_numpy(prim0).cuda()
    host = lambda t: t.cpu().numpy().ravel()  # noqa: E731
    for d in range(3):
        L, R = ts.reconstruct(P, d)
        # the C++ grids start at zero where reconstruct writes nothing
        wrote = ~np.isnan(host(L))
        assert same(F[f"primLeft{d}"][wrote], host(L)[wrote]) and same(F[f"primRight{d}"][wrote], host(R)[wrote])
        # riemannSolver::solve from the reconstructed states = the fused
        # reconstruction + Riemann kernel, on every face it writes
        fused = host(ts.riemannFlux(P, d))
        faces = ~np.isnan(fused)
        assert faces.sum() > 0 and same(F[f"riemann{d}"][faces], fused[faces]), d
    fields = ts.elementFields(P)
    for name, key in (("rho", "rho"), ("tau", "tau"), ("uCov2", "uCov2"), ("bCon0", "bCon0"),
                      ("soundSpeed", "soundSpeed")):
        assert same(F["elem_" + name], host(fields[key])), name
    for d in range(4):
        assert same(F[f"fluxes{d}"], host(ts.elementFluxes(P, d))), d
    # conserved variables of the element = grim_conserved on the batch layout
    U = capi.conserved(p, P, torch.from_numpy(ts.geometry(0)[:16].copy()).cuda(),
                       radius=torch.from_numpy(ts.geometry(0)[22].copy()).cuda())
    assert same(F["fluxes0"], host(U))
    for d in range(3):
        mn, mx = ts.elementCharSpeeds(P, d)
        assert same(F[f"minSpeed{d}"], host(mn)) and same(F[f"maxSpeed{d}"], host(mx))
    assert same(F["fluxesFaceX1"], host(ts.elementFluxes(P, 1, location=1)))
    tau = fields["tau"].contiguous()
    Si = ts.elementImplicitSources(P, tau)
    assert same(F["implicitSources"], host(Si))
    # closed form of computeImplicitSources: g qTilde / tau, zero elsewhere
    g = ts.geometry(0)[15]
    Sih = host(Si).reshape(dof, n)
    np.testing.assert_array_equal(Sih[5], g * prim0[5] / host(tau))
    assert (Sih[:5] == 0).all() and (Sih[7:] == 0).all()
    Pn = prim0.copy()
    Pn[1] *= 1.01
    Pn[2] += 1e-3
    Std = ts.elementTimeDerivSources(P, P, torch.from_numpy(Pn).cuda(), 0.005)
    assert same(F["timeDerivSources"], host(Std))
    sexp = host(ts.computeExplicitSources(P))
    bulkmask = ~np.isnan(sexp)
    assert same(F["explicitSources"][bulkmask], sexp[bulkmask])
    div = host(ts.computeDivOfFluxes(P))
    ok = ~np.isnan(div)
    assert same(F["divFluxes"][ok], div[ok])
    # the standalone geometry equals the context's
    G = ts.geometry(0)
    assert same(F["geom_gCov11"], G[4]) and same(F["geom_gCon23"], G[16 + 4]) and same(F["ts_gCon23"], G[20])
    assert same(F["geom_alpha"], G[14]) and same(F["geom_g"], G[15])
    assert same(F["geom_gamma123"], ts.geometry(3)[1 * 16 + 2 * 4 + 3])
    GX1 = ts.geometry(1)[14]
    assert same(F["ts_gammaRight_alpha"][:-1], GX1[1:])
    # the step written out stage by stage = grim_time_step, bit for bit, and
    # both equal the Python path's two steps
    assert same(F["step_timeStep"], F["step_byStages"])
    dts = F["step_dt"]
    assert dts[0] == dts[1] and dts[2] == dts[3]
    ts.timeStep()
    ts.timeStep()
    assert same(F["step_timeStep"], ts.primOld.ravel())
    # the next half stage through the Python pieces: frozen terms, the
    # residual at the guess, the solve, the residual at the solution
    ts.stepBegin()
    ts.foldRHS(0)
    guess = torch.from_numpy(F["halfStep_guess"].reshape(dof, n).copy()).cuda()
    assert same(F["residual_guess"], host(ts.computeResidual(0, guess)))
    ts.solve(0, prim=guess)
    assert same(F["halfStep_solution"], host(guess))
    res_sol = ts.computeResidual(0, guess)
    assert same(F["residual_solution"], host(res_sol))
    # the solved half stage meets the relative residual test in the bulk
    mask = F["residualMask"]
    it, l2_stats = ts.lastSolveStats()
    l2 = (host(res_sol).reshape(p.nf, n) ** 2).sum(axis=0)
    bulk_zones = mask > 0
    assert (l2_stats[bulk_zones] <= p.nonlinearsolve_atol).mean() > 0.97
    np.testing.assert_allclose(l2[bulk_zones], l2_stats[bulk_zones], rtol=1e-12, atol=1e-300)
    # batchLinearSolve against LAPACK
    A = F["lu_A"].reshape(5, 5, 1000).transpose(2, 0, 1)
    b = F["lu_b"].reshape(5, 1000).T
    x = F["lu_x"].reshape(5, 1000).T
    np.testing.assert_allclose(x, np.linalg.solve(A, b[..., None])[..., 0], rtol=1e-12, atol=1e-14)
    ts.close()
