"""GPU parity of the batched per-zone solve against the oracle.

Every test generates its batch on the GPU with the library's own generator,
copies the inputs to the host, and runs the oracle (grim's algorithm restated
in FP64 C++, -ffp-contract=off) on exactly those inputs.  Bars:
  * residual, FD Jacobian: equal to the oracle up to FP64 rounding
    (relative 1e-14 of the row scale; most entries bit-identical);
  * FD-FP64 Newton: converged primitives within 1e-12 relative of the
    oracle's, all primitives, with identical iteration counts;
  * batched LU: LAPACK (numpy) to 1e-12 (FP64) / 1e-4 (FP32) relative.
Full-size (2^22 zones) checks are size-independent properties: every zone
converged, the known root recovered.
"""
import numpy as np
import pytest

import batchgen

pytestmark = pytest.mark.gpu

SEED = 20261202


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    return torch


@pytest.fixture(scope="module")
def capi(torch_cuda):
    from grim_amd import capi as c

    assert c.device_info().startswith("gfx950")
    return c


def host(t):
    return None if t is None else t.detach().cpu().numpy().copy()


def make_problem(capi, torch, n, emhd, seed=SEED):
    p = capi.default_params(emhd)
    p.maxNonLinearIter = 30
    if emhd:
        p.closure = capi.CLOSURE_PER_ZONE
    b = capi.generate_kerr_schild_batch(p, n, seed)
    rhs = torch.empty((p.nf, n), dtype=torch.float64, device="cuda")
    mask = torch.empty(n, dtype=torch.float64, device="cuda")
    star = capi.make_batch(b["prim_star"], b["geom"], rhs, b["prim_old"], None, b["closure"])
    capi.known_root_rhs(p, star, rhs, mask)
    torch.cuda.synchronize()
    return p, b, rhs, mask


def oracle_params(O, p):
    q = O.Params()
    for name, _ in O.Params._fields_:
        setattr(q, name, getattr(p, name))
    return q


@pytest.mark.parametrize("emhd", [False, True])
def test_generator_matches_numpy(capi, torch_cuda, oracle, emhd):
    n = 4096
    p = capi.default_params(emhd)
    b = capi.generate_kerr_schild_batch(p, n, SEED)
    ref = batchgen.kerr_schild_batch(oracle, n, emhd=emhd, seed=SEED)
    np.testing.assert_allclose(host(b["geom"]), ref["geom"], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(host(b["prim_star"]), ref["prim_star"], rtol=1e-9, atol=1e-15)
    np.testing.assert_allclose(host(b["guess"]), ref["guess"], rtol=1e-9, atol=1e-15)
    if emhd:
        np.testing.assert_allclose(host(b["prim_old"]), ref["prim_old"], rtol=1e-9, atol=1e-15)
        np.testing.assert_allclose(host(b["closure"]), ref["closure"], rtol=1e-12)


@pytest.mark.parametrize("emhd", [False, True])
def test_known_root_rhs_matches_oracle(capi, torch_cuda, oracle, emhd):
    torch = torch_cuda
    n = 8192
    p, b, rhs, mask = make_problem(capi, torch, n, emhd)
    q = oracle_params(oracle, p)
    B = oracle.Batch(host(b["prim_star"]), host(b["geom"]), np.zeros((p.nf, n)), host(b["prim_old"]), None,
                     host(b["closure"]))
    ref = oracle.known_root_rhs(q, B)
    np.testing.assert_allclose(host(rhs), ref, rtol=1e-14, atol=0)
    np.testing.assert_allclose(host(mask), 1.0 / np.abs(ref).max(axis=0), rtol=1e-14)


@pytest.mark.parametrize("emhd", [False, True])
def test_residual_matches_oracle(capi, torch_cuda, oracle, emhd):
    torch = torch_cuda
    n = 16384
    p, b, rhs, mask = make_problem(capi, torch, n, emhd)
    R = torch.empty((p.nf, n), dtype=torch.float64, device="cuda")
    capi.residual(p, capi.make_batch(b["guess"], b["geom"], rhs, b["prim_old"], None, b["closure"],
                                     residual_mask=mask), R)
    torch.cuda.synchronize()
    q = oracle_params(oracle, p)
    ref = oracle.residual(q, oracle.Batch(host(b["guess"]), host(b["geom"]), host(rhs), host(b["prim_old"]),
                                          None, host(b["closure"]), None, 1.0, host(mask)))
    got = host(R)
    # residual rows are (U - rhs) * mask with |rhs| * mask <= 1: absolute 1e-14
    assert np.abs(got - ref).max() <= 1e-14
    print("bit-identical residual entries: %.4f" % np.mean(got == ref))


@pytest.mark.parametrize("emhd", [False, True])
def test_fd_jacobian_matches_oracle(capi, torch_cuda, oracle, emhd):
    torch = torch_cuda
    n = 8192
    p, b, rhs, mask = make_problem(capi, torch, n, emhd)
    J = torch.empty((p.nf * p.nf, n), dtype=torch.float64, device="cuda")
    batch = capi.make_batch(b["guess"], b["geom"], rhs, b["prim_old"], None, b["closure"], residual_mask=mask)
    capi.jacobian(p, batch, capi.JAC_FD, J)
    torch.cuda.synchronize()
    q = oracle_params(oracle, p)
    ref = oracle.fd_jacobian(q, oracle.Batch(host(b["guess"]), host(b["geom"]), host(rhs), host(b["prim_old"]),
                                             None, host(b["closure"]), None, 1.0, host(mask)))
    got = host(J).reshape(p.nf, p.nf, n)
    ref = ref.reshape(p.nf, p.nf, n)
    # forward differences amplify a 1-ulp residual difference by 1/eps
    colscale = np.abs(ref).max(axis=0, keepdims=True) + 1e-300
    err = np.abs(got - ref) / colscale
    print("bit-identical Jacobian entries: %.4f" % np.mean(got == ref))
    assert np.median(err.max(axis=(0, 1))) < 1e-12
    assert np.percentile(err.max(axis=(0, 1)), 99) < 1e-6


@pytest.mark.parametrize("emhd", [False, True])
@pytest.mark.parametrize("semantics", ["per_zone", "global"])
def test_newton_fd_fp64_matches_oracle(capi, torch_cuda, oracle, emhd, semantics):
    torch = torch_cuda
    n = 16384
    p, b, rhs, mask = make_problem(capi, torch, n, emhd)
    sem = capi.NEWTON_PER_ZONE if semantics == "per_zone" else capi.NEWTON_GLOBAL
    if sem == capi.NEWTON_GLOBAL:
        p.maxNonLinearIter = 8
    prim = b["guess"].clone()
    it = torch.zeros(n, dtype=torch.int32, device="cuda")
    ev = torch.zeros(n, dtype=torch.int32, device="cuda")
    l2 = torch.zeros(n, dtype=torch.float64, device="cuda")
    batch = capi.make_batch(prim, b["geom"], rhs, b["prim_old"], None, b["closure"], residual_mask=mask)
    capi.batched_newton(p, batch, capi.JAC_FD, capi.LU_FP64, sem, iterations=it, residual_evals=ev, final_l2=l2)
    torch.cuda.synchronize()
    q = oracle_params(oracle, p)
    OB = oracle.Batch(host(b["guess"]), host(b["geom"]), host(rhs), host(b["prim_old"]), None, host(b["closure"]),
                      None, 1.0, host(mask))
    mode = oracle.NEWTON_PER_ZONE if sem == capi.NEWTON_PER_ZONE else oracle.NEWTON_GLOBAL
    bad, oit, ol2, oev = oracle.newton(q, OB, mode)
    got = host(prim)
    ok = ol2 <= q.nonlinearsolve_atol
    rel = np.abs(got - OB.prim) / np.maximum(np.abs(OB.prim), 1e-300)
    print("%s emhd=%s: oracle unconverged %d, identical iterations %.4f, bit-identical prims %.4f, "
          "max rel diff (converged) %.3e" % (semantics, emhd, bad, np.mean(host(it) == oit),
                                             np.mean(got == OB.prim), rel[:, ok].max()))
    assert np.all(got[p.nf:] == host(b["guess"])[p.nf:])  # B untouched
    if not emhd:
        assert bad == 0
    # Same Newton path and the same answer as the oracle: with no fused
    # multiply-adds and correctly rounded division and square root the two
    # FP64 evaluations agree bit for bit.  (Near-floor EMHD zones have FD
    # Jacobians dominated by rounding, so a 1-ulp difference there could
    # change the path; the bound on iterations leaves room for that.)
    assert np.mean(host(it) == oit) >= 0.999
    assert rel[:, ok].max() <= 1e-12
    np.testing.assert_array_equal(host(l2) <= q.nonlinearsolve_atol, ok)


@pytest.mark.parametrize("lu", ["fp64", "fp32"])
@pytest.mark.parametrize("nvar", [1, 5, 7, 10])
def test_batch_linear_solve_vs_lapack(capi, torch_cuda, lu, nvar):
    torch = torch_cuda
    n = 4096
    rng = np.random.default_rng(nvar)
    A = rng.standard_normal((n, nvar, nvar)) + 3 * np.eye(nvar)
    bvec = rng.standard_normal((n, nvar))
    ref = np.linalg.solve(A, bvec[..., None])[..., 0]
    Ad = torch.tensor(np.ascontiguousarray(A.reshape(n, nvar * nvar).T), device="cuda")
    bd = torch.tensor(np.ascontiguousarray(bvec.T), device="cuda")
    x = torch.empty_like(bd)
    capi.batch_linear_solve(Ad, bd, x, capi.LU_FP64 if lu == "fp64" else capi.LU_FP32)
    torch.cuda.synchronize()
    got = host(x).T
    cond = np.linalg.cond(A)
    tol = (1e-15 if lu == "fp64" else 1e-7) * 10 * cond
    err = np.abs(got - ref).max(axis=1) / np.abs(ref).max(axis=1)
    assert np.all(err <= tol)


def test_conserved_matches_oracle(capi, torch_cuda, oracle):
    torch = torch_cuda
    n = 8192
    p = capi.default_params(False)
    b = capi.generate_kerr_schild_batch(p, n, SEED)
    U = capi.conserved(p, b["prim_star"], b["geom"])
    torch.cuda.synchronize()
    ref = oracle.conserved(oracle_params(oracle, p), host(b["prim_star"]), host(b["geom"]))
    np.testing.assert_allclose(host(U), ref, rtol=1e-14, atol=1e-300)


def test_full_size_fd_fp64_converges(capi, torch_cuda):
    """2^22 zones: every zone reaches sum R^2 <= 1e-20 and the root."""
    torch = torch_cuda
    n = 1 << 22
    p, b, rhs, mask = make_problem(capi, torch, n, False)
    prim = b["guess"].clone()
    l2 = torch.empty(n, dtype=torch.float64, device="cuda")
    it = torch.empty(n, dtype=torch.int32, device="cuda")
    batch = capi.make_batch(prim, b["geom"], rhs, residual_mask=mask)
    capi.batched_newton(p, batch, iterations=it, final_l2=l2)
    assert capi.count_above(l2, p.nonlinearsolve_atol) == 0
    rel = ((prim[:5] - b["prim_star"][:5]).abs() / b["prim_star"][:5].abs()).amax(dim=0)
    assert float(rel.median()) < 1e-9
    assert int(it.max()) <= 10  # a handful of the 4M zones need up to 7 updates
    del b, prim, rhs, mask, l2, it
    torch.cuda.empty_cache()


# ------------------------------------------------------- analytic Jacobian
@pytest.mark.parametrize("emhd", [False, True])
def test_analytic_jacobian_matches_dual_numbers(capi, torch_cuda, oracle, emhd):
    """Hand-written tangent vs the oracle's forward-mode dual numbers (exact)."""
    torch = torch_cuda
    n = 16384
    p, b, rhs, mask = make_problem(capi, torch, n, emhd)
    J = torch.empty((p.nf * p.nf, n), dtype=torch.float64, device="cuda")
    batch = capi.make_batch(b["guess"], b["geom"], rhs, b["prim_old"], None, b["closure"], residual_mask=mask)
    capi.jacobian(p, batch, capi.JAC_ANALYTIC, J)
    torch.cuda.synchronize()
    q = oracle_params(oracle, p)
    ref = oracle.dual_jacobian(q, oracle.Batch(host(b["guess"]), host(b["geom"]), host(rhs), host(b["prim_old"]),
                                               None, host(b["closure"]), None, 1.0,
                                               host(mask))).reshape(p.nf, p.nf, n)
    got = host(J).reshape(p.nf, p.nf, n)
    colscale = np.abs(ref).max(axis=0, keepdims=True) + 1e-300
    err = np.abs(got - ref) / colscale
    worst = err.max(axis=(0, 1))
    print("analytic vs dual (emhd=%s): median %.2e  p99 %.2e  max %.2e" % (
        emhd, np.median(worst), np.percentile(worst, 99), err.max()))
    if not emhd:
        assert err.max() < 1e-9
    else:
        # near-floor zones (rho ~ 1e-8, u ~ 1e-11) lose digits to cancellation
        # in either evaluation; everywhere else the two agree to rounding
        assert np.median(worst) < 1e-12
        assert np.percentile(worst, 99) < 1e-7


@pytest.mark.parametrize("emhd", [False, True])
@pytest.mark.parametrize("mode", ["analytic64", "mixed", "mixed_escalating"])
def test_analytic_newton_converges_to_fd_solution(capi, torch_cuda, oracle, mode, emhd):
    """The analytic Jacobian in FP64, in FP32, and in FP32 for the first two
    updates of each zone then FP64 (GRIM_LU_FP32_FP64) reach grim's FD-FP64
    solution."""
    torch = torch_cuda
    n = 65536
    p, b, rhs, mask = make_problem(capi, torch, n, emhd)
    results = {}
    lu_of = {"analytic64": capi.LU_FP64, "mixed": capi.LU_FP32, "mixed_escalating": capi.LU_FP32_FP64}
    for name, jac, lu in (("fd", capi.JAC_FD, capi.LU_FP64), (mode, capi.JAC_ANALYTIC, lu_of[mode])):
        prim = b["guess"].clone()
        it = torch.zeros(n, dtype=torch.int32, device="cuda")
        l2 = torch.zeros(n, dtype=torch.float64, device="cuda")
        capi.batched_newton(p, capi.make_batch(prim, b["geom"], rhs, b["prim_old"], None, b["closure"],
                                               residual_mask=mask), jac, lu, iterations=it, final_l2=l2,
                            fused=1 if lu != capi.LU_FP64 else 0)
        torch.cuda.synchronize()
        results[name] = (host(prim), host(it), host(l2))
    fd, an = results["fd"], results[mode]
    star = host(b["prim_star"])
    nf = p.nf
    ok_fd = fd[2] <= p.nonlinearsolve_atol
    ok = an[2] <= p.nonlinearsolve_atol
    rel_fd = (np.abs(an[0][:nf] - fd[0][:nf]) / np.maximum(np.abs(fd[0][:nf]), 1e-300)).max(axis=0)
    rel_star = (np.abs(an[0][:nf] - star[:nf]) / np.maximum(np.abs(star[:nf]), 1e-300)).max(axis=0)
    rel_fd_star = (np.abs(fd[0][:nf] - star[:nf]) / np.maximum(np.abs(star[:nf]), 1e-300)).max(axis=0)
    print("%s emhd=%s: converged %.4f (fd %.4f); iterations %s (fd %s); |P-P_fd|/|P_fd| p99 %.2e median %.2e; "
          "|P-P*|/|P*| p99 %.2e median %.2e (fd %.2e, %.2e)" % (
              mode, emhd, ok.mean(), ok_fd.mean(), np.bincount(an[1]).tolist(), np.bincount(fd[1]).tolist(),
              np.percentile(rel_fd[ok & ok_fd], 99), np.median(rel_fd[ok & ok_fd]),
              np.percentile(rel_star[ok], 99), np.median(rel_star[ok]),
              np.percentile(rel_fd_star[ok_fd], 99), np.median(rel_fd_star[ok_fd])))
    assert np.all(an[0][nf:] == host(b["guess"])[nf:])
    if not emhd:
        assert np.all(ok)
        # both satisfy the relative-residual test; their distance is bounded
        # by the conditioning of each zone times the tolerance
        assert np.median(rel_fd) < 1e-10
        assert np.percentile(rel_star, 99) < 10 * np.percentile(rel_fd_star, 99) + 1e-12
    else:
        # the exact Jacobian converges at least where grim's FD Jacobian does
        assert ok.mean() >= ok_fd.mean()
        assert np.median(rel_star[ok]) < 1e-9


# ------------------------------------------------ compaction, autotune, graphs
@pytest.mark.parametrize("case", [("fd", False, 1), ("fd", True, 2), ("mixed", False, 1), ("mixed", False, 2),
                                  ("fd_fused", True, 3)])
def test_compaction_is_bit_identical_to_one_pass(capi, torch_cuda, case):
    """Stopping after k1 updates and resuming the compacted list, or gathering
    each workgroup's unconverged zones onto its first lanes, changes nothing."""
    torch = torch_cuda
    path, emhd, k1 = case
    n = 65536
    p, b, rhs, mask = make_problem(capi, torch, n, emhd)
    jac = capi.JAC_FD if path.startswith("fd") else capi.JAC_ANALYTIC
    lu = capi.LU_FP32 if path == "mixed" else capi.LU_FP64
    fused = 1 if path in ("mixed", "fd_fused") else 0
    out = {}
    for compaction in (0, 1, 2):
        prim = b["guess"].clone()
        it = torch.zeros(n, dtype=torch.int32, device="cuda")
        l2 = torch.zeros(n, dtype=torch.float64, device="cuda")
        ws = torch.empty(capi.workspace_bytes(n), dtype=torch.uint8, device="cuda")
        o = capi.newton_opts(jac, lu, compaction=compaction, fused=fused, phase1_iters=k1, phase2_group=4,
                             workspace=ws)
        capi.batched_newton(p, capi.make_batch(prim, b["geom"], rhs, b["prim_old"], None, b["closure"],
                                               residual_mask=mask), iterations=it, final_l2=l2, opts=o)
        torch.cuda.synchronize()
        out[compaction] = (host(prim), host(it), host(l2))
    for c in (1, 2):
        np.testing.assert_array_equal(out[0][0], out[c][0])
        np.testing.assert_array_equal(out[0][1], out[c][1])
        np.testing.assert_array_equal(out[0][2], out[c][2])


@pytest.mark.parametrize("case", [("fd", False, 0), ("mixed", True, 0), ("mixed", True, 1), ("escalate", True, 1),
                                  ("fd_global", False, 0), ("mixed_noalloc", True, 1)])
def test_guess_input_equals_copy_then_solve(capi, torch_cuda, case):
    """batch.guess: a solve that starts from the guess planes leaves prim
    bitwise as copying the guess into prim's fluid rows and solving in place
    (every row, iteration counts and final sum R^2 equal), reads prim's B
    rows, leaves the guess untouched; zones outside the mask take the
    guess's floored values; phase 2 of a compacted solve resumes from prim."""
    torch = torch_cuda
    path, emhd, compaction = case
    n = 65536
    p, b, rhs, mask = make_problem(capi, torch, n, emhd)
    mask[::7] = 0.
    jac = capi.JAC_FD if path.startswith("fd") else capi.JAC_ANALYTIC
    lu = {"fd": capi.LU_FP64, "fd_global": capi.LU_FP64, "mixed": capi.LU_FP32, "mixed_noalloc": capi.LU_FP32,
          "escalate": capi.LU_FP32_FP64}[path]
    sem = capi.NEWTON_GLOBAL if path == "fd_global" else capi.NEWTON_PER_ZONE
    ws = torch.empty(capi.workspace_bytes(n), dtype=torch.uint8, device="cuda")
    out = []
    guess = b["guess"][: p.nf]
    keep = guess.clone()
    for use_guess in (False, True):
        prim = b["guess"].clone()
        if use_guess:
            prim[: p.nf] = 3. * prim[: p.nf] + 1.  # overwritten by the solve from the guess
        it = torch.zeros(n, dtype=torch.int32, device="cuda")
        l2 = torch.zeros(n, dtype=torch.float64, device="cuda")
        batch = capi.make_batch(prim, b["geom"], rhs, b["prim_old"], None, b["closure"], residual_mask=mask,
                                guess=guess if use_guess else None)
        if path == "mixed_noalloc":
            o = capi.newton_opts(jac, lu, compaction=compaction, phase1_iters=1, phase2_group=2)
            capi.batched_newton_noalloc(p, batch, o, iterations=it, final_l2=l2)
        else:
            o = capi.newton_opts(jac, lu, semantics=sem, compaction=compaction, phase1_iters=1, phase2_group=2,
                                 workspace=ws if compaction else None)
            capi.batched_newton(p, batch, iterations=it, final_l2=l2, opts=o)
        torch.cuda.synchronize()
        out.append((host(prim), host(it), host(l2)))
    assert torch.equal(guess, keep)
    for a, c in zip(out[0], out[1]):
        np.testing.assert_array_equal(a, c)


def test_autotune_returns_a_working_configuration(capi, torch_cuda):
    torch = torch_cuda
    n = 1 << 20
    p, b, rhs, mask = make_problem(capi, torch, n, False)
    prim = b["guess"].clone()
    batch = capi.make_batch(prim, b["geom"], rhs, residual_mask=mask)
    o = capi.autotune_newton(p, batch, capi.newton_opts(capi.JAC_ANALYTIC, capi.LU_FP32, fused=1))
    torch.cuda.synchronize()
    assert torch.equal(prim, b["guess"])  # the batch is left untouched
    assert o.block_size in (64, 128, 256)
    print("autotuned: block %d compaction %d phase1 %d phase2 group %d" % (
        o.block_size, o.compaction, o.phase1_iters, o.phase2_group))
    ws = torch.empty(capi.workspace_bytes(n), dtype=torch.uint8, device="cuda")
    o.workspace = ws.data_ptr()
    o.workspace_bytes = ws.numel()
    l2 = torch.empty(n, dtype=torch.float64, device="cuda")
    capi.batched_newton(p, batch, final_l2=l2, opts=o)
    assert capi.count_above(l2, p.nonlinearsolve_atol) == 0


def test_graph_captured_step_matches_eager(capi, torch_cuda):
    """restore guess + compacted mixed solve captured once, replayed 4 times"""
    torch = torch_cuda
    n = 1 << 18
    p, b, rhs, mask = make_problem(capi, torch, n, False)
    prim = b["guess"].clone()
    l2 = torch.zeros(n, dtype=torch.float64, device="cuda")
    it = torch.zeros(n, dtype=torch.int32, device="cuda")
    ws = torch.empty(capi.workspace_bytes(n), dtype=torch.uint8, device="cuda")
    batch = capi.make_batch(prim, b["geom"], rhs, residual_mask=mask)
    o = capi.newton_opts(capi.JAC_ANALYTIC, capi.LU_FP32, compaction=1, fused=1, phase1_iters=2,
                         phase2_group=4, workspace=ws)

    def step():
        prim[:5].copy_(b["guess"][:5])
        capi.batched_newton(p, batch, iterations=it, final_l2=l2, opts=o)

    step()
    torch.cuda.synchronize()
    eager = (prim.clone(), it.clone(), l2.clone())
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        step()  # warm-up on the capture stream
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        step()
    for r in range(4):
        prim[:5].fill_(0.0)  # the step restores the fluid rows from the guess
        it.fill_(-1)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(prim, eager[0]) and torch.equal(it, eager[1]) and torch.equal(l2, eager[2]), r
        # eager work between replays
        step()
        torch.cuda.synchronize()
        assert torch.equal(prim, eager[0])


def test_emhd_tuned_graph_replay_equals_one_pass_mixed(capi, torch_cuda):
    """The EMHD batch's tuned path: per-zone closure alphas over three decades,
    10 % near-floor zones, maxNonLinearIter 20, analytic Jacobian + FP32 LU,
    in the configuration the autotuner picks, captured in a HIP graph and
    replayed.  Every replay is bitwise the one-pass mixed solve (no
    compaction, default workgroup), and the solution is within the
    known-root bounds of that batch: median error 1e-9, p99 1e-3 (the tail
    is the near-floor zones, whose conditioning times the relative tolerance
    1e-10 allows it), unconverged below 0.1 % of the zones."""
    torch = torch_cuda
    n = 1 << 18
    p, b, rhs, mask = make_problem(capi, torch, n, True)
    p.maxNonLinearIter = 20
    prim = b["guess"].clone()
    it = torch.zeros(n, dtype=torch.int32, device="cuda")
    l2 = torch.zeros(n, dtype=torch.float64, device="cuda")
    ws = torch.empty(capi.workspace_bytes(n), dtype=torch.uint8, device="cuda")
    batch = capi.make_batch(prim, b["geom"], rhs, b["prim_old"], None, b["closure"], residual_mask=mask)
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    tuned = capi.autotune_newton(p, batch, capi.newton_opts(capi.JAC_ANALYTIC, capi.LU_FP32, workspace=ws),
                                 stream=stream)
    torch.cuda.synchronize()
    assert torch.equal(prim, b["guess"])
    print(f"autotuned: block {tuned.block_size} compaction {tuned.compaction} phase1 {tuned.phase1_iters} "
          f"phase2 group {tuned.phase2_group}", flush=True)

    def step(opts):
        prim[: p.nf].copy_(b["guess"][: p.nf])
        capi.batched_newton(p, batch, iterations=it, final_l2=l2, opts=opts, stream=torch.cuda.current_stream())

    with torch.cuda.stream(stream):
        step(capi.newton_opts(capi.JAC_ANALYTIC, capi.LU_FP32))
    torch.cuda.synchronize()
    one_pass = (prim.clone(), it.clone(), l2.clone())
    with torch.cuda.stream(stream):
        step(tuned)  # warm-up on the capture stream
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=stream):
        step(tuned)
    torch.cuda.synchronize()
    for r in range(3):
        prim[: p.nf].fill_(0.0)
        it.fill_(-1)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(prim, one_pass[0]) and torch.equal(it, one_pass[1]) and torch.equal(l2, one_pass[2]), r
        with torch.cuda.stream(stream):
            step(capi.newton_opts(capi.JAC_ANALYTIC, capi.LU_FP32))  # eager work between replays
        torch.cuda.synchronize()
    del g
    star = b["prim_star"][: p.nf]
    scale = star.abs().amax(dim=1, keepdim=True)
    d = (one_pass[0][: p.nf] - star).abs()
    d = torch.cat([d[:5] / star[:5].abs(), d[5:] / scale[5:]]).amax(dim=0)
    unconverged = capi.count_above(one_pass[2], p.nonlinearsolve_atol)
    med, p99 = float(d.median()), float(torch.quantile(d.float(), 0.99))
    print(f"EMHD tuned graph: unconverged {unconverged} of {n}, error vs root median {med:.2e} p99 {p99:.2e}")
    assert unconverged < 1e-3 * n and med < 1e-9 and p99 < 1e-3


# ------------------------------------------------ tangent and adjoint products
@pytest.mark.parametrize("emhd", [False, True])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_adjoint_dot_product_and_dual_numbers(capi, torch_cuda, oracle, emhd, precision):
    """The reverse sweep J^T w against the column-wise tangent J v (dot-product
    test <J v, w> = <v, J^T w> per zone) and against the oracle's
    forward-mode dual-number Jacobian, transposed."""
    torch = torch_cuda
    n = 16384
    p, b, rhs, mask = make_problem(capi, torch, n, emhd)
    prec = capi.LU_FP64 if precision == "fp64" else capi.LU_FP32
    batch = capi.make_batch(b["guess"], b["geom"], rhs, b["prim_old"], None, b["closure"], residual_mask=mask)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(SEED)
    # v in the primitives' own scale, w in the residual rows' scale
    v = (torch.rand((p.nf, n), generator=gen, device="cuda", dtype=torch.float64) - 0.5) * b["guess"][:p.nf].abs()
    w = torch.rand((p.nf, n), generator=gen, device="cuda", dtype=torch.float64) - 0.5
    Jv = torch.empty_like(v)
    JTw = torch.empty_like(w)
    capi.residual_tangent(p, batch, v, Jv, prec)
    capi.residual_adjoint(p, batch, w, JTw, prec)
    torch.cuda.synchronize()
    lhs = (Jv * w).sum(dim=0)
    rhs_ = (v * JTw).sum(dim=0)
    # scale of the sums: the largest single product term of either side
    J = torch.empty((p.nf * p.nf, n), dtype=torch.float64, device="cuda")
    capi.jacobian(p, batch, capi.JAC_ANALYTIC, J)
    torch.cuda.synchronize()
    Jm = host(J).reshape(p.nf, p.nf, n)
    terms = np.abs(Jm) * np.abs(host(w))[:, None, :] * np.abs(host(v))[None, :, :]
    scale = terms.sum(axis=(0, 1)) + 1e-300
    dot_err = np.abs(host(lhs) - host(rhs_)) / scale
    # J^T w against the dual-number Jacobian of the oracle's residual
    q = oracle_params(oracle, p)
    ref = oracle.dual_jacobian(q, oracle.Batch(host(b["guess"]), host(b["geom"]), host(rhs), host(b["prim_old"]),
                                               None, host(b["closure"]), None, 1.0,
                                               host(mask))).reshape(p.nf, p.nf, n)
    ref_JTw = np.einsum("ijz,iz->jz", ref, host(w))
    colscale = (np.abs(ref) * np.abs(host(w))[:, None, :]).sum(axis=0) + 1e-300
    adj_err = (np.abs(host(JTw) - ref_JTw) / colscale).max(axis=0)
    print("%s emhd=%s: dot-product relative error median %.2e p99 %.2e max %.2e; J^T w vs dual median %.2e "
          "p99 %.2e max %.2e" % (precision, emhd, np.median(dot_err), np.percentile(dot_err, 99), dot_err.max(),
                                 np.median(adj_err), np.percentile(adj_err, 99), adj_err.max()))
    assert np.isfinite(host(JTw)).all()
    if precision == "fp64":
        assert dot_err.max() < 1e-12
        if not emhd:
            assert adj_err.max() < 1e-9
        else:
            # near-floor zones lose digits to cancellation in either evaluation
            assert np.median(adj_err) < 1e-12 and np.percentile(adj_err, 99) < 1e-7
    else:
        # FP32 sweep: rounding of single precision over a few hundred operations
        assert np.percentile(dot_err, 99) < 1e-4
        assert np.median(adj_err) < 1e-5
