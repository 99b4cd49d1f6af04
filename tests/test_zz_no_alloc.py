"""The allocation-free batched solve (grim_batched_newton_noalloc) against
grim_batched_newton with a caller workspace, through the C ABI.

Both run the same kernels; the only difference is where the compaction work
list lives (the library's per-device array against the caller's buffer), so
primitives, iteration counts and final residuals must be bit-identical.
Each case runs two calls in a row on the shared array, the second on a
smaller batch, so the second call reads nothing the first one left.
"""
import ctypes as C

import numpy as np
import pytest

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


def problem(capi, torch, n, emhd, seed):
    p = capi.default_params(emhd)
    p.maxNonLinearIter = 30
    if emhd:
        p.closure = capi.CLOSURE_PER_ZONE
    b = capi.generate_kerr_schild_batch(p, n, seed)
    rhs = torch.empty((p.nf, n), dtype=torch.float64, device="cuda")
    mask = torch.empty(n, dtype=torch.float64, device="cuda")
    capi.known_root_rhs(p, capi.make_batch(b["prim_star"], b["geom"], rhs, b["prim_old"], None, b["closure"]),
                        rhs, mask)
    torch.cuda.synchronize()
    return p, b, rhs, mask


def solve(capi, torch, p, b, rhs, mask, opts_kw, noalloc):
    n = rhs.shape[-1]
    prim = b["guess"].clone()
    it = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    ev = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    l2 = torch.full((n,), -1.0, dtype=torch.float64, device="cuda")
    batch = capi.make_batch(prim, b["geom"], rhs, b["prim_old"], None, b["closure"], residual_mask=mask)
    if noalloc:
        o = capi.newton_opts(**opts_kw)
        capi.batched_newton_noalloc(p, batch, o, iterations=it, residual_evals=ev, final_l2=l2)
    else:
        ws = torch.empty(capi.workspace_bytes(n), dtype=torch.uint8, device="cuda")
        o = capi.newton_opts(workspace=ws, **opts_kw)
        capi.batched_newton(p, batch, iterations=it, residual_evals=ev, final_l2=l2, opts=o)
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in (prim, it, ev, l2)]


CASES = {
    # the escalating mixed solve (FP32 phase 1, FP64 phase 2), EMHD: the torus path
    "emhd_escalating": dict(emhd=True, opts=dict(jacobian=1, lu_precision=2, compaction=1, fused=1,
                                                 phase1_iters=2, phase2_group=4)),
    # grim's FD-FP64 solve, compacting after one update, ideal MHD
    "ideal_fd_compacting": dict(emhd=False, opts=dict(jacobian=0, lu_precision=0, compaction=1, phase1_iters=1,
                                                      phase2_group=2)),
    # no compaction: the scratch array is not touched
    "ideal_mixed_one_pass": dict(emhd=False, opts=dict(jacobian=1, lu_precision=1, fused=1)),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_noalloc_matches_batched_newton(capi, torch_cuda, case):
    torch = torch_cuda
    c = CASES[case]
    for call, n in enumerate((200_000, 70_001)):
        print(f"case {case} call {call} n {n}", flush=True)
        p, b, rhs, mask = problem(capi, torch, n, c["emhd"], SEED + call)
        ref = solve(capi, torch, p, b, rhs, mask, c["opts"], noalloc=False)
        got = solve(capi, torch, p, b, rhs, mask, c["opts"], noalloc=True)
        for name, r, g in zip(("prim", "iterations", "residual_evals", "final_l2"), ref, got):
            np.testing.assert_array_equal(g, r, err_msg=f"{case} call {call}: {name}")
        assert (got[1] >= 0).all() and (got[3] >= 0).all()


def test_noalloc_rejects_bad_requests_before_launch(capi, torch_cuda):
    torch = torch_cuda
    p, b, rhs, mask = problem(capi, torch, 4096, False, SEED)
    prim = b["guess"].clone()
    batch = capi.make_batch(prim, b["geom"], rhs, residual_mask=mask)
    # lock-step semantics need per-zone state beyond the scratch array
    with pytest.raises(capi.GrimError) as e:
        capi.batched_newton_noalloc(p, batch, capi.newton_opts(semantics=capi.NEWTON_GLOBAL))
    assert e.value.code == capi.GRIM_ERR_BAD_ARG
    # a caller workspace is not taken
    ws = torch.empty(capi.workspace_bytes(4096), dtype=torch.uint8, device="cuda")
    with pytest.raises(capi.GrimError) as e:
        capi.batched_newton_noalloc(p, batch, capi.newton_opts(compaction=1, workspace=ws))
    assert e.value.code == capi.GRIM_ERR_BAD_ARG
    # more zones than the scratch array holds: rejected on the host, nothing read
    big = capi.ZoneBatch.from_buffer_copy(batch)
    big.n = capi.NOALLOC_MAX_ZONES + 1
    rc = capi.lib().grim_batched_newton_noalloc(C.byref(p), C.byref(big), C.byref(capi.newton_opts(compaction=1)),
                                                None, capi._stream(None))
    assert rc == capi.GRIM_ERR_BAD_ARG
    torch.cuda.synchronize()
    assert torch.equal(prim, b["guess"])
