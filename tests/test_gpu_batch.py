"""GPU: batched solves (cmx_solve_batch) -- several problems advancing through one sequence of batched
kernels (CG, reliable CG), or one Arnoldi step per host read-back (FGMRES-DR) -- against the same
problems solved one at a time.

With the operators' own launch shapes (tune=0) a batched problem runs the kernels of its single solve
with the same work-group decomposition and reduction order, so its solution must equal the single
solve's bit for bit, with the same iteration count.  With tuned shapes the solutions are judged like
every solve: FP64 true residual <= 1e-10 and agreement with an FP64 CG solution to 1e-8.
"""
import sys

import numpy as np
import pytest

from chroma_amd import lib as cmx
from tests.util import rel

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_problem(L, seed, kappa, prec=cmx.HALF):
    U = cmx.gauge_generate(L, kind=1, seed=seed, eps=0.3)
    U = cmx.gauge_apply_bc(L, U, (1, 1, 1, -1))
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_CLOVER, cmx.kappa_to_mass(kappa), 1.0, sloppy_prec=prec)
    Vh = int(np.prod(L)) // 2
    chi = cmx.z2_source((Vh, 4, 3), seed + 1)
    return op, chi


UNEVEN = [((4, 4, 4, 4), 1000, 0.120), ((6, 6, 6, 12), 1001, 0.128), ((8, 8, 8, 16), 1002, 0.134),
          ((4, 4, 4, 4), 1003, 0.131), ((8, 8, 8, 16), 1004, 0.122)]


@pytest.mark.parametrize("solver,sloppy", [(cmx.RELIABLE_CG, cmx.HALF), (cmx.RELIABLE_CG, cmx.SINGLE),
                                           (cmx.CG, cmx.DOUBLE)])
def test_batch_uneven_matches_single(solver, sloppy):
    say("batch uneven", solver, sloppy)
    probs = [make_problem(L, s, k, cmx.HALF if sloppy == cmx.HALF else cmx.SINGLE) for L, s, k in UNEVEN]
    ops = [p[0] for p in probs]
    chis = [p[1] for p in probs]
    kw = dict(solver=solver, rsd=1e-10, maxit=5000, delta=0.1, sloppy_prec=sloppy, true_residual=True)
    single = [op.solve(chi, **kw) for op, chi in probs]
    psis, res, summ = cmx.solve_batch(ops, chis, tune=False, **kw)
    say("  summary", summ)
    for k in range(len(probs)):
        say("  problem", k, "single", single[k][1]["n_count"], "batch", res[k]["n_count"], res[k]["rel_resid"],
            "restarts", res[k]["n_restarts"])
        assert res[k]["n_count"] == single[k][1]["n_count"]
        assert res[k]["n_restarts"] == single[k][1]["n_restarts"]
        assert np.array_equal(psis[k].view(np.uint64), single[k][0].view(np.uint64))
        assert res[k]["rel_resid"] <= 1e-10


def test_batch_graph_matches_eager():
    say("batch graph")
    probs = [make_problem(L, s, k) for L, s, k in UNEVEN[:3]]
    ops = [p[0] for p in probs]
    chis = [p[1] for p in probs]
    kw = dict(solver=cmx.RELIABLE_CG, rsd=1e-10, maxit=5000, sloppy_prec=cmx.HALF, tune=False, true_residual=True)
    p0, r0, _ = cmx.solve_batch(ops, chis, graph=False, **kw)
    for _ in range(2):
        p1, r1, _ = cmx.solve_batch(ops, chis, graph=True, **kw)
        for k in range(len(probs)):
            assert r1[k]["n_count"] == r0[k]["n_count"]
            assert np.array_equal(p1[k].view(np.uint64), p0[k].view(np.uint64))


def test_batch_tuned_uneven():
    """Tuned launch shape over the batch: every solution meets the FP64 true residual and agrees with
    an FP64 CG solution."""
    say("batch tuned")
    probs = [make_problem(L, s, k) for L, s, k in UNEVEN]
    ops = [p[0] for p in probs]
    chis = [p[1] for p in probs]
    psis, res, summ = cmx.solve_batch(ops, chis, solver=cmx.RELIABLE_CG, rsd=1e-10, maxit=5000,
                                      sloppy_prec=cmx.HALF, true_residual=True, tune=True)
    say("  summary", summ)
    assert summ["hop_block"] in (64, 128, 256)
    for k, (op, chi) in enumerate(probs):
        ref, _ = op.solve(chi, solver=cmx.CG, rsd=1e-12, maxit=20000)
        say("  problem", k, res[k]["n_count"], res[k]["rel_resid"], rel(psis[k], ref))
        assert res[k]["rel_resid"] <= 1e-10
        assert rel(psis[k], ref) < 1e-8


def test_batch_sources_on_one_operator():
    """The 12 spin-colour point sources of a propagator (quarkProp4_a) on one operator: the repeated
    operator gets one solver context per source; each solution equals its single solve."""
    L = (8, 8, 8, 16)
    say("batch sources on one operator")
    op, _ = make_problem(L, 1010, 0.115)
    Vh = int(np.prod(L)) // 2
    chis = []
    for s in range(4):
        for c in range(3):
            chi = np.zeros((Vh, 4, 3), complex)
            chi[0, s, c] = 1.0  # the first odd site
            chis.append(chi)
    kw = dict(solver=cmx.RELIABLE_CG, rsd=1e-10, maxit=5000, sloppy_prec=cmx.HALF, true_residual=True)
    psis, res, summ = cmx.solve_batch([op] * 12, chis, tune=False, **kw)
    say("  summary", summ)
    for k in range(12):
        ref, rr = op.solve(chis[k], **kw)
        assert res[k]["rel_resid"] <= 1e-10
        assert res[k]["n_count"] == rr["n_count"]
        assert np.array_equal(psis[k].view(np.uint64), ref.view(np.uint64))


FGMRES_CASES = [(cmx.PRECOND_LEARNED, cmx.HALF, 0, 0), (cmx.PRECOND_LEARNED, cmx.HALF, 0, 1),
                (cmx.PRECOND_LEARNED, cmx.SINGLE, 4, 1), (cmx.PRECOND_LEARNED, cmx.SINGLE, 0, 2),
                (cmx.PRECOND_LEARNED, cmx.HALF, 0, 3), (cmx.PRECOND_LEARNED, cmx.HALF, 0, 4),
                (cmx.PRECOND_MR, cmx.SINGLE, 4, -1), (cmx.PRECOND_NONE, cmx.DOUBLE, 4, -1),
                (cmx.PRECOND_MG, cmx.SINGLE, 0, -1), (cmx.PRECOND_MG, cmx.HALF, 0, -1)]


@pytest.mark.parametrize("precond,sloppy,ndefl,variant", FGMRES_CASES)
def test_batch_fgmres_sources_match_single(precond, sloppy, ndefl, variant):
    """The 12 point sources on one operator through the batched FGMRES-DR (a repeated operator's solves
    share its network, applied to all sources in one launch per kernel with the inference path `variant`:
    0 operands in registers, 1 transports in LDS, 2 unfused, one source at a time, 3 and 4 as 0 and 1 with
    streamed links; or its multigrid; the
    operator applied to all sources in one batched grid): each solution and
    iteration count equals the single solve's bit for bit, and reaches the FP64 true residual 1e-10."""
    L = (8, 8, 8, 16)
    say("batch fgmres", precond, sloppy, ndefl, variant)
    kappa = 0.13
    op, _ = make_problem(L, 1030, kappa, cmx.HALF if sloppy == cmx.HALF else cmx.SINGLE)
    if precond == cmx.PRECOND_LEARNED:
        net = cmx.Net()
        net.init_neumann(4.0 + cmx.kappa_to_mass(kappa))
        net.set_depth(2)
        cmx.net_attach(op, net)
        cmx.net_set_variant(op, sloppy, variant)
    if precond == cmx.PRECOND_MG:  # the clones share the operator's multigrid
        op.mg_setup(blocking=(2, 2, 2, 2), nvec=4, nu_post=3, setup_maxit=30, coarse_tol=0.1, seed=1)
    Vh = int(np.prod(L)) // 2
    chis = []
    for s in range(4):
        for c in range(3):
            chi = np.zeros((Vh, 4, 3), complex)
            chi[0, s, c] = 1.0
            chis.append(chi)
    kw = dict(solver=cmx.FGMRESDR, rsd=1e-10, maxit=3000, sloppy_prec=sloppy, basis_prec=sloppy, n_krylov=16,
              n_defl=ndefl, precond=precond, precond_iters=4, true_residual=True,
              cycle_reduction=3e-3 if sloppy == cmx.HALF else 0.0)
    psis, res, summ = cmx.solve_batch([op] * 12, chis, **kw)
    say("  summary", summ)
    for k in range(12):
        ref, rr = op.solve(chis[k], **kw)
        say("  source", k, "batch", res[k]["n_count"], res[k]["rel_resid"], "single", rr["n_count"])
        assert res[k]["rel_resid"] <= 1e-10
        assert res[k]["n_count"] == rr["n_count"]
        assert np.array_equal(psis[k].view(np.uint64), ref.view(np.uint64))


def test_batch_fgmres_limited_concurrency():
    """12 sources with at most 3 solves in flight: each waiting solve takes over the solver context of a
    finished one; every solution still equals its single solve bit for bit."""
    L = (8, 8, 8, 16)
    say("batch fgmres, 3 in flight")
    kappa = 0.13
    op, _ = make_problem(L, 1031, kappa, cmx.HALF)
    net = cmx.Net()
    net.init_neumann(4.0 + cmx.kappa_to_mass(kappa))
    cmx.net_attach(op, net)
    Vh = int(np.prod(L)) // 2
    chis = [cmx.z2_source((Vh, 4, 3), 40 + k) for k in range(12)]
    kw = dict(solver=cmx.FGMRESDR, rsd=1e-10, maxit=3000, sloppy_prec=cmx.HALF, basis_prec=cmx.HALF, n_krylov=16,
              precond=cmx.PRECOND_LEARNED, cycle_reduction=3e-3)
    psis, res, summ = cmx.solve_batch([op] * 12, chis, max_concurrent=3, **kw)
    say("  summary", summ)
    for k in range(12):
        ref, rr = op.solve(chis[k], **kw)
        assert res[k]["n_count"] == rr["n_count"]
        assert np.array_equal(psis[k].view(np.uint64), ref.view(np.uint64))
        assert res[k]["rel_resid"] <= 1e-10


def test_batch_fgmres_uneven_matches_single():
    """FGMRES-DR (MR preconditioner, deflation) over problems on different lattices, gauge fields and
    masses: each equals its single solve bit for bit."""
    say("batch fgmres uneven")
    probs = [make_problem(L, s, k, cmx.SINGLE) for L, s, k in UNEVEN]
    kw = dict(solver=cmx.FGMRESDR, rsd=1e-10, maxit=3000, sloppy_prec=cmx.SINGLE, n_krylov=12, n_defl=3,
              precond=cmx.PRECOND_MR, precond_iters=3)
    single = [op.solve(chi, **kw) for op, chi in probs]
    psis, res, summ = cmx.solve_batch([p[0] for p in probs], [p[1] for p in probs], **kw)
    say("  summary", summ)
    for k in range(len(probs)):
        say("  problem", k, "single", single[k][1]["n_count"], "batch", res[k]["n_count"], res[k]["rel_resid"])
        assert res[k]["n_count"] == single[k][1]["n_count"]
        assert np.array_equal(psis[k].view(np.uint64), single[k][0].view(np.uint64))
        assert res[k]["rel_resid"] <= 1e-10


def test_batch_rejects_bad_input():
    op, chi = make_problem((4, 4, 4, 4), 1020, 0.12)
    with pytest.raises(cmx.CmxError):
        cmx.solve_batch([op], [chi], solver=cmx.BICGSTAB)
