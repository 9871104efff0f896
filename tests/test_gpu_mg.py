"""GPU: the two-level multigrid preconditioner (cmx_mg_*).

* The null vectors are orthonormal on every (block, chirality) aggregate.
* Restriction and prolongation equal R = P^dag and P computed in numpy from the downloaded null vectors.
* The coarse operator built by probing equals the Galerkin product: A_c x = R M P x for random coarse x,
  with M the oracle's FP64 unpreconditioned clover operator (unprecLinOp, eoprec_linop.h:147-165) and
  R, P from numpy -- an independent composition that a wrong colouring, stencil entry or scatter fails.
  Covered: coarse extents 2 (the +mu and -mu entries on one block) and 4 (face-restricted probes, 2^4
  colours), and odd extents 3 (whole-block probes with 3^3 x 4 colours).
* FGMRES-DR with the V-cycle reaches the FP64 true residual 1e-10 in fewer iterations than without it;
  cmx_mg_update changes the cycle without a new set-up, and restoring the set-up's cycle restores the
  solution bit for bit; the FP16 V-cycle (FP16 copies of the null vectors) does the same.
"""
import sys

import numpy as np
import pytest

import oracle
from chroma_amd import lib as cmx
from tests.util import rel

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def host_blocks(L, B):
    """block index (x fastest over the coarse lattice) of every site in the host order"""
    V = int(np.prod(L))
    lex = np.arange(V)
    x = lex % L[0]
    y = (lex // L[0]) % L[1]
    z = (lex // (L[0] * L[1])) % L[2]
    t = lex // (L[0] * L[1] * L[2])
    par = (x + y + z + t) % 2
    host = par * (V // 2) + lex // 2
    Lc = [L[m] // B[m] for m in range(4)]
    blk = x // B[0] + Lc[0] * (y // B[1] + Lc[1] * (z // B[2] + Lc[2] * (t // B[3])))
    out = np.empty(V, np.int64)
    out[host] = blk
    return out, int(np.prod(Lc))


def make(L, seed, kappa, stout=False):
    if stout:
        U = cmx.gauge_generate(L, kind=2, seed=seed, nstout=6, rho=0.1)
    else:
        U = cmx.gauge_generate(L, kind=1, seed=seed, eps=0.3)
    U = cmx.gauge_apply_bc(L, U, (1, 1, 1, -1))
    mass = cmx.kappa_to_mass(kappa)
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_CLOVER, mass, 1.0, sloppy_prec=cmx.SINGLE)
    return U, mass, op


def numpy_restrict(P, blk, nb, v):
    nvec = P.shape[0]
    out = np.zeros((nb, 2 * nvec), complex)
    for c in range(2):
        sl = slice(2 * c, 2 * c + 2)
        for k in range(nvec):
            prod = np.sum(np.conj(P[k][:, sl, :]) * v[:, sl, :], axis=(1, 2))
            out[:, c * nvec + k] = np.bincount(blk, weights=prod.real, minlength=nb) + \
                1j * np.bincount(blk, weights=prod.imag, minlength=nb)
    return out


def numpy_prolong(P, blk, e):
    nvec = P.shape[0]
    out = np.zeros(P.shape[1:], complex)
    for c in range(2):
        sl = slice(2 * c, 2 * c + 2)
        for k in range(nvec):
            out[:, sl, :] += P[k][:, sl, :] * e[blk, c * nvec + k][:, None, None]
    return out


@pytest.mark.parametrize("L,B", [((8, 8, 8, 16), (2, 2, 2, 2)), ((8, 8, 8, 16), (4, 4, 4, 4)),
                                 ((6, 6, 6, 8), (2, 2, 2, 2))])
def test_mg_prolongator_and_coarse_operator(L, B):
    say("mg setup", L, B)
    U, mass, op = make(L, 3001, 0.128)
    nvec = 4
    info = op.mg_setup(blocking=B, nvec=nvec, setup_maxit=60, seed=5)
    say("  setup", info)
    blk, nb = host_blocks(L, B)
    assert info["n_blocks"] == nb and info["coarse_dof"] == 2 * nvec
    P = op.mg_nullvecs()
    # orthonormal per (block, chirality)
    worst = 0.0
    for c in range(2):
        sl = slice(2 * c, 2 * c + 2)
        for b in range(nb):
            sel = blk == b
            M = P[:, sel][:, :, sl, :].reshape(nvec, -1)
            G = np.conj(M) @ M.T
            worst = max(worst, float(np.max(np.abs(G - np.eye(nvec)))))
    say("  orthonormality error", worst)
    assert worst < 1e-5
    rng = np.random.default_rng(7)
    V = int(np.prod(L))
    v = rng.standard_normal((V, 4, 3)) + 1j * rng.standard_normal((V, 4, 3))
    r_dev = op.mg_restrict(v)
    r_np = numpy_restrict(P, blk, nb, v)
    say("  restrict", rel(r_dev, r_np))
    assert rel(r_dev, r_np) < 1e-5
    e = rng.standard_normal((nb, 2 * nvec)) + 1j * rng.standard_normal((nb, 2 * nvec))
    p_dev = op.mg_prolong(e)
    p_np = numpy_prolong(P, blk, e)
    say("  prolong", rel(p_dev, p_np))
    assert rel(p_dev, p_np) < 1e-5
    # Galerkin: A_c e = R M P e with the oracle's FP64 operator
    ref = oracle.Op(oracle.EO_CLOVER, L, U, mass, 1.0, 1.0)
    Me = ref.piece("unprec", p_np)
    want = numpy_restrict(P, blk, nb, Me)
    got = op.mg_coarse_apply(e)
    say("  coarse operator vs R M P", rel(got, want))
    assert rel(got, want) < 1e-4


def test_mg_half_precision_relaxation():
    """Null vectors relaxed in FP16 storage (Precision HALF): the aggregates stay orthonormal, the coarse
    operator is still R M P of the downloaded vectors, and the V-cycle preconditions FGMRES-DR to 1e-10,
    with the MR smoother and with the learned network as the post-smoother."""
    L, B = (8, 8, 8, 16), (4, 4, 4, 4)
    U, mass, op = make(L, 3003, 0.128)
    nvec = 4
    info = op.mg_setup(blocking=B, nvec=nvec, setup_maxit=60, seed=5, prec=cmx.HALF, nu_pre=0, nu_post=4,
                       coarse_tol=0.1, coarse_maxit=8)
    say("  half-precision setup", info)
    blk, nb = host_blocks(L, B)
    P = op.mg_nullvecs()
    for c in range(2):
        sl = slice(2 * c, 2 * c + 2)
        for b in range(nb):
            M = P[:, blk == b][:, :, sl, :].reshape(nvec, -1)
            assert np.max(np.abs(np.conj(M) @ M.T - np.eye(nvec))) < 1e-5
    rng = np.random.default_rng(11)
    e = rng.standard_normal((nb, 2 * nvec)) + 1j * rng.standard_normal((nb, 2 * nvec))
    ref = oracle.Op(oracle.EO_CLOVER, L, U, mass, 1.0, 1.0)
    want = numpy_restrict(P, blk, nb, ref.piece("unprec", numpy_prolong(P, blk, e)))
    assert rel(op.mg_coarse_apply(e), want) < 1e-4
    Vh = int(np.prod(L)) // 2
    chi = cmx.z2_source((Vh, 4, 3), 4)
    kw = dict(solver=cmx.FGMRESDR, rsd=1e-10, maxit=2000, sloppy_prec=cmx.HALF, n_krylov=8)
    _, r = op.solve(chi, precond=cmx.PRECOND_MG, **kw)
    say("  iterations", r["n_count"], "resid", r["rel_resid"])
    assert r["rel_resid"] <= 1e-10
    # the learned network as the post-smoother (its Neumann start), alone and followed by one MR step
    net = cmx.Net()
    net.set_depth(2)
    net.init_neumann(4.0 + mass)
    cmx.net_attach(op, net)
    for nu in (0, 1):
        op.mg_update(smoother=1, nu_post=nu)
        _, rs = op.solve(chi, precond=cmx.PRECOND_MG, **kw)
        say("  network smoother, nu_post", nu, "iterations", rs["n_count"], "resid", rs["rel_resid"])
        assert rs["rel_resid"] <= 1e-10
    with pytest.raises(cmx.CmxError):
        op.mg_update(smoother=2)


def test_fgmres_mg_preconditioner():
    L = (16, 16, 16, 32)
    say("fgmres + mg", L)
    U, mass, op = make(L, 3002, 0.1575, stout=True)
    Vh = int(np.prod(L)) // 2
    chi = cmx.z2_source((Vh, 4, 3), 9)
    kw = dict(solver=cmx.FGMRESDR, rsd=1e-10, maxit=4000, sloppy_prec=cmx.SINGLE, n_krylov=16)
    psi0, r_none = op.solve(chi, precond=cmx.PRECOND_NONE, **kw)
    info = op.mg_setup(blocking=(4, 4, 4, 4), nvec=8, nu_pre=0, nu_post=4, setup_maxit=300, coarse_tol=0.1,
                       coarse_maxit=12, seed=3)
    say("  setup", info)
    psi, r_mg = op.solve(chi, precond=cmx.PRECOND_MG, **kw)
    say("  iterations: none", r_none["n_count"], r_none["solve_seconds"], "mg", r_mg["n_count"],
        r_mg["solve_seconds"], "resid", r_mg["rel_resid"])
    assert r_mg["rel_resid"] <= 1e-10
    assert rel(psi, psi0) < 1e-8
    assert r_mg["n_count"] < r_none["n_count"] / 2
    # cmx_mg_update: a lighter V-cycle on the same null space, then back to the set-up's cycle (same
    # solution bit for bit: the V-cycle is deterministic)
    op.mg_update(nu_post=2, coarse_maxit=6, coarse_nkrylov=20)
    psi2, r2 = op.solve(chi, precond=cmx.PRECOND_MG, **kw)
    say("  updated cycle: iterations", r2["n_count"], "resid", r2["rel_resid"])
    assert r2["rel_resid"] <= 1e-10 and rel(psi2, psi0) < 1e-8
    op.mg_update(nu_post=4, coarse_maxit=12, coarse_nkrylov=12)
    psi3, r3 = op.solve(chi, precond=cmx.PRECOND_MG, **kw)
    assert r3["n_count"] == r_mg["n_count"] and np.array_equal(psi3, psi)
    with pytest.raises(cmx.CmxError):
        op.mg_update(nvec=4)
    # the FP16 V-cycle (restriction and prolongation with the FP16 copies of the null vectors)
    kw16 = dict(kw, sloppy_prec=cmx.HALF)
    _, r16n = op.solve(chi, precond=cmx.PRECOND_NONE, **kw16)
    psi16, r16 = op.solve(chi, precond=cmx.PRECOND_MG, **kw16)
    say("  FP16: iterations none", r16n["n_count"], "mg", r16["n_count"], "resid", r16["rel_resid"])
    assert r16["rel_resid"] <= 1e-10 and rel(psi16, psi0) < 1e-8
    assert r16["n_count"] < r16n["n_count"] / 2
