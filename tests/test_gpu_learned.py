"""GPU parity of the learned preconditioner (through the C-ABI) against the oracle's numpy restatement,
its training, and its use inside FGMRES-DR.

Tolerances: the network stores transported fields and the contracted weights in BF16 (8-bit mantissa)
with FP32 accumulation, so its output is compared at 2e-2 relative (Frobenius norm) and its gradient
at 5e-2; the solves that use it are judged by the FP64 true residual (1e-10), as every solve is.
"""
import sys

import numpy as np
import pytest

from chroma_amd import lib as cmx
from oracle import learned
from tests.util import rel, random_fermion

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def random_weights(seed, scale=0.1):
    rng = np.random.default_rng(seed)
    w1 = (rng.standard_normal((64, 32)) + 1j * rng.standard_normal((64, 32))) * scale
    w2 = (rng.standard_normal((32, 64)) + 1j * rng.standard_normal((32, 64))) * scale
    w0 = (rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))) * scale
    # the library holds FP32 weights
    return w1.astype(np.complex64).astype(complex), w2.astype(np.complex64).astype(complex), \
        w0.astype(np.complex64).astype(complex)


def make_op(L, seed, prec, stout=False):
    if stout:
        U = cmx.gauge_generate(L, kind=2, seed=seed, nstout=6, rho=0.1)
    else:
        U = cmx.gauge_generate(L, kind=1, seed=seed, eps=0.3)
    U = cmx.gauge_apply_bc(L, U, (1, 1, 1, -1))
    sp = cmx.HALF if prec == cmx.HALF else cmx.SINGLE
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_CLOVER, cmx.kappa_to_mass(0.135), 1.0, sloppy_prec=sp)
    return U, op


@pytest.mark.parametrize("L", [(8, 8, 8, 16), (4, 2, 2, 6)])
@pytest.mark.parametrize("prec", [cmx.DOUBLE, cmx.SINGLE, cmx.HALF])
def test_net_forward(L, prec):
    say("net forward", L, prec)
    U, op = make_op(L, 70, prec)
    Vh = int(np.prod(L)) // 2
    w1, w2, w0 = random_weights(71)
    net = cmx.Net()
    net.set_weights(w1, w2, w0)
    cmx.net_attach(op, net)
    v = random_fermion(Vh, 72)
    want = learned.ptc_apply(L, U, w1, w2, w0, v)
    for variant in (0, 1, 2, 3, 4):  # fused (registers), fused (LDS), separate kernels; 3, 4 streamed links
        cmx.net_set_variant(op, prec, variant)
        got = cmx.net_apply(op, v, prec)
        say("  variant", variant, "rel", rel(got, want))
        assert rel(got, want) < 2e-2


def test_net_forward_neumann():
    """Neumann weights through the device: 1/a + (1/(4a^3)) D_oe D_eo against the oracle network."""
    L = (8, 8, 8, 16)
    U, op = make_op(L, 73, cmx.SINGLE)
    Vh = int(np.prod(L)) // 2
    a = 4.0 + cmx.kappa_to_mass(0.135)
    net = cmx.Net()
    net.init_neumann(a)
    cmx.net_attach(op, net)
    w1, w2, w0 = net.get_weights()
    v = random_fermion(Vh, 74)
    assert rel(cmx.net_apply(op, v, cmx.SINGLE), learned.ptc_apply(L, U, w1, w2, w0, v)) < 1e-2


@pytest.mark.parametrize("prec", [cmx.SINGLE, cmx.HALF])
def test_net_gradient(prec):
    L = (8, 8, 8, 16)
    say("net gradient", prec)
    U, op = make_op(L, 75, prec)
    Vh = int(np.prod(L)) // 2
    w1, w2, w0 = random_weights(76)
    net = cmx.Net()
    net.set_weights(w1, w2, w0)
    cmx.net_attach(op, net)
    v = random_fermion(Vh, 77)
    e = random_fermion(Vh, 78)
    loss, S, S0 = cmx.net_grad(op, v, e, prec)
    l_ref, S_ref, S0_ref = learned.ptc_grad(L, U, w1, w2, w0, v, e)
    say("  loss", loss, l_ref, "S rel", rel(S, S_ref), "S0 rel", rel(S0, S0_ref))
    assert abs(loss - l_ref) < 2e-2 * l_ref
    assert rel(S, S_ref) < 5e-2
    assert rel(S0, S0_ref) < 2e-2


@pytest.mark.parametrize("prec", [cmx.SINGLE, cmx.HALF])
def test_train_and_fgmres_learned(prec):
    """Training from the Neumann start lowers its loss; FGMRES-DR with the network reaches a 1e-10 FP64
    true residual; at FP32 (where the cycles are not limited by the operator's precision) the network
    needs fewer iterations than no preconditioner."""
    L = (8, 8, 8, 16)
    say("train + fgmres", prec)
    U, op = make_op(L, 79, prec, stout=True)
    Vh = int(np.prod(L)) // 2
    a = 4.0 + cmx.kappa_to_mass(0.135)
    chi = cmx.z2_source((Vh, 4, 3), 8)
    kw = dict(solver=cmx.FGMRESDR, rsd=1e-10, maxit=3000, sloppy_prec=prec, n_krylov=16,
              basis_prec=cmx.SINGLE, cycle_reduction=1e-3 if prec == cmx.HALF else 0.0)
    _, r_none = op.solve(chi, precond=cmx.PRECOND_NONE, **kw)
    net = cmx.Net()
    net.init_neumann(a)
    cmx.net_attach(op, net)
    _, r_neu = op.solve(chi, precond=cmx.PRECOND_LEARNED, **kw)
    tr = cmx.net_train(op, n_sources=8, snapshot_every=1, pairs_per_source=4, normalize=False, epochs=4,
                       lr=1e-3, seed=5, prec=prec)
    say("  train", tr)
    assert tr["n_pairs"] > 0 and tr["loss_last"] < tr["loss_first"]
    psi, r_tr = op.solve(chi, precond=cmx.PRECOND_LEARNED, **kw)
    say("  iterations none", r_none["n_count"], "neumann", r_neu["n_count"], "trained", r_tr["n_count"])
    for r in (r_none, r_neu, r_tr):
        assert r["rel_resid"] <= 1e-10
    if prec == cmx.SINGLE:
        assert r_neu["n_count"] < r_none["n_count"]
        assert r_tr["n_count"] < r_none["n_count"]
    psi_cg, _ = op.solve(chi, solver=cmx.CG, rsd=1e-12, maxit=20000)
    assert rel(psi, psi_cg) < 1e-8


@pytest.mark.parametrize("prec", [cmx.SINGLE, cmx.HALF])
def test_deep_net_forward_and_gradient(prec):
    """A chain of two blocks with random weights: the forward application against the oracle's chain, and
    the back-propagated gradient of every block's weights (through the adjoint gather, the adjoint
    mixing and the adjoint scatter kernels) against the oracle's back-propagation."""
    L = (8, 8, 8, 16)
    say("deep net", prec)
    U, op = make_op(L, 81, prec)
    Vh = int(np.prod(L)) // 2
    blocks = [random_weights(82), random_weights(83)]
    net = cmx.Net()
    net.set_depth(2)
    for k, (w1, w2, w0) in enumerate(blocks):
        net.set_block(k, w1, w2, w0)
    cmx.net_attach(op, net)
    v = random_fermion(Vh, 84)
    e = random_fermion(Vh, 85)
    got = cmx.net_apply(op, v, prec)
    want = learned.ptc_apply_deep(L, U, blocks, v)
    say("  forward rel", rel(got, want))
    assert rel(got, want) < 2e-2
    loss, S, S0 = cmx.net_grad(op, v, e, prec)
    l_ref, grads = learned.ptc_grad_deep(L, U, blocks, v, e)
    for k in range(2):
        say("  block", k, "S rel", rel(S[k], grads[k][0]), "S0 rel", rel(S0[k], grads[k][1]))
        assert rel(S[k], grads[k][0]) < 5e-2
        assert rel(S0[k], grads[k][1]) < 2e-2
    assert abs(loss - l_ref) < 2e-2 * l_ref
    # the identity blocks added by set_depth leave a trained network's output unchanged
    net1 = cmx.Net()
    net1.set_weights(*blocks[0])
    cmx.net_attach(op, net1)
    one = cmx.net_apply(op, v, prec)
    net1.set_depth(3)
    assert net1.depth() == 3
    assert rel(cmx.net_apply(op, v, prec), one) < 1e-2


def test_training_gains_over_neumann_start_near_critical():
    """Near the critical kappa of the stout ensemble (kappa 0.158; plain mixed CG needs 1700 iterations at
    32^3x64 and 2200 at 48^3x96): a two-block network trained on two 16^3x32 fields of the recipe needs
    fewer FGMRES-DR iterations on a held-out 24^3x48 field than its untrained start (the Neumann block
    followed by an identity block, the same operator as the one-block Neumann network), and fewer than
    FGMRES-DR without a preconditioner; every solve reaches the FP64 true residual 1e-10."""
    kappa = 0.158
    mass = cmx.kappa_to_mass(kappa)
    say("training gain near critical", kappa)

    def stout(L, seed):
        U = cmx.gauge_stout_smear(L, cmx.gauge_generate(L, kind=2, seed=seed), 6, 0.1)
        return cmx.gauge_apply_bc(L, U, (1, 1, 1, -1))

    L = (24, 24, 24, 48)
    op = cmx.Operator(cmx.Gauge(L, stout(L, 1)), cmx.EO_CLOVER, mass, 1.0, sloppy_prec=cmx.HALF)
    chi = cmx.z2_source((int(np.prod(L)) // 2, 4, 3), 2)
    kw = dict(solver=cmx.FGMRESDR, rsd=1e-10, maxit=5000, sloppy_prec=cmx.HALF, basis_prec=cmx.HALF, n_krylov=16,
              cycle_reduction=3e-3)
    _, r_none = op.solve(chi, precond=cmx.PRECOND_NONE, **kw)
    net = cmx.Net()
    net.set_depth(2)
    net.init_neumann(4.0 + mass)
    cmx.net_attach(op, net)
    _, r_start = op.solve(chi, precond=cmx.PRECOND_LEARNED, **kw)
    TL = (16, 16, 16, 32)
    for k in range(2):
        top = cmx.Operator(cmx.Gauge(TL, stout(TL, 10 + k)), cmx.EO_CLOVER, mass, 1.0, sloppy_prec=cmx.HALF)
        cmx.net_attach(top, net)
        tr = cmx.net_train(top, n_sources=8, snapshot_every=1, pairs_per_source=16, normalize=False, epochs=4,
                           lr=1e-3, seed=100 + k, prec=cmx.HALF, max_pairs=100000)
        say("  train field", k, tr)
    cmx.net_attach(op, net)
    psi, r_tr = op.solve(chi, precond=cmx.PRECOND_LEARNED, **kw)
    say("  iterations: none", r_none["n_count"], "Neumann start", r_start["n_count"], "trained", r_tr["n_count"])
    for r in (r_none, r_start, r_tr):
        assert r["rel_resid"] <= 1e-10
    assert r_tr["n_count"] < r_start["n_count"] < r_none["n_count"]
