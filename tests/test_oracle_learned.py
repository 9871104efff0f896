"""The oracle's learned-preconditioner network, pinned independently of its own formula:
the Neumann-series weights reproduce 1/a + (1/(4a^3)) D_oe D_eo built from the oracle's hopping term
(which fixes the signs of the half-spinor projectors); the gradient agrees with central finite
differences of the loss; the network is gauge covariant; the projectors are the rank-2 spin projectors
of the hopping directions, and the output depends on Wc only through Pi_q Wc Pi_p.  For a chain of
blocks: each block's adjoint satisfies <a, B b> = <B^dag a, b>, the identity block passes its input
through, and back-propagation agrees with finite differences in every block's weights."""
import numpy as np

import oracle
from oracle import learned
from tests.test_oracle import GAMMA
from tests.util import numpy_gauge, random_fermion, random_su3_numpy, rel

L = (4, 4, 4, 4)
V = int(np.prod(L))
Vh = V // 2


def neumann(a):
    W1 = np.zeros((64, 32), complex)
    W2 = np.zeros((32, 64), complex)
    for p in range(8):
        mu, sg = p // 2, (-1.0 if p % 2 == 0 else 1.0)
        P = np.eye(4) + sg * GAMMA[mu]
        W1[p * 4:(p + 1) * 4, p * 4:(p + 1) * 4] = P
        for c in range(8):
            W2[p * 4:(p + 1) * 4, c * 4:(c + 1) * 4] = P / (4 * a ** 3)
    return W1, W2, np.eye(4) / a


def test_neumann_weights_are_two_hopping_terms():
    U = numpy_gauge(L, 60)
    a = 4.0 + 0.3
    W1, W2, W0 = neumann(a)
    v = random_fermion(Vh, 61)
    got = learned.ptc_apply(L, U, W1, W2, W0, v)
    full = np.zeros((V, 4, 3), complex)
    full[Vh:] = v
    de = oracle.dslash(L, U, full, +1, 0)           # D_eo v on the even sites
    de[Vh:] = 0
    dod = oracle.dslash(L, U, de, +1, 1)[Vh:]       # D_oe D_eo v
    want = v / a + dod / (4 * a ** 3)
    assert rel(got, want) < 1e-13


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(62)
    U = numpy_gauge(L, 63)
    Wc = (rng.standard_normal((32, 32)) + 1j * rng.standard_normal((32, 32))) * 0.05
    W1 = np.vstack([np.eye(32), np.zeros((32, 32))])
    W2 = np.hstack([Wc, np.zeros((32, 32))])
    W0 = (rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))) * 0.2
    v = random_fermion(Vh, 64)
    e = random_fermion(Vh, 65)
    loss, S, S0 = learned.ptc_grad(L, U, W1, W2, W0, v, e)
    h = 1e-6

    def loss_at(Wc_, W0_):
        out = learned.ptc_apply(L, U, W1, np.hstack([Wc_, np.zeros((32, 32))]), W0_, v)
        d = out - e
        return float(np.vdot(d, d).real)

    for (i, j) in [(0, 0), (5, 17), (31, 2), (12, 12)]:
        for dz, part in ((h, "re"), (1j * h, "im")):
            Wp, Wm = Wc.copy(), Wc.copy()
            Wp[i, j] += dz
            Wm[i, j] -= dz
            fd = (loss_at(Wp, W0) - loss_at(Wm, W0)) / (2 * h)
            an = 2 * (S[i, j].real if part == "re" else S[i, j].imag)
            assert abs(fd - an) < 1e-5 * max(1.0, abs(an)), (i, j, part, fd, an)
    for (i, j) in [(0, 0), (1, 3), (3, 2)]:
        for dz, part in ((h, "re"), (1j * h, "im")):
            Wp, Wm = W0.copy(), W0.copy()
            Wp[i, j] += dz
            Wm[i, j] -= dz
            fd = (loss_at(Wc, Wp) - loss_at(Wc, Wm)) / (2 * h)
            an = 2 * (S0[i, j].real if part == "re" else S0[i, j].imag)
            assert abs(fd - an) < 1e-5 * max(1.0, abs(an)), (i, j, part, fd, an)


def test_gauge_covariance():
    rng = np.random.default_rng(66)
    U = numpy_gauge(L, 67)
    W1 = (rng.standard_normal((64, 32)) + 1j * rng.standard_normal((64, 32))) * 0.1
    W2 = (rng.standard_normal((32, 64)) + 1j * rng.standard_normal((32, 64))) * 0.1
    W0 = np.eye(4) * 0.3
    v = random_fermion(Vh, 68)
    g = random_su3_numpy(V, 69)
    fwd, _ = learned.neighbour_tables(L)
    Ug = np.stack([np.einsum("kab,kbc,kdc->kad", g, U[mu], g[fwd[mu]].conj()) for mu in range(4)])
    vg = np.einsum("kab,ksb->ksa", g[Vh:], v)
    out = learned.ptc_apply(L, U, W1, W2, W0, v)
    outg = learned.ptc_apply(L, Ug, W1, W2, W0, vg)
    assert rel(outg, np.einsum("kab,ksb->ksa", g[Vh:], out)) < 1e-12


def test_projectors_and_projected_dependence():
    for p in range(8):
        P = learned.projector(p)
        assert np.allclose(P @ P, P) and np.allclose(P, P.conj().T)
        assert abs(np.trace(P).real - 2) < 1e-14
        mu, sg = p // 2, (-1.0 if p % 2 == 0 else 1.0)
        assert np.allclose(2 * P, np.eye(4) + sg * GAMMA[mu])  # the tests' own gammas
    rng = np.random.default_rng(70)
    U = numpy_gauge(L, 71)
    Wc = rng.standard_normal((32, 32)) + 1j * rng.standard_normal((32, 32))
    Pblk = np.zeros((32, 32), complex)
    for p in range(8):
        Pblk[4 * p:4 * p + 4, 4 * p:4 * p + 4] = learned.projector(p)
    W1 = np.vstack([np.eye(32), np.zeros((32, 32))])
    W0 = np.eye(4) * 0.2
    v = random_fermion(Vh, 72)
    a = learned.ptc_apply(L, U, W1, np.hstack([Wc, np.zeros((32, 32))]), W0, v)
    b = learned.ptc_apply(L, U, W1, np.hstack([Pblk @ Wc @ Pblk, np.zeros((32, 32))]), W0, v)
    assert rel(a, b) < 1e-13
    # and the complement does nothing
    c = learned.ptc_apply(L, U, W1, np.hstack([Wc - Pblk @ Wc @ Pblk, np.zeros((32, 32))]), W0 * 0, v)
    assert np.abs(c).max() < 1e-12


def _random_block(rng, scale=0.05):
    Wc = (rng.standard_normal((32, 32)) + 1j * rng.standard_normal((32, 32))) * scale
    W1 = np.vstack([np.eye(32), np.zeros((32, 32))])
    W2 = np.hstack([Wc, np.zeros((32, 32))])
    W0 = (rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))) * 0.3
    return W1, W2, W0


def test_block_adjoint_is_the_adjoint():
    """<a, B(b)> = <B^dag(a), b> for random weights, vectors and a random gauge field: pins the adjoint
    transports (opposite paths, Pi_p U^dag) and the adjoint mixing used in back-propagation."""
    rng = np.random.default_rng(70)
    U = numpy_gauge(L, 71)
    W1, W2, W0 = _random_block(rng, 0.2)
    a = random_fermion(Vh, 72)
    b = random_fermion(Vh, 73)
    lhs = np.vdot(a, learned.ptc_apply(L, U, W1, W2, W0, b))
    rhs = np.vdot(learned.ptc_adjoint(L, U, W1, W2, W0, a), b)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    # planted mistakes fail the identity: the mixing not transposed (Wc in place of Wc^dag), and W0 not
    # conjugated
    Wc = W2 @ W1
    W2bad = np.hstack([Wc.conj().T, np.zeros((32, 32))])
    bad = np.vdot(learned.ptc_adjoint(L, U, W1, W2bad, W0, a), b)
    assert abs(bad - lhs) > 1e-3 * abs(lhs)
    bad = np.vdot(learned.ptc_adjoint(L, U, W1, W2, W0.conj(), a), b)
    assert abs(bad - lhs) > 1e-3 * abs(lhs)


def test_identity_block_passes_through():
    rng = np.random.default_rng(74)
    U = numpy_gauge(L, 75)
    v = random_fermion(Vh, 76)
    W1, W2, W0 = learned.identity_block()
    assert rel(learned.ptc_apply(L, U, W1, W2, W0, v), v) < 1e-15
    blocks = [_random_block(rng), learned.identity_block()]
    assert rel(learned.ptc_apply_deep(L, U, blocks, v), learned.ptc_apply(L, U, *blocks[0], v)) < 1e-15


def test_deep_gradient_matches_finite_differences():
    """Back-propagation through two blocks against central finite differences of |P v - e|^2 in each
    block's Wc and W0."""
    rng = np.random.default_rng(77)
    U = numpy_gauge(L, 78)
    blocks = [_random_block(rng), _random_block(rng)]
    v = random_fermion(Vh, 79)
    e = random_fermion(Vh, 80)
    loss, grads = learned.ptc_grad_deep(L, U, blocks, v, e)
    h = 1e-6

    def loss_at(k, Wc_=None, W0_=None):
        bl = list(blocks)
        W1, W2, W0 = bl[k]
        if Wc_ is not None:
            W2 = np.hstack([Wc_, np.zeros((32, 32))])
        if W0_ is not None:
            W0 = W0_
        bl[k] = (W1, W2, W0)
        d = learned.ptc_apply_deep(L, U, bl, v) - e
        return float(np.vdot(d, d).real)

    for k in (0, 1):
        Wc = blocks[k][1][:, :32]
        S, S0 = grads[k]
        for (i, j) in [(0, 0), (5, 17), (31, 2)]:
            for dz, part in ((h, "re"), (1j * h, "im")):
                Wp, Wm = Wc.copy(), Wc.copy()
                Wp[i, j] += dz
                Wm[i, j] -= dz
                fd = (loss_at(k, Wc_=Wp) - loss_at(k, Wc_=Wm)) / (2 * h)
                an = 2 * (S[i, j].real if part == "re" else S[i, j].imag)
                assert abs(fd - an) < 1e-5 * max(1.0, abs(an)), (k, i, j, part, fd, an)
        W0 = blocks[k][2]
        for (i, j) in [(0, 0), (1, 3)]:
            for dz, part in ((h, "re"), (1j * h, "im")):
                Wp, Wm = W0.copy(), W0.copy()
                Wp[i, j] += dz
                Wm[i, j] -= dz
                fd = (loss_at(k, W0_=Wp) - loss_at(k, W0_=Wm)) / (2 * h)
                an = 2 * (S0[i, j].real if part == "re" else S0[i, j].imag)
                assert abs(fd - an) < 1e-5 * max(1.0, abs(an)), (k, i, j, part, fd, an)
    # a single block is the one-block gradient
    l1, S, S0 = learned.ptc_grad(L, U, *blocks[0], v, e)
    l2, g2 = learned.ptc_grad_deep(L, U, blocks[:1], v, e)
    assert abs(l1 - l2) < 1e-12 * l1 and rel(g2[0][0], S) < 1e-13 and rel(g2[0][1], S0) < 1e-13
