"""The oracle's learned-preconditioner network, pinned independently of its own formula:
the Neumann-series weights reproduce 1/a + (1/(4a^3)) D_oe D_eo built from the oracle's hopping term
(which fixes the signs of the half-spinor projectors); the gradient agrees with central finite
differences of the loss; the network is gauge covariant; the projectors are the rank-2 spin projectors
of the hopping directions, and the output depends on Wc only through Pi_q Wc Pi_p."""
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
