"""The oracle's clover force terms (CloverTermBase::deriv / derivTrLn, clover_term_base_w.h:320-870) and the
even-odd operator's derivative (EvenOddPrecLogDetLinearOperator::deriv, eoprec_logdet_linop.h:114-157),
pinned by central finite differences of the functions they differentiate.

Convention (the one WilsonDslashBase::deriv follows, lwldslash_base_w.h:132-227): ds is the derivative with
respect to the occurrences of U (not of U^dag), so for a link moved as U_mu(y) -> exp(eps X) U_mu(y), X in
su(3),
    d/d eps <chi, M(U) psi> = tr(X U_mu(y) ds(chi, psi, isign)_mu(y)) + conj tr(X U_mu(y) ds(psi, chi, -isign)_mu(y)),
the second term carrying the U^dag occurrences; for the real log det of the clover term it is 2 Re of the
first.  The links tried cover all four directions, both checkerboards and the time boundary, so every one
of the 16 insertions of deriv_loops enters some derivative."""
import numpy as np
import pytest
from scipy.linalg import expm

import oracle
from tests.util import apply_simple_bc, numpy_gauge, random_fermion

L = (4, 4, 4, 4)
V = int(np.prod(L))
Vh = V // 2
MASS = 0.1
LINKS = [(0, 0), (1, 5), (2, Vh + 3), (3, Vh + 17), (3, 2), (0, Vh + 40), (1, V - 1), (2, 77)]


def su3_direction(seed):
    r = np.random.default_rng(seed)
    A = r.standard_normal((3, 3)) + 1j * r.standard_normal((3, 3))
    X = (A - A.conj().T) / 2
    return X - np.trace(X) / 3 * np.eye(3)


def moved(U, mu, y, X, eps):
    Ue = U.copy()
    Ue[mu, y] = expm(eps * X) @ U[mu, y]
    return Ue


def field():
    return apply_simple_bc(L, numpy_gauge(L, 90), (1, 1, 1, -1))


def fd(f, h=1e-5):
    return (f(h) - f(-h)) / (2 * h)


@pytest.mark.parametrize("cb", [0, 1])
def test_clover_deriv_finite_differences(cb):
    U = field()
    chi = random_fermion(V, 91)
    psi = random_fermion(V, 92)
    csw_r, csw_t = 1.0, 0.7  # distinct coefficients, so the time planes are told apart
    ds1 = oracle.clover_deriv(L, U, chi, psi, csw_r, csw_t, cb)
    ds2 = oracle.clover_deriv(L, U, psi, chi, csw_r, csw_t, cb)
    sl = slice(cb * Vh, (cb + 1) * Vh)

    def S(Ue):
        tri = oracle.make_clover(L, Ue, csw_r, csw_t, MASS)
        return np.vdot(chi[sl], oracle.clover_apply(L, tri, psi, cb)[sl])

    for k, (mu, y) in enumerate(LINKS):
        X = su3_direction(100 + k)
        want = fd(lambda e: S(moved(U, mu, y, X, e)))
        got = np.trace(X @ U[mu, y] @ ds1[mu, y]) + np.conj(np.trace(X @ U[mu, y] @ ds2[mu, y]))
        assert abs(got - want) < 1e-7 * max(1.0, abs(want)), (mu, y, got, want)
    # deriv on both checkerboards is the sum (CloverTermBase::deriv, :144-160)
    both = oracle.clover_deriv(L, U, chi, psi, csw_r, csw_t, -1)
    assert np.allclose(both, oracle.clover_deriv(L, U, chi, psi, csw_r, csw_t, 0) +
                       oracle.clover_deriv(L, U, chi, psi, csw_r, csw_t, 1), atol=1e-14, rtol=0)


@pytest.mark.parametrize("cb", [0, 1])
def test_clover_deriv_trln_finite_differences(cb):
    """derivTrLn: d log|det A_cb| = 2 Re tr(X U ds) for the clover term inverted on cb."""
    U = field()
    csw_r, csw_t = 1.0, 0.7
    ds = oracle.clover_deriv_trln(L, U, MASS, csw_r, csw_t, cb)

    def S(Ue):
        tri = oracle.make_clover(L, Ue, csw_r, csw_t, MASS)
        logdet, _ = oracle.clover_invert(L, tri, cb)
        return logdet

    for k, (mu, y) in enumerate(LINKS):
        X = su3_direction(200 + k)
        want = fd(lambda e: S(moved(U, mu, y, X, e)))
        got = 2 * np.trace(X @ U[mu, y] @ ds[mu, y]).real
        assert abs(got - want) < 1e-7 * max(1.0, abs(want)), (mu, y, got, want)


@pytest.mark.parametrize("isign", [+1, -1])
def test_eo_clover_deriv_finite_differences(isign):
    U = field()
    chi = random_fermion(V, 93)
    psi = random_fermion(V, 94)
    chi[:Vh] = 0
    psi[:Vh] = 0
    ds1 = oracle.eo_clover_deriv(L, U, MASS, chi, psi, isign)
    ds2 = oracle.eo_clover_deriv(L, U, MASS, psi, chi, -isign)

    def S(Ue):
        return np.vdot(chi[Vh:], oracle.Op(oracle.EO_CLOVER, L, Ue, MASS, 1.0, 1.0)(psi, isign)[Vh:])

    for k, (mu, y) in enumerate(LINKS):
        X = su3_direction(300 + k)
        want = fd(lambda e: S(moved(U, mu, y, X, e)))
        got = np.trace(X @ U[mu, y] @ ds1[mu, y]) + np.conj(np.trace(X @ U[mu, y] @ ds2[mu, y]))
        assert abs(got - want) < 1e-7 * max(1.0, abs(want)), (mu, y, got, want)
