"""The oracle's flexible solver (FGMRES-DR) and MR preconditioner, pinned against dense linear algebra:
the exact solution, GMRES's minimal-residual property over the Krylov space, harmonic Ritz values
from their defining generalised eigenproblem, and MR's one-step optimality."""
import numpy as np
import pytest

from oracle import krylov
from tests.test_oracle import dense_dslash
from tests.util import numpy_gauge, random_fermion, rel


def wilson_system(seed=40, L=(2, 2, 2, 4), mass=0.2):
    V = int(np.prod(L))
    U = numpy_gauge(L, seed)
    M = (4 + mass) * np.eye(12 * V) - 0.5 * dense_dslash(L, U, +1)
    A = lambda x: (M @ x.reshape(-1)).reshape(x.shape)
    return V, M, A


def test_fgmres_exact_solution_no_precond():
    V, M, A = wilson_system()
    chi = random_fermion(V, 41)
    psi, n, r, cycles = krylov.fgmres_dr(A, chi, np.zeros_like(chi), 1e-11, n_krylov=10, max_iter=2000)
    exact = np.linalg.solve(M, chi.reshape(-1)).reshape(chi.shape)
    assert rel(psi, exact) < 1e-9
    assert r <= 1e-11 * np.linalg.norm(chi)
    assert cycles > 1  # restarted


def test_first_cycle_is_minimal_residual():
    # one cycle of m steps from psi0 minimises |chi - A(psi0 + y)| over y in K_m(A, r0)
    V, M, A = wilson_system(seed=42)
    chi = random_fermion(V, 43)
    psi0 = 0.1 * random_fermion(V, 44)
    m = 6
    psi, n, r, cycles = krylov.fgmres_dr(A, chi, psi0, 1e-30, n_krylov=m, max_iter=m)
    assert n == m and cycles == 1
    r0 = (chi - A(psi0)).reshape(-1)
    K = [r0]
    for _ in range(m - 1):
        K.append(M @ K[-1])
    K = np.stack(K, axis=1)
    y, *_ = np.linalg.lstsq(M @ K, r0, rcond=None)
    best = np.linalg.norm(r0 - M @ K @ y)
    assert abs(r - best) <= 1e-9 * np.linalg.norm(r0)


def test_harmonic_ritz_values_of_the_deflated_restart():
    V, M, A = wilson_system(seed=45)
    chi = random_fermion(V, 46)
    rec = {}
    m, k = 8, 3
    krylov.fgmres_dr(A, chi, np.zeros_like(chi), 1e-10, n_krylov=m, n_defl=k, max_iter=m + 2, record=rec)
    cyc = rec["cycles"][0]
    Vm = np.stack([v.reshape(-1) for v in cyc["V"][:m]], axis=1)
    AV = M @ Vm
    # harmonic Ritz pairs: (AV)^H (AV g - theta V g) = 0
    lhs = AV.conj().T @ AV
    rhs = AV.conj().T @ Vm
    theta = np.linalg.eigvals(np.linalg.solve(rhs, lhs))
    theta = theta[np.argsort(np.abs(theta))][:k]
    got = rec["harmonic_ritz"][0]
    assert np.allclose(np.sort_complex(got), np.sort_complex(theta), rtol=1e-8, atol=1e-10)


@pytest.mark.parametrize("n_defl", [0, 2, 4])
def test_fgmres_dr_with_mr_preconditioner(n_defl):
    V, M, A = wilson_system(seed=47, mass=0.05)
    chi = random_fermion(V, 48)
    prec = lambda z0, v: krylov.mr(A, v, z0, 1e-12, 3)[0]
    psi, n, r, cycles = krylov.fgmres_dr(A, chi, np.zeros_like(chi), 1e-10, n_krylov=8, n_defl=n_defl,
                                         max_iter=3000, precond=prec)
    exact = np.linalg.solve(M, chi.reshape(-1)).reshape(chi.shape)
    assert rel(psi, exact) < 1e-8


def test_deflation_reduces_iterations():
    # near-critical mass (smallest |eigenvalue| of M ~ 0.2): the kept harmonic Ritz vectors pay off
    V, M, A = wilson_system(seed=49, mass=-1.8, L=(2, 2, 4, 4))
    chi = random_fermion(V, 50)
    n0 = krylov.fgmres_dr(A, chi, np.zeros_like(chi), 1e-10, n_krylov=8, n_defl=0, max_iter=20000)[1]
    n4 = krylov.fgmres_dr(A, chi, np.zeros_like(chi), 1e-10, n_krylov=8, n_defl=4, max_iter=20000)[1]
    assert n4 < n0


def test_deflation_after_short_cycles():
    """Cycles that end with fewer Arnoldi vectors than n_defl (cycle_reduction stops each cycle after a
    factor 2): the deflated restart keeps only as many harmonic Ritz vectors as the cycle has, and the
    solve still reaches the dense solution."""
    V, M, A = wilson_system(seed=53, mass=0.05)
    chi = random_fermion(V, 54)
    rec = {}
    psi, n, r, cycles = krylov.fgmres_dr(A, chi, np.zeros_like(chi), 1e-10, n_krylov=12, n_defl=8,
                                         max_iter=5000, record=rec, cycle_reduction=0.5)
    dims = [c["dim"] - c["nd"] for c in rec["cycles"]]
    assert min(dims) < 8  # some cycles are shorter than n_defl
    assert max(c["nd"] for c in rec["cycles"]) > 0  # and deflation does take place
    for a, b in zip(rec["cycles"], rec["cycles"][1:]):
        assert b["nd"] <= a["dim"]
    exact = np.linalg.solve(M, chi.reshape(-1)).reshape(chi.shape)
    assert rel(psi, exact) < 1e-8
    assert r <= 1e-10 * np.linalg.norm(chi)


def test_mr_step_is_optimal():
    # one MR step minimises |r - a A r| over complex a
    V, M, A = wilson_system(seed=51)
    chi = random_fermion(V, 52)
    psi1, k = krylov.mr(A, chi, np.zeros_like(chi), 1e-30, 1)
    r1 = np.linalg.norm(chi - A(psi1))
    r = chi.reshape(-1)
    Ar = M @ r
    a = np.vdot(Ar, r) / np.vdot(Ar, Ar)
    assert abs(r1 - np.linalg.norm(r - a * Ar)) < 1e-12 * np.linalg.norm(r)
    for da in (1e-3, 1e-3j):
        assert np.linalg.norm(r - (a + da) * Ar) > r1
