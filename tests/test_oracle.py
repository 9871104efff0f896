"""The oracle against what pins it (CPU only):

* the reference's RNG-free regression golden prec_wilson-twisted (prop_corr to the metric file's 1e-5,
  n_count 17 per solve);
* the free-field closed form of the Wilson operator (plane waves, periodic / antiperiodic / twisted);
* an independent dense construction of the Wilson-clover matrix from the DR gamma matrices (numpy);
* gamma5-hermiticity, gauge covariance, clover identities (unit gauge, A A^-1 = 1, log det);
* Schur-complement / EO-reconstruction identities and a dense solve for the solvers.

Planted-mistake checks at the end show that each pinning test fails for a plausible oracle error
(introduced here, on the CPU, in a numpy copy of the relevant formula).
"""
import json
import os

import numpy as np
import pytest

import oracle
from tests.util import numpy_gauge, random_fermion, rel

HERE = os.path.dirname(os.path.abspath(__file__))

# DeGrand-Rossi gammas (docs/notes/gamma_conventions.tex:74-79), written independently of oracle.cpp
s1 = np.array([[0, 1], [1, 0]], complex)
s2 = np.array([[0, -1j], [1j, 0]], complex)
s3 = np.array([[1, 0], [0, -1]], complex)
Z2 = np.zeros((2, 2), complex)
I2 = np.eye(2)
GAMMA = [np.block([[Z2, 1j * s1], [-1j * s1, Z2]]), np.block([[Z2, -1j * s2], [1j * s2, Z2]]),
         np.block([[Z2, 1j * s3], [-1j * s3, Z2]]), np.block([[Z2, I2], [I2, Z2]])]
G5 = GAMMA[0] @ GAMMA[1] @ GAMMA[2] @ GAMMA[3]


def coords(L):
    V = int(np.prod(L))
    return np.array([oracle.site_coords(L, s) for s in range(V)])


def index_of(L, x):
    cb = int(sum(x)) & 1
    lex = ((x[3] * L[2] + x[2]) * L[1] + x[1]) * L[0] + x[0]
    return cb * (int(np.prod(L)) // 2) + lex // 2


def dense_dslash(L, U, isign):
    """D(isign) as a dense (12V x 12V) matrix built from the formula of lwldslash_w.h:21-43."""
    V = int(np.prod(L))
    X = coords(L)
    D = np.zeros((V, 4, 3, V, 4, 3), complex)
    for s in range(V):
        for mu in range(4):
            xp = X[s].copy(); xp[mu] = (xp[mu] + 1) % L[mu]
            xm = X[s].copy(); xm[mu] = (xm[mu] - 1) % L[mu]
            sp, sm = index_of(L, xp), index_of(L, xm)
            D[s, :, :, sp, :, :] += np.einsum("ab,ij->aibj", np.eye(4) - isign * GAMMA[mu], U[mu, s])
            D[s, :, :, sm, :, :] += np.einsum("ab,ij->aibj", np.eye(4) + isign * GAMMA[mu], U[mu, sm].conj().T)
    return D.reshape(12 * V, 12 * V)


def field_strength_numpy(L, U):
    """F_{mu nu} = (Q - Q^dag)/8 with Q the four clover leaves of mesfield.cc:15-21 (numpy)."""
    V = int(np.prod(L))
    X = coords(L)

    def sh(s, mu, d):
        x = X[s].copy(); x[mu] = (x[mu] + d) % L[mu]
        return index_of(L, x)

    F = {}
    for mu in range(3):
        for nu in range(mu + 1, 4):
            f = np.zeros((V, 3, 3), complex)
            for s in range(V):
                u = lambda m, t: U[m, t]
                d = lambda m, t: U[m, t].conj().T
                Q = u(mu, s) @ u(nu, sh(s, mu, 1)) @ d(mu, sh(s, nu, 1)) @ d(nu, s)
                xm = sh(s, mu, -1)
                Q += u(nu, s) @ d(mu, sh(xm, nu, 1)) @ d(nu, xm) @ u(mu, xm)
                Q += d(mu, xm) @ d(nu, sh(xm, nu, -1)) @ u(mu, sh(xm, nu, -1)) @ u(nu, sh(s, nu, -1))
                xn = sh(s, nu, -1)
                Q += d(nu, xn) @ u(mu, xn) @ u(nu, sh(xn, mu, 1)) @ d(mu, s)
                f[s] = (Q - Q.conj().T) / 8
            F[(mu, nu)] = f
    return F


def tri_to_dense(tri):
    """PrimitiveClovTriang (72 reals) -> dense 12x12 per site (block b on components 6b..6b+5)."""
    V = tri.shape[0]
    A = np.zeros((V, 12, 12), complex)
    for b in range(2):
        diag = tri[:, 6 * b:6 * b + 6]
        off = tri[:, 12 + 30 * b:12 + 30 * b + 30].reshape(V, 15, 2)
        off = off[..., 0] + 1j * off[..., 1]
        for i in range(6):
            A[:, 6 * b + i, 6 * b + i] = diag[:, i]
            for j in range(i):
                e = i * (i - 1) // 2 + j
                A[:, 6 * b + i, 6 * b + j] = off[:, e]
                A[:, 6 * b + j, 6 * b + i] = off[:, e].conj()
    return A


# ---------------------------------------------------------------------------------------------------
def test_gamma_algebra():
    for m in range(4):
        for n in range(4):
            ac = GAMMA[m] @ GAMMA[n] + GAMMA[n] @ GAMMA[m]
            assert np.allclose(ac, 2 * np.eye(4) * (m == n))
    assert np.allclose(G5, np.diag([1, 1, -1, -1]))


def test_twisted_golden():
    """tests/chroma/hadron/propagator/prec_wilson-twisted: prop_corr (metric 1e-5) and n_count."""
    g = json.load(open(os.path.join(HERE, "golden", "prec_wilson_twisted.json")))
    L = tuple(g["nrow"])
    V = int(np.prod(L))
    U = np.zeros((4, V, 3, 3), complex)
    U[:] = np.eye(3)
    for ph, d in zip(g["phases_by_pi"], g["phases_dir"]):
        U[d] *= np.exp(1j * np.pi * ph / L[d])
    op = oracle.Op(oracle.EO_WILSON, L, U, 1 / (2 * g["kappa"]) - 4)
    ncs, res, corr, _ = op.quarkprop4_point(index_of(L, g["t_srce"]), oracle.CG, g["RsdCG"], g["MaxCG"])
    gold = np.array(g["prop_corr"])
    assert np.max(np.abs(corr - gold) / gold) < g["tolerance"]
    assert all(abs(n - g["n_count"]) <= 1 for n in ncs)  # the golden is an FP32 build


@pytest.mark.parametrize("bc", ["periodic", "antiperiodic", "twisted"])
def test_free_field_closed_form(bc):
    """U = 1 (with BCs): plane waves psi = e^{ip.x} u are eigenvectors and
    D(isign) psi = sum_mu [2 cos q_mu - 2 i isign gamma_mu sin q_mu] psi, q = p + (twist phase).
    Antiperiodic time (link -1 on the last slice, simple_fermbc.h:87-104) admits p_t = pi(2n+1)/L_t."""
    L = (4, 4, 6, 8)
    V = int(np.prod(L))
    X = coords(L)
    U = np.zeros((4, V, 3, 3), complex)
    U[:] = np.eye(3)
    n = np.array([1, 2, 1, 3])
    p = 2 * np.pi * n / np.array(L)
    theta = np.zeros(4)
    if bc == "antiperiodic":
        U[3, X[:, 3] == L[3] - 1] *= -1
        p[3] = np.pi * (2 * n[3] + 1) / L[3]
    if bc == "twisted":
        for d, ph in zip((0, 1, 2), (0.2, 0.4, 0.3)):
            U[d] *= np.exp(1j * np.pi * ph / L[d])
            theta[d] = np.pi * ph / L[d]
    u = np.random.default_rng(1).standard_normal((4, 3)) + 0j
    psi = np.exp(1j * X @ p)[:, None, None] * u[None]
    q = p + theta
    for isign in (+1, -1):
        full = np.zeros_like(psi)
        for cb in (0, 1):
            oracle.dslash(L, U, psi, isign, cb, out=full)
        Mspin = sum(2 * np.cos(q[m]) * np.eye(4) - 2j * isign * GAMMA[m] * np.sin(q[m]) for m in range(4))
        expect = np.exp(1j * X @ p)[:, None, None] * np.einsum("ab,bc->ac", Mspin, u)[None]
        assert rel(full, expect) < 1e-13


@pytest.mark.parametrize("L", [(2, 2, 2, 2), (2, 2, 4, 2)])
def test_dslash_dense_bruteforce(L):
    V = int(np.prod(L))
    U = numpy_gauge(L, 3)
    psi = random_fermion(V, 4)
    for isign in (+1, -1):
        D = dense_dslash(L, U, isign)
        ref = (D @ psi.reshape(-1)).reshape(V, 4, 3)
        got = np.zeros_like(psi)
        for cb in (0, 1):
            oracle.dslash(L, U, psi, isign, cb, out=got)
        assert rel(got, ref) < 1e-13


def test_gamma5_hermiticity_and_gauge_covariance():
    L = (4, 4, 4, 4)
    V = int(np.prod(L))
    U = numpy_gauge(L, 5)
    chi, psi = random_fermion(V, 6), random_fermion(V, 7)
    m = oracle.Op(oracle.UNPREC_WILSON, L, U, 0.1)
    # <chi, M(+) psi> = <M(-) chi, psi>
    assert abs(np.vdot(chi, m(psi, +1)) - np.vdot(m(chi, -1), psi)) < 1e-12 * V
    # gamma5 M gamma5 = M^dag
    g5 = np.diag(G5)
    assert rel(g5[None, :, None] * m(g5[None, :, None] * psi, +1), m(psi, -1)) < 1e-13
    # gauge covariance: U -> g(x) U g^dag(x+mu), psi -> g psi
    from tests.util import random_su3_numpy
    g = random_su3_numpy(V, 8)
    X = coords(L)
    Ug = np.empty_like(U)
    for mu in range(4):
        xp = X.copy(); xp[:, mu] = (xp[:, mu] + 1) % L[mu]
        idx = np.array([index_of(L, x) for x in xp])
        Ug[mu] = g @ U[mu] @ g[idx].conj().transpose(0, 2, 1)
    mg = oracle.Op(oracle.UNPREC_WILSON, L, Ug, 0.1)
    gpsi = np.einsum("vij,vsj->vsi", g, psi)
    assert rel(mg(gpsi), np.einsum("vij,vsj->vsi", g, m(psi))) < 1e-13


def test_mesfield_against_numpy_leaves():
    L = (2, 4, 2, 4)
    U = numpy_gauge(L, 9)
    F = oracle.mesfield(L, U)
    Fn = field_strength_numpy(L, U)
    for k, (mu, nu) in enumerate([(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]):
        assert np.max(np.abs(F[k] - Fn[(mu, nu)])) < 1e-14


def test_clover_dense_construction():
    """A = (Nd+m) 1 - sum_{mu<nu} c_{mu nu} gamma_mu gamma_nu (x) F_{mu nu} with c = csw/2: the packing of
    makeClovSiteLoop (clover_term_qdp_w.h:398-522) against a dense construction from the gammas."""
    L = (2, 4, 2, 4)
    U = numpy_gauge(L, 10)
    mass, cR, cT = 0.07, 1.3, 0.9
    tri = oracle.make_clover(L, U, cR, cT, mass)
    Fn = field_strength_numpy(L, U)
    V = int(np.prod(L))
    A = np.zeros((V, 12, 12), complex)
    A[:] = (4 + mass) * np.eye(12)
    for (mu, nu), f in Fn.items():
        c = 0.5 * (cT if 3 in (mu, nu) else cR)
        A -= c * np.einsum("ab,vij->vaibj", GAMMA[mu] @ GAMMA[nu], f).reshape(V, 12, 12)
    assert np.max(np.abs(tri_to_dense(tri) - A)) < 1e-13


def quda_packed_from_dense(A):
    """QUDAPackedClovSite from dense Hermitian 12x12 site matrices (index spin*3+colour): per chiral
    block b (spins 2b, 2b+1) the real diagonal, then the strict lower triangle column by column (col 0:
    rows 1..5, col 1: rows 2..5, ...), (re, im) pairs."""
    V = A.shape[0]
    out = np.zeros((V, 72))
    for b in range(2):
        blk = A[:, 6 * b:6 * b + 6, 6 * b:6 * b + 6]
        out[:, 36 * b:36 * b + 6] = np.real(np.diagonal(blk, axis1=1, axis2=2))
        low = [blk[:, row, col] for col in range(6) for row in range(col + 1, 6)]
        low = np.stack(low, axis=1)
        out[:, 36 * b + 6:36 * b + 36:2] = low.real
        out[:, 36 * b + 7:36 * b + 36:2] = low.imag
    return out


def test_pack_quda_wire_format():
    """packForQUDA (clover_term_qdp_w.h:2171-2228) against the dense clover term of the gammas and the
    field strength: diag1/offDiag1/diag2/offDiag2 in the column-major lower-triangle order of
    QUDAPackedClovSite, at the full-lattice index of each site of the checkerboard, the other
    checkerboard untouched; the same for the inverse on both checkerboards (choles(0), choles(1))."""
    L = (2, 4, 2, 4)
    U = numpy_gauge(L, 15)
    mass, cR, cT = 0.07, 1.3, 0.9
    tri = oracle.make_clover(L, U, cR, cT, mass)
    Fn = field_strength_numpy(L, U)
    V = int(np.prod(L))
    Vh = V // 2
    A = np.zeros((V, 12, 12), complex)
    A[:] = (4 + mass) * np.eye(12)
    for (mu, nu), f in Fn.items():
        c = 0.5 * (cT if 3 in (mu, nu) else cR)
        A -= c * np.einsum("ab,vij->vaibj", GAMMA[mu] @ GAMMA[nu], f).reshape(V, 12, 12)
    want = quda_packed_from_dense(A)
    out = np.full((V, 72), np.nan)
    oracle.pack_quda(tri, 0, out)
    assert np.isnan(out[Vh:]).all()
    assert np.max(np.abs(out[:Vh] - want[:Vh])) < 1e-13
    oracle.pack_quda(tri, 1, out)
    assert np.max(np.abs(out - want)) < 1e-13
    inv = tri.copy()
    oracle.clover_invert(L, inv, 0)
    oracle.clover_invert(L, inv, 1)
    pinv = oracle.pack_quda(inv, 1, oracle.pack_quda(inv, 0))
    assert np.max(np.abs(pinv - quda_packed_from_dense(np.linalg.inv(A)))) < 1e-12
    # planted mistakes: the lower triangle row by row (the PrimitiveClovTriang order itself), and the
    # upper triangle column by column (the conjugates), are other wire formats and fail
    bad = want.copy()
    for b in range(2):
        low = np.stack([A[:, 6 * b + row, 6 * b + col] for row in range(6) for col in range(row)], axis=1)
        bad[:, 36 * b + 6:36 * b + 36:2] = low.real
        bad[:, 36 * b + 7:36 * b + 36:2] = low.imag
    assert np.max(np.abs(out - bad)) > 1e-3
    conj = want.copy()
    for b in range(2):
        conj[:, 36 * b + 7:36 * b + 36:2] *= -1
    assert np.max(np.abs(out - conj)) > 1e-3


def test_clover_unit_gauge_and_inverse():
    L = (4, 4, 4, 4)
    V = int(np.prod(L))
    U = np.zeros((4, V, 3, 3), complex)
    U[:] = np.eye(3)
    mass = 0.25
    tri = oracle.make_clover(L, U, 1.0, 1.0, mass)
    assert np.allclose(tri_to_dense(tri), (4 + mass) * np.eye(12))
    ld, nn = oracle.clover_invert(L, tri, 0)
    # log det closed form: 12 Vh log|Nd + m| (cf. symm_prec_tests.cc:633-653)
    assert abs(ld - 12 * (V // 2) * np.log(4 + mass)) < 1e-10 and nn == 0
    U = numpy_gauge(L, 11)
    tri = oracle.make_clover(L, U, 1.5, 1.5, -0.2)
    inv = tri.copy()
    ld, nn = oracle.clover_invert(L, inv, 0)
    A, Ai = tri_to_dense(tri)[: V // 2], tri_to_dense(inv)[: V // 2]
    assert np.max(np.abs(A @ Ai - np.eye(12))) < 1e-12
    sign, logabs = np.linalg.slogdet(A)
    assert abs(ld - logabs.sum()) < 1e-9
    # A is Hermitian
    assert np.allclose(A, A.conj().transpose(0, 2, 1))


def test_clover_apply_matches_dense():
    L = (4, 4, 2, 4)
    V = int(np.prod(L))
    U = numpy_gauge(L, 12)
    tri = oracle.make_clover(L, U, 1.0, 1.0, 0.1)
    psi = random_fermion(V, 13)
    out = np.zeros_like(psi)
    for cb in (0, 1):
        oracle.clover_apply(L, tri, psi, cb, out=out)
    ref = np.einsum("vab,vb->va", tri_to_dense(tri), psi.reshape(V, 12)).reshape(V, 4, 3)
    assert rel(out, ref) < 1e-14


@pytest.mark.parametrize("kind", [oracle.EO_WILSON, oracle.EO_CLOVER])
def test_schur_and_reconstruction(kind):
    """EO pieces and the Schur complement: solving the preconditioned system and reconstructing the
    even part solves the unpreconditioned system (eoprec_fermact_qprop.cc:41-80, eoprec_linop.h:147-165)."""
    L = (4, 4, 4, 4)
    V = int(np.prod(L))
    U = numpy_gauge(L, 14)
    op = oracle.Op(kind, L, U, 0.2, 1.0, 1.0)
    chi = random_fermion(V, 15)
    psi, nc, resid = op.qprop(chi, oracle.CG, 1e-12, 2000)
    assert resid / np.linalg.norm(chi) < 1e-10
    r = op.piece("unprec", psi) - chi
    assert np.linalg.norm(r) / np.linalg.norm(chi) < 1e-10
    # M_oo = A_oo - D_oe A_ee^-1 D_eo on the odd sites
    t = random_fermion(V, 16)
    t[: V // 2] = 0
    m = op(t)[V // 2:]
    s = op.piece("oddEven", op.piece("evenEvenInv", op.piece("evenOdd", t)))
    assert rel(m, (op.piece("oddOdd", t) - s)[V // 2:]) < 1e-13


def test_prec_vs_unprec_propagator():
    """prop_corr of the EO-preconditioned and unpreconditioned Wilson solves agree (the invariant of
    prec_wilson.out.xml:487 vs unprec_wilson.out.xml:487)."""
    L = (4, 4, 4, 4)
    U = numpy_gauge(L, 17)
    m = 0.3
    a = oracle.Op(oracle.EO_WILSON, L, U, m).quarkprop4_point(0, oracle.CG, 1e-11, 3000)[2]
    b = oracle.Op(oracle.UNPREC_WILSON, L, U, m)
    V = int(np.prod(L))
    corr = np.zeros(L[3])
    X = coords(L)
    for col in range(3):
        for sp in range(4):
            chi = np.zeros((V, 4, 3), complex)
            chi[0, sp, col] = 1
            psi = b.solve(chi, solver=oracle.CG, rsd=1e-11, maxit=3000)[0]
            np.add.at(corr, X[:, 3], np.sum(np.abs(psi.reshape(V, 12)) ** 2, axis=1))
    assert np.max(np.abs(a - corr) / corr) < 1e-8


def test_cg_against_dense_solve_and_trace():
    L = (2, 2, 2, 4)
    V = int(np.prod(L))
    U = numpy_gauge(L, 18)
    mass = 0.4
    op = oracle.Op(oracle.UNPREC_WILSON, L, U, mass)
    D = dense_dslash(L, U, +1)
    M = (4 + mass) * np.eye(12 * V) - 0.5 * D
    chi = random_fermion(V, 19)
    psi, nc, resid, trace = op.solve(chi, solver=oracle.CG, rsd=1e-12, maxit=500)
    exact = np.linalg.solve(M, chi.reshape(-1)).reshape(V, 4, 3)
    assert rel(psi, exact) < 1e-9
    # CG residual norms decrease to the target; exact arithmetic terminates in <= 12V steps
    assert trace[-1] <= 1e-24 * np.vdot(chi, chi).real * 1e2
    assert nc <= 12 * V


def test_bicgstab_and_reliable_cg():
    L = (4, 4, 4, 4)
    V = int(np.prod(L))
    U = numpy_gauge(L, 20)
    op = oracle.Op(oracle.EO_CLOVER, L, U, 0.05, 1.0, 1.0)
    chi = random_fermion(V, 21)
    chi[: V // 2] = 0
    ref = op.solve(chi, solver=oracle.CG, rsd=1e-12, maxit=5000)[0]
    for solver in (oracle.BICGSTAB, oracle.RELIABLE_CG, oracle.RELIABLE_BICGSTAB):
        psi, nc, resid, _ = op.solve(chi, solver=solver, rsd=1e-10, maxit=5000, delta=0.1)
        assert nc > 0
        assert rel(psi[V // 2:], ref[V // 2:]) < 1e-8


def test_reliable_bicgstab_against_dense_solve():
    # FP32 inner iteration with FP64 reliable updates reaches the FP64 solution of M psi = chi
    L = (2, 2, 2, 4)
    V = int(np.prod(L))
    U = numpy_gauge(L, 30)
    mass = 0.3
    op = oracle.Op(oracle.UNPREC_WILSON, L, U, mass)
    M = (4 + mass) * np.eye(12 * V) - 0.5 * dense_dslash(L, U, +1)
    chi = random_fermion(V, 31)
    exact = np.linalg.solve(M, chi.reshape(-1)).reshape(V, 4, 3)
    for delta in (0.1, 0.01):
        psi, nc, resid, trace = op.solve(chi, solver=oracle.RELIABLE_BICGSTAB, rsd=1e-11, maxit=2000, delta=delta)
        assert nc > 0
        # an FP32 solution vector alone is good to ~1e-7; 1e-9 needs the FP64 grouped x updates
        assert rel(psi, exact) < 1e-9
        assert resid <= 1e-10 * np.linalg.norm(chi)
        assert trace[-1] < 1e-22 * np.vdot(chi, chi).real * 1e2


@pytest.mark.parametrize("solver", ["CG", "BICGSTAB", "RELIABLE_CG", "RELIABLE_BICGSTAB"])
def test_mdagm_solvers_against_dense_normal_equations(solver):
    L = (2, 2, 2, 4)
    V = int(np.prod(L))
    U = numpy_gauge(L, 32)
    mass = 0.3
    op = oracle.Op(oracle.UNPREC_WILSON, L, U, mass)
    M = (4 + mass) * np.eye(12 * V) - 0.5 * dense_dslash(L, U, +1)
    chi = random_fermion(V, 33)
    exact = np.linalg.solve(M.conj().T @ M, chi.reshape(-1)).reshape(V, 4, 3)
    guess = 0.01 * random_fermion(V, 34)
    psi, nc, resid = op.solve_mdagm(chi, psi=guess, solver=getattr(oracle, solver), rsd=1e-11, maxit=3000)
    assert nc > 0
    assert rel(psi, exact) < 1e-8
    r = chi.reshape(-1) - M.conj().T @ (M @ psi.reshape(-1))
    assert abs(resid - np.linalg.norm(r)) <= 1e-6 * np.linalg.norm(chi)


# ---- planted mistakes: each pinning check above fails for a plausible error -----------------------
def test_planted_mistakes_are_caught():
    L = (2, 2, 2, 2)
    V = int(np.prod(L))
    U = numpy_gauge(L, 22)
    psi = random_fermion(V, 23)
    good = dense_dslash(L, U, +1) @ psi.reshape(-1)
    # (a) sign of the projector flipped: (1 + gamma) instead of (1 - gamma) in the forward hop
    bad = (dense_dslash(L, U, -1) @ psi.reshape(-1))
    assert rel(bad, good) > 1e-2
    # (b) link not daggered in the backward hop
    Ub = U.copy()
    X = coords(L)
    D_bad = np.zeros((V, 4, 3, V, 4, 3), complex)
    for s in range(V):
        for mu in range(4):
            xp = X[s].copy(); xp[mu] = (xp[mu] + 1) % L[mu]
            xm = X[s].copy(); xm[mu] = (xm[mu] - 1) % L[mu]
            D_bad[s, :, :, index_of(L, xp)] += np.einsum("ab,ij->aibj", np.eye(4) - GAMMA[mu], Ub[mu, s])
            D_bad[s, :, :, index_of(L, xm)] += np.einsum("ab,ij->aibj", np.eye(4) + GAMMA[mu], Ub[mu, index_of(L, xm)])
    assert rel(D_bad.reshape(12 * V, 12 * V) @ psi.reshape(-1), good) > 1e-2
    # (c) clover coefficient without the factor 1/2, or with the wrong sign, breaks the dense check
    tri = oracle.make_clover(L, U, 1.0, 1.0, 0.0)
    Fn = field_strength_numpy(L, U)
    for c_bad in (1.0, -0.5):
        A = np.zeros((V, 12, 12), complex)
        A[:] = 4 * np.eye(12)
        for (mu, nu), f in Fn.items():
            A -= c_bad * np.einsum("ab,vij->vaibj", GAMMA[mu] @ GAMMA[nu], f).reshape(V, 12, 12)
        assert np.max(np.abs(tri_to_dense(tri) - A)) > 1e-3


def test_minvcg2_against_dense_solves():
    """MInvCG2 (minvcg2.cc:72-355): every shifted solution equals the dense solution of
    (M^dag M + shift) psi = chi; the shift-0 iteration count is CG's on M^dag M; the returned residual
    is the shifted system's."""
    L = (2, 2, 2, 4)
    V = int(np.prod(L))
    U = numpy_gauge(L, 40)
    mass = 0.2
    op = oracle.Op(oracle.UNPREC_WILSON, L, U, mass)
    M = (4 + mass) * np.eye(12 * V) - 0.5 * dense_dslash(L, U, +1)
    A = M.conj().T @ M
    chi = random_fermion(V, 41)
    shifts = [0.0, 0.01, 0.25, 1.5]
    psi, nc, resid = op.solve_multishift(chi, shifts, 1e-12, maxit=2000)
    for s, sh in enumerate(shifts):
        exact = np.linalg.solve(A + sh * np.eye(12 * V), chi.reshape(-1)).reshape(V, 4, 3)
        assert rel(psi[s], exact) < 1e-9, (s, rel(psi[s], exact))
        r = chi.reshape(-1) - (A + sh * np.eye(12 * V)) @ psi[s].reshape(-1)
        assert abs(resid[s] - np.linalg.norm(r)) < 1e-12 * np.linalg.norm(chi)
    _, nc_cg, _ = op.solve_mdagm(chi, solver=oracle.CG, rsd=1e-12, maxit=2000)
    assert abs(nc - nc_cg) <= 1
    # shifts in any order, with per-shift targets: the loose shift stops early and meets only its own target
    psi2, nc2, resid2 = op.solve_multishift(chi, [0.25, 0.0], [1e-4, 1e-12], maxit=2000)
    assert nc2 == nc
    assert resid2[0] <= 1e-4 * np.linalg.norm(chi) and rel(psi2[1], psi[0]) < 1e-12


def test_minvcg2_zero_source():
    L = (2, 2, 2, 4)
    op = oracle.Op(oracle.UNPREC_WILSON, L, numpy_gauge(L, 42), 0.2)
    psi, nc, _ = op.solve_multishift(np.zeros((int(np.prod(L)), 4, 3), complex), [0.0, 0.1], 1e-10)
    assert nc == 0 and not np.any(psi)


@pytest.mark.parametrize("isign", [+1, -1])
@pytest.mark.parametrize("cb", [0, 1])
def test_dslash_deriv_is_the_link_derivative(isign, cb):
    """WilsonDslashBase::deriv (lwldslash_base_w.h:132-227): with chi supported on cb, the forward hops
    into cb are the only terms of <chi, D psi> that contain the links U_mu(x), x on cb, and they are
    linear in them, so sum_{x in cb, mu} tr(U_mu(x) ds_mu(x)) = <chi, D(U) psi> - <chi, D(U') psi>
    with U' = U with those links set to zero; ds is zero off cb."""
    L = (4, 4, 4, 8)
    V = int(np.prod(L))
    Vh = V // 2
    U = numpy_gauge(L, 50)
    chi = random_fermion(V, 51)
    chi[(1 - cb) * Vh:(2 - cb) * Vh] = 0
    psi = random_fermion(V, 52)
    ds = oracle.dslash_deriv(L, chi, psi, isign, cb)
    assert not np.any(ds[:, (1 - cb) * Vh:(2 - cb) * Vh])
    U0 = U.copy()
    U0[:, cb * Vh:(cb + 1) * Vh] = 0
    full = np.vdot(chi, oracle.dslash(L, U, psi, isign, cb))
    rest = np.vdot(chi, oracle.dslash(L, U0, psi, isign, cb))
    lin = np.einsum("mvab,mvba->", U, ds)
    assert abs(lin - (full - rest)) < 1e-12 * abs(full)
    # both checkerboards at once is the sum (lwldslash_base_w.h:110-124)
    both = oracle.dslash_deriv(L, chi, psi, isign, -1)
    assert np.allclose(both, ds + oracle.dslash_deriv(L, chi, psi, isign, 1 - cb), atol=0, rtol=0)


# ---------------------------------------------------------------------------------------------------
# Stout smearing (Stouting::smear_links, stout_utils.cc:936-962; staples and Q of getQsandCs, :105-180)
STOUT_X0, STOUT_MU0, STOUT_EPS, STOUT_H = (1, 2, 1, 2), 1, 0.4, np.array([1.0, 1.0, -2.0])


def _one_link_field(L):
    """Unit field except U_mu0(x0) = diag(exp(i eps h)), h traceless but with sin(eps h) not traceless."""
    V = int(np.prod(L))
    U = np.zeros((4, V, 3, 3), complex)
    U[:, :] = np.eye(3)
    U[STOUT_MU0, index_of(L, STOUT_X0)] = np.diag(np.exp(1j * STOUT_EPS * STOUT_H))
    return U


def _one_link_closed_form(L, rho, dirs=15):
    """One stout step of _one_link_field by hand.  Every link stays diagonal, so exp(iQ) is a phase per
    colour.  For U_mu0(x0) all six staples are 1: C = 6 rho, Omega = 6 rho D^dag, and
    U' = exp(i (eps h - 6 rho (sin(eps h) - mean)));  a link with one staple equal to D (or D^dag) and
    five equal to 1 gets exp(+i rho (sin(eps h) - mean)) (or exp(-i ...)).  That is U_nu(x0) and
    U_nu(x0-nu+mu0), U_mu0(x0+nu) and U_mu0(x0-nu) with D; U_nu(x0+mu0) and U_nu(x0-nu) with D^dag."""
    U = _one_link_field(L)
    sn = np.sin(STOUT_EPS * STOUT_H)
    ph = rho * (sn - sn.mean())
    x0, mu0 = np.array(STOUT_X0), STOUT_MU0

    def at(mu, shifts, sign):
        x = x0.copy()
        for m, d in shifts:
            x[m] = (x[m] + d) % L[m]
        U[mu, index_of(L, x)] = np.diag(np.exp(1j * sign * ph))

    # with only the directions in `dirs` smeared, U_mu0(x0) has 2 staples per other smeared direction and
    # links of unsmeared directions stay 1 (StoutLinkSmearing smear_dirs, stout_link_smearing.cc:62-120)
    others = [nu for nu in range(4) if nu != mu0 and dirs >> nu & 1]
    U[mu0, index_of(L, x0)] = np.diag(np.exp(1j * (STOUT_EPS * STOUT_H - 2 * len(others) * ph)))
    for nu in others:
        at(nu, [], +1)
        at(nu, [(mu0, 1), (nu, -1)], +1)
        at(mu0, [(nu, 1)], +1)
        at(mu0, [(nu, -1)], +1)
        at(nu, [(mu0, 1)], -1)
        at(nu, [(nu, -1)], -1)
    return U


@pytest.mark.parametrize("rho,dirs", [(0.1, 15), (0.22, 15), (0.1, 7), (0.15, 0b1010)])
def test_stout_one_link_closed_form(rho, dirs):
    """All directions, spatial only (no_smear_dir = 3), and mu0 with t only."""
    L = (4, 4, 4, 4)
    out = oracle.stout(L, _one_link_field(L), rho, 1, dirs)
    assert np.max(np.abs(out - _one_link_closed_form(L, rho, dirs))) < 1e-14


def test_stout_gauge_covariance_unitarity_and_rho_zero():
    """stout(g U g^dag(x+mu)) = g stout(U) g^dag(x+mu) for a random gauge transform g (two steps on a Haar
    field): a staple on the wrong sites or with a missing dagger breaks it.  The result stays in SU(3),
    rho = 0 returns the field unchanged, and smearing raises the plaquette of a rough field."""
    L = (4, 4, 4, 6)
    V = int(np.prod(L))
    U = numpy_gauge(L, 60)
    from tests.util import random_su3_numpy
    g = random_su3_numpy(V, 61)
    X = coords(L)

    def transform(W):
        out = np.empty_like(W)
        for mu in range(4):
            xp = X.copy(); xp[:, mu] = (xp[:, mu] + 1) % L[mu]
            idx = np.array([index_of(L, x) for x in xp])
            out[mu] = g @ W[mu] @ g[idx].conj().transpose(0, 2, 1)
        return out

    S = oracle.stout(L, U, 0.1, 2)
    assert rel(oracle.stout(L, transform(U), 0.1, 2), transform(S)) < 1e-13
    eye = np.eye(3)
    assert np.max(np.abs(S @ S.conj().transpose(0, 1, 3, 2) - eye)) < 1e-13
    assert np.max(np.abs(np.linalg.det(S) - 1)) < 1e-13
    assert np.array_equal(oracle.stout(L, U, 0.0), U)

    def plaq(W):
        acc = 0.0
        for mu in range(3):
            for nu in range(mu + 1, 4):
                xm = X.copy(); xm[:, mu] = (xm[:, mu] + 1) % L[mu]
                xn = X.copy(); xn[:, nu] = (xn[:, nu] + 1) % L[nu]
                im = np.array([index_of(L, x) for x in xm])
                inn = np.array([index_of(L, x) for x in xn])
                P = W[mu] @ W[nu][im] @ W[mu][inn].conj().transpose(0, 2, 1) @ W[nu].conj().transpose(0, 2, 1)
                acc += np.trace(P, axis1=1, axis2=2).real.sum() / 3
        return acc / (6 * V)

    p = [plaq(oracle.stout(L, U, 0.1, n)) for n in (0, 1, 3)]
    assert abs(p[0]) < 0.1 and p[0] < p[1] < p[2]


def test_stout_planted_mistakes_are_caught():
    """A numpy copy of one stout step with a planted mistake fails the closed form each time: the
    backward staple built at x+nu, no traceless projection, Q without the factor 1/2, Q of the wrong sign."""
    from scipy.linalg import expm
    L = (4, 4, 4, 4)
    V = int(np.prod(L))
    U = _one_link_field(L)
    X = coords(L)
    want = _one_link_closed_form(L, 0.1)

    def nb(s, mu, d):
        x = X[s].copy(); x[mu] = (x[mu] + d) % L[mu]
        return index_of(L, x)

    def step(mistake):
        out = np.empty_like(U)
        dg = lambda m: m.conj().T
        for mu in range(4):
            for s in range(V):
                C = np.zeros((3, 3), complex)
                for nu in range(4):
                    if nu == mu:
                        continue
                    C += U[nu, s] @ U[mu, nb(s, nu, 1)] @ dg(U[nu, nb(s, mu, 1)])
                    y = nb(s, nu, 1 if mistake == "backward_site" else -1)
                    C += dg(U[nu, y]) @ U[mu, y] @ U[nu, nb(y, mu, 1)]
                Om = 0.1 * C @ dg(U[mu, s])
                T = dg(Om) - Om
                if mistake != "no_traceless":
                    T -= np.trace(T) / 3 * np.eye(3)
                Q = (1.0 if mistake == "no_half" else 0.5) * 1j * T
                if mistake == "sign":
                    Q = -Q
                out[mu, s] = expm(1j * Q) @ U[mu, s]
        return out

    assert np.max(np.abs(step(None) - want)) < 1e-13
    for mistake in ("backward_site", "no_traceless", "no_half", "sign"):
        assert np.max(np.abs(step(mistake) - want)) > 1e-3, mistake


# ---------------------------------------------------------------------------------------------------
# Meson two-point functions (mesons_w.cc:97-180 with SftMom phases, sftmom.cc:405-460)
def _gamma_n(n):
    """Gamma(n) = gamma_1^n0 gamma_2^n1 gamma_3^n2 gamma_4^n3 (docs/notes/chroma_gamma_matrices.tex)."""
    G = np.eye(4, dtype=complex)
    for mu in range(4):
        if n >> mu & 1:
            G = G @ GAMMA[mu]
    return G


def _mesons_numpy(L, S1, S2, origin, moms, mistake=None):
    V = int(np.prod(L))
    X = coords(L)
    g5 = np.kron(G5, np.eye(3))
    anti = S2 if mistake == "no_g5" else g5 @ S2 @ g5
    out = np.zeros((16, len(moms), L[3]), complex)
    teff = (X[:, 3] - origin[3]) % L[3]
    for n in range(16):
        G = np.kron(_gamma_n(n), np.eye(3))
        B = G @ S1 @ G
        c = np.einsum("vij,vij->v", anti if mistake == "no_conj" else anti.conj(), B)
        for m, p in enumerate(moms):
            pdx = sum(2 * np.pi * p[k] * (X[:, k] - origin[k]) / L[k] for k in range(3))
            out[n, m] = np.bincount(teff, weights=(np.exp(1j * pdx) * c).real, minlength=L[3]) + \
                1j * np.bincount(teff, weights=(np.exp(1j * pdx) * c).imag, minlength=L[3])
    return out


def test_mesons_against_numpy_contraction():
    """Random propagators, an origin off zero and three sink momenta: the oracle equals a numpy
    contraction written from mesons_w.cc:128-142 with the tests' own gamma matrices."""
    L = (4, 4, 2, 6)
    V = int(np.prod(L))
    r = np.random.default_rng(70)
    S1 = r.standard_normal((V, 12, 12)) + 1j * r.standard_normal((V, 12, 12))
    S2 = r.standard_normal((V, 12, 12)) + 1j * r.standard_normal((V, 12, 12))
    origin, moms = (1, 3, 0, 2), ((0, 0, 0), (1, 0, 0), (-1, 2, 1))
    got = oracle.mesons(L, S1, S2, origin, moms)
    assert rel(got, _mesons_numpy(L, S1, S2, origin, moms)) < 1e-13
    # gamma_value 15 is gamma_5: with S1 = S2 the pion is sum |S|^2 per slice (prop_corr)
    assert np.allclose(_gamma_n(15), G5)
    pion = oracle.mesons(L, S1, None, origin)[15, 0]
    X = coords(L)
    want = np.bincount((X[:, 3] - origin[3]) % L[3], weights=np.sum(np.abs(S1) ** 2, axis=(1, 2)), minlength=L[3])
    assert np.max(np.abs(pion - want)) < 1e-12 * np.max(want)
    # summing the projection over every spatial momentum returns Vs times the correlator at the origin's
    # spatial site (sum_p exp(i p.(x - o)) = Vs delta_{x,o}), which fixes 2 pi / L_k
    allp = [(a, b, c) for a in range(L[0]) for b in range(L[1]) for c in range(L[2])]
    tot = oracle.mesons(L, S1, S2, origin, allp).sum(axis=1)
    site0 = np.zeros((16, L[3]), complex)
    for t in range(L[3]):
        x = [origin[0], origin[1], origin[2], (origin[3] + t) % L[3]]
        s = index_of(L, x)
        for n in range(16):
            G = np.kron(_gamma_n(n), np.eye(3))
            g5 = np.kron(G5, np.eye(3))
            site0[n, t] = np.sum((g5 @ S2[s] @ g5).conj() * (G @ S1[s] @ G))
    assert rel(tot, np.prod(L[:3]) * site0) < 1e-12


def test_mesons_pion_is_the_golden_prop_corr():
    """prec_wilson-twisted golden: the pion correlator of the oracle's propagator is the golden prop_corr
    (t_srce = 0, so t_eff = t); planted mistakes (no gamma_5 sandwich on the anti-quark, no complex
    conjugate) move it off."""
    g = json.load(open(os.path.join(HERE, "golden", "prec_wilson_twisted.json")))
    L = tuple(g["nrow"])
    V = int(np.prod(L))
    U = np.zeros((4, V, 3, 3), complex)
    U[:] = np.eye(3)
    for ph, d in zip(g["phases_by_pi"], g["phases_dir"]):
        U[d] *= np.exp(1j * np.pi * ph / L[d])
    op = oracle.Op(oracle.EO_WILSON, L, U, 1 / (2 * g["kappa"]) - 4)
    _, _, _, prop = op.quarkprop4_point(index_of(L, g["t_srce"]), oracle.CG, g["RsdCG"], g["MaxCG"], want_prop=True)
    gold = np.array(g["prop_corr"])
    m = oracle.mesons(L, prop, None, tuple(g["t_srce"]))
    assert np.max(np.abs(m[15, 0].real - gold) / gold) < g["tolerance"]
    assert np.max(np.abs(m[15, 0].imag)) < 1e-12 * np.max(gold)
    for mistake in ("no_g5", "no_conj"):
        bad = _mesons_numpy(L, prop, prop, tuple(g["t_srce"]), ((0, 0, 0),), mistake)[15, 0]
        assert np.max(np.abs(bad - gold) / gold) > 1e-3, mistake
