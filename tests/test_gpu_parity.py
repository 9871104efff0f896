"""GPU parity: the HIP library (through its C-ABI) against the oracle on the same seeded inputs.

Tolerances (relative, Frobenius norm over the output): FP64 1e-12, FP32 1e-6, FP16 storage 2e-3.
Layout conversion in FP64 is checked bit for bit.
"""
import sys

import numpy as np
import pytest

import oracle
from chroma_amd import lib as cmx
from tests.util import rel, random_fermion

pytestmark = pytest.mark.gpu

TOL = {cmx.DOUBLE: 1e-12, cmx.SINGLE: 1e-6, cmx.HALF: 2e-3}
LATS = [(4, 4, 4, 4), (4, 4, 4, 8), (8, 8, 8, 16), (6, 4, 8, 10)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gauge_field(L, seed, bc=(1, 1, 1, -1), twist=None):
    U = cmx.gauge_generate(L, kind=1, seed=seed, eps=0.3)
    return cmx.gauge_apply_bc(L, U, boundary=bc, twist_by_pi=twist)


@pytest.mark.parametrize("prec", [cmx.DOUBLE, cmx.SINGLE, cmx.HALF])
def test_layout_roundtrip(prec):
    L = (4, 4, 4, 8)
    say("layout roundtrip", L, prec)
    U = gauge_field(L, 1)
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_WILSON, 0.1)
    V = int(np.prod(L))
    psi = random_fermion(V // 2, 3)
    out = op.layout_roundtrip(psi, prec)
    if prec == cmx.DOUBLE:
        assert np.array_equal(out.view(np.uint64), psi.view(np.uint64))  # bit exact
    else:
        assert rel(out, psi) < {cmx.SINGLE: 1e-7, cmx.HALF: 1e-3}[prec]


def test_fp16_store_flushes_subnormal_sites():
    """FP16 storage of sites whose largest component is below the FP32 normal range (the decaying front
    of a heavy-quark propagator on a large lattice): stored as zero, never as inf/NaN; other sites as usual."""
    L = (4, 4, 4, 8)
    say("fp16 subnormal sites")
    U = gauge_field(L, 1)
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_WILSON, 0.1, sloppy_prec=cmx.HALF)
    Vh = int(np.prod(L)) // 2
    psi = random_fermion(Vh, 4)
    psi[1::3] *= 1e-40   # every component subnormal in FP32
    psi[2::7] *= 1e-30   # small but normal
    out = op.layout_roundtrip(psi, cmx.HALF)
    assert np.isfinite(out).all()
    assert np.all(out[1::3] == 0)
    keep = np.ones(Vh, bool)
    keep[1::3] = False
    for k in np.nonzero(keep)[0][:50]:
        assert rel(out[k], psi[k]) < 1e-3


@pytest.mark.parametrize("L", LATS)
@pytest.mark.parametrize("prec,recon", [(cmx.DOUBLE, 18), (cmx.SINGLE, 18), (cmx.SINGLE, 12), (cmx.HALF, 12)])
def test_dslash(L, prec, recon):
    say("dslash", L, prec, recon)
    V = int(np.prod(L))
    U = gauge_field(L, 11)
    sp = cmx.HALF if prec == cmx.HALF else cmx.SINGLE
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_WILSON, 0.1, sloppy_prec=sp, sloppy_recon=recon)
    psi = random_fermion(V, 5)
    for cb in (0, 1):
        for isign in (+1, -1):
            ref = oracle.dslash(L, U, psi, isign, cb)
            Vh = V // 2
            src = psi[(1 - cb) * Vh:(2 - cb) * Vh]
            got = op.dslash(src, isign, cb, prec)
            e = rel(got, ref[cb * Vh:(cb + 1) * Vh])
            assert e < TOL[prec], (cb, isign, e)


@pytest.mark.parametrize("twist", [None, (0.2, 0.4, 0.0)])
def test_dslash_boundary_phases_recon12(twist):
    """12-real reconstruction with antiperiodic time and twisted phases (det(U) != 1 on those links)."""
    L = (4, 4, 6, 8)
    say("dslash bc", twist)
    V = int(np.prod(L))
    U = gauge_field(L, 21, bc=(1, 1, 1, -1), twist=twist)
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_WILSON, 0.1, sloppy_prec=cmx.SINGLE, sloppy_recon=12)
    psi = random_fermion(V, 6)
    ref = oracle.dslash(L, U, psi, +1, 1)
    got = op.dslash(psi[: V // 2], +1, 1, cmx.SINGLE)
    assert rel(got, ref[V // 2:]) < 1e-6


@pytest.mark.parametrize("L", [(4, 4, 4, 4), (4, 4, 4, 8), (8, 8, 8, 16)])
def test_clover_build_and_inverse(L):
    say("clover", L)
    V = int(np.prod(L))
    U = gauge_field(L, 31)
    mass, csw = cmx.kappa_to_mass(0.125), 1.0
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_CLOVER, mass, csw)
    ref = oracle.make_clover(L, U, csw, csw, mass)
    got = op.clover(0)
    assert np.max(np.abs(got - ref)) < 1e-13 * np.max(np.abs(ref))
    inv_ref = ref.copy()
    ld_ref, nn_ref = oracle.clover_invert(L, inv_ref, 0)
    inv = op.clover(1)
    Vh = V // 2
    assert np.max(np.abs(inv[:Vh] - inv_ref[:Vh])) < 1e-12 * np.max(np.abs(inv_ref[:Vh]))
    ld, nn = op.clover_logdet()
    assert abs(ld - ld_ref) < 1e-10 * max(1.0, abs(ld_ref)) and nn == nn_ref


def test_clover_pack_quda_and_per_direction_links():
    """QUDAPackedClovSite wire format (packForQUDA, clover_term_qdp_w.h:2171-2228) of A and A^-1 on both
    checkerboards against the oracle's restatement on the oracle's clover term, at full-lattice site
    indices with the other checkerboard untouched; bit for bit against the same restatement applied
    to the library's own FP64 clover download.  The operator is built from four per-direction link
    arrays (QUDA_QDP_GAUGE_ORDER) and equals the one built from the contiguous array bit for bit."""
    L = (4, 4, 4, 8)
    say("clover pack quda")
    V = int(np.prod(L))
    Vh = V // 2
    U = gauge_field(L, 33)
    mass, csw = cmx.kappa_to_mass(0.125), 1.0
    op = cmx.Operator(cmx.Gauge(L, [U[mu].copy() for mu in range(4)], per_direction=True), cmx.EO_CLOVER, mass, csw)
    op_c = cmx.Operator(cmx.Gauge(L, U), cmx.EO_CLOVER, mass, csw)
    psi = random_fermion(Vh, 34)
    assert np.array_equal(op(psi).view(np.uint64), op_c(psi).view(np.uint64))
    ref = oracle.make_clover(L, U, csw, csw, mass)
    out = np.full((V, 72), np.nan)
    op.clover_pack_quda(0, 0, out)
    assert np.isnan(out[Vh:]).all() and not np.isnan(out[:Vh]).any()
    op.clover_pack_quda(0, 1, out)
    want = oracle.pack_quda(ref, 1, oracle.pack_quda(ref, 0))
    assert np.max(np.abs(out - want)) < 1e-13 * np.max(np.abs(want))
    own = op.clover(0)
    assert np.array_equal(out, oracle.pack_quda(own, 1, oracle.pack_quda(own, 0)))
    inv_ref = ref.copy()
    oracle.clover_invert(L, inv_ref, 0)
    oracle.clover_invert(L, inv_ref, 1)
    outi = op.clover_pack_quda(1, 1, op.clover_pack_quda(1, 0))
    wanti = oracle.pack_quda(inv_ref, 1, oracle.pack_quda(inv_ref, 0))
    assert np.max(np.abs(outi - wanti)) < 1e-12 * np.max(np.abs(wanti))


@pytest.mark.parametrize("kind", [cmx.EO_WILSON, cmx.EO_CLOVER])
@pytest.mark.parametrize("prec", [cmx.DOUBLE, cmx.SINGLE, cmx.HALF])
def test_eo_operator(kind, prec):
    L = (8, 8, 8, 16)
    say("eo op", kind, prec)
    V = int(np.prod(L))
    U = gauge_field(L, 41)
    mass = cmx.kappa_to_mass(0.13)
    okind = {cmx.EO_WILSON: oracle.EO_WILSON, cmx.EO_CLOVER: oracle.EO_CLOVER}[kind]
    ref_op = oracle.Op(okind, L, U, mass, 1.0, 1.0)
    sp = cmx.HALF if prec == cmx.HALF else cmx.SINGLE
    op = cmx.Operator(cmx.Gauge(L, U), kind, mass, 1.0, sloppy_prec=sp)
    psi = random_fermion(V, 7)
    for isign in (+1, -1):
        ref = ref_op(psi, isign)[V // 2:]
        got = op(psi[V // 2:], isign, prec)
        assert rel(got, ref) < TOL[prec], (isign, rel(got, ref))


@pytest.mark.parametrize("kind", [cmx.EO_WILSON, cmx.EO_CLOVER])
def test_eo_pieces(kind):
    L = (4, 4, 4, 8)
    say("eo pieces", kind)
    V = int(np.prod(L))
    U = gauge_field(L, 51)
    mass = cmx.kappa_to_mass(0.12)
    okind = {cmx.EO_WILSON: oracle.EO_WILSON, cmx.EO_CLOVER: oracle.EO_CLOVER}[kind]
    ref_op = oracle.Op(okind, L, U, mass, 1.2, 1.2)
    op = cmx.Operator(cmx.Gauge(L, U), kind, mass, 1.2)
    psi = random_fermion(V, 8)
    Vh = V // 2
    for name, sl in [("evenEven", slice(0, Vh)), ("evenEvenInv", slice(0, Vh)), ("evenOdd", slice(0, Vh)),
                     ("oddEven", slice(Vh, V)), ("oddOdd", slice(Vh, V)), ("unprec", slice(0, V))]:
        for isign in (+1, -1):
            ref = ref_op.piece(name, psi, isign)[sl]
            got = cmx_piece = op.piece(name, psi, isign)[sl]
            assert rel(got, ref) < 1e-12, (name, isign, rel(cmx_piece, ref))


def test_unprec_wilson_operator():
    L = (8, 8, 8, 16)
    say("unprec wilson")
    V = int(np.prod(L))
    U = gauge_field(L, 61)
    mass = cmx.kappa_to_mass(0.115)
    ref_op = oracle.Op(oracle.UNPREC_WILSON, L, U, mass)
    op = cmx.Operator(cmx.Gauge(L, U), cmx.UNPREC_WILSON, mass)
    psi = random_fermion(V, 9)
    for isign in (+1, -1):
        for prec, tol in [(cmx.DOUBLE, 1e-12), (cmx.SINGLE, 1e-6), (cmx.HALF, 2e-3)]:
            assert rel(op(psi, isign, prec), ref_op(psi, isign)) < tol


def test_cg_unprec_wilson_config1_small():
    """Unpreconditioned Wilson CGNE at 4^3x8, kappa = 0.115, FP64, point source at the origin."""
    L = (4, 4, 4, 8)
    say("cg unprec")
    V = int(np.prod(L))
    U = gauge_field(L, 1)
    mass = cmx.kappa_to_mass(0.115)
    ref_op = oracle.Op(oracle.UNPREC_WILSON, L, U, mass)
    op = cmx.Operator(cmx.Gauge(L, U), cmx.UNPREC_WILSON, mass)
    chi = np.zeros((V, 4, 3), np.complex128)
    chi[0, 0, 0] = 1.0
    psi_ref, n_ref, r_ref, _ = ref_op.solve(chi, solver=oracle.CG, rsd=1e-10, maxit=2000)
    psi, res = op.solve(chi, solver=cmx.CG, rsd=1e-10, maxit=2000)
    assert abs(res["n_count"] - n_ref) <= 1, (res["n_count"], n_ref)
    assert res["rel_resid"] < 1e-10 * 2
    assert rel(psi, psi_ref) < 1e-8


@pytest.mark.parametrize("kind", [cmx.EO_WILSON, cmx.EO_CLOVER])
def test_cg_eo(kind):
    L = (8, 8, 8, 16)
    say("cg eo", kind)
    V = int(np.prod(L))
    U = gauge_field(L, 71)
    mass = cmx.kappa_to_mass(0.125)
    okind = {cmx.EO_WILSON: oracle.EO_WILSON, cmx.EO_CLOVER: oracle.EO_CLOVER}[kind]
    ref_op = oracle.Op(okind, L, U, mass, 1.0, 1.0)
    op = cmx.Operator(cmx.Gauge(L, U), kind, mass, 1.0)
    chi_full = random_fermion(V, 12)
    chi_full[: V // 2] = 0
    psi_ref, n_ref, r_ref, _ = ref_op.solve(chi_full, solver=oracle.CG, rsd=1e-10, maxit=5000)
    psi, res = op.solve(chi_full[V // 2:], solver=cmx.CG, rsd=1e-10, maxit=5000)
    assert abs(res["n_count"] - n_ref) <= 1, (res["n_count"], n_ref)
    assert rel(psi, psi_ref[V // 2:]) < 1e-8
    assert res["rel_resid"] < 1e-9


def test_bicgstab_eo_clover():
    L = (8, 8, 8, 16)
    say("bicgstab")
    V = int(np.prod(L))
    U = gauge_field(L, 81)
    mass = cmx.kappa_to_mass(0.125)
    ref_op = oracle.Op(oracle.EO_CLOVER, L, U, mass, 1.0, 1.0)
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_CLOVER, mass, 1.0)
    chi_full = random_fermion(V, 13)
    chi_full[: V // 2] = 0
    psi_ref, n_ref, r_ref, _ = ref_op.solve(chi_full, solver=oracle.BICGSTAB, rsd=1e-10, maxit=2000)
    psi, res = op.solve(chi_full[V // 2:], solver=cmx.BICGSTAB, rsd=1e-10, maxit=2000)
    assert abs(res["n_count"] - n_ref) <= 2, (res["n_count"], n_ref)
    assert rel(psi, psi_ref[V // 2:]) < 1e-8


@pytest.mark.parametrize("sloppy", [cmx.SINGLE, cmx.HALF])
def test_reliable_cg_mixed(sloppy):
    """Mixed precision with FP64 reliable updates reaches a 1e-10 FP64 true residual."""
    L = (8, 8, 8, 16)
    say("reliable cg", sloppy)
    V = int(np.prod(L))
    U = cmx.gauge_apply_bc(L, cmx.gauge_generate(L, kind=2, seed=91, nstout=6, rho=0.1), (1, 1, 1, -1))
    mass = cmx.kappa_to_mass(0.1350)
    ref_op = oracle.Op(oracle.EO_CLOVER, L, U, mass, 1.0, 1.0)
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_CLOVER, mass, 1.0, sloppy_prec=sloppy)
    chi_full = cmx.z2_source((V, 4, 3), 2)
    chi_full[: V // 2] = 0
    psi_ref, n_ref, r_ref, _ = ref_op.solve(chi_full, solver=oracle.CG, rsd=1e-11, maxit=10000)
    # reference semantics: stops on the normal-equation residual (reliable_cg.cc:156)
    psi0, res0 = op.solve(chi_full[V // 2:], solver=cmx.RELIABLE_CG, rsd=1e-10, maxit=20000, delta=0.1,
                          sloppy_prec=sloppy)
    assert res0["n_restarts"] == 0 and res0["rel_resid"] < 1e-9
    # true-residual mode: FP64 |chi - A psi| <= 1e-10 |chi|
    psi, res = op.solve(chi_full[V // 2:], solver=cmx.RELIABLE_CG, rsd=1e-10, maxit=20000, delta=0.1,
                        sloppy_prec=sloppy, true_residual=True)
    say("  n_count", res["n_count"], "oracle CG", n_ref, "rel resid", res["rel_resid"], "updates", res["n_reliable"],
        "restarts", res["n_restarts"])
    assert res["rel_resid"] <= 1e-10
    assert rel(psi, psi_ref[V // 2:]) < 1e-8


@pytest.mark.parametrize("sloppy", [cmx.DOUBLE, cmx.SINGLE, cmx.HALF])
def test_reliable_bicgstab_mixed(sloppy):
    """Reliable BiCGStab (FP64 / FP32 / FP16-storage inner iteration, FP64 reliable updates) against the
    oracle's single/double RelInvBiCGStab: iteration count, solution, and a 1e-10 FP64 true residual."""
    L = (8, 8, 8, 16)
    say("reliable bicgstab", sloppy)
    V = int(np.prod(L))
    U = cmx.gauge_apply_bc(L, cmx.gauge_generate(L, kind=2, seed=93, nstout=6, rho=0.1), (1, 1, 1, -1))
    mass = cmx.kappa_to_mass(0.1350)
    ref_op = oracle.Op(oracle.EO_CLOVER, L, U, mass, 1.0, 1.0)
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_CLOVER, mass, 1.0, sloppy_prec=cmx.HALF if sloppy == cmx.HALF else cmx.SINGLE)
    chi_full = cmx.z2_source((V, 4, 3), 4)
    chi_full[: V // 2] = 0
    psi_cg, _, _, _ = ref_op.solve(chi_full, solver=oracle.CG, rsd=1e-12, maxit=10000)
    psi_ref, n_ref, r_ref, _ = ref_op.solve(chi_full, solver=oracle.RELIABLE_BICGSTAB, rsd=1e-10, maxit=5000,
                                            delta=0.1)
    psi0, res0 = op.solve(chi_full[V // 2:], solver=cmx.RELIABLE_BICGSTAB, rsd=1e-10, maxit=5000, delta=0.1,
                          sloppy_prec=sloppy)
    say("  n_count", res0["n_count"], "oracle", n_ref, "rel resid", res0["rel_resid"], "updates", res0["n_reliable"])
    assert res0["n_reliable"] > 0
    if sloppy == cmx.SINGLE:  # the same arithmetic as the oracle up to the order of the sums
        assert abs(res0["n_count"] - n_ref) <= max(3, n_ref // 10), (res0["n_count"], n_ref)
    assert res0["rel_resid"] < 1e-9
    psi, res = op.solve(chi_full[V // 2:], solver=cmx.RELIABLE_BICGSTAB, rsd=1e-10, maxit=5000, delta=0.1,
                        sloppy_prec=sloppy, true_residual=True)
    assert res["rel_resid"] <= 1e-10
    assert rel(psi, psi_cg[V // 2:]) < 1e-8


@pytest.mark.parametrize("solver", [cmx.CG, cmx.BICGSTAB, cmx.RELIABLE_CG, cmx.RELIABLE_BICGSTAB])
def test_mdagm_solvers(solver):
    """MdagM system solvers (M^dag M psi = chi) against the oracle's, from a non-zero initial guess."""
    L = (8, 8, 8, 16)
    say("mdagm", solver)
    V = int(np.prod(L))
    U = gauge_field(L, 95)
    mass = cmx.kappa_to_mass(0.125)
    ref_op = oracle.Op(oracle.EO_CLOVER, L, U, mass, 1.0, 1.0)
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_CLOVER, mass, 1.0)
    chi_full = random_fermion(V, 14)
    chi_full[: V // 2] = 0
    guess = 0.05 * random_fermion(V, 15)
    guess[: V // 2] = 0
    osolver = {cmx.CG: oracle.CG, cmx.BICGSTAB: oracle.BICGSTAB, cmx.RELIABLE_CG: oracle.RELIABLE_CG,
               cmx.RELIABLE_BICGSTAB: oracle.RELIABLE_BICGSTAB}[solver]
    psi_ref, n_ref, r_ref = ref_op.solve_mdagm(chi_full, psi=guess, solver=osolver, rsd=1e-10, maxit=5000)
    psi, res = op.solve(chi_full[V // 2:], psi=guess[V // 2:], solver=solver, rsd=1e-10, maxit=5000,
                        sloppy_prec=cmx.SINGLE, mdagm=True)
    say("  n_count", res["n_count"], "oracle", n_ref, "resid", res["resid"], "oracle", r_ref)
    if solver == cmx.CG:
        assert abs(res["n_count"] - n_ref) <= 1, (res["n_count"], n_ref)
    assert rel(psi, psi_ref[V // 2:]) < 1e-7
    # the two-step forms bound |chi - M^dag M psi| by about (1 + cond M) x RsdTarget
    assert res["rel_resid"] < (1e-9 if solver in (cmx.CG, cmx.RELIABLE_CG) else 1e-8)


def _odd_op(ref_op, V):
    def A(x):
        full = np.zeros((V, 4, 3), np.complex128)
        full[V // 2:] = x
        return ref_op(full)[V // 2:]
    return A


@pytest.mark.parametrize("n_defl,precond", [(0, 0), (3, 0), (0, 1), (3, 1)])
def test_fgmres_dr_vs_oracle(n_defl, precond):
    """FGMRES-DR with an FP64 basis against the oracle's restatement: cycle structure, iteration count
    and solution (the library orthogonalises by classical Gram-Schmidt twice, the oracle by modified
    Gram-Schmidt, equal to rounding)."""
    from oracle import krylov
    L = (8, 8, 8, 16)
    say("fgmres-dr", n_defl, precond)
    V = int(np.prod(L))
    U = gauge_field(L, 97)
    mass = cmx.kappa_to_mass(0.125)
    ref_op = oracle.Op(oracle.EO_CLOVER, L, U, mass, 1.0, 1.0)
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_CLOVER, mass, 1.0)
    chi = random_fermion(V, 16)[V // 2:]
    A = _odd_op(ref_op, V)
    pre = None if precond == 0 else (lambda z0, v: krylov.mr(A, v, z0, 1e-12, 3)[0])
    psi_ref, n_ref, r_ref, cyc_ref = krylov.fgmres_dr(A, chi, np.zeros_like(chi), 1e-10, n_krylov=12,
                                                      n_defl=n_defl, max_iter=3000, precond=pre)
    psi, res = op.solve(chi, solver=cmx.FGMRESDR, rsd=1e-10, maxit=3000, sloppy_prec=cmx.DOUBLE, n_krylov=12,
                        n_defl=n_defl, precond=precond, precond_iters=3, precond_rsd=1e-12, reorth=2, verbose=True)
    say("  n_count", res["n_count"], "oracle", n_ref, "cycles", res["n_reliable"], "oracle", cyc_ref,
        "rel resid", res["rel_resid"])
    assert abs(res["n_count"] - n_ref) <= 1, (res["n_count"], n_ref)
    assert res["rel_resid"] <= 1e-10
    assert rel(psi, psi_ref) < 1e-8


@pytest.mark.parametrize("basis", [cmx.DOUBLE, cmx.HALF])
def test_fgmres_dr_deflation_after_short_cycles(basis):
    """NDefl = 8 with cycles that end after a residual reduction of 2 (fewer Arnoldi vectors than NDefl):
    the deflated restart keeps at most as many harmonic Ritz vectors as the cycle built.  FP64 basis:
    against the oracle's restatement; FP16 operator and basis: reaches the FP64 true residual 1e-10 and
    agrees with the oracle's solution."""
    from oracle import krylov
    L = (8, 8, 8, 16)
    say("fgmres-dr short cycles", basis)
    V = int(np.prod(L))
    U = gauge_field(L, 96)
    mass = cmx.kappa_to_mass(0.125)
    ref_op = oracle.Op(oracle.EO_CLOVER, L, U, mass, 1.0, 1.0)
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_CLOVER, mass, 1.0, sloppy_prec=cmx.HALF)
    chi = random_fermion(V, 17)[V // 2:]
    A = _odd_op(ref_op, V)
    rec = {}
    psi_ref, n_ref, r_ref, cyc_ref = krylov.fgmres_dr(A, chi, np.zeros_like(chi), 1e-10, n_krylov=12, n_defl=8,
                                                      max_iter=3000, record=rec, cycle_reduction=0.5)
    assert min(c["dim"] - c["nd"] for c in rec["cycles"]) < 8
    sp = cmx.DOUBLE if basis == cmx.DOUBLE else cmx.HALF
    psi, res = op.solve(chi, solver=cmx.FGMRESDR, rsd=1e-10, maxit=3000, sloppy_prec=sp, basis_prec=basis,
                        n_krylov=12, n_defl=8, reorth=2, cycle_reduction=0.5)
    say("  n_count", res["n_count"], "oracle", n_ref, "cycles", res["n_reliable"], "oracle", cyc_ref,
        "rel resid", res["rel_resid"])
    assert res["rel_resid"] <= 1e-10
    assert rel(psi, psi_ref) < 1e-8
    if basis == cmx.DOUBLE:
        assert abs(res["n_count"] - n_ref) <= 2, (res["n_count"], n_ref)


@pytest.mark.parametrize("sloppy", [cmx.SINGLE, cmx.HALF])
@pytest.mark.parametrize("n_defl", [0, 4])
def test_fgmres_dr_sloppy_basis(sloppy, n_defl):
    """FP32 / FP16 Krylov basis with the FP64 residual recomputed every cycle reaches a 1e-10 FP64 true
    residual; the solution agrees with the FP64 CG solution."""
    L = (8, 8, 8, 16)
    say("fgmres-dr sloppy", sloppy, n_defl)
    V = int(np.prod(L))
    U = cmx.gauge_apply_bc(L, cmx.gauge_generate(L, kind=2, seed=99, nstout=6, rho=0.1), (1, 1, 1, -1))
    mass = cmx.kappa_to_mass(0.1350)
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_CLOVER, mass, 1.0, sloppy_prec=cmx.HALF if sloppy == cmx.HALF else cmx.SINGLE)
    chi = cmx.z2_source((V // 2, 4, 3), 6)
    psi_cg, r_cg = op.solve(chi, solver=cmx.CG, rsd=1e-12, maxit=20000)
    psi, res = op.solve(chi, solver=cmx.FGMRESDR, rsd=1e-10, maxit=5000, sloppy_prec=sloppy, n_krylov=16,
                        n_defl=n_defl, precond=cmx.PRECOND_MR, precond_iters=4, precond_rsd=1e-12)
    say("  n_count", res["n_count"], "cycles", res["n_reliable"], "rel resid", res["rel_resid"],
        "seconds", res["solve_seconds"])
    assert res["rel_resid"] <= 1e-10
    assert rel(psi, psi_cg) < 1e-8


@pytest.mark.parametrize("n_defl", [0, 4])
def test_fgmres_dr_wide_basis_and_repeat(n_defl):
    """FP16 operator with an FP32 Krylov basis, plain and deflated restarts (the deflated basis is
    formed at the basis precision), solved twice on the operator's persistent workspace with identical
    results."""
    L = (8, 8, 8, 16)
    say("fgmres-dr wide basis", n_defl)
    V = int(np.prod(L))
    U = cmx.gauge_apply_bc(L, cmx.gauge_generate(L, kind=2, seed=98, nstout=6, rho=0.1), (1, 1, 1, -1))
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_CLOVER, cmx.kappa_to_mass(0.1350), 1.0, sloppy_prec=cmx.HALF)
    chi = cmx.z2_source((V // 2, 4, 3), 7)
    psi_cg, _ = op.solve(chi, solver=cmx.CG, rsd=1e-12, maxit=20000)
    kw = dict(solver=cmx.FGMRESDR, rsd=1e-10, maxit=5000, sloppy_prec=cmx.HALF, basis_prec=cmx.SINGLE,
              n_krylov=16, n_defl=n_defl, cycle_reduction=1e-3)
    psi1, r1 = op.solve(chi, **kw)
    psi2, r2 = op.solve(chi, **kw)
    say("  n_count", r1["n_count"], "cycles", r1["n_reliable"], "rel resid", r1["rel_resid"])
    assert r1["rel_resid"] <= 1e-10
    assert rel(psi1, psi_cg) < 1e-8
    assert r2["n_count"] == r1["n_count"]
    assert np.array_equal(psi2.view(np.uint64), psi1.view(np.uint64))


@pytest.mark.parametrize("L", [(8, 8, 8, 16), (16, 16, 8, 16), (6, 4, 8, 10)])
@pytest.mark.parametrize("nrhs", [1, 5, 12])
def test_dslash_multi_source(L, nrhs):
    """The multi-source hop (links loaded once per group of sources) against the oracle, per source."""
    say("dslash multi", L, nrhs)
    V = int(np.prod(L))
    Vh = V // 2
    U = gauge_field(L, 21)
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_WILSON, 0.1, sloppy_prec=cmx.HALF, sloppy_recon=12)
    psis = np.stack([random_fermion(V, 100 + k) for k in range(nrhs)])
    refs = {}
    for fast, block, group in [(1, 256, 0), (3, 256, 0), (1, 256, 1), (1, 64, 2), (3, 128, 3), (3, 256, 4)]:
        # link loads default / streamed; sources per lane: tuned, 1..4
        op.set_hop_config(cmx.HALF, block, 0, 1, 0, fast)
        op.set_multi_group(group)
        for cb in (0, 1):
            for isign in (+1, -1):
                src = np.ascontiguousarray(psis[:, (1 - cb) * Vh:(2 - cb) * Vh])
                got = op.dslash_multi(src, isign, cb, cmx.HALF)
                for k in range(nrhs):
                    if (cb, isign, k) not in refs:
                        refs[cb, isign, k] = oracle.dslash(L, U, psis[k], isign, cb)[cb * Vh:(cb + 1) * Vh]
                    assert rel(got[k], refs[cb, isign, k]) < TOL[cmx.HALF], (fast, block, group, cb, isign, k)
                # every grouping performs the single-source arithmetic per source
                single = op.dslash(src[0], isign, cb, cmx.HALF)
                assert np.array_equal(got[0].view(np.uint64), single.view(np.uint64))


@pytest.mark.parametrize("L", [(8, 8, 8, 16), (16, 16, 8, 16), (6, 4, 8, 10), (12, 12, 12, 24)])
@pytest.mark.parametrize("nrhs", [1, 5, 12])
def test_dslash_multi_source_mfma(L, nrhs):
    """The multi-source hop with the link products on the matrix cores (cmx_op_set_multi_group 5): every
    source against the oracle at the FP16 tolerance and against the single-source kernel (the same FP16
    operands; products summed in FP32 in another order), both site orders, both link cache policies."""
    say("dslash multi mfma", L, nrhs)
    V = int(np.prod(L))
    Vh = V // 2
    U = gauge_field(L, 22)
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_WILSON, 0.1, sloppy_prec=cmx.HALF, sloppy_recon=12)
    psis = np.stack([random_fermion(V, 200 + k) for k in range(nrhs)])
    op.set_multi_group(5)
    tiled_ok = L[3] % 8 == 0
    for fast, tiled in [(1, 0), (3, 0)] + ([(1, 1), (3, 1)] if tiled_ok else []):
        op.set_hop_config(cmx.HALF, 256, tiled, 1, 0, fast, target=1)
        for cb in (0, 1):
            for isign in (+1, -1):
                src = np.ascontiguousarray(psis[:, (1 - cb) * Vh:(2 - cb) * Vh])
                got = op.dslash_multi(src, isign, cb, cmx.HALF)
                for k in range(nrhs):
                    ref = oracle.dslash(L, U, psis[k], isign, cb)[cb * Vh:(cb + 1) * Vh]
                    assert rel(got[k], ref) < TOL[cmx.HALF], (fast, tiled, cb, isign, k)
                    single = op.dslash(src[k], isign, cb, cmx.HALF)
                    assert rel(got[k], single) < 1e-3, (fast, tiled, cb, isign, k)


@pytest.mark.parametrize("solver,sloppy", [(cmx.CG, cmx.DOUBLE), (cmx.BICGSTAB, cmx.DOUBLE),
                                           (cmx.RELIABLE_CG, cmx.HALF), (cmx.RELIABLE_CG, cmx.SINGLE)])
def test_graph_replay_matches_eager(solver, sloppy):
    """Iterations replayed from captured HIP graphs give the eager result bit for bit (same kernels,
    same deterministic reductions), with the same iteration count; run twice to replay each graph."""
    L = (8, 8, 8, 16)
    say("graph replay", solver, sloppy)
    V = int(np.prod(L))
    U = cmx.gauge_apply_bc(L, cmx.gauge_generate(L, kind=2, seed=91, nstout=6, rho=0.1), (1, 1, 1, -1))
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_CLOVER, cmx.kappa_to_mass(0.1350), 1.0,
                      sloppy_prec=cmx.HALF if sloppy == cmx.HALF else cmx.SINGLE)
    chi = cmx.z2_source((V // 2, 4, 3), 2)
    kw = dict(solver=solver, rsd=1e-10, maxit=20000, delta=0.1, sloppy_prec=sloppy)
    psi0, r0 = op.solve(chi, graph=False, **kw)
    for _ in range(2):
        psi1, r1 = op.solve(chi, graph=True, **kw)
        say("  eager", r0["n_count"], r0["solve_seconds"], "graph", r1["n_count"], r1["solve_seconds"])
        assert r1["n_count"] == r0["n_count"]
        assert np.array_equal(psi1.view(np.uint64), psi0.view(np.uint64))


def test_twisted_golden_through_gpu():
    """prec_wilson-twisted (tests/chroma/hadron/propagator/prec_wilson-twisted.{ini,out}.xml) on the GPU."""
    from chroma_amd import qprop as qp
    import json
    import os
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "prec_wilson_twisted.json")))
    L = tuple(gold["nrow"])
    say("twisted golden")
    V = int(np.prod(L))
    U = cmx.gauge_generate(L, kind=0)
    U = cmx.gauge_apply_bc(L, U, (1, 1, 1, 1), twist_by_pi=gold["phases_by_pi"], twist_dir=gold["phases_dir"])
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_WILSON, cmx.kappa_to_mass(gold["kappa"]))
    qprop = qp.PrecFermActQprop(op, cmx.CG, rsd=gold["RsdCG"], maxit=gold["MaxCG"])
    prop, results = qp.quark_prop4(qprop, V, 0)
    corr = qp.prop_corr(L, prop)
    assert np.max(np.abs(corr - np.array(gold["prop_corr"])) / np.array(gold["prop_corr"])) < gold["tolerance"]
    assert all(abs(r["n_count"] - gold["n_count"]) <= 1 for r in results)


@pytest.mark.parametrize("prec", [cmx.DOUBLE, cmx.SINGLE, cmx.HALF])
@pytest.mark.parametrize("cfg", [(64, 0, 1, 0, 0), (64, 1, 1, 0, 0), (64, 1, 2, 1, 0), (128, 1, 4, 1, 0),
                                 (128, 0, 1, 1, 0), (128, 1, 8, 0, 0), (256, 1, 2, 1, 0), (64, 0, 1, 0, 1),
                                 (128, 0, 1, 0, 1), (256, 0, 1, 0, 1), (64, 1, 4, 0, 1), (256, 1, 2, 0, 1),
                                 (128, 1, 8, 0, 1), (256, 0, 1, 0, 3), (128, 1, 8, 0, 3), (64, 1, 2, 0, 2),
                                 (128, 0, 1, 0, 2)])
def test_hop_launch_shapes(prec, cfg):
    """Every launch shape the tuner may pick (block size, tiled work-group order, prefetch, plane-grid
    kernel with or without streamed link loads) gives the same operator as the oracle; 16x16x8x16 admits the tiling for blocks of 64 and
    128, and its antiperiodic links admit the plane-grid kernel."""
    L = (16, 16, 8, 16)
    say("launch shape", prec, cfg)
    V = int(np.prod(L))
    U = gauge_field(L, 101)
    mass = cmx.kappa_to_mass(0.13)
    ref_op = oracle.Op(oracle.EO_CLOVER, L, U, mass, 1.0, 1.0)
    sp = cmx.HALF if prec == cmx.HALF else cmx.SINGLE
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_CLOVER, mass, 1.0, sloppy_prec=sp)
    op.set_hop_config(prec, *cfg)
    assert op.link_recon(cmx.DOUBLE) == 12
    assert op.hop_config(prec)["fast"] == cfg[4]  # SU(3) links: every precision admits the plane kernel
    psi = random_fermion(V, 17)
    for isign in (+1, -1):
        ref = ref_op(psi, isign)[V // 2:]
        assert rel(op(psi[V // 2:], isign, prec), ref) < TOL[prec]
    ref = oracle.dslash(L, U, psi, +1, 1)[V // 2:]
    assert rel(op.dslash(psi[: V // 2], +1, 1, prec), ref) < TOL[prec]


def test_fp64_links_not_su3_keep_18_reals():
    """FP64 links that are not SU(3) to rounding (here perturbed by 1e-9) are read as 18 reals by the
    site-linear kernel, also when the plane-grid kernel is asked for; the operator matches the oracle."""
    L = (8, 8, 8, 16)
    say("fp64 non-su3 links")
    V = int(np.prod(L))
    U = gauge_field(L, 103)
    U = U + 1e-9 * np.random.default_rng(5).standard_normal(U.shape)
    mass = cmx.kappa_to_mass(0.13)
    ref_op = oracle.Op(oracle.EO_CLOVER, L, U, mass, 1.0, 1.0)
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_CLOVER, mass, 1.0, sloppy_prec=cmx.SINGLE)
    assert op.link_recon(cmx.DOUBLE) == 18
    op.set_hop_config(cmx.DOUBLE, 128, 1, 8, 0, 3)
    assert op.hop_config(cmx.DOUBLE)["fast"] == 2
    psi = random_fermion(V, 19)
    for isign in (+1, -1):
        assert rel(op(psi[V // 2:], isign, cmx.DOUBLE), ref_op(psi, isign)[V // 2:]) < TOL[cmx.DOUBLE]


def test_autotune_keeps_results():
    L = (16, 16, 8, 16)
    say("autotune")
    V = int(np.prod(L))
    U = gauge_field(L, 102)
    mass = cmx.kappa_to_mass(0.13)
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_CLOVER, mass, 1.0, sloppy_prec=cmx.HALF, autotune=True)
    ref = oracle.Op(oracle.EO_CLOVER, L, U, mass, 1.0, 1.0)
    psi = random_fermion(V, 18)
    for prec in (cmx.DOUBLE, cmx.HALF):
        assert rel(op(psi[V // 2:], +1, prec), ref(psi)[V // 2:]) < TOL[prec]


@pytest.mark.parametrize("isign", [+1, -1])
@pytest.mark.parametrize("cb", [0, 1, -1])
def test_dslash_deriv(isign, cb):
    """WilsonDslashBase::deriv (lwldslash_base_w.h:132-227) on the device against the oracle (FP64)."""
    L = (4, 4, 6, 8)
    say("dslash deriv", isign, cb)
    V = int(np.prod(L))
    U = gauge_field(L, 70)
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_WILSON, 0.1)
    chi = random_fermion(V, 71)
    psi = random_fermion(V, 72)
    got = op.dslash_deriv(chi, psi, isign, cb)
    ref = oracle.dslash_deriv(L, chi, psi, isign, cb)
    assert rel(got, ref) < 1e-13
    if cb >= 0:
        assert not np.any(got[:, (1 - cb) * (V // 2):(2 - cb) * (V // 2)])
