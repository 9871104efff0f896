"""GPU: every benchmark configuration at its stated lattice size, through the C-ABI.

c1  8^3x16 unpreconditioned Wilson CGNE, kappa 0.115, FP64, point source at the origin: iteration count
    within 1 of the oracle's InvCG2 (invcg2.cc:70-232) and the solution to 1e-8.
c2  32^3x64 EO clover (Haar + 6 stout steps, rho 0.1), csw 1, kappa 0.1350: the operator in FP64 / FP32 /
    FP16 storage against the oracle (1e-12 / 1e-6 / 2e-3), and mixed FP16-storage reliable CG
    (reliable_cg.cc:10-188) to the FP64 true residual 1e-10, the residual recomputed by the oracle.
c3A 48^3x96 EO Wilson (exp(0.3 i H) field), kappa 0.115: the 12 spin-colour point sources of
    quarkProp4_a (quarkprop4_w.cc:70-117) batched, each equal bit for bit to its sequential solve and at
    the FP64 true residual 1e-10 (one recomputed by the oracle).
c3B 64 solves on 8^4 / 12^3x24 / 16^3x32, own field (seed 1000+i) and kappa ~ U[0.120, 0.134] (seed 999):
    batched equal to sequential bit for bit; the 22 8^4 solutions against the oracle's CG to 1e-8.
c4  32^3x64 clover kappa 0.1355: FGMRES-DR(16) with the trained network to the FP64 true residual 1e-10,
    fewer iterations than without the network, solution equal to FP64 CG's to 1e-8.
c5  64^3x128 clover kappa 0.1357: one Z2 source, plain mixed CG and the learned FGMRES-DR, each to the
    FP64 true residual 1e-10.

Each test prints its case before it starts; sizes are the configurations' own.
"""
import sys
import time

import numpy as np
import pytest

import oracle
from chroma_amd import lib as cmx
from tests.util import rel, random_fermion

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def stout_field(L, seed):
    """Haar SU(3) start, 6 isotropic stout steps with rho = 0.1 (device), antiperiodic in time."""
    U = cmx.gauge_stout_smear(L, cmx.gauge_generate(L, kind=2, seed=seed), 6, 0.1)
    return cmx.gauge_apply_bc(L, U, (1, 1, 1, -1))


def c1_field(L, seed):
    """exp(0.3 i H) links, periodic in space, antiperiodic in time."""
    return cmx.gauge_apply_bc(L, cmx.gauge_generate(L, kind=1, seed=seed, eps=0.3), (1, 1, 1, -1))


def point_sources(Vh):
    """The 12 spin-colour unit sources on the first odd site."""
    out = []
    for s in range(4):
        for c in range(3):
            chi = np.zeros((Vh, 4, 3), np.complex128)
            chi[0, s, c] = 1.0
            out.append(chi)
    return out


def oracle_rel_resid(ref_op, chi_odd, psi_odd):
    """|chi - A psi| / |chi| in FP64 by the oracle's operator (full-lattice arrays, odd half used)."""
    Vh = chi_odd.shape[0]
    full = np.zeros((2 * Vh, 4, 3), np.complex128)
    full[Vh:] = psi_odd
    r = chi_odd - ref_op(full)[Vh:]
    return float(np.linalg.norm(r) / np.linalg.norm(chi_odd))


def test_c1_wilson_cgne_8x8x8x16():
    L = (8, 8, 8, 16)
    say("c1: unprec Wilson CGNE", L)
    V = int(np.prod(L))
    U = c1_field(L, 1)
    mass = cmx.kappa_to_mass(0.115)
    ref_op = oracle.Op(oracle.UNPREC_WILSON, L, U, mass)
    op = cmx.Operator(cmx.Gauge(L, U), cmx.UNPREC_WILSON, mass)
    chi = np.zeros((V, 4, 3), np.complex128)
    chi[0, 0, 0] = 1.0
    psi_ref, n_ref, _, _ = ref_op.solve(chi, solver=oracle.CG, rsd=1e-10, maxit=5000)
    psi, res = op.solve(chi, solver=cmx.CG, rsd=1e-10, maxit=5000)
    say("  n_count", res["n_count"], "oracle", n_ref, "rel resid", res["rel_resid"], "err", rel(psi, psi_ref))
    assert abs(res["n_count"] - n_ref) <= 1, (res["n_count"], n_ref)
    assert rel(psi, psi_ref) < 1e-8
    # the FP64 true residual 1e-10 of the configuration (restarted CG where the normal-equation stop is short)
    psi_t, res_t = op.solve(chi, solver=cmx.CG, rsd=1e-10, maxit=5000, true_residual=True)
    assert res_t["rel_resid"] <= 1e-10
    assert rel(psi_t, psi_ref) < 1e-8


@pytest.fixture(scope="module")
def c2_setup():
    L = (32, 32, 32, 64)
    t0 = time.time()
    U = stout_field(L, 1)
    mass = cmx.kappa_to_mass(0.1350)
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_CLOVER, mass, 1.0, sloppy_prec=cmx.HALF)
    ref_op = oracle.Op(oracle.EO_CLOVER, L, U, mass, 1.0, 1.0)
    say(f"  c2 setup {time.time() - t0:.1f} s, plaquette {cmx.plaquette(L, U):.4f}")
    del U
    return L, op, ref_op


def test_c2_eo_clover_operator_32x32x32x64(c2_setup):
    L, op, ref_op = c2_setup
    say("c2: EO clover operator", L)
    V = int(np.prod(L))
    psi = random_fermion(V, 5)
    for isign in (+1, -1):
        ref = ref_op(psi, isign)[V // 2:]
        for prec, tol in [(cmx.DOUBLE, 1e-12), (cmx.SINGLE, 1e-6), (cmx.HALF, 2e-3)]:
            err = rel(op(psi[V // 2:], isign, prec), ref)
            say("  isign", isign, "prec", prec, "rel err", err)
            assert err < tol


def test_c2_mixed_reliable_cg_32x32x32x64(c2_setup):
    L, op, ref_op = c2_setup
    say("c2: mixed FP16/FP32 reliable CG", L)
    Vh = int(np.prod(L)) // 2
    chi = cmx.z2_source((Vh, 4, 3), 2)
    psi, res = op.solve(chi, solver=cmx.RELIABLE_CG, rsd=1e-10, maxit=20000, delta=0.1, sloppy_prec=cmx.HALF,
                        true_residual=True)
    psi64, res64 = op.solve(chi, solver=cmx.CG, rsd=1e-12, maxit=20000, true_residual=True)
    r_or = oracle_rel_resid(ref_op, chi, psi)
    say("  mixed", res["n_count"], "it, resid", res["rel_resid"], "oracle resid", r_or, "updates",
        res["n_reliable"], "| FP64 CG", res64["n_count"], "it, diff", rel(psi, psi64))
    assert res["rel_resid"] <= 1e-10
    assert r_or <= 1.05e-10
    assert rel(psi, psi64) < 1e-8


def test_c3a_twelve_sources_48x48x48x96():
    L = (48, 48, 48, 96)
    say("c3A: 12 point sources batched", L)
    Vh = int(np.prod(L)) // 2
    U = c1_field(L, 1)
    mass = cmx.kappa_to_mass(0.115)
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_WILSON, mass, sloppy_prec=cmx.HALF)
    ref_op = oracle.Op(oracle.EO_WILSON, L, U, mass)
    del U
    chis = point_sources(Vh)
    kw = dict(solver=cmx.RELIABLE_CG, rsd=1e-10, maxit=5000, delta=0.1, sloppy_prec=cmx.HALF, true_residual=True)
    seq = [op.solve(c, **kw) for c in chis]
    psis, res, summ = cmx.solve_batch([op] * 12, chis, tune=False, **kw)
    say("  batched", summ)
    for k in range(12):
        say("  source", k, "seq", seq[k][1]["n_count"], seq[k][1]["rel_resid"], "batch", res[k]["n_count"],
            res[k]["rel_resid"])
        assert res[k]["rel_resid"] <= 1e-10
        assert res[k]["n_count"] == seq[k][1]["n_count"]
        assert np.array_equal(psis[k].view(np.uint64), seq[k][0].view(np.uint64))
    r_or = oracle_rel_resid(ref_op, chis[0], psis[0])
    say("  oracle true residual of source 0", r_or)
    assert r_or <= 1.05e-10


def test_c3b_64_uneven_volumes():
    say("c3B: 64 solves on uneven volumes")
    rng = np.random.default_rng(999)
    kap = rng.uniform(0.120, 0.134, 64)
    lats = [(8, 8, 8, 8), (12, 12, 12, 24), (16, 16, 16, 32)]
    ops, chis, fields = [], [], []
    for i in range(64):
        L = lats[i % 3]
        U = c1_field(L, 1000 + i)
        ops.append(cmx.Operator(cmx.Gauge(L, U), cmx.EO_WILSON, cmx.kappa_to_mass(kap[i]), sloppy_prec=cmx.HALF))
        fields.append(U if i % 3 == 0 else None)
        chi = np.zeros((int(np.prod(L)) // 2, 4, 3), np.complex128)
        chi[0, 0, 0] = 1.0
        chis.append(chi)
    kw = dict(solver=cmx.RELIABLE_CG, rsd=1e-10, maxit=5000, delta=0.1, sloppy_prec=cmx.HALF, true_residual=True)
    seq = [op.solve(c, **kw) for op, c in zip(ops, chis)]
    psis, res, summ = cmx.solve_batch(ops, chis, tune=False, **kw)
    say("  batched", summ)
    for k in range(64):
        assert res[k]["rel_resid"] <= 1e-10, (k, res[k])
        assert res[k]["n_count"] == seq[k][1]["n_count"], k
        assert np.array_equal(psis[k].view(np.uint64), seq[k][0].view(np.uint64)), k
    worst = 0.0
    for k in range(0, 64, 3):
        L = lats[0]
        V = int(np.prod(L))
        ref_op = oracle.Op(oracle.EO_WILSON, L, fields[k], cmx.kappa_to_mass(kap[k]))
        chi_full = np.zeros((V, 4, 3), np.complex128)
        chi_full[V // 2:] = chis[k]
        psi_ref, _, _, _ = ref_op.solve(chi_full, solver=oracle.CG, rsd=1e-12, maxit=5000)
        worst = max(worst, rel(psis[k], psi_ref[V // 2:]))
    say("  worst 8^4 error against the oracle", worst)
    assert worst < 1e-8


@pytest.fixture(scope="module")
def trained_net():
    """The network trained on four 16^3x32 fields of the stout recipe (seeds 10-13) at kappa 0.1355."""
    mass = cmx.kappa_to_mass(0.1355)
    net = cmx.Net()
    net.init_neumann(4.0 + mass)
    TL = (16, 16, 16, 32)
    for k in range(4):
        top = cmx.Operator(cmx.Gauge(TL, stout_field(TL, 10 + k)), cmx.EO_CLOVER, mass, 1.0, sloppy_prec=cmx.HALF)
        cmx.net_attach(top, net)
        t = cmx.net_train(top, n_sources=8, snapshot_every=1, pairs_per_source=4, normalize=False, epochs=4,
                          lr=1e-3, seed=100 + k, prec=cmx.HALF)
        say("  train field", k, t)
    return net


FGMRES = dict(solver=cmx.FGMRESDR, rsd=1e-10, maxit=5000, sloppy_prec=cmx.HALF, basis_prec=cmx.HALF, n_krylov=16,
              cycle_reduction=3e-3)


def test_c4_learned_32x32x32x64(trained_net):
    L = (32, 32, 32, 64)
    say("c4: learned FGMRES-DR", L)
    Vh = int(np.prod(L)) // 2
    op = cmx.Operator(cmx.Gauge(L, stout_field(L, 1)), cmx.EO_CLOVER, cmx.kappa_to_mass(0.1355), 1.0,
                      sloppy_prec=cmx.HALF)
    cmx.net_attach(op, trained_net)
    chi = cmx.z2_source((Vh, 4, 3), 2)
    psi_l, rl = op.solve(chi, precond=cmx.PRECOND_LEARNED, **FGMRES)
    psi_n, rn = op.solve(chi, precond=cmx.PRECOND_NONE, **FGMRES)
    psi64, r64 = op.solve(chi, solver=cmx.CG, rsd=1e-12, maxit=20000, true_residual=True)
    say("  learned", rl["n_count"], "it", rl["rel_resid"], "| no network", rn["n_count"], "it", rn["rel_resid"],
        "| diff to FP64 CG", rel(psi_l, psi64))
    assert rl["rel_resid"] <= 1e-10 and rn["rel_resid"] <= 1e-10
    assert rl["n_count"] < rn["n_count"]
    assert rel(psi_l, psi64) < 1e-8


def test_c5_combination_64x64x64x128(trained_net):
    L = (64, 64, 64, 128)
    say("c5: plain and learned", L)
    Vh = int(np.prod(L)) // 2
    t0 = time.time()
    op = cmx.Operator(cmx.Gauge(L, stout_field(L, 1)), cmx.EO_CLOVER, cmx.kappa_to_mass(0.1357), 1.0,
                      sloppy_prec=cmx.HALF)
    say(f"  setup {time.time() - t0:.1f} s")
    cmx.net_attach(op, trained_net)
    chi = cmx.z2_source((Vh, 4, 3), 2)
    psi_p, rp = op.solve(chi, solver=cmx.RELIABLE_CG, rsd=1e-10, maxit=20000, delta=0.1, sloppy_prec=cmx.HALF,
                         true_residual=True)
    psi_l, rl = op.solve(chi, precond=cmx.PRECOND_LEARNED, **FGMRES)
    say("  plain", rp["n_count"], "it", rp["rel_resid"], f"{rp['solve_seconds']:.3f} s | learned", rl["n_count"],
        "it", rl["rel_resid"], f"{rl['solve_seconds']:.3f} s | diff", rel(psi_l, psi_p))
    assert rp["rel_resid"] <= 1e-10 and rl["rel_resid"] <= 1e-10
    # two solutions with residuals <= 1e-10 of an operator of modest condition number
    assert rel(psi_l, psi_p) < 1e-7
