"""GPU: a zero source, as the reference's solvers meet it.  InvCG2 (invcg2.cc:70-232) and FGMRES-DR
(syssolver_linop_fgmres_dr.cc:78-774) stop before the first iteration with psi = 0; BiCGStab
(invbicgstab.cc:60-70, rho = 0), reliable CG (reliable_cg.cc:53-70, |A p| = 0) and reliable BiCGStab
end in QDP_abort, which the library reports as an error.  The oracle shows the same split (n_count 0
against -1), and the operator solves normally afterwards."""
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


def test_zero_source():
    L = (4, 4, 4, 8)
    V = int(np.prod(L))
    Vh = V // 2
    U = cmx.gauge_apply_bc(L, cmx.gauge_generate(L, kind=1, seed=4, eps=0.3), (1, 1, 1, -1))
    mass, csw = cmx.kappa_to_mass(0.12), 1.0
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_CLOVER, mass, csw)
    ref = oracle.Op(oracle.EO_CLOVER, L, U, mass, csw, csw)
    zero = np.zeros((Vh, 4, 3), complex)
    full = np.zeros((V, 4, 3), complex)
    for solver, osolver, sloppy in [(cmx.CG, oracle.CG, cmx.DOUBLE), (cmx.BICGSTAB, oracle.BICGSTAB, cmx.DOUBLE),
                                    (cmx.RELIABLE_CG, oracle.RELIABLE_CG, cmx.HALF),
                                    (cmx.RELIABLE_BICGSTAB, oracle.RELIABLE_BICGSTAB, cmx.HALF)]:
        say("zero source", solver)
        _, n_ref, _, _ = ref.solve(full, solver=osolver, rsd=1e-10, maxit=50)
        if n_ref == 0:
            psi, res = op.solve(zero, solver=solver, rsd=1e-10, maxit=50, sloppy_prec=sloppy)
            assert res["n_count"] == 0 and not np.any(psi)
        else:
            assert n_ref == -1
            with pytest.raises(cmx.CmxError):
                op.solve(zero, solver=solver, rsd=1e-10, maxit=50, sloppy_prec=sloppy)
    say("zero source FGMRES-DR")
    psi, res = op.solve(zero, solver=cmx.FGMRESDR, rsd=1e-10, maxit=50, n_krylov=8, sloppy_prec=cmx.HALF,
                        basis_prec=cmx.HALF)
    assert res["n_count"] == 0 and not np.any(psi)
    # the operator is intact: a normal solve follows and matches the oracle's
    chi = np.random.default_rng(5).standard_normal((Vh, 4, 3)) + 0j
    psi, res = op.solve(chi, solver=cmx.BICGSTAB, rsd=1e-10, maxit=1000)
    full[Vh:] = chi
    psi_ref, _, _, _ = ref.solve(full, solver=oracle.BICGSTAB, rsd=1e-10, maxit=1000)
    assert rel(psi, psi_ref[Vh:]) < 1e-8
