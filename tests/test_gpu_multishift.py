"""GPU: multishift CG (MInvCG2, minvcg2.cc:72-355, through cmx_solve_multishift) against the oracle's
restatement and against single-shift solves."""
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


def field(L, seed):
    U = cmx.gauge_generate(L, kind=1, seed=seed, eps=0.3)
    return cmx.gauge_apply_bc(L, U, (1, 1, 1, -1))


@pytest.mark.parametrize("kind", [cmx.EO_CLOVER, cmx.EO_WILSON])
def test_multishift_vs_oracle(kind):
    L = (4, 4, 4, 8)
    say("multishift vs oracle", kind)
    V = int(np.prod(L))
    U = field(L, 60)
    mass = cmx.kappa_to_mass(0.12)
    csw = 1.0 if kind == cmx.EO_CLOVER else 0.0
    op = cmx.Operator(cmx.Gauge(L, U), kind, mass, csw)
    ref_op = oracle.Op(oracle.EO_CLOVER if kind == cmx.EO_CLOVER else oracle.EO_WILSON, L, U, mass, csw, csw)
    rng = np.random.default_rng(61)
    chi = rng.standard_normal((V // 2, 4, 3)) + 1j * rng.standard_normal((V // 2, 4, 3))
    shifts = [0.0, 0.002, 0.05, 0.4]
    psi, res = op.solve_multishift(chi, shifts, 1e-10, maxit=2000)
    full = np.zeros((V, 4, 3), complex)
    full[V // 2:] = chi
    ref, nc, rres = ref_op.solve_multishift(full, shifts, 1e-10, maxit=2000)
    chin = np.linalg.norm(chi)
    say("  n_count", res[0]["n_count"], "oracle", nc, [r["rel_resid"] for r in res])
    assert abs(res[0]["n_count"] - nc) <= 1
    for s in range(len(shifts)):
        assert rel(psi[s], ref[s][V // 2:]) < 1e-8
        assert res[s]["rel_resid"] < 1e-9
        assert abs(res[s]["resid"] - rres[s]) < 1e-9 * chin
    # the zero shift is the M^dag M solve
    single, r1 = op.solve(chi, solver=cmx.CG, rsd=1e-12, maxit=5000, mdagm=True)
    assert rel(psi[0], single) < 1e-8


def test_multishift_per_shift_targets_and_order():
    L = (4, 4, 4, 8)
    say("multishift targets")
    V = int(np.prod(L))
    op = cmx.Operator(cmx.Gauge(L, field(L, 62)), cmx.EO_CLOVER, cmx.kappa_to_mass(0.12), 1.0)
    chi = cmx.z2_source((V // 2, 4, 3), 63)
    psi_a, res_a = op.solve_multishift(chi, [0.0, 0.3], [1e-10, 1e-10])
    psi_b, res_b = op.solve_multishift(chi, [0.3, 0.0], [1e-4, 1e-10])
    assert res_b[0]["n_count"] == res_a[0]["n_count"]
    assert np.array_equal(psi_b[1].view(np.uint64), psi_a[0].view(np.uint64))  # the same Krylov iteration
    assert res_b[0]["rel_resid"] <= 1e-4 and res_b[0]["rel_resid"] > 1e-10


def test_multishift_zero_source_and_bad_input():
    L = (4, 4, 4, 4)
    op = cmx.Operator(cmx.Gauge(L, field(L, 64)), cmx.EO_WILSON, 0.1)
    psi, res = op.solve_multishift(np.zeros((int(np.prod(L)) // 2, 4, 3), complex), [0.0, 0.1], 1e-10)
    assert res[0]["n_count"] == 0 and not np.any(psi)
    with pytest.raises(cmx.CmxError):
        op.solve_multishift(np.ones((int(np.prod(L)) // 2, 4, 3), complex), np.zeros(65), 1e-10)
    with pytest.raises(cmx.CmxError):
        op.solve_multishift(np.ones((int(np.prod(L)) // 2, 4, 3), complex), [0.0], 0.0)
