"""GPU: meson two-point functions on the device (mesons(), mesons_w.cc:97-180, through cmx_meson_2pt)
against the oracle, and the pion of a device propagator against its prop_corr."""
import sys
import time

import numpy as np
import pytest

import oracle
from chroma_amd import lib as cmx
from tests.util import rel

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("L,origin,same", [((4, 4, 4, 8), (0, 0, 0, 0), False), ((6, 4, 2, 8), (3, 1, 1, 5), False),
                                           ((4, 4, 4, 4), (2, 0, 3, 1), True)])
def test_mesons_vs_oracle(L, origin, same):
    say("mesons vs oracle", L, origin, same)
    V = int(np.prod(L))
    r = np.random.default_rng(80)
    S1 = r.standard_normal((V, 12, 12)) + 1j * r.standard_normal((V, 12, 12))
    S2 = None if same else r.standard_normal((V, 12, 12)) + 1j * r.standard_normal((V, 12, 12))
    moms = ((0, 0, 0), (1, 0, 0), (0, -1, 1), (1, 1, 1))
    got = cmx.meson_2pt(L, S1, S2, origin, moms)
    want = oracle.mesons(L, S1, S2, origin, moms)
    assert rel(got, want) < 1e-13


def test_mesons_of_device_propagator():
    """The pion (gamma_value 15) of a clover propagator from the device solver equals its prop_corr,
    and a0 (gamma_value 0) matches the oracle's contraction of the same propagator."""
    from chroma_amd import qprop as qp
    L = (4, 4, 4, 8)
    say("mesons of a device propagator", L)
    U = cmx.gauge_apply_bc(L, cmx.gauge_generate(L, kind=1, seed=9, eps=0.3), (1, 1, 1, -1))
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_CLOVER, cmx.kappa_to_mass(0.12), 1.0)
    prop, _ = qp.quark_prop4(qp.PrecFermActQprop(op, rsd=1e-10), int(np.prod(L)), 0)
    m = cmx.meson_2pt(L, prop)
    pc = qp.prop_corr(L, prop)
    assert np.max(np.abs(m[15, 0].real - pc) / pc) < 1e-12
    assert rel(m, oracle.mesons(L, prop)) < 1e-13
    with pytest.raises(cmx.CmxError):
        cmx.meson_2pt(L, prop, origin=(0, 0, 0, 8))


def test_mesons_timing():
    L = (24, 24, 24, 48)
    V = int(np.prod(L))
    say("mesons timing", L)
    r = np.random.default_rng(81)
    S = (r.standard_normal((V, 12, 12)) + 1j * r.standard_normal((V, 12, 12))).astype(np.complex128)
    cmx.meson_2pt((4, 4, 4, 4), S[:256])
    t = time.perf_counter()
    m = cmx.meson_2pt(L, S, None, (0, 0, 0, 0), ((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)))
    say(f"16 gammas x 4 momenta at {L} (with the {S.nbytes / 1e9:.1f} GB host copy): "
        f"{time.perf_counter() - t:.3f} s")
    assert np.all(m[15, 0].real > 0)
