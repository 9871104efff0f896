"""GPU: stout smearing on the device (Stouting::smear_links, stout_utils.cc:936-962, through
cmx_gauge_stout_smear) against the oracle and against the host generator's stout steps."""
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


@pytest.mark.parametrize("L,kind,rho,nstep,dirs", [((4, 4, 4, 8), 2, 0.1, 2, 15), ((6, 4, 8, 4), 1, 0.22, 3, 15),
                                                   ((4, 4, 4, 4), 2, 0.0, 1, 15), ((4, 6, 4, 8), 2, 0.1, 2, 7),
                                                   ((4, 4, 4, 8), 1, 0.15, 1, 0b1010)])
def test_stout_vs_oracle(L, kind, rho, nstep, dirs):
    say("stout vs oracle", L, kind, rho, nstep, dirs)
    U = cmx.gauge_generate(L, kind=kind, seed=31, eps=0.5)
    got = cmx.gauge_stout_smear(L, U, nstep, rho, dirs)
    assert rel(got, oracle.stout(L, U, rho, nstep, dirs)) < 1e-13
    for mu in range(4):  # unsmeared directions are copied bit for bit
        if not dirs >> mu & 1:
            assert got[mu].tobytes() == U[mu].tobytes()
    uu = np.einsum("mvij,mvkj->mvik", got, got.conj())
    assert np.max(np.abs(uu - np.eye(3))) < 1e-13


def test_stout_vs_host_generator_and_timing():
    L = (8, 8, 8, 16)
    say("stout vs host generator", L)
    U0 = cmx.gauge_generate(L, kind=2, seed=7)
    U6 = cmx.gauge_generate(L, kind=2, seed=7, nstout=6, rho=0.1)
    got = cmx.gauge_stout_smear(L, U0, 6, 0.1)
    assert rel(got, U6) < 1e-13
    assert abs(cmx.plaquette(L, got) - cmx.plaquette(L, U6)) < 1e-13
    assert cmx.gauge_stout_smear(L, U0, 0, 0.1).tobytes() == U0.tobytes()
    with pytest.raises(cmx.CmxError):
        cmx.gauge_stout_smear((4, 4, 4, 5), np.zeros((4, 320, 3, 3), complex), 1, 0.1)
    with pytest.raises(cmx.CmxError):
        cmx.gauge_stout_smear(L, U0, 1, 0.1, 16)
    L = (32, 32, 32, 64)
    U = cmx.gauge_generate(L, kind=2, seed=8)
    cmx.gauge_stout_smear((8, 8, 8, 8), U[:, :4096], 1, 0.1)  # warm the module
    t = time.perf_counter()
    S = cmx.gauge_stout_smear(L, U, 6, 0.1)
    say(f"6 stout steps at {L} on the device (with the 2 x 1.2 GB host copies): {time.perf_counter() - t:.3f} s,"
        f" plaquette {cmx.plaquette(L, S):.6f}")
