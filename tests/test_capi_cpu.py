"""CPU-side checks of the product library: it loads, exports every symbol its header declares, fails
cleanly without a device, and its host-side input generators have the properties they promise."""
import ctypes
import subprocess
import sys

import numpy as np

from chroma_amd import lib as cmx
from tests.util import rel


def test_library_exports_every_declared_symbol():
    declared = cmx.declared_symbols()
    assert len(declared) >= 20
    L = cmx.lib()
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing


def test_calls_before_init_fail_cleanly():
    h = ctypes.c_void_p()
    dims = (ctypes.c_int * 4)(4, 4, 4, 4)
    U = np.zeros(4 * 256 * 18)
    rc = cmx.lib().cmx_gauge_create(dims, U.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(h))
    assert rc == cmx.CMX_ERR_STATE
    assert b"cmx_init" in cmx.lib().cmx_last_error()
    assert cmx.lib().cmx_op_subset_doubles(None) == -1


def test_generator_su3_and_plaquette():
    L = (4, 4, 4, 8)
    for kind, nstout in [(0, 0), (1, 0), (2, 0), (2, 2)]:
        U = cmx.gauge_generate(L, kind=kind, seed=3, eps=0.3, nstout=nstout, rho=0.1)
        uu = np.einsum("mvij,mvkj->mvik", U, U.conj())
        assert np.max(np.abs(uu - np.eye(3))) < 1e-13
        assert np.max(np.abs(np.linalg.det(U) - 1)) < 1e-13
    assert cmx.plaquette(L, cmx.gauge_generate(L, kind=0)) == 1.0
    # small eps -> plaquette close to one; the generator is deterministic in its seed
    a = cmx.gauge_generate(L, kind=1, seed=5, eps=0.05)
    assert 0.99 < cmx.plaquette(L, a) < 1.0
    assert np.array_equal(a, cmx.gauge_generate(L, kind=1, seed=5, eps=0.05))
    assert not np.array_equal(a, cmx.gauge_generate(L, kind=1, seed=6, eps=0.05))


def test_haar_moments_and_stout_smoothing():
    L = (8, 8, 8, 8)
    U = cmx.gauge_generate(L, kind=2, seed=11)
    tr = np.trace(U, axis1=2, axis2=3).ravel()
    # Haar SU(3): E[tr U] = 0, E[|tr U|^2] = 1 (16384 samples: 5 sigma bounds)
    assert abs(tr.mean()) < 5 / np.sqrt(tr.size)
    assert abs(np.mean(np.abs(tr) ** 2) - 1) < 5 * np.sqrt(2 / tr.size) * 1.5
    assert abs(cmx.plaquette(L, U)) < 0.02
    p = [cmx.plaquette(L, cmx.gauge_generate(L, kind=2, seed=11, nstout=n, rho=0.1)) for n in (1, 3, 6)]
    assert p[0] < p[1] < p[2] and 0.55 < p[2] < 0.75


def test_boundary_conditions():
    L = (4, 4, 4, 6)
    V = int(np.prod(L))
    U = cmx.gauge_generate(L, kind=1, seed=2)
    import oracle
    X = np.array([oracle.site_coords(L, s) for s in range(V)])
    Ub = cmx.gauge_apply_bc(L, U.copy(), (1, 1, 1, -1))
    expect = U.copy()
    expect[3, X[:, 3] == L[3] - 1] *= -1
    assert np.array_equal(Ub, expect)
    Ut = cmx.gauge_apply_bc(L, U.copy(), (1, 1, 1, 1), twist_by_pi=(0.2, 0.4, 0.0), twist_dir=(0, 1, 2))
    for d, ph in zip((0, 1, 2), (0.2, 0.4, 0.0)):
        assert rel(Ut[d], U[d] * np.exp(1j * np.pi * ph / L[d])) < 1e-15


def test_z2_source():
    z = cmx.z2_source((64, 4, 3), 9)
    v = np.concatenate([z.real.ravel(), z.imag.ravel()])
    assert np.allclose(np.abs(v), 1 / np.sqrt(2))
    assert abs(np.mean(v > 0) - 0.5) < 0.1
    assert np.array_equal(z, cmx.z2_source((64, 4, 3), 9))


def test_kappa_to_mass():
    assert abs(cmx.kappa_to_mass(0.1) - 1.0) < 1e-15
    assert abs(cmx.kappa_to_mass(0.115) - 0.347826086956522) < 1e-14


def test_bench_refuses_multi_gpu():
    import os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    p = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert p.returncode == 2 and "one GPU" in p.stderr


def test_learned_net_neumann_initialisation():
    """The library's Neumann weights equal the ones built from the DR gammas in the oracle tests
    (host-side call: no device needed)."""
    from tests.test_oracle_learned import neumann
    a = 4.0 + cmx.kappa_to_mass(0.135)
    net = cmx.Net()
    net.init_neumann(a)
    w1, w2, w0 = net.get_weights()
    e1, e2, e0 = neumann(a)
    # the factorisation W2 W1 is not unique; the contracted weights are what the network applies
    assert np.allclose(w2 @ w1, e2 @ e1, rtol=1e-5, atol=1e-8)
    assert np.allclose(w0, e0, rtol=1e-6)
    # set / get round trip
    rng = np.random.default_rng(3)
    r1 = rng.standard_normal((64, 32)) + 1j * rng.standard_normal((64, 32))
    net.set_weights(r1, e2, e0)
    g1, _, _ = net.get_weights()
    assert np.allclose(g1, r1.astype(np.complex64), rtol=0, atol=0)


def test_generator_stout_matches_oracle():
    """The host generator's stout steps (Taylor exponential with scaling and squaring, re-projected onto
    SU(3)) against the oracle's (plain Taylor series, no re-projection)."""
    import oracle
    L = (4, 4, 6, 4)
    U0 = cmx.gauge_generate(L, kind=2, seed=21)
    U2 = cmx.gauge_generate(L, kind=2, seed=21, nstout=2, rho=0.12)
    assert rel(U2, oracle.stout(L, U0, 0.12, 2)) < 1e-13
