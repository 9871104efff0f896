"""Shared helpers for the tests: seeded inputs in the library's host layout and error measures."""
import numpy as np


def rel(a, b):
    """|a - b| / |b| (Frobenius over all components)."""
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def random_fermion(V, seed):
    r = np.random.default_rng(seed)
    return (r.standard_normal((V, 4, 3)) + 1j * r.standard_normal((V, 4, 3))) / np.sqrt(2)


def random_su3_numpy(n, seed):
    """Haar SU(3) matrices from numpy's QR (independent of the library's generator)."""
    r = np.random.default_rng(seed)
    z = (r.standard_normal((n, 3, 3)) + 1j * r.standard_normal((n, 3, 3))) / np.sqrt(2)
    q, rr = np.linalg.qr(z)
    d = np.diagonal(rr, axis1=1, axis2=2)
    q = q * (d / np.abs(d))[:, None, :]
    det = np.linalg.det(q)
    return q / (det ** (1.0 / 3.0))[:, None, None]


def numpy_gauge(L, seed):
    V = int(np.prod(L))
    return random_su3_numpy(4 * V, seed).reshape(4, V, 3, 3)


def apply_simple_bc(L, U, boundary):
    """SimpleFermBC::modify (simple_fermbc.h:87-104) in numpy, via site coordinates."""
    import oracle
    U = U.copy()
    V = U.shape[1]
    coords = np.array([oracle.site_coords(L, s) for s in range(V)])
    for mu in range(4):
        if boundary[mu] != 1:
            U[mu, coords[:, mu] == L[mu] - 1] *= boundary[mu]
    return U
