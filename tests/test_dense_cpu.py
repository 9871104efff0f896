"""The library's small dense algebra (chroma_amd/csrc/dense.h), used by FGMRES-DR in place of LAPACK,
against numpy's LAPACK routines: the QR factors equal zgeqrf/zungqr's (reflector convention and the
phases of the completion columns included, which the reference's incremental residual test
|Re g_{j+1}| depends on, syssolver_linop_fgmres_dr.cc:172), the eigenpairs satisfy A w = lambda w with
numpy's eigenvalues, and the LU solve equals numpy.linalg.solve."""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "chroma_amd", "csrc")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("dense") / "dense_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, os.path.join(HERE, "dense_check.cpp"), "-o", exe],
                   check=True)
    return exe


def run(exe, A):
    m, n = A.shape
    txt = f"{m} {n}\n" + "\n".join(f"{float(z.real)!r} {float(z.imag)!r}" for z in A.ravel())
    out = subprocess.run([exe], input=txt, capture_output=True, text=True, check=True).stdout.split("\n")
    res, k = {}, 0
    while k < len(out) and out[k]:
        tag, r, c = out[k].split()
        r, c = int(r), int(c)
        vals = [complex(float(a), float(b)) for a, b in (ln.split() for ln in out[k + 1:k + 1 + r * c])]
        res[tag] = np.array(vals).reshape(r, c)
        k += 1 + r * c
    return res


@pytest.mark.parametrize("m,n,seed", [(5, 5, 1), (7, 4, 2), (9, 8, 3), (2, 1, 4), (13, 12, 5)])
def test_qr_matches_lapack_convention(driver, m, n, seed):
    r = np.random.default_rng(seed)
    A = r.standard_normal((m, n)) + 1j * r.standard_normal((m, n))
    got = run(driver, A)
    Q, R = np.linalg.qr(A, mode="complete")
    assert np.abs(got["Q"] - Q).max() < 1e-13
    assert np.abs(got["R"] - R[:n]).max() < 1e-13
    assert np.abs(np.diag(got["R"]).imag).max() == 0.0


def test_qr_identity_reflector(driver):
    """A column already reduced (zero below a real diagonal): the reflector is the identity (tau = 0),
    as in zlarfg."""
    A = np.array([[2.0, 1 + 1j], [0.0, 3.0], [0.0, 0.5j]], complex)
    got = run(driver, A)
    Q, R = np.linalg.qr(A, mode="complete")
    assert np.abs(got["Q"] - Q).max() < 1e-14
    assert np.abs(got["R"] - R[:2]).max() < 1e-14


@pytest.mark.parametrize("n,seed", [(4, 6), (8, 7), (16, 8)])
def test_eig_and_lu_solve(driver, n, seed):
    r = np.random.default_rng(seed)
    A = r.standard_normal((n, n)) + 1j * r.standard_normal((n, n))
    got = run(driver, A)
    lam, W = got["E"][0], got["W"]
    ref = np.linalg.eigvals(A)
    for z in ref:
        assert np.min(np.abs(lam - z)) < 1e-10 * max(1.0, abs(z))
    assert np.abs(A @ W - W * lam[None, :]).max() < 1e-10
    assert np.abs(np.linalg.norm(W, axis=0) - 1).max() < 1e-12
    b = np.arange(1, n + 1).astype(complex)
    assert np.abs(got["X"][0] - np.linalg.solve(A, b)).max() < 1e-11
