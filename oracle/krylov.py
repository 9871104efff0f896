"""ORACLE -- test infrastructure only.

numpy restatement of the flexible solver FGMRES-DR (syssolver_linop_fgmres_dr.cc:78-774, Givens
rotations syssolver_linop_fgmres_dr.h:32-140) and its MR preconditioner (InvMR, invmr.cc:55-160),
over any linear operator given as a Python callable on odd-checkerboard arrays.

The small dense algebra (LU solve, QR, eigen-decomposition) is numpy.linalg; the reference calls
LAPACK zgetrf/zgetrs/zgeev/zgeqrf/zunmqr/zungqr at the same points.

Three choices differ from the reference on purpose (the library makes the same ones):
  * in a deflated cycle only the current cycle's Givens rotations are applied to a new column
    (the reference also applies the previous cycle's rotations of the deflated columns, :151-153);
  * every solve starts with Z = 0 (the reference keeps Z across solves, zeroing it only at
    construction, :197-225); within a solve, a plain restart passes the previous cycle's Z_j to the
    preconditioner as its initial guess, as the reference does;
  * a deflated cycle that fails to halve the true residual is followed by a plain restart.
"""
import numpy as np


def _dot(a, b):
    return np.vdot(a.reshape(-1), b.reshape(-1))


def _norm(a):
    return float(np.sqrt(np.vdot(a.reshape(-1), a.reshape(-1)).real))


def mr(A, chi, psi, rsd, maxit, omega=1.0):
    """InvMR (invmr.cc:55-160): psi is the initial guess; returns (psi, k)."""
    psi = psi.copy()
    rsd_sq = rsd * rsd * _norm(chi) ** 2
    r = chi - A(psi)
    cp = _norm(r) ** 2
    k = 0
    while k < maxit and cp > rsd_sq:
        k += 1
        Mr = A(r)
        c = _dot(Mr, r)
        d = _norm(Mr) ** 2
        a = omega * c / d
        psi = psi + a * r
        r = r - a * Mr
        cp = _norm(r) ** 2
    return psi, k


class Givens:
    """Rotation with a real non-negative sine (syssolver_linop_fgmres_dr.h:32-118)."""

    def __init__(self, col, f, g):
        self.col = col
        if f == 0:
            self.c, self.s, self.r = 0.0, 1.0, g
        elif g == 0:
            self.s, self.c, self.r = 0.0, np.conj(f) / abs(f), abs(f)
        else:
            t = np.sqrt(abs(f) ** 2 + abs(g) ** 2)
            self.r, self.c, self.s = t, f / t, g / t

    def apply(self, v):
        j = self.col
        a, b = v[j], v[j + 1]
        v[j] = np.conj(self.c) * a + np.conj(self.s) * b
        v[j + 1] = -self.s * a + self.c * b


def fgmres_dr(A, chi, psi, rsd, n_krylov, n_defl=0, max_iter=1000, precond=None, record=None, cycle_reduction=0.0):
    """LinOpSysSolverFGMRESDR::operator() (syssolver_linop_fgmres_dr.cc:459-774).

    A: callable on arrays shaped like chi.  precond(z0, v) -> z (None: identity).
    cycle_reduction > 0 ends a cycle once its residual estimate has fallen by that factor below the
    residual the cycle started from (the library's option; 0 is the reference's rule).  A cycle that
    ends with fewer Arnoldi vectors than n_defl deflates only as many harmonic Ritz vectors as it has.
    Returns (psi, n_count, |chi - A psi|, cycles).  `record` (a dict) receives the Hessenberg matrix,
    the bases and residual norms of each cycle, for tests.
    """
    shape = chi.shape
    nk, nd_max = n_krylov, n_defl
    total = nk + nd_max
    psi = psi.copy()
    norm_rhs = _norm(chi)
    target = norm_rhs * rsd
    r = chi - A(psi)
    r_norm = _norm(r)
    V = [np.zeros(shape, complex) for _ in range(total + 1)]
    Z = [np.zeros(shape, complex) for _ in range(total)]
    H = np.zeros((total + 1, total), complex)
    R = np.zeros((total + 1, total), complex)
    g = np.zeros(total + 1, complex)
    c = np.zeros(total + 1, complex)
    eta = np.zeros(total, complex)
    rots = [None] * total
    Qk = None
    iters, cycles, prev_dim = 0, 0, nk
    finished = r_norm <= target
    plain_next, r_prev = True, r_norm
    while not finished:
        cycles += 1
        nd = min(nd_max, prev_dim)
        if plain_next or nd == 0:
            c[:] = 0
            g[:] = 0
            c[0] = g[0] = r_norm
            H[:] = 0
            R[:] = 0
            V[0] = r / r_norm
            nd = 0
        else:
            m = prev_dim
            Hm = H[:m, :m].copy()
            hm = np.conj(H[m, :m])
            f = np.linalg.solve(Hm.conj().T, hm)
            Ht = Hm + np.outer(f, np.conj(hm))
            evals, evecs = np.linalg.eig(Ht)
            order = np.argsort(np.abs(evals) ** 2, kind="stable")
            G = np.zeros((m + 1, nd + 1), complex)
            G[:m, :nd] = evecs[:, order[:nd]]
            G[:, nd] = c[:m + 1] - H[:m + 1, :m] @ eta[:m]
            Q, _ = np.linalg.qr(G)
            Hn = Q.conj().T @ H[:m + 1, :m] @ Q[:m, :nd]
            newV = [sum(Q[j, i] * V[j] for j in range(m + 1)) for i in range(nd + 1)]
            newZ = [sum(Q[j, i] * Z[j] for j in range(m)) for i in range(nd)]
            for i in range(nd + 1):
                V[i] = newV[i]
            for i in range(nd):
                Z[i] = newZ[i]
            for i in range(nd, total):
                Z[i] = np.zeros(shape, complex)
            H[:] = 0
            R[:] = 0
            c[:] = 0
            g[:] = 0
            for i in range(nd + 1):
                c[i] = g[i] = _dot(V[i], r)
            H[:nd + 1, :nd] = Hn
            Qk, Rk = np.linalg.qr(Hn, mode="complete")
            R[:nd, :nd] = Rk[:nd, :nd]
            g[:nd + 1] = Qk.conj().T @ g[:nd + 1]
            if record is not None:
                record.setdefault("harmonic_ritz", []).append(evals[order[:nd]])
        cycle_stop = max(target, cycle_reduction * r_norm if cycle_reduction > 0 else 0.0)
        dim = nd
        for j in range(nd, nd + nk):
            Z[j] = V[j].copy() if precond is None else precond(Z[j], V[j])
            w = A(Z[j])
            for i in range(j + 1):  # modified Gram-Schmidt (:98-104)
                H[i, j] = _dot(V[i], w)
                w = w - H[i, j] * V[i]
            wnorm = _norm(w)
            H[j + 1, j] = wnorm
            if abs(wnorm) < 1e-14:
                dim = j
                break
            V[j + 1] = w / wnorm
            col = H[:j + 2, j].copy()
            if nd > 0:
                col[:nd + 1] = Qk.conj().T @ col[:nd + 1]
            for i in range(nd, j):
                rots[i].apply(col)
            rots[j] = Givens(j, col[j], col[j + 1])
            col[j], col[j + 1] = rots[j].r, 0
            rots[j].apply(g)
            R[:j + 2, j] = col
            dim = j + 1
            if abs(g[j + 1].real) <= cycle_stop:  # (:172)
                break
        if dim > 0:
            for row in range(dim - 1, -1, -1):
                s = g[row] - R[row, row + 1:dim] @ eta[row + 1:dim]
                eta[row] = s / R[row, row]
            psi = psi + sum(eta[j] * Z[j] for j in range(dim))
        if record is not None:
            record.setdefault("cycles", []).append(dict(nd=nd, dim=dim, r_in=r.copy(),
                                                       Z=[Z[j].copy() for j in range(dim)],
                                                       V=[V[j].copy() for j in range(dim + 1)],
                                                       H=H[:dim + 1, :dim].copy()))
        r = chi - A(psi)
        r_norm = _norm(r)
        iters += dim - nd
        plain_next = nd > 0 and r_norm > 0.5 * r_prev
        r_prev = r_norm
        finished = r_norm <= target or iters >= max_iter or (dim == nd and not plain_next)
        prev_dim = dim
    return psi, iters, r_norm, cycles
