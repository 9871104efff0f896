"""ORACLE -- test infrastructure only.

numpy restatement of the learned preconditioner (the library's gauge-equivariant parallel-transport
network) and of its loss gradient, in FP64, on odd-checkerboard vectors:

    P(v)(x) = W0 v(x) + sum_q U_q(x) Pi_q g_q(x+q),        x odd,  q in {+-mu}
    g_q(y)  = sum_c W2[q][c] h_c(y),   h_c(y) = sum_p W1[c][p] U_p(y) (Pi_p v)(y+p),     y even

U_{+mu}(y) = U_mu(y), U_{-mu}(y) = U_mu(y-mu)^dag; Pi_p = (1 + sigma_p gamma_mu) / 2 with sigma = -1 on the
forward and +1 on the backward path: the rank-2 spin projector of the hopping term's direction
(lwldslash_w.h:252-295), so every transport moves a half spinor, as the hopping term's does.  Weights:
W1 complex[64, 32] indexed ((c*4 + s), (p*4 + t)), W2 complex[32, 64] indexed ((q*4 + s), (c*4 + t)),
W0 complex[4, 4]; path p = 2 mu + (0 forward, 1 backward).  Spin matrices act on spin, the identity on
colour.  The network depends on Wc = W2 W1 only through Pi_q Wc Pi_p.

Host layout as in oracle/__init__.py: full lattice checkerboard-major, x-fastest; gauge
complex128[4, V, 3, 3] with the boundary conditions applied.
"""
import numpy as np

# DeGrand-Rossi gamma matrices (the spin basis of the fermion layout), as in oracle/oracle.cpp
_I = 1j
GAMMA = np.array([
    [[0, 0, 0, _I], [0, 0, _I, 0], [0, -_I, 0, 0], [-_I, 0, 0, 0]],
    [[0, 0, 0, -1], [0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0]],
    [[0, 0, _I, 0], [0, 0, 0, -_I], [-_I, 0, 0, 0], [0, _I, 0, 0]],
    [[0, 0, 1, 0], [0, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0]],
], dtype=complex)


def projector(p):
    """Pi_p = (1 + sigma_p gamma_mu) / 2, sigma = -1 forward (p even), +1 backward."""
    mu, sg = p // 2, (-1.0 if p % 2 == 0 else 1.0)
    return (np.eye(4) + sg * GAMMA[mu]) / 2


def neighbour_tables(L):
    """fwd[mu][s], bwd[mu][s]: full-lattice index of s +- mu."""
    L = tuple(int(v) for v in L)
    V = int(np.prod(L))
    Vh = V // 2
    s = np.arange(V)
    cb = s // Vh
    i = s % Vh
    Lxh = L[0] // 2
    xh = i % Lxh
    r = i // Lxh
    y = r % L[1]
    r //= L[1]
    z = r % L[2]
    t = r // L[2]
    x = 2 * xh + ((cb + y + z + t) & 1)
    X = np.stack([x, y, z, t], axis=1)

    def index(C):
        c = (C.sum(axis=1)) & 1
        return c * Vh + (((C[:, 3] * L[2] + C[:, 2]) * L[1] + C[:, 1]) * L[0] + C[:, 0]) // 2

    fwd, bwd = [], []
    for mu in range(4):
        C = X.copy()
        C[:, mu] = (C[:, mu] + 1) % L[mu]
        fwd.append(index(C))
        C = X.copy()
        C[:, mu] = (C[:, mu] - 1) % L[mu]
        bwd.append(index(C))
    return np.array(fwd), np.array(bwd)


def _transports(L, U, v_full, sites):
    """t[p][k] = U_p(y) (Pi_p v)(y+p) for y = sites[k]: array [8, len(sites), 4, 3]."""
    fwd, bwd = neighbour_tables(L)
    out = np.zeros((8, len(sites), 4, 3), complex)
    for mu in range(4):
        yf = fwd[mu][sites]
        pv = np.einsum("st,ktc->ksc", projector(2 * mu), v_full[yf])
        out[2 * mu] = np.einsum("kab,ksb->ksa", U[mu, sites], pv)
        yb = bwd[mu][sites]
        pv = np.einsum("st,ktc->ksc", projector(2 * mu + 1), v_full[yb])
        out[2 * mu + 1] = np.einsum("kba,ksb->ksa", U[mu, yb].conj(), pv)
    return out


def ptc_apply(L, U, W1, W2, W0, v_odd, return_parts=False):
    """P(v) on the odd checkerboard (v_odd: [Vh, 4, 3])."""
    V = int(np.prod(L))
    Vh = V // 2
    full = np.zeros((V, 4, 3), complex)
    full[Vh:] = v_odd
    even = np.arange(Vh)
    t = _transports(L, U, full, even)                       # [8, Vh, 4, 3]
    tk = t.transpose(1, 0, 2, 3).reshape(Vh, 32, 3)          # (p*4 + t) components
    Wc = W2 @ W1
    g = np.einsum("qp,kpc->kqc", Wc, tk).reshape(Vh, 8, 4, 3)  # g_q(y)
    gfull = np.zeros((8, V, 4, 3), complex)
    for q in range(8):  # Pi_q g_q
        gfull[q, :Vh] = np.einsum("st,ktc->ksc", projector(q), g[:, q])
    fwd, bwd = neighbour_tables(L)
    odd = np.arange(Vh, V)
    out = np.einsum("st,kta->ksa", W0, v_odd)
    for mu in range(4):
        yf = fwd[mu][odd]
        out += np.einsum("kab,ksb->ksa", U[mu, odd], gfull[2 * mu, yf])
        yb = bwd[mu][odd]
        out += np.einsum("kba,ksb->ksa", U[mu, yb].conj(), gfull[2 * mu + 1, yb])
    if return_parts:
        return out, tk
    return out


def ptc_block_grad(L, U, W1, W2, W0, v_odd, delta):
    """(dL/d conj(Wc) [32, 32], dL/d conj(W0) [4, 4]) of one block B with input v and output gradient
    delta = dL/d conj(B(v))."""
    V = int(np.prod(L))
    Vh = V // 2
    full = np.zeros((V, 4, 3), complex)
    full[Vh:] = v_odd
    tk = _transports(L, U, full, np.arange(Vh)).transpose(1, 0, 2, 3).reshape(Vh, 32, 3)
    dgk = _scatter_adjoint(L, U, delta)
    S = np.einsum("kqc,kpc->qp", dgk, tk.conj())
    S0 = np.einsum("ksc,ktc->st", delta, v_odd.conj())
    return S, S0


def _scatter_adjoint(L, U, delta):
    """dg_q(y) = Pi_q U_q(x)^dag delta(x) for y = x + q: the adjoint of the scatter, as [Vh, 32, 3]."""
    V = int(np.prod(L))
    Vh = V // 2
    fwd, bwd = neighbour_tables(L)
    odd = np.arange(Vh, V)
    dg = np.zeros((8, V, 4, 3), complex)
    for mu in range(4):
        yf = fwd[mu][odd]
        np.add.at(dg[2 * mu], yf, np.einsum("kba,ksb->ksa", U[mu, odd].conj(), delta))
        yb = bwd[mu][odd]
        np.add.at(dg[2 * mu + 1], yb, np.einsum("kab,ksb->ksa", U[mu, yb], delta))
    for q in range(8):  # the adjoint of Pi_q (Hermitian)
        dg[q] = np.einsum("st,ktc->ksc", projector(q), dg[q])
    return dg[:, :Vh].transpose(1, 0, 2, 3).reshape(Vh, 32, 3)


def ptc_grad(L, U, W1, W2, W0, v_odd, e_odd):
    """(|P v - e|^2, dL/d conj(Wc) [32, 32], dL/d conj(W0) [4, 4]) with Wc = W2 W1."""
    out = ptc_apply(L, U, W1, W2, W0, v_odd)
    delta = out - e_odd
    loss = float(np.vdot(delta, delta).real)
    S, S0 = ptc_block_grad(L, U, W1, W2, W0, v_odd, delta)
    return loss, S, S0


def ptc_adjoint(L, U, W1, W2, W0, d_odd):
    """B^dag d = W0^dag d + G^dag(Wc^dag S^dag d) for one block: S^dag the adjoint of the scatter, Wc^dag
    the adjoint mixing, G^dag(w)(x) = sum_p Pi_p U_p(x-p)^dag w_p(x-p) the adjoint of the transports."""
    V = int(np.prod(L))
    Vh = V // 2
    dgk = _scatter_adjoint(L, U, d_odd)
    Wc = W2 @ W1
    dtk = np.einsum("qp,kqc->kpc", Wc.conj(), dgk).reshape(Vh, 8, 4, 3)
    fwd, bwd = neighbour_tables(L)
    even = np.arange(Vh)
    full = np.zeros((V, 4, 3), complex)
    for mu in range(4):
        w = np.einsum("st,ktc->ksc", projector(2 * mu), dtk[:, 2 * mu])
        np.add.at(full, fwd[mu][even], np.einsum("kba,ksb->ksa", U[mu, even].conj(), w))
        yb = bwd[mu][even]
        w = np.einsum("st,ktc->ksc", projector(2 * mu + 1), dtk[:, 2 * mu + 1])
        np.add.at(full, yb, np.einsum("kab,ksb->ksa", U[mu, yb], w))
    return np.einsum("ts,kta->ksa", W0.conj(), d_odd) + full[Vh:]


def ptc_apply_deep(L, U, blocks, v_odd):
    """P(v) = B_{n-1}(... B_0(v)) for blocks = [(W1, W2, W0), ...]."""
    out = v_odd
    for W1, W2, W0 in blocks:
        out = ptc_apply(L, U, W1, W2, W0, out)
    return out


def ptc_grad_deep(L, U, blocks, v_odd, e_odd):
    """(|P v - e|^2, [(dL/d conj(Wc_k), dL/d conj(W0_k)) per block k]): back-propagation through the
    chain, the output gradient of block k carried to its input by B_k^dag."""
    us = [v_odd]
    for W1, W2, W0 in blocks[:-1]:
        us.append(ptc_apply(L, U, W1, W2, W0, us[-1]))
    W1, W2, W0 = blocks[-1]
    delta = ptc_apply(L, U, W1, W2, W0, us[-1]) - e_odd
    loss = float(np.vdot(delta, delta).real)
    grads = [None] * len(blocks)
    for k in range(len(blocks) - 1, -1, -1):
        W1, W2, W0 = blocks[k]
        grads[k] = ptc_block_grad(L, U, W1, W2, W0, us[k], delta)
        if k > 0:
            delta = ptc_adjoint(L, U, W1, W2, W0, delta)
    return loss, grads


def identity_block():
    """The identity block (W0 = 1, W2 = 0, W1 channel c < 8 carrying path c with 1 -+ gamma_mu)."""
    W1 = np.zeros((64, 32), complex)
    for p in range(8):
        W1[p * 4:p * 4 + 4, p * 4:p * 4 + 4] = 2 * projector(p)
    return W1, np.zeros((32, 64), complex), np.eye(4, dtype=complex)
