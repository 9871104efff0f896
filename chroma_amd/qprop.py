"""Propagator drivers above the C-ABI, mirroring Chroma's callers of the solver.

* ``PrecFermActQprop`` -- EvenOddPrecWilsonTypeFermAct::qprop's solver object
  (lib/actions/ferm/qprop/eoprec_fermact_qprop.cc:16-80): even-odd source preparation, the odd solve
  through the device solver, even reconstruction and the full unpreconditioned residual.
* ``quark_prop4`` -- quarkProp4_a (lib/actions/ferm/qprop/quarkprop4_w.cc:31-117): the 12 spin x colour
  solves with source normalisation.
* ``prop_corr`` -- the propagator sanity correlator of inline_propagator_w.cc:383-395.

Fields are host arrays in the library's host layout (complex128[V, 4, 3], checkerboard-major); every
operator application and solve runs on the device.
"""
import numpy as np

from . import lib as cmx


class PrecFermActQprop:
    """SystemSolver over the full lattice built from an even-odd operator and its odd solver."""

    def __init__(self, op, solver=cmx.CG, rsd=1e-8, maxit=1000, delta=0.1, sloppy_prec=cmx.SINGLE):
        assert op.kind in (cmx.EO_WILSON, cmx.EO_CLOVER)
        self.op, self.solver, self.rsd, self.maxit = op, solver, rsd, maxit
        self.delta, self.sloppy_prec = delta, sloppy_prec
        self.Vh = op.V // 2

    def __call__(self, psi, chi):
        """Returns (psi, result) with result = {n_count, resid} (eoprec_fermact_qprop.cc:41-80)."""
        op, Vh = self.op, self.Vh
        # chi_tmp_o = chi_o - D_oe A_ee^-1 chi_e
        t1 = op.piece("evenEvenInv", chi)
        t2 = op.piece("oddEven", t1)
        chi_tmp_o = chi[Vh:] - t2[Vh:]
        psi_o, res = op.solve(chi_tmp_o, psi[Vh:], solver=self.solver, rsd=self.rsd, maxit=self.maxit,
                              delta=self.delta, sloppy_prec=self.sloppy_prec)
        out = np.zeros_like(chi)
        out[Vh:] = psi_o
        # psi_e = A_ee^-1 (chi_e - D_eo psi_o)
        t1 = op.piece("evenOdd", out)
        t2 = np.zeros_like(chi)
        t2[:Vh] = chi[:Vh] - t1[:Vh]
        out[:Vh] = op.piece("evenEvenInv", t2)[:Vh]
        r = op.piece("unprec", out) - chi
        return out, dict(n_count=res["n_count"], resid=float(np.sqrt(np.vdot(r, r).real)),
                         odd_rel_resid=res["rel_resid"], solve_seconds=res["solve_seconds"])


def quark_prop4(qprop, V, src_site=0):
    """12 point-source solves (QUARK_SPIN_TYPE_FULL).  Returns (prop[V,12,12], list of results)."""
    prop = np.zeros((V, 12, 12), np.complex128)
    results = []
    for color_source in range(3):
        for spin_source in range(4):
            chi = np.zeros((V, 4, 3), np.complex128)
            chi[src_site, spin_source, color_source] = 1.0
            nrm = np.sqrt(np.vdot(chi, chi).real)
            fact = 1.0 / nrm if nrm != 0 else 1.0
            chi *= fact
            psi, res = qprop(np.zeros_like(chi), chi)
            psi /= fact
            prop[:, :, spin_source * 3 + color_source] = psi.reshape(V, 12)
            results.append(res)
    return prop, results


def site_time(L):
    """Time coordinate of every site in the host (checkerboard-major) order."""
    L = [int(v) for v in L]
    V = int(np.prod(L))
    Vh = V // 2
    s = np.arange(V)
    r = s % Vh
    r = r // (L[0] // 2)
    r = r // L[1]
    r = r // L[2]
    return r


def prop_corr(L, prop):
    """sum over each time slice of localNorm2(propagator) (inline_propagator_w.cc:388-391)."""
    t = site_time(L)
    per_site = np.sum(np.abs(prop.reshape(prop.shape[0], -1)) ** 2, axis=1)
    return np.bincount(t, weights=per_site, minlength=int(L[3]))
