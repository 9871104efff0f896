"""ORACLE -- test infrastructure only.

ctypes wrapper over ``liboracle.so``, the CPU restatement of Chroma's even-odd Wilson-clover path
(``oracle.cpp``).  Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, as the checker or as the timed CPU baseline.

Host layout (identical to the product C-ABI): full-lattice arrays, checkerboard-major, x-fastest
within a checkerboard; fermions ``complex128[V, 4, 3]``; gauge ``complex128[4, V, 3, 3]``.

Every oracle function is pinned by the golden vectors, closed forms and dense constructions in
``tests/test_oracle.py``.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

UNPREC_WILSON, EO_WILSON, EO_CLOVER = 0, 1, 2
CG, BICGSTAB, RELIABLE_CG, RELIABLE_BICGSTAB = 0, 1, 2, 3

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)


def lib():
    global _LIB
    if _LIB is None:
        # CMX_ORACLE_LIB: another build of the same sources (the sanitizer build, oracle/_san)
        path = os.environ.get("CMX_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_op_create.restype = ctypes.c_void_p
        L.orc_op_create.argtypes = [ctypes.c_int, _ip, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.orc_op_destroy.argtypes = [ctypes.c_void_p]
        L.orc_op_apply.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int]
        L.orc_op_apply_f32.argtypes = [ctypes.c_void_p, _fp, _fp, ctypes.c_int]
        L.orc_op_eo_piece.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_int]
        L.orc_dslash.argtypes = [_ip, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int]
        L.orc_dslash_deriv.argtypes = [_ip, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int]
        L.orc_clover_deriv.argtypes = [_ip, _dp, ctypes.c_double, ctypes.c_double, _dp, _dp, ctypes.c_int, _dp]
        L.orc_clover_deriv_trln.argtypes = [_ip, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp]
        L.orc_eo_clover_deriv.argtypes = [_ip, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp,
                                          ctypes.c_int, _dp]
        L.orc_dslash_f32.argtypes = [_ip, _fp, _fp, _fp, ctypes.c_int, ctypes.c_int]
        L.orc_mesfield.argtypes = [_ip, _dp, _dp]
        L.orc_stout.argtypes = [_ip, _dp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.orc_mesons.argtypes = [_ip, _dp, _dp, _ip, ctypes.c_int, _ip, _dp]
        L.orc_make_clover.argtypes = [_ip, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]
        L.orc_clover_invert.argtypes = [_ip, _dp, ctypes.c_int, _dp, _lp]
        L.orc_clover_apply.argtypes = [_ip, _dp, _dp, _dp, ctypes.c_int]
        L.orc_op_logdet_even.argtypes = [ctypes.c_void_p, _dp, _lp]
        L.orc_solve.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_int,
                                ctypes.c_double, _ip, _dp, _dp, ctypes.c_int, _ip]
        L.orc_solve_mdagm.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_int,
                                      ctypes.c_double, _ip, _dp]
        L.orc_minvcg2.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, _dp,
                                  ctypes.POINTER(ctypes.c_int), _dp]
        L.orc_qprop.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_int,
                                ctypes.c_double, _ip, _dp]
        L.orc_quarkprop4_point.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_long, ctypes.c_double,
                                           ctypes.c_int, ctypes.c_double, _ip, _dp, _dp, _dp]
        L.orc_time.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_site_coords.argtypes = [_ip, ctypes.c_long, _ip]
        _LIB = L
    return _LIB


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())


def _dims(L):
    return (ctypes.c_int * 4)(*[int(v) for v in L])


def _d(a):
    assert a.flags.c_contiguous and a.dtype in (np.complex128, np.float64)
    return a.ctypes.data_as(_dp)


def _f(a):
    assert a.flags.c_contiguous and a.dtype in (np.complex64, np.float32)
    return a.ctypes.data_as(_fp)


def set_threads(n=0):
    return lib().orc_set_threads(int(n))


def fermion_zeros(L):
    return np.zeros((int(np.prod(L)), 4, 3), np.complex128)


def dslash_deriv(L, chi, psi, isign, cb=-1):
    """WilsonDslashBase::deriv (lwldslash_base_w.h:132-227): ds_u[4, V, 3, 3]; cb -1 sums both."""
    ds = np.zeros((4, int(np.prod(L)), 3, 3), np.complex128)
    _check(lib().orc_dslash_deriv(_dims(L), _d(ds), _d(np.ascontiguousarray(chi, np.complex128)),
                                  _d(np.ascontiguousarray(psi, np.complex128)), int(isign), int(cb)))
    return ds


def clover_deriv(L, U, chi, psi, csw_r=1.0, csw_t=None, cb=-1):
    """CloverTermBase::deriv (clover_term_base_w.h:691-745): ds[4, V, 3, 3] of chi^dag A' psi on cb (-1: both)."""
    ds = np.zeros((4, int(np.prod(L)), 3, 3), complex)
    _check(lib().orc_clover_deriv(_dims(L), _d(np.ascontiguousarray(U, np.complex128)), float(csw_r),
                                  float(csw_r if csw_t is None else csw_t), _d(np.ascontiguousarray(chi, np.complex128)),
                                  _d(np.ascontiguousarray(psi, np.complex128)), int(cb), _d(ds)))
    return ds


def clover_deriv_trln(L, U, mass, csw_r=1.0, csw_t=None, cb=0):
    """CloverTermBase::derivTrLn (clover_term_base_w.h:825-870) of the clover term inverted on cb."""
    ds = np.zeros((4, int(np.prod(L)), 3, 3), complex)
    _check(lib().orc_clover_deriv_trln(_dims(L), _d(np.ascontiguousarray(U, np.complex128)), float(csw_r),
                                       float(csw_r if csw_t is None else csw_t), float(mass), int(cb), _d(ds)))
    return ds


def eo_clover_deriv(L, U, mass, chi, psi, isign=+1, csw_r=1.0, csw_t=None):
    """EvenOddPrecLogDetLinearOperator::deriv (eoprec_logdet_linop.h:114-157) of the EO clover operator:
    ds of chi^dag M' psi, chi and psi full-lattice (odd half used)."""
    ds = np.zeros((4, int(np.prod(L)), 3, 3), complex)
    _check(lib().orc_eo_clover_deriv(_dims(L), _d(np.ascontiguousarray(U, np.complex128)), float(mass), float(csw_r),
                                     float(csw_r if csw_t is None else csw_t),
                                     _d(np.ascontiguousarray(chi, np.complex128)),
                                     _d(np.ascontiguousarray(psi, np.complex128)), int(isign), _d(ds)))
    return ds


def dslash(L, U, psi, isign, cb, out=None):
    """Hopping term D(isign) psi on output checkerboard ``cb``; other sites of ``out`` untouched."""
    if U.dtype == np.complex64:
        chi = np.zeros_like(psi) if out is None else out
        _check(lib().orc_dslash_f32(_dims(L), _f(U), _f(chi), _f(psi), int(isign), int(cb)))
        return chi
    chi = np.zeros_like(psi) if out is None else out
    _check(lib().orc_dslash(_dims(L), _d(U), _d(chi), _d(psi), int(isign), int(cb)))
    return chi


def stout(L, U, rho, nstep=1, dirs=15):
    """nstep stout steps (Stouting::smear_links, stout_utils.cc:936-962) over the directions in the bit
    mask dirs: complex128[4, V, 3, 3]."""
    U = np.ascontiguousarray(U, np.complex128)
    out = np.empty_like(U)
    _check(lib().orc_stout(_dims(L), _d(U), _d(out), float(rho), int(nstep), int(dirs)))
    return out


def mesons(L, prop1, prop2=None, origin=(0, 0, 0, 0), moms=((0, 0, 0),)):
    """mesons() (mesons_w.cc:97-180): complex128[16, nmom, L_t] (gamma insertion, sink momentum, t_eff)."""
    p1 = np.ascontiguousarray(prop1, np.complex128)
    p2 = p1 if prop2 is None else np.ascontiguousarray(prop2, np.complex128)
    mom = np.ascontiguousarray(moms, np.int32).reshape(-1, 3)
    out = np.zeros((16, mom.shape[0], int(L[3])), np.complex128)
    _check(lib().orc_mesons(_dims(L), _d(p1), _d(p2), (ctypes.c_int * 4)(*origin), int(mom.shape[0]),
                            mom.ctypes.data_as(_ip), _d(out)))
    return out


def mesfield(L, U):
    V = int(np.prod(L))
    F = np.zeros((6, V, 3, 3), np.complex128)
    _check(lib().orc_mesfield(_dims(L), _d(U), _d(F)))
    return F


def make_clover(L, U, csw_r, csw_t, mass):
    """Packed clover term, float64[V, 72] laid out as {diag[2][6], offd[2][15][2]}."""
    V = int(np.prod(L))
    tri = np.zeros((V, 72), np.float64)
    _check(lib().orc_make_clover(_dims(L), _d(U), float(csw_r), float(csw_t), float(mass), _d(tri)))
    return tri


def clover_invert(L, tri, cb):
    """In-place LDL^dag inversion on checkerboard cb; returns (sum log|d|, #sites with negative pivots)."""
    ld = ctypes.c_double()
    nn = ctypes.c_long()
    _check(lib().orc_clover_invert(_dims(L), _d(tri), int(cb), ctypes.byref(ld), ctypes.byref(nn)))
    return ld.value, nn.value


def clover_apply(L, tri, psi, cb, out=None):
    chi = np.zeros_like(psi) if out is None else out
    _check(lib().orc_clover_apply(_dims(L), _d(tri), _d(chi), _d(psi), int(cb)))
    return chi


def pack_quda(tri, cb, out=None):
    """QDPCloverTermT::packForQUDA (clover_term_qdp_w.h:2171-2228, qudaPackSiteLoop :2176-2224) from a
    PrimitiveClovTriang array tri float64[V, 72] ({diag[2][6], offd[2][15][2]}, full-lattice host order):
    QUDAPackedClovSite {diag1[6], offDiag1[15][2], diag2[6], offDiag2[15][2]} (:26-32) written at the
    full-lattice index of every site of checkerboard cb; the other checkerboard's rows of out are left
    as they are.  Within a chiral block the off-diagonal entries go column by column down the strict
    lower triangle, target index t, from the triangle's source index row*(row-1)/2 + col (:2191-2199)."""
    V = tri.shape[0]
    Vh = V // 2
    out = np.zeros((V, 72), np.float64) if out is None else out
    sites = slice(cb * Vh, (cb + 1) * Vh)
    for b in range(2):
        q0 = 36 * b
        out[sites, q0:q0 + 6] = tri[sites, 6 * b:6 * b + 6]
        t = 0
        for col in range(5):
            for row in range(col + 1, 6):
                src = row * (row - 1) // 2 + col
                out[sites, q0 + 6 + 2 * t] = tri[sites, 12 + 30 * b + 2 * src]
                out[sites, q0 + 6 + 2 * t + 1] = tri[sites, 12 + 30 * b + 2 * src + 1]
                t += 1
    return out


def site_coords(L, idx):
    x = (ctypes.c_int * 4)()
    lib().orc_site_coords(_dims(L), int(idx), x)
    return tuple(x)


class Op:
    """An oracle linear operator (kind = UNPREC_WILSON, EO_WILSON or EO_CLOVER)."""

    def __init__(self, kind, L, U, mass, csw_r=0.0, csw_t=0.0):
        self.L = tuple(int(v) for v in L)
        self.V = int(np.prod(L))
        self.kind = kind
        U = np.ascontiguousarray(U, np.complex128)
        self.h = lib().orc_op_create(int(kind), _dims(L), _d(U), float(mass), float(csw_r), float(csw_t))
        if not self.h:
            raise RuntimeError("oracle: " + lib().orc_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_op_destroy(self.h)
            self.h = None

    def __call__(self, psi, isign=+1, out=None):
        if psi.dtype == np.complex64:
            chi = np.zeros_like(psi) if out is None else out
            _check(lib().orc_op_apply_f32(self.h, _f(chi), _f(psi), int(isign)))
            return chi
        chi = np.zeros_like(psi) if out is None else out
        _check(lib().orc_op_apply(self.h, _d(chi), _d(psi), int(isign)))
        return chi

    def piece(self, which, psi, isign=+1):
        names = {"evenEven": 0, "evenEvenInv": 1, "evenOdd": 2, "oddEven": 3, "oddOdd": 4, "unprec": 5}
        chi = np.zeros_like(psi)
        _check(lib().orc_op_eo_piece(self.h, names[which], _d(chi), _d(psi), int(isign)))
        return chi

    def logdet_even(self):
        ld = ctypes.c_double()
        nn = ctypes.c_long()
        _check(lib().orc_op_logdet_even(self.h, ctypes.byref(ld), ctypes.byref(nn)))
        return ld.value, nn.value

    def solve(self, chi, psi=None, solver=CG, rsd=1e-8, maxit=1000, delta=0.1, trace_cap=100000):
        psi = np.zeros_like(chi) if psi is None else psi.copy()
        nc = ctypes.c_int()
        rs = ctypes.c_double()
        tr = np.zeros(trace_cap, np.float64)
        tl = ctypes.c_int()
        _check(lib().orc_solve(self.h, int(solver), _d(psi), _d(chi), float(rsd), int(maxit), float(delta),
                               ctypes.byref(nc), ctypes.byref(rs), _d(tr), int(trace_cap), ctypes.byref(tl)))
        return psi, nc.value, rs.value, tr[: min(tl.value, trace_cap)]

    def solve_mdagm(self, chi, psi=None, solver=CG, rsd=1e-8, maxit=1000, delta=0.1):
        """psi = (M^dag M)^-1 chi with the MdagM system solvers; returns (psi, n_count, |chi - M^dag M psi|)."""
        psi = np.zeros_like(chi) if psi is None else psi.copy()
        nc = ctypes.c_int()
        rs = ctypes.c_double()
        _check(lib().orc_solve_mdagm(self.h, int(solver), _d(psi), _d(chi), float(rsd), int(maxit), float(delta),
                                     ctypes.byref(nc), ctypes.byref(rs)))
        return psi, nc.value, rs.value

    def solve_multishift(self, chi, shifts, rsd, maxit=1000):
        """MInvCG2 (minvcg2.cc:72-355): psi[s] = (M^dag M + shifts[s])^-1 chi; returns
        (psi[nshift, ...], n_count, resid[nshift])."""
        shifts = np.ascontiguousarray(shifts, np.float64)
        rsd = np.ascontiguousarray(np.broadcast_to(rsd, shifts.shape), np.float64)
        psi = np.zeros((len(shifts),) + chi.shape, np.complex128)
        nc = ctypes.c_int()
        res = np.zeros(len(shifts))
        _check(lib().orc_minvcg2(self.h, len(shifts), _d(shifts), _d(rsd), int(maxit), _d(psi),
                                 _d(np.ascontiguousarray(chi, np.complex128)), ctypes.byref(nc), _d(res)))
        return psi, nc.value, res

    def qprop(self, chi, solver=CG, rsd=1e-8, maxit=1000, delta=0.1):
        psi = np.zeros_like(chi)
        nc = ctypes.c_int()
        rs = ctypes.c_double()
        _check(lib().orc_qprop(self.h, int(solver), _d(psi), _d(chi), float(rsd), int(maxit), float(delta),
                               ctypes.byref(nc), ctypes.byref(rs)))
        return psi, nc.value, rs.value

    def quarkprop4_point(self, src_site=0, solver=CG, rsd=1e-8, maxit=1000, delta=0.1, want_prop=False):
        ncs = (ctypes.c_int * 12)()
        res = np.zeros(12, np.float64)
        corr = np.zeros(self.L[3], np.float64)
        prop = np.zeros((self.V, 12, 12), np.complex128) if want_prop else None
        _check(lib().orc_quarkprop4_point(self.h, int(solver), int(src_site), float(rsd), int(maxit), float(delta),
                                          ncs, _d(res), _d(corr), _d(prop) if want_prop else None))
        return list(ncs), res, corr, prop

    def time(self, what, n):
        t = ctypes.c_double()
        _check(lib().orc_time(self.h, int(what), int(n), ctypes.byref(t)))
        return t.value
