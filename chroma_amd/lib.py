"""ctypes binding of libchroma_mi355x.so (declared in include/chroma_mi355x.h).

There is no fallback path: if the HIP library is missing, every entry point raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libchroma_mi355x.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "chroma_mi355x.h")

CMX_OK = 0
CMX_ERR_ARG, CMX_ERR_HIP, CMX_ERR_STATE, CMX_ERR_NOCONV, CMX_ERR_BREAKDOWN, CMX_ERR_UNSUPPORTED = 1, 2, 3, 4, 5, 6
DOUBLE, SINGLE, HALF = 0, 1, 2
RECONS_18, RECONS_12 = 18, 12
UNPREC_WILSON, EO_WILSON, EO_CLOVER = 0, 1, 2
CG, BICGSTAB, RELIABLE_CG, RELIABLE_BICGSTAB, FGMRESDR = 0, 1, 2, 3, 4
PRECOND_NONE, PRECOND_MR, PRECOND_LEARNED, PRECOND_MG = 0, 1, 2, 3
PLUS, MINUS = 1, -1


class CmxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[cmx error {code}] {msg}")
        self.code = code


class OpParams(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("mass", ctypes.c_double), ("clov_coeff_r", ctypes.c_double),
                ("clov_coeff_t", ctypes.c_double), ("sloppy_prec", ctypes.c_int), ("sloppy_recon", ctypes.c_int),
                ("autotune_dslash", ctypes.c_int)]


class SolverParams(ctypes.Structure):
    _fields_ = [("solver", ctypes.c_int), ("max_iter", ctypes.c_int), ("rsd_target", ctypes.c_double),
                ("delta", ctypes.c_double), ("sloppy_prec", ctypes.c_int), ("check_every", ctypes.c_int),
                ("verbose", ctypes.c_int), ("true_residual", ctypes.c_int), ("n_krylov", ctypes.c_int),
                ("n_defl", ctypes.c_int), ("precond", ctypes.c_int), ("precond_iters", ctypes.c_int),
                ("precond_rsd", ctypes.c_double), ("precond_omega", ctypes.c_double), ("reorth", ctypes.c_int),
                ("basis_prec", ctypes.c_int), ("cycle_reduction", ctypes.c_double), ("graph", ctypes.c_int)]


class MGParams(ctypes.Structure):
    """cmx_mg_params: the two-level form of MULTIGRIDSolverParams (quda_multigrid_params.h:12-113)."""
    _fields_ = [("blocking", ctypes.c_int * 4), ("nvec", ctypes.c_int), ("nu_pre", ctypes.c_int),
                ("nu_post", ctypes.c_int), ("omega", ctypes.c_double), ("coarse_tol", ctypes.c_double),
                ("coarse_maxit", ctypes.c_int), ("coarse_nkrylov", ctypes.c_int), ("setup_maxit", ctypes.c_int),
                ("seed", ctypes.c_ulonglong), ("verbose", ctypes.c_int), ("prec", ctypes.c_int),
                ("smoother", ctypes.c_int)]


class MGResult(ctypes.Structure):
    _fields_ = [("n_blocks", ctypes.c_int), ("coarse_dof", ctypes.c_int), ("n_probes", ctypes.c_int),
                ("setup_seconds", ctypes.c_double), ("nullspace_seconds", ctypes.c_double),
                ("coarse_op_seconds", ctypes.c_double), ("nullvec_rel_residual_max", ctypes.c_double)]


class TrainParams(ctypes.Structure):
    _fields_ = [("n_sources", ctypes.c_int), ("snapshot_every", ctypes.c_int), ("solve_rsd", ctypes.c_double),
                ("epochs", ctypes.c_int), ("lr", ctypes.c_double), ("seed", ctypes.c_ulonglong),
                ("max_pairs", ctypes.c_int), ("prec", ctypes.c_int), ("verbose", ctypes.c_int),
                ("pairs_per_source", ctypes.c_int), ("normalize", ctypes.c_int)]


class TrainResult(ctypes.Structure):
    _fields_ = [("n_pairs", ctypes.c_int), ("n_steps", ctypes.c_int), ("loss_first", ctypes.c_double),
                ("loss_last", ctypes.c_double), ("gen_seconds", ctypes.c_double), ("train_seconds", ctypes.c_double)]


class SolverResult(ctypes.Structure):
    _fields_ = [("n_count", ctypes.c_int), ("resid", ctypes.c_double), ("rel_resid", ctypes.c_double),
                ("n_reliable", ctypes.c_int), ("solve_seconds", ctypes.c_double),
                ("total_seconds", ctypes.c_double), ("gflops", ctypes.c_double), ("n_restarts", ctypes.c_int)]


class BatchParams(ctypes.Structure):
    _fields_ = [("tune", ctypes.c_int), ("graph", ctypes.c_int), ("max_concurrent", ctypes.c_int)]


class BatchResult(ctypes.Structure):
    _fields_ = [("solve_seconds", ctypes.c_double), ("total_seconds", ctypes.c_double), ("n_iter", ctypes.c_int),
                ("n_stages", ctypes.c_int), ("blocks_per_iter", ctypes.c_long), ("hop_block", ctypes.c_int),
                ("hop_fast", ctypes.c_int)]


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)
_vp = ctypes.c_void_p
_L = None


def lib():
    """Loads the HIP library; raises if it has not been built."""
    global _L
    if _L is not None:
        return _L
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} not found: build it with `python -m chroma_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "cmx_init": (ctypes.c_int, [ctypes.c_int]),
        "cmx_finalize": (ctypes.c_int, []),
        "cmx_last_error": (ctypes.c_char_p, []),
        "cmx_version": (ctypes.c_int, []),
        "cmx_gauge_create": (ctypes.c_int, [_ip, _dp, ctypes.POINTER(_vp)]),
        "cmx_gauge_create_dirs": (ctypes.c_int, [_ip, ctypes.POINTER(_dp), ctypes.POINTER(_vp)]),
        "cmx_gauge_destroy": (ctypes.c_int, [_vp]),
        "cmx_op_create": (ctypes.c_int, [_vp, ctypes.POINTER(OpParams), ctypes.POINTER(_vp)]),
        "cmx_op_destroy": (ctypes.c_int, [_vp]),
        "cmx_op_subset_doubles": (ctypes.c_long, [_vp]),
        "cmx_op_apply": (ctypes.c_int, [_vp, _dp, _dp, ctypes.c_int, ctypes.c_int]),
        "cmx_dslash_apply": (ctypes.c_int, [_vp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "cmx_op_eo_piece": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, ctypes.c_int]),
        "cmx_dslash_multi": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "cmx_dslash_deriv": (ctypes.c_int, [_vp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int]),
        "cmx_clover_deriv": (ctypes.c_int, [_vp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int]),
        "cmx_clover_deriv_trln": (ctypes.c_int, [_vp, _dp]),
        "cmx_op_deriv": (ctypes.c_int, [_vp, _dp, _dp, _dp, ctypes.c_int]),
        "cmx_clover_download": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
        "cmx_clover_logdet": (ctypes.c_int, [_vp, _dp, _lp]),
        "cmx_clover_pack_quda": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp]),
        "cmx_solve": (ctypes.c_int, [_vp, ctypes.POINTER(SolverParams), _dp, _dp, ctypes.c_int,
                                     ctypes.POINTER(SolverResult)]),
        "cmx_solve_mdagm": (ctypes.c_int, [_vp, ctypes.POINTER(SolverParams), _dp, _dp, ctypes.c_int,
                                           ctypes.POINTER(SolverResult)]),
        "cmx_solve_multishift": (ctypes.c_int, [_vp, ctypes.POINTER(SolverParams), ctypes.c_int, _dp, _dp, _dp, _dp,
                                                ctypes.POINTER(SolverResult)]),
        "cmx_solve_batch": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.POINTER(SolverParams),
                                           ctypes.POINTER(BatchParams), ctypes.POINTER(_dp), ctypes.POINTER(_dp),
                                           ctypes.c_int, ctypes.POINTER(SolverResult), ctypes.POINTER(BatchResult)]),
        "cmx_net_create": (ctypes.c_int, [ctypes.POINTER(_vp)]),
        "cmx_net_destroy": (ctypes.c_int, [_vp]),
        "cmx_net_init_neumann": (ctypes.c_int, [_vp, ctypes.c_double]),
        "cmx_net_get_weights": (ctypes.c_int, [_vp, _dp, _dp, _dp]),
        "cmx_net_get_block": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp]),
        "cmx_net_set_block": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp]),
        "cmx_net_set_depth": (ctypes.c_int, [_vp, ctypes.c_int]),
        "cmx_net_depth": (ctypes.c_int, [_vp, _ip]),
        "cmx_net_set_weights": (ctypes.c_int, [_vp, _dp, _dp, _dp]),
        "cmx_op_set_net": (ctypes.c_int, [_vp, _vp]),
        "cmx_net_apply": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp]),
        "cmx_op_set_net_variant": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
        "cmx_net_grad": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, _dp, _dp]),
        "cmx_net_train": (ctypes.c_int, [_vp, ctypes.POINTER(TrainParams), ctypes.POINTER(TrainResult)]),
        "cmx_mg_params_default": (ctypes.c_int, [ctypes.POINTER(MGParams)]),
        "cmx_mg_setup": (ctypes.c_int, [_vp, ctypes.POINTER(MGParams), ctypes.POINTER(MGResult)]),
        "cmx_mg_update": (ctypes.c_int, [_vp, ctypes.POINTER(MGParams)]),
        "cmx_mg_get_nullvecs": (ctypes.c_int, [_vp, _dp]),
        "cmx_mg_get_coarse_op": (ctypes.c_int, [_vp, _dp]),
        "cmx_mg_restrict": (ctypes.c_int, [_vp, _dp, _dp]),
        "cmx_mg_prolong": (ctypes.c_int, [_vp, _dp, _dp]),
        "cmx_mg_coarse_apply": (ctypes.c_int, [_vp, _dp, _dp]),
        "cmx_op_time": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]),
        "cmx_layout_roundtrip": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp]),
        "cmx_op_get_hop_config": (ctypes.c_int, [_vp, ctypes.c_int, _ip]),
        "cmx_op_link_recon": (ctypes.c_int, [_vp, ctypes.c_int, _ip]),
        "cmx_op_set_multi_group": (ctypes.c_int, [_vp, ctypes.c_int]),
        "cmx_op_set_hop_config": (ctypes.c_int, [_vp, ctypes.c_int, _ip]),
        "cmx_op_autotune": (ctypes.c_int, [_vp]),
        "cmx_gauge_generate": (ctypes.c_int, [_ip, ctypes.c_int, ctypes.c_ulonglong, ctypes.c_double, ctypes.c_int,
                                              ctypes.c_double, _dp]),
        "cmx_gauge_apply_bc": (ctypes.c_int, [_ip, _dp, _ip, _dp, _ip]),
        "cmx_plaquette": (ctypes.c_double, [_ip, _dp]),
        "cmx_gauge_stout_smear": (ctypes.c_int, [_ip, _dp, ctypes.c_int, ctypes.c_double]),
        "cmx_gauge_stout_smear_dirs": (ctypes.c_int, [_ip, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_int]),
        "cmx_meson_2pt": (ctypes.c_int, [_ip, _dp, _dp, _ip, ctypes.c_int, _ip, _dp]),
        "cmx_meson_2pt_noalloc": (ctypes.c_int, [_ip, _vp, _vp, _ip, ctypes.c_int, _vp, _vp, _vp]),
        "cmx_fill_z2": (ctypes.c_int, [ctypes.c_long, ctypes.c_ulonglong, _dp]),
        "cmx_field_write_scidac": (ctypes.c_int, [ctypes.c_char_p, _ip, ctypes.c_int, ctypes.c_int, _dp,
                                                  ctypes.c_char_p, ctypes.c_char_p]),
        "cmx_field_read_scidac": (ctypes.c_int, [ctypes.c_char_p, _ip, ctypes.c_int, _dp, ctypes.c_char_p,
                                                 ctypes.c_int, ctypes.c_char_p, ctypes.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _L = L
    return L


def check(rc):
    if rc != CMX_OK:
        raise CmxError(rc, lib().cmx_last_error().decode())


def _dims(L):
    return (ctypes.c_int * 4)(*[int(v) for v in L])


def _d(a):
    assert a.flags.c_contiguous and a.dtype in (np.float64, np.complex128), (a.dtype, a.flags)
    return a.ctypes.data_as(_dp)


_initialised = False


def init(device=0):
    global _initialised
    if not _initialised:
        check(lib().cmx_init(int(device)))
        _initialised = True


# ---- host-side synthetic inputs -----------------------------------------------------------------
def gauge_generate(L, kind=1, seed=1, eps=0.3, nstout=0, rho=0.1):
    """kind 0 unit, 1 exp(i eps H), 2 Haar; then nstout stout steps.  complex128[4, V, 3, 3]."""
    V = int(np.prod(L))
    U = np.zeros((4, V, 3, 3), np.complex128)
    check(lib().cmx_gauge_generate(_dims(L), int(kind), int(seed), float(eps), int(nstout), float(rho), _d(U)))
    return U


def gauge_apply_bc(L, U, boundary=(1, 1, 1, 1), twist_by_pi=None, twist_dir=(0, 1, 2)):
    U = np.ascontiguousarray(U, np.complex128)
    b = (ctypes.c_int * 4)(*boundary)
    if twist_by_pi is None:
        check(lib().cmx_gauge_apply_bc(_dims(L), _d(U), b, None, None))
    else:
        tw = np.ascontiguousarray(twist_by_pi, np.float64)
        td = (ctypes.c_int * 3)(*twist_dir)
        check(lib().cmx_gauge_apply_bc(_dims(L), _d(U), b, _d(tw), td))
    return U


def gauge_stout_smear(L, U, nstep, rho, smear_dirs=15):
    """nstep stout steps on the device (Stouting::smear_links, stout_utils.cc:936-962) over the directions
    whose bit is set in smear_dirs (15: all, 7: spatial); returns a new complex128[4, V, 3, 3]."""
    init()
    out = np.array(U, np.complex128, order="C", copy=True)
    if smear_dirs == 15:
        check(lib().cmx_gauge_stout_smear(_dims(L), _d(out), int(nstep), float(rho)))
    else:
        check(lib().cmx_gauge_stout_smear_dirs(_dims(L), _d(out), int(nstep), float(rho), int(smear_dirs)))
    return out


def meson_2pt(L, prop1, prop2=None, origin=(0, 0, 0, 0), moms=((0, 0, 0),)):
    """mesons() (mesons_w.cc:97-180) on the device: complex128[16, len(moms), L_t], gamma insertion n,
    sink momentum m, time counted from origin[3]; mesons() writes the real part.  Propagators are
    complex128[V, 12, 12] (sink spin*3+colour, source spin*3+colour)."""
    init()
    p1 = np.ascontiguousarray(prop1, np.complex128)
    p2 = p1 if prop2 is None else np.ascontiguousarray(prop2, np.complex128)
    V = int(np.prod(L))
    if p1.shape != (V, 12, 12) or p2.shape != (V, 12, 12):
        raise ValueError("meson_2pt: propagators must be complex[V, 12, 12]")
    mom = np.ascontiguousarray(moms, np.int32).reshape(-1, 3)
    out = np.zeros((16, mom.shape[0], int(L[3])), np.complex128)
    org = (ctypes.c_int * 4)(*origin)
    check(lib().cmx_meson_2pt(_dims(L), _d(p1), _d(p2), org, int(mom.shape[0]),
                              mom.ctypes.data_as(_ip), _d(out)))
    return out


SCIDAC_PROPAGATOR, SCIDAC_FERMION = 0, 1


def field_write_scidac(path, L, field, kind=SCIDAC_PROPAGATOR, prec=DOUBLE, file_xml="", record_xml=""):
    """SciDAC single-file output of a propagator (complex[V, 12, 12]) or fermion (complex[V, 4, 3]) in the
    host site order (cmx_field_write_scidac; QIO_WRITE_NAMED_OBJECT's file).  Host only."""
    V = int(np.prod(L))
    f = np.ascontiguousarray(field, np.complex128)
    if f.size != V * (144 if kind == SCIDAC_PROPAGATOR else 12):
        raise ValueError("field_write_scidac: field size does not match the lattice and kind")
    check(lib().cmx_field_write_scidac(os.fsencode(path), _dims(L), int(kind), int(prec), _d(f),
                                       file_xml.encode(), record_xml.encode()))


def field_read_scidac(path, L, kind=SCIDAC_PROPAGATOR, xml_cap=1 << 20):
    """Reads a SciDAC single-file propagator or fermion written in double or single precision; checks
    dims, datatype and checksum.  Returns (field, file_xml, record_xml)."""
    V = int(np.prod(L))
    out = np.zeros((V, 12, 12) if kind == SCIDAC_PROPAGATOR else (V, 4, 3), np.complex128)
    fx = ctypes.create_string_buffer(xml_cap)
    rx = ctypes.create_string_buffer(xml_cap)
    check(lib().cmx_field_read_scidac(os.fsencode(path), _dims(L), int(kind), _d(out), fx, xml_cap, rx, xml_cap))
    return out, fx.value.decode(), rx.value.decode()


def plaquette(L, U):
    return lib().cmx_plaquette(_dims(L), _d(np.ascontiguousarray(U, np.complex128)))


def z2_source(shape, seed):
    out = np.zeros(shape, np.complex128)
    check(lib().cmx_fill_z2(out.size * 2, int(seed), _d(out)))
    return out


def kappa_to_mass(kappa):
    """kappaToMass, lib/io/param_io.cc:12-15."""
    return 1.0 / (2.0 * kappa) - 4.0


# ---- device objects -------------------------------------------------------------------------------
class Gauge:
    """Device gauge field from host links: one array U[4, V, 3, 3], or (per_direction=True) a sequence
    of four arrays U[mu][V, 3, 3] handed over as separate pointers (QUDA_QDP_GAUGE_ORDER)."""

    def __init__(self, L, U, per_direction=False):
        init()
        self.L = tuple(int(v) for v in L)
        h = _vp()
        if per_direction:
            dirs = [np.ascontiguousarray(u, np.complex128) for u in U]
            assert len(dirs) == 4
            ptrs = (_dp * 4)(*[_d(u) for u in dirs])
            check(lib().cmx_gauge_create_dirs(_dims(L), ptrs, ctypes.byref(h)))
        else:
            U = np.ascontiguousarray(U, np.complex128)
            check(lib().cmx_gauge_create(_dims(L), _d(U), ctypes.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _L is not None:
            _L.cmx_gauge_destroy(self.h)
            self.h = None


class Operator:
    """A device fermion operator (LinearOperator<LatticeFermion>, lib/linearop.h:25-49)."""

    def __init__(self, gauge, kind, mass, csw_r=0.0, csw_t=None, sloppy_prec=SINGLE, sloppy_recon=RECONS_12,
                 autotune=False):
        self.gauge = gauge
        self.L = gauge.L
        self.V = int(np.prod(self.L))
        self.kind = kind
        p = OpParams(int(kind), float(mass), float(csw_r), float(csw_r if csw_t is None else csw_t),
                     int(sloppy_prec), int(sloppy_recon), int(bool(autotune)))
        h = _vp()
        check(lib().cmx_op_create(gauge.h, ctypes.byref(p), ctypes.byref(h)))
        self.h = h
        self.nsub = lib().cmx_op_subset_doubles(h) // 24

    def __del__(self):
        if getattr(self, "h", None) and _L is not None:
            _L.cmx_op_destroy(self.h)
            self.h = None

    def subset_zeros(self):
        return np.zeros((self.nsub, 4, 3), np.complex128)

    def __call__(self, psi, isign=PLUS, prec=DOUBLE):
        psi = np.ascontiguousarray(psi, np.complex128)
        chi = np.zeros_like(psi)
        check(lib().cmx_op_apply(self.h, _d(chi), _d(psi), int(isign), int(prec)))
        return chi

    def dslash(self, psi_cb, isign, cb, prec=DOUBLE):
        psi_cb = np.ascontiguousarray(psi_cb, np.complex128)
        chi = np.zeros_like(psi_cb)
        check(lib().cmx_dslash_apply(self.h, _d(chi), _d(psi_cb), int(isign), int(cb), int(prec)))
        return chi

    def dslash_multi(self, psis_cb, isign, cb, prec=HALF):
        """D on n sources at once: psis_cb complex[n, Vh, 4, 3]."""
        psis_cb = np.ascontiguousarray(psis_cb, np.complex128)
        chi = np.zeros_like(psis_cb)
        check(lib().cmx_dslash_multi(self.h, int(psis_cb.shape[0]), _d(chi), _d(psis_cb), int(isign), int(cb),
                                     int(prec)))
        return chi

    def dslash_deriv(self, chi_full, psi_full, isign=PLUS, cb=-1):
        """WilsonDslashBase::deriv (lwldslash_base_w.h:132-227): ds_u complex[4, V, 3, 3] from full-lattice
        chi, psi; cb 0/1 one checkerboard, -1 both."""
        ds = np.zeros((4, self.V, 3, 3), np.complex128)
        check(lib().cmx_dslash_deriv(self.h, _d(ds), _d(np.ascontiguousarray(chi_full, np.complex128)),
                                     _d(np.ascontiguousarray(psi_full, np.complex128)), int(isign), int(cb)))
        return ds

    def clover_deriv(self, chi_full, psi_full, isign=PLUS, cb=-1):
        """CloverTermBase::deriv: ds[4, V, 3, 3] of chi^dag A' psi over cb (-1: both checkerboards)."""
        ds = np.zeros((4, self.V, 3, 3), np.complex128)
        check(lib().cmx_clover_deriv(self.h, _d(ds), _d(np.ascontiguousarray(chi_full, np.complex128)),
                                     _d(np.ascontiguousarray(psi_full, np.complex128)), int(isign), int(cb)))
        return ds

    def clover_deriv_trln(self):
        """CloverTermBase::derivTrLn of the inverted even-even clover term (the even-even log det force)."""
        ds = np.zeros((4, self.V, 3, 3), np.complex128)
        check(lib().cmx_clover_deriv_trln(self.h, _d(ds)))
        return ds

    def deriv(self, chi_odd, psi_odd, isign=PLUS):
        """EvenOddPrecLogDetLinearOperator::deriv: ds[4, V, 3, 3] of chi^dag M'(isign) psi (odd arrays)."""
        ds = np.zeros((4, self.V, 3, 3), np.complex128)
        check(lib().cmx_op_deriv(self.h, _d(ds), _d(np.ascontiguousarray(chi_odd, np.complex128)),
                                 _d(np.ascontiguousarray(psi_odd, np.complex128)), int(isign)))
        return ds

    def piece(self, which, psi_full, isign=PLUS):
        names = {"evenEven": 0, "evenEvenInv": 1, "evenOdd": 2, "oddEven": 3, "oddOdd": 4, "unprec": 5}
        psi_full = np.ascontiguousarray(psi_full, np.complex128)
        chi = np.zeros_like(psi_full)
        check(lib().cmx_op_eo_piece(self.h, names[which], _d(chi), _d(psi_full), int(isign)))
        return chi

    def clover(self, which=0):
        tri = np.zeros((self.V, 72), np.float64)
        check(lib().cmx_clover_download(self.h, int(which), _d(tri)))
        return tri

    def clover_logdet(self):
        ld = ctypes.c_double()
        nn = ctypes.c_long()
        check(lib().cmx_clover_logdet(self.h, ctypes.byref(ld), ctypes.byref(nn)))
        return ld.value, nn.value

    def clover_pack_quda(self, which, cb, out=None):
        """QUDAPackedClovSite array over the full lattice (V x 72 doubles, host site order); only the
        sites of checkerboard cb are written (packForQUDA, clover_term_qdp_w.h:2171-2228)."""
        out = np.zeros((self.V, 72), np.float64) if out is None else out
        assert out.shape == (self.V, 72) and out.dtype == np.float64 and out.flags.c_contiguous
        check(lib().cmx_clover_pack_quda(self.h, int(which), int(cb), _d(out)))
        return out

    # ---- multigrid (cmx_mg_*) ----
    def mg_setup(self, params=None, **kw):
        """Builds the two-level multigrid on this operator (cmx_mg_setup); params: MGParams (defaults of
        mg_params()), keyword overrides applied on top.  Returns the set-up summary."""
        p = mg_params(**kw) if params is None else params
        for k, v in kw.items():
            if params is not None:
                _mg_set(p, k, v)
        r = MGResult()
        check(lib().cmx_mg_setup(self.h, ctypes.byref(p), ctypes.byref(r)))
        self._mg = (r.n_blocks, r.coarse_dof, p.nvec)
        self._mg_par = MGParams.from_buffer_copy(p)
        return {f: getattr(r, f) for f, _ in MGResult._fields_}

    def mg_update(self, params=None, **kw):
        """Changes the V-cycle of the set-up multigrid (cmx_mg_update): nu_pre, nu_post, omega, coarse_tol,
        coarse_maxit, coarse_nkrylov; params: MGParams (default: those of the last mg_setup or mg_update),
        keyword overrides on top."""
        base = params if params is not None else getattr(self, "_mg_par", None)
        if base is None:
            raise CmxError(CMX_ERR_STATE, "mg_update: mg_setup has not been called on this operator")
        p = MGParams.from_buffer_copy(base)
        for k, v in kw.items():
            _mg_set(p, k, v)
        check(lib().cmx_mg_update(self.h, ctypes.byref(p)))
        self._mg_par = p

    def mg_nullvecs(self):
        nb, d, nvec = self._mg
        out = np.zeros((nvec, self.V, 4, 3), np.complex128)
        check(lib().cmx_mg_get_nullvecs(self.h, _d(out)))
        return out

    def mg_coarse_op(self):
        nb, d, _ = self._mg
        out = np.zeros((nb, 9, d, d), np.complex128)
        check(lib().cmx_mg_get_coarse_op(self.h, _d(out)))
        return out

    def mg_restrict(self, full):
        nb, d, _ = self._mg
        f = np.ascontiguousarray(full, np.complex128)
        out = np.zeros((nb, d), np.complex128)
        check(lib().cmx_mg_restrict(self.h, _d(f), _d(out)))
        return out

    def mg_prolong(self, coarse):
        c = np.ascontiguousarray(coarse, np.complex128)
        out = np.zeros((self.V, 4, 3), np.complex128)
        check(lib().cmx_mg_prolong(self.h, _d(c), _d(out)))
        return out

    def mg_coarse_apply(self, coarse):
        c = np.ascontiguousarray(coarse, np.complex128)
        out = np.zeros_like(c)
        check(lib().cmx_mg_coarse_apply(self.h, _d(c), _d(out)))
        return out

    def layout_roundtrip(self, psi_cb, prec):
        psi_cb = np.ascontiguousarray(psi_cb, np.complex128)
        out = np.zeros_like(psi_cb)
        check(lib().cmx_layout_roundtrip(self.h, int(prec), _d(psi_cb), _d(out)))
        return out

    def hop_config(self, prec, target=1):
        """Launch shape of the EO operator's kernels (target 1) or of the plain hopping term (2);
        `fast` is the kernel admitted for this operator (bit 0 plane-grid, bit 1 streamed loads)."""
        cfg = (ctypes.c_int * 6)(0, 0, 0, 0, 0, int(target))
        check(lib().cmx_op_get_hop_config(self.h, int(prec), cfg))
        return dict(block=cfg[0], tiled=cfg[1], zb=cfg[2], prefetch=cfg[3], fast=cfg[4])

    def set_multi_group(self, group):
        """Kernel of dslash_multi: 1..4 sources per lane (vector units), 5 matrix cores, 0 timed at first use."""
        check(lib().cmx_op_set_multi_group(self.h, int(group)))

    def link_recon(self, prec):
        """Reals stored per link at `prec` (12 or 18)."""
        r = ctypes.c_int()
        check(lib().cmx_op_link_recon(self.h, int(prec), ctypes.byref(r)))
        return r.value

    def set_hop_config(self, prec, block=256, tiled=0, zb=1, prefetch=0, fast=0, target=0):
        """fast: 0 site-linear kernel, 1 plane-grid kernel; +2 with the read-once loads (links, clover,
        xpay input) streamed non-temporally."""
        cfg = (ctypes.c_int * 6)(int(block), int(tiled), int(zb), int(prefetch), int(fast), int(target))
        check(lib().cmx_op_set_hop_config(self.h, int(prec), cfg))

    def autotune(self):
        check(lib().cmx_op_autotune(self.h))

    def time(self, what, prec, n):
        ms = ctypes.c_double()
        check(lib().cmx_op_time(self.h, int(what), int(prec), int(n), ctypes.byref(ms)))
        return ms.value

    def solve(self, chi, psi=None, solver=CG, rsd=1e-10, maxit=10000, delta=0.1, sloppy_prec=SINGLE,
              check_every=0, verbose=False, true_residual=False, mdagm=False, n_krylov=16, n_defl=0,
              precond=PRECOND_NONE, precond_iters=4, precond_rsd=1e-12, precond_omega=1.0, reorth=1,
              basis_prec=-1, cycle_reduction=0.0, graph=False):
        """SystemSolver<T>::operator()(psi, chi) (lib/syssolver.h:46); chi/psi: [nrhs,] subset arrays.

        mdagm=True solves M^dag M psi = chi (MdagMSystemSolver, lib/syssolver.h:72-84)."""
        chi = np.ascontiguousarray(chi, np.complex128)
        single = chi.ndim == 3
        chis = chi[None] if single else chi
        psis = np.zeros_like(chis) if psi is None else np.ascontiguousarray(psi, np.complex128).reshape(chis.shape).copy()
        nrhs = chis.shape[0]
        p = SolverParams(int(solver), int(maxit), float(rsd), float(delta), int(sloppy_prec), int(check_every),
                         int(bool(verbose)), int(bool(true_residual)), int(n_krylov), int(n_defl), int(precond),
                         int(precond_iters), float(precond_rsd), float(precond_omega), int(reorth), int(basis_prec),
                         float(cycle_reduction), int(bool(graph)))
        res = (SolverResult * nrhs)()
        fn = lib().cmx_solve_mdagm if mdagm else lib().cmx_solve
        check(fn(self.h, ctypes.byref(p), _d(psis), _d(chis), nrhs, res))
        out = [dict(n_count=r.n_count, resid=r.resid, rel_resid=r.rel_resid, n_reliable=r.n_reliable,
                    solve_seconds=r.solve_seconds, total_seconds=r.total_seconds, gflops=r.gflops,
                    n_restarts=r.n_restarts) for r in res]
        return (psis[0], out[0]) if single else (psis, out)

    def solve_multishift(self, chi, shifts, rsd=1e-8, maxit=1000, verbose=False):
        """(M^dag M + shifts[s])^-1 chi for every shift with one Krylov space (MInvCG2, minvcg2.cc:72-355),
        FP64.  rsd: one relative target or one per shift.  Returns (psi[nshift, ...], [result per shift])."""
        chi = np.ascontiguousarray(chi, np.complex128)
        shifts = np.ascontiguousarray(shifts, np.float64)
        rsd = np.ascontiguousarray(np.broadcast_to(np.asarray(rsd, np.float64), shifts.shape))
        ns = len(shifts)
        psis = np.zeros((ns,) + chi.shape, np.complex128)
        p = SolverParams(CG, int(maxit), float(rsd.max()), 0.1, DOUBLE, 0, int(bool(verbose)), 0, 16, 0, 0, 4,
                         1e-12, 1.0, 1, -1, 0.0, 0)
        res = (SolverResult * ns)()
        check(lib().cmx_solve_multishift(self.h, ctypes.byref(p), ns, _d(shifts), _d(rsd), _d(psis), _d(chi), res))
        return psis, [dict(n_count=r.n_count, resid=r.resid, rel_resid=r.rel_resid, solve_seconds=r.solve_seconds,
                           total_seconds=r.total_seconds) for r in res]


def solve_batch(ops, chis, psis=None, solver=RELIABLE_CG, rsd=1e-10, maxit=10000, delta=0.1, sloppy_prec=HALF,
                check_every=0, mdagm=False, true_residual=False, tune=True, graph=False, n_krylov=16, n_defl=0,
                precond=PRECOND_NONE, precond_iters=4, precond_rsd=1e-12, precond_omega=1.0, reorth=1,
                basis_prec=-1, cycle_reduction=0.0, max_concurrent=0):
    """n independent solves advanced together (cmx_solve_batch): CG and reliable CG launch every kernel
    once for all problems; FGMRES-DR advances all solves one Arnoldi step per host read-back.
    ops: Operators (an Operator may repeat: the sources of one propagator); chis: subset arrays, one per
    problem.  Returns (psis, per-problem results, batch summary)."""
    n = len(ops)
    assert len(chis) == n
    chis = [np.ascontiguousarray(c, np.complex128) for c in chis]
    psis = [np.zeros_like(c) for c in chis] if psis is None else \
        [np.ascontiguousarray(x, np.complex128).copy() for x in psis]
    p = SolverParams(int(solver), int(maxit), float(rsd), float(delta), int(sloppy_prec), int(check_every),
                     0, int(bool(true_residual)), int(n_krylov), int(n_defl), int(precond), int(precond_iters),
                     float(precond_rsd), float(precond_omega), int(reorth), int(basis_prec), float(cycle_reduction),
                     0)
    bp = BatchParams(int(bool(tune)), int(bool(graph)), int(max_concurrent))
    hs = (_vp * n)(*[o.h for o in ops])
    pp = (_dp * n)(*[_d(x) for x in psis])
    cp = (_dp * n)(*[_d(c) for c in chis])
    res = (SolverResult * n)()
    br = BatchResult()
    check(lib().cmx_solve_batch(hs, n, ctypes.byref(p), ctypes.byref(bp), pp, cp, int(bool(mdagm)), res,
                                ctypes.byref(br)))
    out = [dict(n_count=r.n_count, resid=r.resid, rel_resid=r.rel_resid, n_reliable=r.n_reliable,
                solve_seconds=r.solve_seconds, total_seconds=r.total_seconds, gflops=r.gflops,
                n_restarts=r.n_restarts) for r in res]
    summary = dict(solve_seconds=br.solve_seconds, total_seconds=br.total_seconds, n_iter=br.n_iter,
                   n_stages=br.n_stages, blocks_per_iter=br.blocks_per_iter, hop_block=br.hop_block,
                   hop_fast=br.hop_fast)
    return psis, out, summary


class Net:
    """Learned preconditioner: a gauge-equivariant parallel-transport network (see include/chroma_mi355x.h).

    Weights: W1 complex[64, 32] over ((channel, spin), (path, spin)), W2 complex[32, 64], W0 complex[4, 4];
    path p = 2 mu + (0 forward, 1 backward)."""

    def __init__(self):
        h = _vp()
        check(lib().cmx_net_create(ctypes.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _L is not None:
            _L.cmx_net_destroy(self.h)
            self.h = None

    def init_neumann(self, diag):
        check(lib().cmx_net_init_neumann(self.h, float(diag)))

    def set_depth(self, nblocks):
        """1..4 blocks applied one after another; blocks added start as the identity."""
        check(lib().cmx_net_set_depth(self.h, int(nblocks)))

    def depth(self):
        n = ctypes.c_int()
        check(lib().cmx_net_depth(self.h, ctypes.byref(n)))
        return n.value

    def get_block(self, blk):
        w1 = np.zeros((64, 32), np.complex128)
        w2 = np.zeros((32, 64), np.complex128)
        w0 = np.zeros((4, 4), np.complex128)
        check(lib().cmx_net_get_block(self.h, int(blk), _d(w1), _d(w2), _d(w0)))
        return w1, w2, w0

    def set_block(self, blk, w1, w2, w0):
        w1 = np.ascontiguousarray(w1, np.complex128)
        w2 = np.ascontiguousarray(w2, np.complex128)
        w0 = np.ascontiguousarray(w0, np.complex128)
        assert w1.shape == (64, 32) and w2.shape == (32, 64) and w0.shape == (4, 4)
        check(lib().cmx_net_set_block(self.h, int(blk), _d(w1), _d(w2), _d(w0)))

    def get_weights(self):
        return self.get_block(0)

    def set_weights(self, w1, w2, w0):
        self.set_block(0, w1, w2, w0)


def net_attach(op, net):
    check(lib().cmx_op_set_net(op.h, net.h if net is not None else None))
    op.net = net


def net_set_variant(op, prec, variant):
    check(lib().cmx_op_set_net_variant(op.h, int(prec), int(variant)))


def net_apply(op, v, prec=SINGLE):
    v = np.ascontiguousarray(v, np.complex128)
    out = np.zeros_like(v)
    check(lib().cmx_net_apply(op.h, int(prec), _d(out), _d(v)))
    return out


def net_grad(op, v, e, prec=SINGLE):
    """(|P v - e|^2, d/d conj(W2 W1) complex[32, 32], d/d conj(W0) complex[4, 4]) of a one-block network;
    for a deeper network the gradients are complex[nblocks, 32, 32] and complex[nblocks, 4, 4]."""
    v = np.ascontiguousarray(v, np.complex128)
    e = np.ascontiguousarray(e, np.complex128)
    nb = op.net.depth() if getattr(op, "net", None) is not None else 1
    loss = ctypes.c_double()
    gwc = np.zeros((nb, 32, 32), np.complex128)
    gw0 = np.zeros((nb, 4, 4), np.complex128)
    check(lib().cmx_net_grad(op.h, int(prec), _d(v), _d(e), ctypes.byref(loss), _d(gwc), _d(gw0)))
    if nb == 1:
        return loss.value, gwc[0], gw0[0]
    return loss.value, gwc, gw0


def _mg_set(p, k, v):
    if k == "blocking":
        for i in range(4):
            p.blocking[i] = int(v[i])
    else:
        setattr(p, k, type(getattr(p, k))(v))


def mg_params(**kw):
    """cmx_mg_params with the defaults of cmx_mg_params_default (the reader's), keyword overrides applied."""
    p = MGParams()
    check(lib().cmx_mg_params_default(ctypes.byref(p)))
    for k, v in kw.items():
        _mg_set(p, k, v)
    return p


def mg_params_from_xml(xml):
    """MULTIGRIDSolverParams read from a <MULTIGRIDParams> element, with the element names and defaults of
    quda_multigrid_params.cc:44-235, restricted to two levels (one Blocking entry; per-level arrays give
    their first-level value, or their coarse-level value for the coarse solver's CoarseResidual and
    MaxCoarseIterations; a single value is broadcast as the reader does).  Precision (HALF, SINGLE,
    DOUBLE; the reader's DEFAULT keeps the default) sets the null-space relaxation's precision.  Elements
    without a counterpart here (Reconstruct, SchwarzType, SetupOnGPU, ...) are accepted and ignored."""
    import xml.etree.ElementTree as ET
    root = ET.fromstring(xml) if isinstance(xml, str) else xml
    if root.tag != "MULTIGRIDParams":
        found = root.find(".//MULTIGRIDParams")
        if found is None:
            raise ValueError("mg_params_from_xml: no MULTIGRIDParams element")
        root = found

    def arr(name, conv):
        e = root.find(name)
        return None if e is None or e.text is None else [conv(t) for t in e.text.split()]

    p = mg_params()
    blk = root.find("Blocking")
    if blk is None:
        raise ValueError("mg_params_from_xml: Blocking is required")
    elems = [[int(t) for t in e.text.split()] for e in blk.findall("elem")]
    if len(elems) != 1:
        raise ValueError(f"mg_params_from_xml: {len(elems)} blockings; this multigrid has two levels (one blocking)")
    _mg_set(p, "blocking", elems[0])
    levels = 2

    def level_value(name, conv, level):
        v = arr(name, conv)
        if v is None:
            return None
        if len(v) == 1:
            return v[0]
        if name in ("NullVectors", "Pre-SmootherApplications", "Post-SmootherApplications",
                    "MaxIterSubspaceCreate") and len(v) == levels - 1:
            return v[level]
        if len(v) != levels:
            raise ValueError(f"mg_params_from_xml: {name} has {len(v)} entries for {levels} levels")
        return v[level]

    for name, field, conv, level in (("NullVectors", "nvec", int, 0),
                                     ("Pre-SmootherApplications", "nu_pre", int, 0),
                                     ("Post-SmootherApplications", "nu_post", int, 0),
                                     ("RelaxationOmegaMG", "omega", float, 0),
                                     ("CoarseResidual", "coarse_tol", float, 1),
                                     ("MaxCoarseIterations", "coarse_maxit", int, 1),
                                     ("MaxIterSubspaceCreate", "setup_maxit", int, 0)):
        v = level_value(name, conv, level)
        if v is not None:
            _mg_set(p, field, v)
    for name, field in (("PrecondGCRNKrylov", "coarse_nkrylov"),):
        e = root.find(name)
        if e is not None:
            _mg_set(p, field, int(e.text))
    e = root.find("Verbosity")
    if e is not None:
        p.verbose = 1 if e.text.strip().lower() == "true" else 0
    e = root.find("Precision")
    if e is not None and e.text is not None:
        names = {"HALF": HALF, "SINGLE": SINGLE, "DOUBLE": DOUBLE}
        t = e.text.strip().upper()
        if t in names:
            p.prec = names[t]
        elif t != "DEFAULT":
            raise ValueError(f"mg_params_from_xml: Precision {e.text.strip()}")
    return p


def net_train(op, n_sources=4, snapshot_every=4, solve_rsd=1e-9, epochs=4, lr=1e-3, seed=1, max_pairs=64,
              prec=SINGLE, verbose=False, pairs_per_source=0, normalize=True):
    p = TrainParams(int(n_sources), int(snapshot_every), float(solve_rsd), int(epochs), float(lr), int(seed),
                    int(max_pairs), int(prec), int(bool(verbose)), int(pairs_per_source), int(bool(normalize)))
    r = TrainResult()
    check(lib().cmx_net_train(op.h, ctypes.byref(p), ctypes.byref(r)))
    return dict(n_pairs=r.n_pairs, n_steps=r.n_steps, loss_first=r.loss_first, loss_last=r.loss_last,
                gen_seconds=r.gen_seconds, train_seconds=r.train_seconds)


def declared_symbols():
    """Function names declared in include/chroma_mi355x.h."""
    import re
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|long|double|const char\*)\s+(cmx_\w+)\s*\(", txt, re.M)))
