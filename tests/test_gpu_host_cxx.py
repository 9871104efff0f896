"""GPU: the C++ host layer over the C-ABI (chroma_amd/host/mi355x_solver.{h,cpp}: SystemSolver,
EvenOddPrecLinearOperator blocks, PrecFermActQprop, quarkProp4) through its propagator driver, run as
a child process the way a Chroma propagator task would run: the reference's twisted-BC golden, and
agreement with the Python path over the same C-ABI."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from chroma_amd import lib as cmx
from chroma_amd import qprop as qp

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "chroma_amd", "host", "mi355x_prop")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run_driver(args):
    say("driver", " ".join(args))
    p = subprocess.run([DRIVER] + args, capture_output=True, text=True, timeout=200)
    assert p.returncode == 0, p.stderr
    return json.loads(p.stdout.strip().splitlines()[-1])


def test_twisted_golden_through_cxx_host():
    """prec_wilson-twisted (tests/chroma/hadron/propagator/prec_wilson-twisted.{ini,out,metric}.xml):
    prop_corr to the metric's tolerance and 17 CG iterations per solve."""
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "prec_wilson_twisted.json")))
    args = ["--nrow"] + [str(v) for v in gold["nrow"]] + ["--cfg", "unit", "--fermact", "WILSON",
            "--kappa", repr(gold["kappa"]), "--boundary"] + [str(v) for v in gold["boundary"]] + \
           ["--twist"] + [repr(v) for v in gold["phases_by_pi"]] + \
           ["--solver", "CG", "--rsd", repr(gold["RsdCG"]), "--maxit", str(gold["MaxCG"]), "--sloppy", "DOUBLE"]
    assert gold["phases_dir"] == [0, 1, 2]
    out = run_driver(args)
    corr = np.array(out["prop_corr"])
    ref = np.array(gold["prop_corr"])
    assert np.max(np.abs(corr - ref) / ref) < gold["tolerance"]
    assert all(abs(n - gold["n_count"]) <= 1 for n in out["n_count"])
    assert abs(out["ncg_had"] - gold["ncg_had"]) <= 12


def test_cxx_host_matches_python_path():
    """EO clover, weak field, reliable CG with FP16 inner vectors: the C++ driver and the Python drivers
    over the same C-ABI give the same iteration counts and the same prop_corr."""
    L = (8, 8, 8, 16)
    kappa, csw = 0.12, 1.0
    out = run_driver(["--nrow"] + [str(v) for v in L] + ["--cfg", "weak", "--seed", "3", "--eps", "0.3",
                      "--fermact", "CLOVER", "--kappa", repr(kappa), "--csw", repr(csw), "--boundary", "1", "1", "1",
                      "-1", "--solver", "RELIABLE_CG", "--rsd", "1e-10", "--maxit", "5000", "--sloppy", "HALF",
                      "--delta", "0.1"])
    U = cmx.gauge_apply_bc(L, cmx.gauge_generate(L, kind=1, seed=3, eps=0.3), (1, 1, 1, -1))
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_CLOVER, cmx.kappa_to_mass(kappa), csw, sloppy_prec=cmx.HALF)
    qprop = qp.PrecFermActQprop(op, cmx.RELIABLE_CG, rsd=1e-10, maxit=5000, delta=0.1, sloppy_prec=cmx.HALF)
    prop, results = qp.quark_prop4(qprop, int(np.prod(L)), 0)
    corr = qp.prop_corr(L, prop)
    say("  cxx", out["n_count"], "python", [r["n_count"] for r in results])
    assert out["n_count"] == [r["n_count"] for r in results]
    assert np.max(np.abs(np.array(out["prop_corr"]) - corr) / corr) < 1e-12
    # the driver's mesons: the pion is prop_corr (source at t = 0), all 16 equal the Python binding's
    mes = np.array(out["mesons"])
    assert np.max(np.abs(mes[15] - corr) / corr) < 1e-12
    assert np.max(np.abs(mes - cmx.meson_2pt(L, prop)[:, 0].real)) < 1e-12 * np.max(np.abs(mes))
    # unpreconditioned residual of every solve, relative to the unit source
    assert max(out["resid"]) < 1e-8


def test_cxx_host_smeared_field_matches_python():
    """The driver's device stout smearing (--smear: 2 spatial steps) and the Python binding's give the
    same propagator: same iteration counts and prop_corr to the last bits."""
    L = (4, 4, 4, 8)
    out = run_driver(["--nrow"] + [str(v) for v in L] + ["--cfg", "haar", "--seed", "5", "--smear", "2", "0.1", "7",
                      "--fermact", "CLOVER", "--kappa", "0.11", "--csw", "1.0", "--boundary", "1", "1", "1", "-1",
                      "--solver", "CG", "--rsd", "1e-10", "--maxit", "5000", "--sloppy", "DOUBLE"])
    U = cmx.gauge_stout_smear(L, cmx.gauge_generate(L, kind=2, seed=5), 2, 0.1, 7)
    U = cmx.gauge_apply_bc(L, U, (1, 1, 1, -1))
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_CLOVER, cmx.kappa_to_mass(0.11), 1.0)
    prop, results = qp.quark_prop4(qp.PrecFermActQprop(op, cmx.CG, rsd=1e-10, maxit=5000, sloppy_prec=cmx.DOUBLE),
                                   int(np.prod(L)), 0)
    assert out["n_count"] == [r["n_count"] for r in results]
    corr = qp.prop_corr(L, prop)
    assert np.max(np.abs(np.array(out["prop_corr"]) - corr) / corr) < 1e-12


def test_cxx_host_batched_propagator():
    """quarkProp4 with the 12 odd solves batched (one launch per kernel for all 12) equals the 12 solves
    one after another: same iteration counts and the same prop_corr to the last bit."""
    args = ["--nrow", "8", "8", "8", "16", "--cfg", "weak", "--seed", "4", "--fermact", "CLOVER", "--kappa", "0.125",
            "--csw", "1.0", "--solver", "RELIABLE_CG", "--rsd", "1e-10", "--maxit", "5000", "--sloppy", "HALF"]
    seq = run_driver(args)
    bat = run_driver(args + ["--batch"])
    assert bat["n_count"] == seq["n_count"]
    assert bat["prop_corr"] == seq["prop_corr"]
    assert max(bat["resid"]) < 1e-8


def test_cxx_host_multishift_matches_python():
    """MdagMMultiSysSolverCG through the C++ host layer equals the Python binding of the same C-ABI."""
    L = (8, 8, 8, 16)
    shifts = [0.0, 0.01, 0.1, 1.0]
    out = run_driver(["--nrow"] + [str(v) for v in L] + ["--cfg", "weak", "--seed", "5", "--fermact", "CLOVER",
                      "--kappa", "0.12", "--csw", "1.0", "--rsd", "1e-10", "--maxit", "3000",
                      "--multishift", ",".join(repr(s) for s in shifts)])
    U = cmx.gauge_apply_bc(L, cmx.gauge_generate(L, kind=1, seed=5, eps=0.3), (1, 1, 1, -1))
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_CLOVER, cmx.kappa_to_mass(0.12), 1.0)
    Vh = int(np.prod(L)) // 2
    chi = np.zeros((Vh, 4, 3), complex)
    chi[0, 0, 0] = 1.0
    psi, res = op.solve_multishift(chi, shifts, 1e-10, maxit=3000)
    assert out["n_count"] == res[0]["n_count"]
    for s in range(len(shifts)):
        assert out["resid"][s] == res[s]["resid"]
        assert abs(out["psi_norm2"][s] - np.vdot(psi[s], psi[s]).real) <= 1e-12 * out["psi_norm2"][s]
        assert res[s]["rel_resid"] < 1e-9


def test_cxx_host_mre_chrono_predictor():
    """MinimalResidualExtrapolation4DChronoPredictor (mre_extrap_predictor.h:104-230) in the C++ host layer:
    on sources chi_k = e0 + 0.1 k e1 every chi_k (k >= 2) lies in the span of the earlier sources, so the
    predicted guess is already the solution (to the earlier solves' accuracy) and the M^dag M CG that
    starts from it needs almost no iterations, against the full count from zero."""
    out = run_driver(["--nrow", "8", "8", "8", "16", "--cfg", "weak", "--seed", "6", "--fermact", "CLOVER",
                      "--kappa", "0.12", "--csw", "1.0", "--solver", "CG", "--rsd", "1e-11", "--maxit", "5000",
                      "--sloppy", "DOUBLE", "--chrono", "4"])
    say("  chrono", out)
    assert out["guess_rel_resid"][0] == 1.0                 # empty buffer: zero guess
    assert out["n_count_chrono"][0] == out["n_count_zero"][0]
    for k in range(2, 5):
        assert out["guess_rel_resid"][k] < 1e-8
        assert out["n_count_chrono"][k] <= out["n_count_zero"][k] // 3
