// Point-source propagator on the MI355X library through the C++ host layer: the path a Chroma
// propagator task takes for an even-odd Wilson or clover action (quarkProp4_a with the qprop solver
// object, quarkprop4_w.cc:31-117, eoprec_fermact_qprop.cc:41-80), with the propagator correlator of
// inline_propagator_w.cc:383-395.  The gauge field is one of the library's synthetic fields.
// Prints one JSON line: per-solve iteration counts and residuals, their sum (ncg_had), prop_corr and the
// zero-momentum point-sink mesons of the 16 gamma insertions (mesprop of mesons_w.cc:160-167).
//
//   mi355x_prop --nrow 4 4 4 8 --cfg unit --fermact WILSON --kappa 0.1 --boundary 1 1 1 1
//               --twist 0.2 0.4 0.0 --solver CG --rsd 1e-8 --maxit 1000 [--batch]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mi355x_solver.h"

using namespace cmx_chroma;

namespace {
void usage() {
  std::printf(
      "usage: mi355x_prop [--nrow Lx Ly Lz Lt] [--cfg unit|weak|haar] [--seed N] [--eps E] [--nstout N --rho R]\n"
      "                   [--smear N R DIRS]   (then N stout steps on the device, rho R, direction bit mask DIRS)\n"
      "                   [--boundary bx by bz bt] [--twist px py pz] [--fermact WILSON|CLOVER] [--kappa K]\n"
      "                   [--csw C] [--solver CG|BICGSTAB|RELIABLE_CG|RELIABLE_BICGSTAB] [--rsd R] [--maxit N]\n"
      "                   [--sloppy DOUBLE|SINGLE|HALF] [--delta D] [--true-residual] [--src-site S] [--device D]\n"
      "                   [--batch]   (the 12 solves advanced together: one launch per kernel for all)\n"
      "                   [--multishift s0,s1,...]   (instead: (M^dag M + s) psi = chi_o for a point source)\n"
      "                   [--chrono N]   (instead: 5 M^dag M solves on chi_k = e0 + 0.1 k e1, MRE predictor of N)\n");
}
}  // namespace

int main(int argc, char** argv) {
  int nrow[4] = {4, 4, 4, 8}, boundary[4] = {1, 1, 1, -1}, nstout = 0, device = 0;
  std::string cfg = "weak";
  unsigned long long seed = 1;
  double eps = 0.3, rho = 0.1, twist[3] = {0, 0, 0};
  bool twisted = false;
  FermActParams fp;
  double kappa = 0.12, csw = 1.0;
  SysSolverMI355XParams sp;
  long src_site = 0;
  bool batched = false;
  std::vector<double> shifts;  // --multishift: the RHMC solves on the odd part of the first point source
  int chrono = 0;              // --chrono N: a sequence of M^dag M solves with the MRE predictor
  int smear_n = 0, smear_dirs = 15;  // --smear N R DIRS: StoutLinkSmearing on the device
  double smear_rho = 0.0;
  for (int a = 1; a < argc; ++a) {
    const std::string k = argv[a];
    auto need = [&](int n) {
      if (a + n >= argc) {
        usage();
        std::exit(2);
      }
    };
    if (k == "--help" || k == "-h") {
      usage();
      return 0;
    } else if (k == "--nrow") {
      need(4);
      for (int m = 0; m < 4; ++m) nrow[m] = std::atoi(argv[++a]);
    } else if (k == "--cfg") {
      need(1);
      cfg = argv[++a];
    } else if (k == "--seed") {
      need(1);
      seed = std::strtoull(argv[++a], nullptr, 10);
    } else if (k == "--eps") {
      need(1);
      eps = std::atof(argv[++a]);
    } else if (k == "--nstout") {
      need(1);
      nstout = std::atoi(argv[++a]);
    } else if (k == "--rho") {
      need(1);
      rho = std::atof(argv[++a]);
    } else if (k == "--boundary") {
      need(4);
      for (int m = 0; m < 4; ++m) boundary[m] = std::atoi(argv[++a]);
    } else if (k == "--twist") {
      need(3);
      for (int m = 0; m < 3; ++m) twist[m] = std::atof(argv[++a]);
      twisted = true;
    } else if (k == "--fermact") {
      need(1);
      fp.FermAct = argv[++a];
    } else if (k == "--kappa") {
      need(1);
      kappa = std::atof(argv[++a]);
    } else if (k == "--csw") {
      need(1);
      csw = std::atof(argv[++a]);
    } else if (k == "--solver") {
      need(1);
      sp.SolverType = argv[++a];
    } else if (k == "--rsd") {
      need(1);
      sp.RsdTarget = std::atof(argv[++a]);
    } else if (k == "--maxit") {
      need(1);
      sp.MaxIter = std::atoi(argv[++a]);
    } else if (k == "--sloppy") {
      need(1);
      sp.SloppyPrecision = argv[++a];
    } else if (k == "--delta") {
      need(1);
      sp.Delta = std::atof(argv[++a]);
    } else if (k == "--smear") {
      need(3);
      smear_n = std::atoi(argv[++a]);
      smear_rho = std::atof(argv[++a]);
      smear_dirs = std::atoi(argv[++a]);
    } else if (k == "--batch") {
      batched = true;
    } else if (k == "--chrono") {
      need(1);
      chrono = std::atoi(argv[++a]);
    } else if (k == "--multishift") {
      need(1);
      for (char* t = std::strtok(argv[++a], ","); t; t = std::strtok(nullptr, ",")) shifts.push_back(std::atof(t));
    } else if (k == "--true-residual") {
      sp.TrueResidual = true;
    } else if (k == "--src-site") {
      need(1);
      src_site = std::atol(argv[++a]);
    } else if (k == "--device") {
      need(1);
      device = std::atoi(argv[++a]);
    } else {
      std::fprintf(stderr, "unknown option %s\n", k.c_str());
      usage();
      return 2;
    }
  }
  const int kind = cfg == "unit" ? 0 : (cfg == "weak" ? 1 : (cfg == "haar" ? 2 : -1));
  if (kind < 0) {
    std::fprintf(stderr, "unknown --cfg %s\n", cfg.c_str());
    return 2;
  }
  const long V = (long)nrow[0] * nrow[1] * nrow[2] * nrow[3];
  try {
    check(cmx_init(device), "cmx_init");
    std::vector<double> links((size_t)4 * V * 18);
    check(cmx_gauge_generate(nrow, kind, seed, eps, nstout, rho, links.data()), "cmx_gauge_generate");
    if (smear_n > 0)
      check(cmx_gauge_stout_smear_dirs(nrow, links.data(), smear_n, smear_rho, smear_dirs), "cmx_gauge_stout_smear_dirs");
    const int tdir[3] = {0, 1, 2};
    check(cmx_gauge_apply_bc(nrow, links.data(), boundary, twisted ? twist : nullptr, twisted ? tdir : nullptr),
          "cmx_gauge_apply_bc");
    GaugeField u(nrow, links);
    fp.Mass = kappaToMass(kappa);
    if (fp.FermAct == "CLOVER") fp.clovCoeffR = fp.clovCoeffT = csw;
    if (sp.SloppyPrecision != "DOUBLE") fp.SloppyPrecision = sp.SloppyPrecision;
    EvenOddPrecLinOp A(u, fp);
    if (chrono > 0) {  // HMC-like sequence of M^dag M solves with and without the chronological predictor
      MdagMSysSolverMI355X solve(A, sp);
      MREChronoPredictor pred(chrono);
      std::vector<int> nc_chrono, nc_zero;
      std::vector<double> guess_rel;
      for (int k = 0; k < 5; ++k) {
        LatticeFermion chi(24 * V, 0.0), psi(24 * V, 0.0), guess, mg(24 * V, 0.0), mmg(24 * V, 0.0);
        chi[(size_t)(V / 2) * 24] = 1.0;
        chi[(size_t)(V / 2 + 1) * 24] = 0.1 * k;
        pred(guess, A, chi);
        A(mg, guess, CMX_PLUS);
        A(mmg, mg, CMX_MINUS);
        double rr = 0, cc = 0;
        for (long q = 24 * (V / 2); q < 24 * V; ++q) rr += (chi[q] - mmg[q]) * (chi[q] - mmg[q]), cc += chi[q] * chi[q];
        guess_rel.push_back(std::sqrt(rr / cc));
        nc_chrono.push_back(solve(psi, chi, pred).n_count);
        LatticeFermion psi0(24 * V, 0.0);
        nc_zero.push_back(solve(psi0, chi).n_count);
      }
      std::printf("{\"n_count_chrono\": [");
      for (size_t q = 0; q < nc_chrono.size(); ++q) std::printf("%s%d", q ? ", " : "", nc_chrono[q]);
      std::printf("], \"n_count_zero\": [");
      for (size_t q = 0; q < nc_zero.size(); ++q) std::printf("%s%d", q ? ", " : "", nc_zero[q]);
      std::printf("], \"guess_rel_resid\": [");
      for (size_t q = 0; q < guess_rel.size(); ++q) std::printf("%s%.17g", q ? ", " : "", guess_rel[q]);
      std::printf("]}\n");
      cmx_finalize();
      return 0;
    }
    if (!shifts.empty()) {  // MdagMMultiSysSolverCG on the odd checkerboard
      LatticeFermion chi(24 * V, 0.0);
      chi[(size_t)(V / 2 + src_site) * 24] = 1.0;
      MdagMMultiSysSolverCGMI355X msolve(A, {sp.RsdTarget}, sp.MaxIter);
      std::vector<LatticeFermion> psi;
      const SystemSolverResults_t r = msolve(psi, shifts, chi);
      std::printf("{\"n_count\": %d, \"resid\": [", r.n_count);
      for (size_t s = 0; s < shifts.size(); ++s) std::printf("%s%.17g", s ? ", " : "", msolve.shiftResiduals()[s]);
      std::printf("], \"psi_norm2\": [");
      for (size_t s = 0; s < shifts.size(); ++s) {
        double n2 = 0;
        for (double v : psi[s]) n2 += v * v;
        std::printf("%s%.17g", s ? ", " : "", n2);
      }
      std::printf("]}\n");
      cmx_finalize();
      return 0;
    }
    LinOpSysSolverMI355X invA(A, sp);
    PrecFermActQprop qprop(A, invA);
    std::vector<Complex> prop;
    std::vector<SystemSolverResults_t> res;
    const int ncg_had = quarkProp4(prop, qprop, V, src_site, &res, batched);
    const std::vector<double> corr = propCorr(nrow, prop);
    // point-sink mesons at zero momentum, time from the source (mesons_w.cc:97-180)
    int origin[4];
    {
      const long Vh = V / 2, cb = src_site >= Vh ? 1 : 0;
      long r = src_site - cb * Vh;
      const int xh = (int)(r % (nrow[0] / 2));
      r /= nrow[0] / 2;
      origin[1] = (int)(r % nrow[1]);
      r /= nrow[1];
      origin[2] = (int)(r % nrow[2]);
      origin[3] = (int)(r / nrow[2]);
      origin[0] = 2 * xh + (int)((cb + origin[1] + origin[2] + origin[3]) & 1);
    }
    const int p0[3] = {0, 0, 0};
    std::vector<double> mes((size_t)16 * nrow[3] * 2);
    const double* pp = reinterpret_cast<const double*>(prop.data());
    check(cmx_meson_2pt(nrow, pp, pp, origin, 1, p0, mes.data()), "cmx_meson_2pt");
    std::printf("{\"nrow\": [%d, %d, %d, %d], \"n_count\": [", nrow[0], nrow[1], nrow[2], nrow[3]);
    for (size_t k = 0; k < res.size(); ++k) std::printf("%s%d", k ? ", " : "", res[k].n_count);
    std::printf("], \"resid\": [");
    for (size_t k = 0; k < res.size(); ++k) std::printf("%s%.17g", k ? ", " : "", res[k].resid);
    std::printf("], \"ncg_had\": %d, \"prop_corr\": [", ncg_had);
    for (size_t t = 0; t < corr.size(); ++t) std::printf("%s%.17g", t ? ", " : "", corr[t]);
    std::printf("], \"mesons\": [");
    for (int n = 0; n < 16; ++n) {
      std::printf("%s[", n ? ", " : "");
      for (int t = 0; t < nrow[3]; ++t) std::printf("%s%.17g", t ? ", " : "", mes[((size_t)n * nrow[3] + t) * 2]);
      std::printf("]");
    }
    std::printf("]}\n");
    cmx_finalize();
  } catch (const SolverAbort& e) {
    std::fprintf(stderr, "mi355x_prop: %s\n", e.what());
    return 1;
  }
  return 0;
}
