// ORACLE -- test infrastructure only.
//
// CPU restatement of Chroma's even-odd Wilson-clover fermion-matrix path and the Krylov solvers
// that drive it.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// it, and only as the checker or as the timed CPU baseline -- never as part of the product path.
//
// Host layout (shared with the C-ABI of the product library):
//   lattice dims L[4] = {Lx, Ly, Lz, Lt}, all even; V = Lx*Ly*Lz*Lt, Vh = V/2
//   checkerboard cb(x) = (x+y+z+t) & 1
//   full site index = cb*Vh + (((t*Lz + z)*Ly + y)*Lx + x)/2     (x-fastest within a checkerboard)
//   fermion   : V x [4 spin][3 colour][re,im]  (spin-major, as QDP++ Fermion elem(spin).elem(colour))
//   gauge     : [4 mu][V][3][3][re,im]        (row-major colour matrix elem(i,j))
//   clover    : V x {diag[2][6], offd[2][15][re,im]}  (PrimitiveClovTriang, clover_term_qdp_w.h:19-24)
#pragma once
#include <complex>
#include <cstdint>
#include <vector>

namespace orc {

template <typename R> using cplx = std::complex<R>;

struct Geom {
  int L[4];
  long V, Vh;
  std::vector<long> nb;  // nb[site*8 + 2*mu + 0] = x+mu, [.. + 1] = x-mu
  explicit Geom(const int dims[4]);
  void coords(long idx, int x[4]) const;
  long index(const int x[4]) const;
  int parity(long idx) const { return idx >= Vh ? 1 : 0; }
  long fwd(long s, int mu) const { return nb[s * 8 + 2 * mu]; }
  long bwd(long s, int mu) const { return nb[s * 8 + 2 * mu + 1]; }
};

// DeGrand-Rossi gamma matrices gamma_mu, mu = 0..3  (docs/notes/gamma_conventions.tex:74-79)
extern const std::complex<double> kGamma[4][4][4];

template <typename R> struct Clov {  // PrimitiveClovTriang<R>
  R diag[2][6];
  cplx<R> offd[2][15];
};

}  // namespace orc
