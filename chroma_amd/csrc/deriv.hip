// Link derivative of the Wilson hopping term, WilsonDslashBase::deriv (lwldslash_base_w.h:132-227),
// in FP64: for every output site x on checkerboard cb and direction mu,
//   ds_mu(x)_{ab} = sum_spin [(1 - isign gamma_mu) psi](x+mu)_{spin,a} conj(chi(x)_{spin,b}),
// the traceSpin(outerProduct(shift(temp, FORWARD, mu), chi)) of the reference (isotropic weights).  One
// lane per site; the projection is the hopping term's half-spinor form, the lower spins rebuilt.
#include "blas.h"
#include "hop_kernel.h"

namespace cmx {

namespace {

template <int MU, int S>
__device__ __forceinline__ void deriv_dir(const Geom& g, const VRef& psi, const double (&c)[24], int i, const int (&x)[4],
                                          double2* ds) {
  int xf[4] = {x[0], x[1], x[2], x[3]};
  xf[MU] = (x[MU] + 1 == g.L[MU]) ? 0 : x[MU] + 1;
  const int jf = ((xf[3] * g.L[2] + xf[2]) * g.L[1] + xf[1]) * g.Lxh + (xf[0] >> 1);
  double p[24], h[12], t[24];
  SpinorIO<P64>::load(psi.v, psi.nrm, psi.stride, jf, p);
  project<MU, S>(p, h);
#pragma unroll
  for (int k = 0; k < 24; ++k) t[k] = 0;
  accumulate<MU, S>(t, h);
  double2* d = ds + ((size_t)MU * g.Vh + i) * 9;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      double re = 0, im = 0;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const double tr = t[2 * (3 * s + a)], ti = t[2 * (3 * s + a) + 1];
        const double cr = c[2 * (3 * s + b)], ci = c[2 * (3 * s + b) + 1];
        re += tr * cr + ti * ci;  // t conj(c)
        im += ti * cr - tr * ci;
      }
      d[3 * a + b] = make_double2(re, im);
    }
}

template <int ISIGN>
__global__ void __launch_bounds__(256) k_hop_deriv(Geom g, int cb, VRef chi, VRef psi, double2* ds) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.Vh) return;
  int x[4];
  {
    int r = i;
    const int xh = r % g.Lxh;
    r /= g.Lxh;
    x[1] = r % g.L[1];
    r /= g.L[1];
    x[2] = r % g.L[2];
    x[3] = r / g.L[2];
    x[0] = 2 * xh + ((cb + x[1] + x[2] + x[3]) & 1);
  }
  double c[24];
  SpinorIO<P64>::load(chi.v, chi.nrm, chi.stride, i, c);
  deriv_dir<0, -ISIGN>(g, psi, c, i, x, ds);
  deriv_dir<1, -ISIGN>(g, psi, c, i, x, ds);
  deriv_dir<2, -ISIGN>(g, psi, c, i, x, ds);
  deriv_dir<3, -ISIGN>(g, psi, c, i, x, ds);
}

}  // namespace

// chi: FP64 fermion on checkerboard cb, psi: FP64 fermion on 1 - cb, ds: 4 * Vh * 9 double2 (direction
// major, then site of cb, then the 3x3 row-major)
void hop_deriv(const Geom& g, int cb, int isign, VRef chi, VRef psi, void* ds, hipStream_t s) {
  CMX_REQUIRE(cb == 0 || cb == 1, CMX_ERR_ARG, "dslash_deriv: cb must be 0 or 1");
  const dim3 grid((g.Vh + 255) / 256);
  double2* d = static_cast<double2*>(ds);
  if (isign == 1)
    hipLaunchKernelGGL(k_hop_deriv<1>, grid, dim3(256), 0, s, g, cb, chi, psi, d);
  else
    hipLaunchKernelGGL(k_hop_deriv<-1>, grid, dim3(256), 0, s, g, cb, chi, psi, d);
}

}  // namespace cmx
