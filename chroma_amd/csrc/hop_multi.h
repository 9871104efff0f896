// Multi-source hopping term (hop_multi.hip).
#pragma once
#include "blas.h"
#include "cmx_internal.h"

namespace cmx {

constexpr int MULTI_MAXR = 12;  // the 12 spin-colour sources of a propagator

struct MultiHopLaunch {
  HopLaunch h;              // geometry, links, output parity, isign, scale b, launch shape
  int nrhs;
  int group;                // sources per kernel (links loaded once per group): 1..4
  VRef in[MULTI_MAXR];      // inputs (other checkerboard), FP16 with norms
  VRef out[MULTI_MAXR];     // outputs (checkerboard h.cb)
};

// out[k] = b D(isign) in[k], k < nrhs; FP16 storage, 12-real links, real boundary phases
void launch_hop_multi(const MultiHopLaunch& M, hipStream_t s);
// the same with the link products on the matrix cores (hop_mfma.hip); M.group is not used
constexpr int MULTI_MFMA = 5;
void launch_hop_mfma(const MultiHopLaunch& M, hipStream_t s);

}  // namespace cmx
