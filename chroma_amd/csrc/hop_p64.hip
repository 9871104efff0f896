// Hopping-term kernels instantiated for FP64 storage: the site-linear kernel (hop_kernel.h) with 18- or
// 12-real links, and the plane-grid kernel (hop_fast.h) with 12-real links.
#include "hop_fast.h"

namespace cmx {
void launch_hop_p64(const HopLaunch& h, hipStream_t s, const BatchGrid* bg) {
  if (h.fast & 1) {
    launch_hop_fast_t<P64>(h, s, bg);
    return;
  }
  if (h.recon == 12)
    launch_hop_t<P64, 12, false>(h, s, bg);
  else
    launch_hop_t<P64, 18, false>(h, s, bg);
}
}  // namespace cmx
