// Hopping-term kernels instantiated for FP16 storage (kernel template in hop_kernel.h).
#include "hop_fast.h"

namespace cmx {
void launch_hop_p16(const HopLaunch& h, hipStream_t s, const BatchGrid* bg) {
  if (h.fast & 1) {
    launch_hop_fast_t<P16>(h, s, bg);
    return;
  }
  if (h.prefetch) launch_hop_t<P16, 12, true>(h, s, bg); else launch_hop_t<P16, 12, false>(h, s, bg);
}
}  // namespace cmx
