// Hopping-term kernels instantiated for FP16 storage (kernel template in hop_kernel.h).
#include "hop_fast.h"

namespace cmx {
void launch_hop_p16(const HopLaunch& h, hipStream_t s) {
  if (h.fast) {
    launch_hop_fast_t<P16>(h, s);
    return;
  }
  if (h.prefetch) launch_hop_t<P16, 12, true>(h, s); else launch_hop_t<P16, 12, false>(h, s);
}
}  // namespace cmx
