// Hopping-term kernels instantiated for FP16 storage (kernel template in hop_kernel.h).
#include "hop_kernel.h"

namespace cmx {
void launch_hop_p16(const HopLaunch& h, hipStream_t s) {
  if (h.prefetch) launch_hop_t<P16, 12, true>(h, s); else launch_hop_t<P16, 12, false>(h, s);
}
}  // namespace cmx
