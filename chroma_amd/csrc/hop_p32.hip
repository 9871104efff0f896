// Hopping-term kernels instantiated for FP32 storage (kernel template in hop_kernel.h).
#include "hop_kernel.h"

namespace cmx {
void launch_hop_p32(const HopLaunch& h, hipStream_t s) {
  if (h.recon == 18) {
    if (h.prefetch) launch_hop_t<P32, 18, true>(h, s); else launch_hop_t<P32, 18, false>(h, s);
  } else {
    if (h.prefetch) launch_hop_t<P32, 12, true>(h, s); else launch_hop_t<P32, 12, false>(h, s);
  }
}
}  // namespace cmx
