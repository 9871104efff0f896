// Hopping-term kernels instantiated for FP32 storage (kernel template in hop_kernel.h).
#include "hop_fast.h"

namespace cmx {
void launch_hop_p32(const HopLaunch& h, hipStream_t s, const BatchGrid* bg) {
  if (h.fast & 1) {
    launch_hop_fast_t<P32>(h, s, bg);
    return;
  }
  if (h.recon == 18) {
    if (h.prefetch) launch_hop_t<P32, 18, true>(h, s, bg); else launch_hop_t<P32, 18, false>(h, s, bg);
  } else {
    if (h.prefetch) launch_hop_t<P32, 12, true>(h, s, bg); else launch_hop_t<P32, 12, false>(h, s, bg);
  }
}
}  // namespace cmx
