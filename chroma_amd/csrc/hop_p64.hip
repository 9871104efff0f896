// Hopping-term kernels instantiated for FP64 storage (kernel template in hop_kernel.h).
#include "hop_kernel.h"

namespace cmx {
void launch_hop_p64(const HopLaunch& h, hipStream_t s, const BatchGrid* bg) { launch_hop_t<P64, 18, false>(h, s, bg); }
}  // namespace cmx
