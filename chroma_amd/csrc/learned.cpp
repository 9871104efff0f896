// Learned preconditioner entry used by the flexible solver (placeholder until the network kernels
// are in place).
#include "op.h"

namespace cmx {

void learned_precond_apply(Op& op, int prec, const Vec& in, Vec& out) {
  (void)op, (void)prec, (void)in, (void)out;
  throw Error(CMX_ERR_UNSUPPORTED, "learned preconditioner: no network loaded for this operator");
}

}  // namespace cmx
