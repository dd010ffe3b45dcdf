// adjoint.hip -- differentiable detected flux and its adjoint (being written).
#include "ctraj_internal.hpp"

extern "C" {
int32_t ctraj_smooth_objective(ctraj_ctx *c, const ctraj_smooth *) {
  return ctraj::fail(c, CTRAJ_EINVAL, "smooth objective not available in this build");
}
}
