// Trajectory adjoint (placeholder).
#include "common.hpp"

namespace aster {
int traj_forward(aster_ctx* c, double, int64_t, int64_t, int32_t, const int64_t*, double*) {
  return fail(c, ASTER_ERR_INVALID, "adjoint not built");
}
int traj_backward(aster_ctx* c, const double*, double*, double*, double*) {
  return fail(c, ASTER_ERR_INVALID, "adjoint not built");
}
}  // namespace aster
