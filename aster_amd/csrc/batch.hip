// Batched ensembles (placeholder until the persistent kernel lands).
#include "common.hpp"

namespace aster {
int batch_step(aster_ctx* c, int, double, int64_t, int, int) { return fail(c, ASTER_ERR_INVALID, "batch not built"); }
int batch_energy(aster_ctx* c, double*, double*) { return fail(c, ASTER_ERR_INVALID, "batch not built"); }
int batch_autotune(aster_ctx* c, int, double, int32_t*, int32_t*) { return fail(c, ASTER_ERR_INVALID, "batch not built"); }
}  // namespace aster
extern "C" {
int aster_batch_set(aster_ctx* c, int64_t, const int64_t*, const double*, const double*, const double*) {
  return aster::fail(c, ASTER_ERR_INVALID, "batch not built");
}
int aster_batch_get(aster_ctx* c, double*, double*) { return aster::fail(c, ASTER_ERR_INVALID, "batch not built"); }
}
