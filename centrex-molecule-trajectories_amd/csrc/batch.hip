// batch.hip -- persistent batched scan over many beamline configurations (being written).
#include "ctraj_internal.hpp"

extern "C" {
int32_t ctraj_batch_create(ctraj_ctx *c, const ctraj_elem *, int32_t, const ctraj_batch_cfg *, int32_t,
                           const ctraj_phys *, ctraj_batch **out) {
  if (out) *out = nullptr;
  return ctraj::fail(c, CTRAJ_EINVAL, "batched scan not available in this build");
}
int32_t ctraj_batch_run(ctraj_batch *, const ctraj_batch_tune *, int64_t *, uint64_t *, float *) {
  return ctraj::fail(nullptr, CTRAJ_EINVAL, "batched scan not available in this build");
}
int32_t ctraj_batch_destroy(ctraj_batch *) { return CTRAJ_OK; }
}
