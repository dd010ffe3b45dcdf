// sort.hip -- ascending radix sort of device int64 keys (hipCUB / rocPRIM), used to put the
// compacted indices of kept molecules back into molecule order.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "ctraj_internal.hpp"

using namespace ctraj;

extern "C" int32_t ctraj_sort_i64(ctraj_ctx *c, int64_t *keys, int64_t n) {
  if (!c || (n > 0 && !keys)) return fail(c, CTRAJ_EINVAL, "null argument");
  if (n < 0 || n > INT32_MAX) return fail(c, CTRAJ_EINVAL, "n must be 0..2^31-1");
  if (n < 2) return CTRAJ_OK;
  HIPCHK(c, hipSetDevice(c->device));
  int rc = ensure_buf(c, c->sort_alt, sizeof(int64_t) * (size_t)n);
  if (rc != CTRAJ_OK) return rc;
  hipcub::DoubleBuffer<int64_t> db(keys, (int64_t *)c->sort_alt.p);
  size_t tmp = 0;
  HIPCHK(c, hipcub::DeviceRadixSort::SortKeys(nullptr, tmp, db, (int)n, 0, 64, c->stream));
  if ((rc = ensure_buf(c, c->sort_tmp, tmp > 0 ? tmp : 1)) != CTRAJ_OK) return rc;
  HIPCHK(c, hipcub::DeviceRadixSort::SortKeys(c->sort_tmp.p, tmp, db, (int)n, 0, 64, c->stream));
  if (db.Current() != keys)
    HIPCHK(c, hipMemcpyAsync(keys, db.Current(), sizeof(int64_t) * (size_t)n, hipMemcpyDeviceToDevice, c->stream));
  return CTRAJ_OK;
}
