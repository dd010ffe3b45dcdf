// resample.hip -- stored trajectories evaluated at one z on the device:
// find_radial_pos_dist / find_vel_dist (post_processing.py:20-140).
//
// Each molecule's rows are scanned by one thread in the order they were appended.  The rule is the
// reference's, comparison for comparison, so NaN and repeated z behave as there:
//   * a molecule whose last row lies before z is skipped (post_processing.py:42-43);
//   * otherwise the first row whose z equals z is returned as stored (:52-56);
//   * otherwise the row before the first row with !(z_row < z) is advanced by take_timestep with its
//     stored acceleration (:61-72, :9-17); that row is the last one when the first row already has
//     !(z_row < z) (Python's index -1).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "ctraj_internal.hpp"

namespace {

__global__ void __launch_bounds__(256) k_sample_rows(ctraj_sample S) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < S.n; j += stride) {
    const int64_t r0 = S.row_off[j], r1 = S.row_off[j + 1];
    double o[6];
    bool valid = false;
    if (r1 > r0) {
      const double zq = S.z;
      const double *last = S.rows + (r1 - 1) * 10;
      if (!(last[2] < zq)) {
        valid = true;
        int64_t hit = -1, first_ge = -1;
        for (int64_t r = r0; r < r1; ++r) {
          const double zr = S.rows[r * 10 + 2];
          if (zr == zq) {
            hit = r;
            break;
          }
          if (first_ge < 0 && !(zr < zq)) first_ge = r;
        }
        if (hit >= 0) {
          const double *p = S.rows + hit * 10;
          o[0] = p[0]; o[1] = p[1]; o[2] = p[3]; o[3] = p[4]; o[4] = p[5]; o[5] = p[9];
        } else {
          // the last row satisfies !(z_row < z) here, so first_ge was found
          const int64_t nidx = first_ge > r0 ? first_ge - 1 : r1 - 1;
          const double *p = S.rows + nidx * 10;
          const double dt = (zq - p[2]) / p[5];
          const double dt2 = dt * dt;
          o[0] = (p[0] + p[3] * dt) + (p[6] * dt2) / 2.0;
          o[1] = (p[1] + p[4] * dt) + (p[7] * dt2) / 2.0;
          o[2] = p[3] + p[6] * dt;
          o[3] = p[4] + p[7] * dt;
          o[4] = p[5] + p[8] * dt;
          o[5] = p[9] + dt;
        }
      }
    }
    if (!valid)
      for (int f = 0; f < 6; ++f) o[f] = __builtin_nan("");
    for (int f = 0; f < 6; ++f)
      if (S.out[f]) S.out[f][j] = o[f];
    if (S.valid) S.valid[j] = valid ? 1 : 0;
  }
}

}  // namespace

extern "C" int32_t ctraj_sample_rows(ctraj_ctx *c, const ctraj_sample *S) {
  using namespace ctraj;
  if (!c || !S) return fail(c, CTRAJ_EINVAL, "null argument");
  if (S->n < 0) return fail(c, CTRAJ_EINVAL, "n < 0");
  if (S->n == 0) return CTRAJ_OK;
  if (!S->row_off || !S->rows) return fail(c, CTRAJ_EINVAL, "missing row offsets or rows");
  HIPCHK(c, hipSetDevice(c->device));
  {  // every row a kernel thread reads must lie in `rows`
    std::vector<int64_t> off((size_t)S->n + 1);
    HIPCHK(c, hipMemcpyAsync(off.data(), S->row_off, sizeof(int64_t) * off.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (off[0] < 0 || off[S->n] > S->n_rows) return fail(c, CTRAJ_EINVAL, "row offsets outside [0, n_rows]");
    for (int64_t j = 0; j < S->n; ++j)
      if (off[j + 1] < off[j]) return fail(c, CTRAJ_EINVAL, "row offsets must be non-decreasing");
  }
  const int grid = (int)std::min<int64_t>((S->n + 255) / 256, (int64_t)c->n_cu * 8);
  hipLaunchKernelGGL(k_sample_rows, dim3(grid), dim3(256), 0, c->stream, *S);
  HIPCHK(c, hipGetLastError());
  return CTRAJ_OK;
}
