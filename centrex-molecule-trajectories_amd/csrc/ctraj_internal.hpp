// ctraj_internal.hpp -- context object, kernel argument blocks, Philox source (shared by the TUs).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/ctraj.h"

// Grow-only device buffer owned by a context.
struct ctraj_devbuf {
  void *p = nullptr;
  size_t bytes = 0;
};

struct ctraj_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int n_cu = 0, lds_max = 0, arch = 0, wave = 64;
  // beamline of ctraj_propagate / ctraj_record_trajectories / ctraj_smooth_objective
  ctraj_elem *d_elems = nullptr;
  int n_elems = 0;
  ctraj_phys phys{};
  std::vector<ctraj_elem> h_elems;
  int force = CTRAJ_FORCE_WSPLINE_HEX, integ = CTRAJ_INTEG_RK4;
  bool have_beamline = false;
  // W' spline tables [state][seg][4]
  double *d_ws = nullptr;
  int ws_nseg = 0, ws_states = 0;
  double ws_end[CTRAJ_MAX_STATES] = {};  // W' at each state table's end (A + B + C of its last segment)
  // r tables, concatenated [knot][4]
  double *d_rt = nullptr;
  int rt_total = 0;
  int rt_off[8] = {}, rt_n[8] = {};
  double rt_invh[8] = {};
  std::vector<double> rt_host[8];
  void *d_counter = nullptr;  // work counter of the lane-refill kernel
  void *d_run = nullptr;      // device copy of the last ctraj_run given to the lane-refill kernel
  ctraj_run last_run{};
  bool have_last_run = false;
  hipEvent_t ev[64];
  // CTRAJ_RUN_PHASE_TIMES: per launch, events before the first kernel, between the kernels and after the
  // last one, in a ring of 64 launches; launches [pt_read, pt_next) are not yet read
  hipEvent_t pt_ev[64][4] = {};
  bool pt_split[64] = {};  // the launch ran as separate kernels of one slice (events 0-3), else events 0-1 bracket it
  int64_t pt_next = 0, pt_read = 0;
  // smooth objective: per-wave checkpoint/recompute store and per-item partial sums
  ctraj_devbuf sm_ws, sm_part;
  // radix sort scratch (ctraj_sort_i64)
  ctraj_devbuf sort_tmp, sort_alt;
  // propagation in separate kernels: the lens-entry queue, its claim and count words
  ctraj_devbuf pipe, pipe_cnt;
  // lens kernel claim order: per-entry lifetime buckets, per-workgroup bucket counts, slot permutation
  ctraj_devbuf ord;
  // this device's scratch array of ctraj_propagate_noalloc (device address), or nullptr
  char *scratch = nullptr;
};

namespace ctraj {

struct KBeam {
  const ctraj_elem *elems;
  int n_elems;
  double g;
  int n_labels, label_detected, n_states, cap;
};

struct KTab {
  const double *ws;
  int ws_nseg, ws_states;
  double umax;  // largest double below ws_nseg
  const double *rt;
  int rt_total;
};

// ---- host helpers (ctraj.hip) ----
int fail(ctraj_ctx *ctx, int code, const std::string &msg);
int hipfail(ctraj_ctx *ctx, hipError_t e, const char *what);
KTab ctx_ktab(const ctraj_ctx *ctx);
size_t ktab_lds_bytes(const KTab &k);
int check_elems(ctraj_ctx *ctx, const ctraj_elem *e, int n, const ctraj_phys *p, int *force, int *integ);
int check_tables(ctraj_ctx *ctx, const ctraj_elem *e, int n, int n_states);
void resolve_elems(const ctraj_ctx *ctx, const ctraj_elem *e, int n, ctraj_elem *out);
// The largest double whose correctly rounded square root is <= d: for r2 >= 0, sqrt(r2) > d exactly
// when r2 > circ_r2_threshold(d).  Device element tables keep it in a circular aperture's p[1].
double circ_r2_threshold(double d);
int check_phys(ctraj_ctx *ctx, const ctraj_phys *p);
KBeam make_kbeam(const ctraj_elem *d_elems, int n, const ctraj_phys &p);
int lds_limit(ctraj_ctx *ctx, const void *kern, size_t bytes);
int grid_for(ctraj_ctx *ctx, const void *kern, int block, size_t lds, int64_t n, int req);
// EINVAL unless `block` threads fit the kernel's launch bounds
int block_ok(ctraj_ctx *ctx, const void *kern, int block);
// make b hold at least `bytes` (grow-only; the old buffer is released first)
int ensure_buf(ctraj_ctx *ctx, ctraj_devbuf &b, size_t bytes);

#define HIPCHK(ctx, call)                                         \
  do {                                                            \
    hipError_t _e = (call);                                       \
    if (_e != hipSuccess) return ::ctraj::hipfail(ctx, _e, #call); \
  } while (0)

// ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) ----
__host__ __device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    const uint32_t n3 = (uint32_t)p0;
    c[0] = n0;
    c[1] = n1;
    c[2] = n2;
    c[3] = n3;
  }
}

// One Box-Muller pair for molecule i, dimension pair p: counter (i_lo, i_hi, p, 'CTRA'),
// key (seed_lo, seed_hi).  Uniforms take 53 bits each: ((hi << 21) | (lo >> 11) + 0.5) 2^-53.
__device__ inline void gaussian_pair(uint64_t seed, uint64_t i, uint32_t p, double &z0, double &z1) {
  uint32_t c[4] = {(uint32_t)i, (uint32_t)(i >> 32), p, 0x43545241u};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const uint64_t m1 = ((uint64_t)c[0] << 21) | (c[1] >> 11);
  const uint64_t m2 = ((uint64_t)c[2] << 21) | (c[3] >> 11);
  const double u1 = ((double)m1 + 0.5) * 0x1p-53;
  const double u2 = ((double)m2 + 0.5) * 0x1p-53;
  const double r = sqrt(-2.0 * log(u1));
  const double th = 6.283185307179586 * u2;
  z0 = r * cos(th);
  z1 = r * sin(th);
}

// One point of the uniform disc source for molecule i: counter (i_lo, i_hi, 'DISC', 'CTRA'), key
// (seed_lo, seed_hi), uniforms as in gaussian_pair; theta = 2 pi u1, then r = radius sqrt(u2)
// (CeNTREXPositionDistribution.draw's order, distributions.py:112-119).
__device__ inline void disc_point(uint64_t seed, uint64_t i, double radius, double &x, double &y) {
  uint32_t c[4] = {(uint32_t)i, (uint32_t)(i >> 32), 0x43534944u, 0x43545241u};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const uint64_t m1 = ((uint64_t)c[0] << 21) | (c[1] >> 11);
  const uint64_t m2 = ((uint64_t)c[2] << 21) | (c[3] >> 11);
  const double u1 = ((double)m1 + 0.5) * 0x1p-53;
  const double u2 = ((double)m2 + 0.5) * 0x1p-53;
  const double th = 6.283185307179586 * u1;
  const double r = sqrt(u2) * radius;
  x = r * cos(th);
  y = r * sin(th);
}

}  // namespace ctraj
