// ctraj_wave.hpp -- LDS table staging and wavefront helpers shared by the kernel files.
#pragma once

#include "ctraj_device.hpp"
#include "ctraj_internal.hpp"

// ---------------------------------------------------------------------------------------------
// LDS staging of the lens tables
// ---------------------------------------------------------------------------------------------
namespace ctraj {

// The W' spline is kept in LDS as two arrays over all segments of all states: (A, B) pairs
// [state][seg] (16-byte stride, at the LDS base) and C [state][seg] after them; the global table is
// [seg][4] = {A, B, C, 0}.  r tables keep [knot][4].  The LDS base must be 16-byte aligned: every
// kernel declares its dynamic LDS with that alignment.
__device__ inline Tables stage_tables(double *lds, const KTab &kt) {
  Tables T;
  const int nseg_all = kt.ws_states * kt.ws_nseg;
  const int nws = nseg_all * 3;
  const int nrt = kt.rt_total * 4;
  double2 *lds_ab = reinterpret_cast<double2 *>(lds);
  double *lds_c = lds + 2 * nseg_all;
  for (int i = threadIdx.x; i < nseg_all; i += blockDim.x) {
    lds_ab[i] = reinterpret_cast<const double2 *>(kt.ws)[2 * i];
    lds_c[i] = kt.ws[4 * i + 2];
  }
  const double2 *src_rt = reinterpret_cast<const double2 *>(kt.rt);
  double2 *dst_rt = reinterpret_cast<double2 *>(lds + nws + (nws & 1));
  for (int i = threadIdx.x; i < nrt / 2; i += blockDim.x) dst_rt[i] = src_rt[i];
  const lds_double *base = (const lds_double *)lds;
  T.ws_ab = (const lds_dbl2 *)lds_ab;
  T.ws_c = base + 2 * nseg_all;
  T.nseg = kt.ws_nseg;
  T.umax = kt.umax;
  T.rt = base + nws + (nws & 1);
  return T;
}

__device__ __forceinline__ int lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

__device__ __forceinline__ int lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// A lens-entry queue entry is 64 bytes, four 16-byte vectors: (x, y) (z, vx) (vy, vz) (t, key) with
// key = molecule index within the slice (bits 0-39) | status bits << 40 | lens outcome << 48 | lens state
// << 56; the walk to the lens queues a molecule with status and outcome 0, and the lens kernel writes
// the molecule's state where the lens left it back into the same entry with both set.  Each field is
// kept as its bit pattern.
constexpr int64_t QKEY_INDEX = ((int64_t)1 << 40) - 1;

__device__ __forceinline__ void queue_put(longlong2 *q, int64_t e, const MState &m, int64_t i, int st) {
  longlong2 *p = q + 4 * e;
  p[0] = make_longlong2(__double_as_longlong(m.x), __double_as_longlong(m.y));
  p[1] = make_longlong2(__double_as_longlong(m.z), __double_as_longlong(m.vx));
  p[2] = make_longlong2(__double_as_longlong(m.vy), __double_as_longlong(m.vz));
  p[3] = make_longlong2(__double_as_longlong(m.t), (long long)(i | ((int64_t)st << 56)));
}

__device__ __forceinline__ void queue_put_done(longlong2 *q, int64_t e, double x, double y, double z, double vx,
                                               double vy, double vz, double t, int64_t key) {
  longlong2 *p = q + 4 * e;
  p[0] = make_longlong2(__double_as_longlong(x), __double_as_longlong(y));
  p[1] = make_longlong2(__double_as_longlong(z), __double_as_longlong(vx));
  p[2] = make_longlong2(__double_as_longlong(vy), __double_as_longlong(vz));
  p[3] = make_longlong2(__double_as_longlong(t), (long long)key);
}

__device__ __forceinline__ void queue_get_done(const longlong2 *q, int64_t e, MState &m, int64_t &key) {
  const longlong2 *p = q + 4 * e;
  const longlong2 a = p[0], b = p[1], c = p[2], d = p[3];
  m.x = __longlong_as_double(a.x);
  m.y = __longlong_as_double(a.y);
  m.z = __longlong_as_double(b.x);
  m.vx = __longlong_as_double(b.y);
  m.vy = __longlong_as_double(c.x);
  m.vz = __longlong_as_double(c.y);
  m.t = __longlong_as_double(d.x);
  key = (int64_t)d.y;
}

__device__ __forceinline__ void queue_get(const longlong2 *q, int64_t e, MState &m, int64_t &i, int &st) {
  const longlong2 *p = q + 4 * e;
  const longlong2 a = p[0], b = p[1], c = p[2], d = p[3];
  m.x = __longlong_as_double(a.x);
  m.y = __longlong_as_double(a.y);
  m.z = __longlong_as_double(b.x);
  m.vx = __longlong_as_double(b.y);
  m.vy = __longlong_as_double(c.x);
  m.vz = __longlong_as_double(c.y);
  m.t = __longlong_as_double(d.x);
  i = (int64_t)d.y & QKEY_INDEX;
  st = (int)((uint64_t)d.y >> 56);
}

// Wave-aggregated histogram: one LDS atomic per distinct key in the wavefront (its lowest lane
// adds the group's count).  The first molecule index of each label is an LDS atomic minimum per
// lane, so lanes need not hold their molecules in index order (the queue-fed phases do not).
__device__ __forceinline__ void wave_hist(uint32_t *s_hist, unsigned long long *s_first, int key, int label,
                                          int64_t gidx) {
  if (s_first) atomicMin(&s_first[label], (unsigned long long)gidx);
  uint64_t todo = __ballot(1);
  const int lane = lane_id();
  while (todo) {
    const int leader = __ffsll((unsigned long long)todo) - 1;
    const int lk = __shfl(key, leader);
    const uint64_t same = __ballot(key == lk) & todo;
    if (lane == leader) atomicAdd(&s_hist[lk], (uint32_t)__popcll(same));
    todo &= ~same;
  }
}

// Lane 0's value as a wave-uniform (scalar) value; every lane must be active.  Unlike a shuffle, the
// compiler knows the result is uniform, so what is computed from it stays in scalar registers.
__device__ __forceinline__ unsigned long long uniform_u64(unsigned long long v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((unsigned long long)hi << 32) | lo;
}

// Sum over the 64 lanes by an xor butterfly: every lane receives the same, bitwise identical
// total (each level adds the same two partial sums, only in swapped order), so the result does not
// depend on which lane is read.  All 64 lanes must be active.
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 1; o < 64; o <<= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// A copy of `p` the compiler cannot see through: loads made through it stay where they are
// written instead of being hoisted out of a walker loop (which would keep dozens of pointers live
// in SGPRs and spill them).
template <class T>
__device__ __forceinline__ const T *opaque(const T *p) {
  uint64_t v = (uint64_t)p;
  asm volatile("" : "+s"(v));
  return (const T *)v;
}

}  // namespace ctraj
