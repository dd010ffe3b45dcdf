// ctraj_sinks.hpp -- trajectory-row sinks for the record kernel.
#pragma once

#include "ctraj_device.hpp"

namespace ctraj {

__device__ __forceinline__ double nan64() { return __builtin_nan(""); }

// Every row of the molecule's Trajectory (molecule.py:111-144): buf[f][row][slot] for the ten
// fields x, y, z, vx, vy, vz, ax, ay, az, t.  Rows past `cap` are counted but not stored; rows after
// the molecule's last row are NaN.
struct RowSink {
  static constexpr bool kRows = true;
  double *buf;
  int64_t n, slot;
  int cap;
  int count;
  __device__ __forceinline__ void put(int f, int r, double v) { buf[((size_t)f * cap + r) * n + slot] = v; }
  __device__ __forceinline__ void row(const MState &m) {
    if (count < cap) {
      put(0, count, m.x); put(1, count, m.y); put(2, count, m.z);
      put(3, count, m.vx); put(4, count, m.vy); put(5, count, m.vz);
      put(6, count, m.ax); put(7, count, m.ay); put(8, count, m.az);
      put(9, count, m.t);
    }
    ++count;
  }
  // rows past the end are NaN, like the reference's np.full(nan) trajectory arrays
  __device__ __forceinline__ void finish(const MState &) {
    for (int r = count; r < cap; ++r)
      for (int f = 0; f < 10; ++f) put(f, r, nan64());
  }
};

// Samples at z_k = z_start + k*z_step by the rule of find_radial_pos_dist / find_vel_dist
// (post_processing.py:40-75): a sample beyond the last row is skipped (:42); a z_k equal to a row's
// z returns the first such row (:52-56); otherwise the row before the first row with z >= z_k is
// advanced by take_timestep (:9-17, :61-72) with its stored acceleration; z_k before the first row
// takes index -1, i.e. the last row (:61).  Streaming form: a sample is decided at the first row
// with z >= z_k; if that row is past z_k the decision waits one more row, because a zero-length
// face step (an L = 0 aperture) can land the next row back on z_k exactly.
// buf[f][k][slot] (MOL false) or buf[slot][k][f] (MOL true): x, y, vx, vy, vz, t; every sample is
// written (NaN where the rule gives none).  In the molecule-major layout a lane writes one
// contiguous stream (48 B per sample, three 16-byte stores), so partial cache lines are completed
// by the same lane's next samples however far the lanes of a wavefront drift apart in z.
template <bool MOL>
struct ZGridSinkT {
  static constexpr bool kRows = true;
  double *buf;
  int64_t n, slot;
  int cap;
  double z_start, z_step;
  int k;       // next undecided sample
  int kdefer;  // samples [0, kdefer) lie before the first row
  int pend_lo, pend_hi;  // samples waiting on the next row; default: extrapolate from pend_row
  int written;
  bool have_prev;
  MState prev, pend_row;
  __device__ __forceinline__ double zk(int i) const { return z_start + (double)i * z_step; }
  __device__ __forceinline__ void put6(int i, double x, double y, double vx, double vy, double vz, double t) {
    if constexpr (MOL) {
      double2 *q = reinterpret_cast<double2 *>(buf + ((size_t)slot * cap + i) * 6);
      q[0] = make_double2(x, y);
      q[1] = make_double2(vx, vy);
      q[2] = make_double2(vz, t);
    } else {
      const size_t s0 = (size_t)i * n + slot, fs = (size_t)cap * n;
      buf[s0] = x; buf[s0 + fs] = y; buf[s0 + 2 * fs] = vx;
      buf[s0 + 3 * fs] = vy; buf[s0 + 4 * fs] = vz; buf[s0 + 5 * fs] = t;
    }
  }
  __device__ __forceinline__ void put_nan(int i) {
    const double q = nan64();
    put6(i, q, q, q, q, q, q);
  }
  __device__ __forceinline__ void exact(int i, const MState &m) {
    put6(i, m.x, m.y, m.vx, m.vy, m.vz, m.t);
    ++written;
  }
  __device__ __forceinline__ void extrap(int i, const MState &p) {
    const double dt = (zk(i) - p.z) / p.vz;
    const double dt2 = dt * dt;
    put6(i, (p.x + p.vx * dt) + (p.ax * dt2) / 2.0, (p.y + p.vy * dt) + (p.ay * dt2) / 2.0, p.vx + p.ax * dt,
         p.vy + p.ay * dt, p.vz + p.az * dt, p.t + dt);
    ++written;
  }
  __device__ __forceinline__ void init() {
    k = 0; kdefer = 0; pend_lo = pend_hi = 0; written = 0; have_prev = false;
  }
  __device__ __forceinline__ void row(const MState &m) {
    if (!have_prev) {
      while (k < cap && zk(k) < m.z) ++k;
      kdefer = k;
      if (k < cap && zk(k) == m.z) { exact(k, m); ++k; }
      have_prev = true;
      prev = m;
      return;
    }
    for (int i = pend_lo; i < pend_hi; ++i) {
      if (zk(i) == m.z) exact(i, m);
      else extrap(i, pend_row);
    }
    pend_lo = pend_hi = k;
    while (k < cap) {
      const double z = zk(k);
      if (z < m.z) { ++k; pend_hi = k; }
      else if (z == m.z) { exact(k, m); ++k; }
      else break;
    }
    if (pend_hi > pend_lo) pend_row = prev;
    prev = m;
  }
  __device__ __forceinline__ void finish(const MState &last) {
    for (int i = pend_lo; i < pend_hi; ++i) extrap(i, pend_row);
    pend_lo = pend_hi = k;
    // post_processing.py:42: nothing beyond the last row's z
    for (int i = k - 1; i >= kdefer && zk(i) > last.z; --i) {
      put_nan(i);
      --written;
    }
    for (int i = 0; i < kdefer; ++i) {
      if (last.z < zk(i)) {
        put_nan(i);
        continue;
      }
      extrap(i, last);
    }
    // samples past the last row's z were never reached
    for (int i = k; i < cap; ++i) put_nan(i);
  }
};
using ZGridSink = ZGridSinkT<false>;

}  // namespace ctraj
