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
// buf[f][k][slot]: x, y, vx, vy, vz, t; every sample is written (NaN where the rule gives none).
struct ZGridSink {
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
  __device__ __forceinline__ void put(int f, int i, double v) { buf[((size_t)f * cap + i) * n + slot] = v; }
  __device__ __forceinline__ void exact(int i, const MState &m) {
    put(0, i, m.x); put(1, i, m.y); put(2, i, m.vx); put(3, i, m.vy); put(4, i, m.vz); put(5, i, m.t);
    ++written;
  }
  __device__ __forceinline__ void extrap(int i, const MState &p) {
    const double dt = (zk(i) - p.z) / p.vz;
    const double dt2 = dt * dt;
    put(0, i, (p.x + p.vx * dt) + (p.ax * dt2) / 2.0);
    put(1, i, (p.y + p.vy * dt) + (p.ay * dt2) / 2.0);
    put(2, i, p.vx + p.ax * dt);
    put(3, i, p.vy + p.ay * dt);
    put(4, i, p.vz + p.az * dt);
    put(5, i, p.t + dt);
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
      for (int f = 0; f < 6; ++f) put(f, i, nan64());
      --written;
    }
    for (int i = 0; i < kdefer; ++i) {
      if (last.z < zk(i)) {
        for (int f = 0; f < 6; ++f) put(f, i, nan64());
        continue;
      }
      extrap(i, last);
    }
    // samples past the last row's z were never reached
    for (int i = k; i < cap; ++i)
      for (int f = 0; f < 6; ++f) put(f, i, nan64());
  }
};

}  // namespace ctraj
