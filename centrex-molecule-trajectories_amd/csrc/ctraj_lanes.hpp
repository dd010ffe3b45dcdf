// ctraj_lanes.hpp -- persistent lane-refill walker for beamlines with one lens.
//
// In the ES-lens beamline most molecules stop inside the lens after a variable number of steps,
// so a wavefront that walks 64 molecules start to finish keeps most lanes idle while its longest
// survivor crosses the lens.  Here a wavefront instead keeps its lanes busy:
//   * lanes are in one of the phases NEED (no molecule), LENS (integrating), POST (left the lens,
//     elements after it still to walk) or DONE;
//   * every iteration advances all LENS lanes by one integration step;
//   * when at least `refill_min` lanes are NEED/POST (or none is in the lens), a convergent refill
//     section walks the POST lanes to their end, records finished molecules, and hands NEED lanes
//     new molecule indices from the wavefront's claim [next, end), refilled in chunks from a global
//     atomic counter; new molecules walk the elements before the lens and the lens entrance.
// Each molecule's arithmetic is exactly the one of walk() in ctraj_device.hpp, so results are
// bit-identical to the simple kernel and the oracle; only the order of work changes.
#pragma once

#include "ctraj_device.hpp"

namespace ctraj {

enum : int { PH_NEED = 0, PH_LENS = 1, PH_POST = 2, PH_DONE = 3 };

struct LaneMol {
  MState m;
  int64_t gidx;
  WsRef wsb;  // LDS spline of the molecule's state
  int st;
  int label;
  uint32_t status;
  int phase;
  int fin;  // 1: finished molecule waiting to be recorded; 2: same, stopped by the lens bore
  // lens integration state
  int i, nfull;
  double h, hp, dzs, r2;  // current step, final partial step, z advance per full step, x^2 + y^2
};

// Walk elements [e0, e1) (no lens among them) from the molecule's current row.  Returns false if
// the molecule stopped (label set).
template <int KM = KM_ALL, bool SE = false>
__device__ __forceinline__ bool walk_plain(const ctraj_elem *__restrict__ elems, int e0, int e1, double g, MState &m,
                                           int &label, WalkCounters &wc) {
  NullSink sk;
  for (int e = e0; e < e1; ++e) {
    const ctraj_elem E = elem_get<SE>(elems, e);
    const int lhit = SE ? stop_label(E.label_hit) : E.label_hit;
    if (E.kind == CTRAJ_CIRC) {
      const double thr2 = E.p[1];  // the squared-radius threshold of d/2 (beyond_r2)
      drift_dt(m, time_to(m, E.z0), g);
      ++wc.faces;
      if (beyond_r2(m.x, m.y, thr2)) { label = lhit; return false; }
      drift_dt(m, time_to(m, E.z1), g);
      ++wc.faces;
      if (beyond_r2(m.x, m.y, thr2)) { label = lhit; return false; }
    } else if ((KM & KM_RECT) && E.kind == CTRAJ_RECT) {
      const double x1 = E.p[0], x2 = E.p[1], y1 = E.p[2], y2 = E.p[3];
      drift_dt(m, time_to(m, E.z0), g);
      ++wc.faces;
      if (!((x1 < m.x && m.x < x2) && (y1 < m.y && m.y < y2))) { label = lhit; return false; }
      drift_dt(m, time_to(m, E.z1), g);
      ++wc.faces;
      if (!((x1 < m.x && m.x < x2) && (y1 < m.y && m.y < y2))) { label = lhit; return false; }
    } else if ((KM & KM_PLATES) && E.kind == CTRAJ_PLATES) {
      const double x1 = E.p[0], x2 = E.p[1];
      drift_dt(m, time_to(m, E.z0), g);
      ++wc.faces;
      if (!(x1 < m.x && m.x < x2)) { label = lhit; return false; }
      double dt = time_to(m, E.z1);
      double xe = m.x;
      if (dt != 0.0) xe = (m.x + m.vx * dt) + (m.ax * (dt * dt)) * 0.5;
      const bool inside = (x1 < xe && xe < x2);
      if (!inside) {
        if (m.vx < 0.0) dt = (x1 - m.x) / m.vx;
        else if (m.vx > 0.0) dt = (x2 - m.x) / m.vx;
      }
      drift_dt(m, dt, g);
      ++wc.faces;
      if (!inside) { label = lhit; return false; }
    }
  }
  (void)sk;
  return true;
}

}  // namespace ctraj
