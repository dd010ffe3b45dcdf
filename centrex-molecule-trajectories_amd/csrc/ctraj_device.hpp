// ctraj_device.hpp -- device-side building blocks of the gfx950 trajectory engine.
//
// One thread owns one molecule.  Its phase space lives in VGPRs as FP64; the element table is
// wave-uniform (scalar loads); lens tables are staged once per workgroup into LDS.
//
// Floating-point contract: the library is compiled with -ffp-contract=off, so every a*b+c below
// that is not written as __builtin_fma rounds twice, exactly as numpy does in the reference.
// The explicit FMAs (config-mode lens only) are mirrored one for one by the CPU oracle, which
// makes the device results reproducible bit for bit on the host.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ctraj.h"

namespace ctraj {

// Walker instantiation without lens code (beamlines that contain no lens).
constexpr int CTRAJ_INTEG_NONE = -1;

// Molecule state plus the acceleration stored with its current trajectory row (molecule.py:46-50).
struct MState {
  double x, y, z, vx, vy, vz, t;
  double ax, ay, az;
};

// LDS-qualified doubles: arithmetic on these pointers stays 32-bit and loads become ds_read.
typedef __attribute__((address_space(3))) double lds_double;
typedef double dbl2_t __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) dbl2_t lds_dbl2;

// Lens tables as seen from one workgroup (staged in LDS).  The W' spline segment k of a state is
// W'(k + s) = (A s + B) s + C.  (A, B) pairs and C sit in two arrays so that a look-up is one
// 16-byte and one 8-byte LDS read at 16- and 8-byte strides: random segments across a wavefront
// then cost about what conflict-free reads do, where one 24-byte record per segment (two 8-byte
// reads and one 8-byte read) costs half again as much.
struct Tables {
  const lds_dbl2 *ws_ab;   // W' spline (A, B): [state][nseg], 16-byte aligned
  const lds_double *ws_c;  // W' spline C: [state][nseg]
  int nseg;
  double umax;           // largest double below nseg: |E| beyond the table is clamped there
  const lds_double *rt;  // r tables: [knot][4] = {r, a, slope, 0}, all tables back to back
};

// Where one r table sits in Tables::rt.  The host resolves a lens's table id into these before
// launch: ctraj_elem.table = knot offset, .reserved = knot count, .p[6] = 1/(r_1 - r_0), and the
// first and last knots .p[4] = r_0, .p[5] = r_last (slots the spline forces use for their scales).
struct RDesc {
  int off, n;
  double invh, r0, rlast;
};

// ---------------------------------------------------------------------------------------------
// Constant-acceleration drift: Molecule.update_trajectory(delta_t) (molecule.py:58-68) using
// Molecule.x(dt) = x + v*dt + a*dt**2/2 and v(dt) = v + a*dt (molecule.py:26-44), with `a` the
// acceleration stored on the current row.  `if not delta_t` returns the current state (:31,41).
// The new row stores a = (0, -g, 0) (:58).
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ void drift_dt(MState &m, double dt, double g) {
  if (dt != 0.0) {
    const double dt2 = dt * dt;
    const double nx = (m.x + m.vx * dt) + (m.ax * dt2) * 0.5;
    const double ny = (m.y + m.vy * dt) + (m.ay * dt2) * 0.5;
    const double nz = (m.z + m.vz * dt) + (m.az * dt2) * 0.5;
    m.vx = m.vx + m.ax * dt;
    m.vy = m.vy + m.ay * dt;
    m.vz = m.vz + m.az * dt;
    m.x = nx;
    m.y = ny;
    m.z = nz;
  }
  m.t = m.t + dt;
  m.ax = 0.0;
  m.ay = -g;
  m.az = 0.0;
}

// time to the plane z = zf: delta_t = (z - x_z) / v_z (apertures.py:103)
__device__ __forceinline__ double time_to(const MState &m, double zf) { return (zf - m.z) / m.vz; }

__device__ __forceinline__ double rho_of(double x, double y) { return __builtin_sqrt(x * x + y * y); }

// rho_of(x, y) > d without the square root: the host stores thr2 = the largest double whose
// correctly rounded square root is <= d (circ_r2_threshold), and sqrt is monotonic, so
// sqrt(r2) > d exactly when r2 > thr2 (also for NaN, infinite and negative d).
__device__ __forceinline__ bool beyond_r2(double x, double y, double thr2) { return x * x + y * y > thr2; }

// ---------------------------------------------------------------------------------------------
// Lens forces.  Each returns the acceleration (ax, ay) at transverse position (x, y); a_z is 0.
// ---------------------------------------------------------------------------------------------

// np.interp on an ascending knot table (the interp1d of electrostatic_lens.py:209, evaluated at
// :217): j with r_j <= r < r_{j+1}; r == r_j returns a_j; r == r_last returns a_last; r below or
// above the table is flagged (the reference raises ValueError there) and returns a_0 or a_last; a
// NaN radius returns itself.
// The common case runs without a branch: the segment is guessed from the first knot spacing (the
// guess clamped to the table in floating point, so a NaN or infinite radius gives a valid index),
// its record (r_j, a_j, slope_j) and r_{j+1} are read in one LDS round trip and the interpolation
// is made; it stands when r_j <= r < r_{j+1}, which holds on every segment of a uniform grid except
// where rounding moves a knot across the guessed boundary.  Everything else -- that case, radii off
// the table, NaN -- takes one rarely entered branch that follows the rules above in order.
__device__ __forceinline__ double rtable_eval(const Tables &T, const RDesc &D, double r, uint32_t &status) {
  const lds_double *tb = T.rt + 4 * D.off;
  const int n = D.n;
  int j = (int)__builtin_fmin(__builtin_fmax((r - D.r0) * D.invh, 0.0), (double)(n - 2));
  const lds_dbl2 *rec = (const lds_dbl2 *)(tb + 4 * j);
  const dbl2_t ra = rec[0];
  const double slope = tb[4 * j + 2], rn = tb[4 * j + 4];
  // all three reads in flight together (the compiler would otherwise defer the slope's read into
  // the branch of r != r_j, a second round trip)
  asm volatile("" : : "v"(ra.x), "v"(ra.y), "v"(slope), "v"(rn));
  double res = r == ra.x ? ra.y : slope * (r - ra.x) + ra.y;
  if (!((ra.x <= r) & (r < rn))) {
    if (r < D.r0) {
      status |= CTRAJ_ST_OOT;
      res = tb[1];
    } else if (!(r < D.rlast)) {
      if (r == D.rlast) {
        res = tb[4 * (n - 1) + 1];
      } else if (r != r) {
        res = r;  // NaN propagates
      } else {
        status |= CTRAJ_ST_OOT;
        res = tb[4 * (n - 1) + 1];
      }
    } else {  // inside the table, one knot off the guess
      while (j > 0 && r < tb[4 * j]) --j;
      while (j < n - 2 && r >= tb[4 * (j + 1)]) ++j;
      const double rj = tb[4 * j], aj = tb[4 * j + 1];
      res = r == rj ? aj : tb[4 * j + 2] * (r - rj) + aj;
    }
  }
  return res;
}

// ElectrostaticLens.lens_acceleration (electrostatic_lens.py:168-228) with a tabulated a_r(r):
// a = (a_r x / r, a_r y / r - g, 0), and (0, -g, 0) on axis.  The quotients are formed whatever r
// and replaced by 0 on axis (a select, not a branch).
__device__ __forceinline__ void force_rtable(const Tables &T, const RDesc &D, double g, double x, double y,
                                             double &ax, double &ay, uint32_t &status) {
  const double r = rho_of(x, y);
  const double ar = rtable_eval(T, D, r, status);
  double fx = (ar * x) / r, fy = (ar * y) / r;
  asm volatile("" : "+v"(fx), "+v"(fy));
  const bool axis = r == 0.0;
  ax = axis ? 0.0 : fx;
  ay = (axis ? 0.0 : fy) - g;
}

// One lens state's W' spline in LDS.
struct WsRef {
  const lds_dbl2 *ab;
  const lds_double *c;
};

__device__ __forceinline__ WsRef ws_of(const Tables &T, int state) {
  const int o = state * T.nseg;
  return WsRef{T.ws_ab + o, T.ws_c + o};
}

// W'(u) from a state's LDS spline; u = |E|/dE >= 0.  u is clamped to umax, the largest double below
// nseg, so k = trunc(u) <= nseg - 1 and s = u - k = fract(u) exactly; the force saturates beyond the
// table end.
// min(u, umax) as one v_min_f64.  __builtin_fmin would also canonicalise umax on every call (a
// second FP64 instruction per look-up) because the compiler cannot see that the table bound, a
// normal double, needs none; u is an arithmetic result and umax a normal number, so the IEEE-mode
// minimum is fmin's result, including umax for a NaN u.  umax is wave-uniform and read from scalar
// registers (a VOP3 operand may be one scalar register pair), so no vector copy of it is made per
// look-up.
__device__ __forceinline__ double clamp_to(double u, double umax) {
  double r;
  asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(u), "s"(umax));
  return r;
}

__device__ __forceinline__ double wspline_eval(const WsRef &w, double umax, double u) {
  u = clamp_to(u, umax);
  const unsigned k = (unsigned)(int)u;
  const double s = __builtin_amdgcn_fract(u);
  const dbl2_t ab = w.ab[k];
  const double c = w.c[k];
  return __builtin_fma(__builtin_fma(ab.x, s, ab.y), s, c);
}

// Hexapole: |E| = cE r^2, a = -W'(|E|) 2 cE (x, y) / m - g y_hat.  r2 = x^2 + y^2 is passed in
// so that the bore test of the previous step can share it.
__device__ __forceinline__ void force_hex_r2(const WsRef &wsb, double umax, double ks, double cu, double g, double x,
                                             double y, double r2, double &ax, double &ay) {
  const double f = wspline_eval(wsb, umax, r2 * cu) * ks;
  ax = f * x;
  ay = __builtin_fma(f, y, -g);
}

// Linear-field model |E| = cE r (the reference's field assumption, electrostatic_lens.py:199-200):
// a = -W'(|E|) cE (x, y) / (m r) - g y_hat, and (0, -g) on axis.
__device__ __forceinline__ void force_lin(const WsRef &wsb, double umax, double ks, double cu, double g, double x,
                                          double y, double &ax, double &ay) {
  const double r = rho_of(x, y);
  double fx = 0.0, fy = 0.0;
  if (r != 0.0) {
    const double f = (wspline_eval(wsb, umax, r * cu) * ks) / r;
    fx = f * x;
    fy = f * y;
  }
  ax = fx;
  ay = fy - g;
}

// A lens force; `wsb` is the LDS spline of the molecule's state (ws_of).
template <int FORCE>
struct Force;

template <>
struct Force<CTRAJ_FORCE_RTABLE> {
  RDesc d;
  double g;
  __device__ __forceinline__ void operator()(const Tables &T, const WsRef & /*wsb*/, double x, double y, double &ax,
                                             double &ay, uint32_t &st) const {
    force_rtable(T, d, g, x, y, ax, ay, st);
  }
};

template <>
struct Force<CTRAJ_FORCE_WSPLINE_HEX> {
  double ks, cu, g;
  double gk;  // -g/ks when the lens folds its force scale into the step constants (hex_folds), else 0
  __device__ __forceinline__ void operator()(const Tables &T, const WsRef &wsb, double x, double y, double &ax,
                                             double &ay, uint32_t &) const {
    force_hex_r2(wsb, T.umax, ks, cu, g, x, y, __builtin_fma(x, x, y * y), ax, ay);
  }
};

template <>
struct Force<CTRAJ_FORCE_WSPLINE_LIN> {
  double ks, cu, g;
  __device__ __forceinline__ void operator()(const Tables &T, const WsRef &wsb, double x, double y, double &ax,
                                             double &ay, uint32_t &) const {
    force_lin(wsb, T.umax, ks, cu, g, x, y, ax, ay);
  }
};

// ---------------------------------------------------------------------------------------------
// Trajectory sinks.  The walker reports every row the reference would append to its Trajectory
// arrays (molecule.py:133-144, electrostatic_lens.py:106-111).  NullSink compiles to nothing.
// ---------------------------------------------------------------------------------------------
struct NullSink {
  static constexpr bool kRows = false;
  __device__ __forceinline__ void row(const MState &) {}
};

// ---------------------------------------------------------------------------------------------
// Element walk: Beamline.propagate_through (beamline.py:20-35) over the z0-sorted element table.
// Returns with `label` set (label_detected if the molecule survives) and m at its final row.
// ---------------------------------------------------------------------------------------------
struct WalkCounters {
  uint32_t faces;
  uint32_t steps;
};

// A hexapole lens with ks = -2 cE/m folds ks into its RK4 step constants (rk4_step_hex_k) when
// |ks| lies in [2^-500, 2^500]: then -g/ks and ks times any step size the lens uses are finite
// normal numbers.  The host calls the same function to choose the kernels.
__host__ __device__ __forceinline__ bool hex_folds(double ks) {
  const double a = __builtin_fabs(ks);
  return a >= 0x1p-500 && a <= 0x1p500;
}

template <int FORCE>
__device__ __forceinline__ Force<FORCE> make_force(const ctraj_elem &E, double g);

template <>
__device__ __forceinline__ Force<CTRAJ_FORCE_RTABLE> make_force<CTRAJ_FORCE_RTABLE>(const ctraj_elem &E, double g) {
  return Force<CTRAJ_FORCE_RTABLE>{RDesc{E.table, E.reserved, E.p[6], E.p[4], E.p[5]}, g};
}
template <>
__device__ __forceinline__ Force<CTRAJ_FORCE_WSPLINE_HEX> make_force<CTRAJ_FORCE_WSPLINE_HEX>(const ctraj_elem &E,
                                                                                           double g) {
  const double ks = E.p[4];
  return Force<CTRAJ_FORCE_WSPLINE_HEX>{ks, E.p[5], g, hex_folds(ks) ? -g / ks : 0.0};
}
template <>
__device__ __forceinline__ Force<CTRAJ_FORCE_WSPLINE_LIN> make_force<CTRAJ_FORCE_WSPLINE_LIN>(const ctraj_elem &E,
                                                                                           double g) {
  return Force<CTRAJ_FORCE_WSPLINE_LIN>{E.p[4], E.p[5], g};
}

// One step of the reference's lens integrator (electrostatic_lens.py:91-111): RK4 with the second
// stage at the full Euler point x + dt k1; the new row stores a = l1.
template <int FORCE>
__device__ __forceinline__ void ref_step(const Force<FORCE> &F, const Tables &T, const WsRef &wsb, double dt,
                                         MState &m, uint32_t &status) {
  const double x = m.x, y = m.y, z = m.z;
  const double k1x = m.vx, k1y = m.vy, k1z = m.vz;
  double l1x, l1y;
  F(T, wsb, x, y, l1x, l1y, status);
  const double k2x = k1x + (dt * l1x) / 2.0;
  const double k2y = k1y + (dt * l1y) / 2.0;
  const double k2z = k1z + (dt * 0.0) / 2.0;
  double l2x, l2y;
  F(T, wsb, x + dt * k1x, y + dt * k1y, l2x, l2y, status);
  const double k3x = k1x + (dt * l2x) / 2.0;
  const double k3y = k1y + (dt * l2y) / 2.0;
  const double k3z = k1z + (dt * 0.0) / 2.0;
  double l3x, l3y;
  F(T, wsb, x + (dt * k2x) / 2.0, y + (dt * k2y) / 2.0, l3x, l3y, status);
  const double k4x = k1x + dt * l3x;
  const double k4y = k1y + dt * l3y;
  const double k4z = k1z + dt * 0.0;
  double l4x, l4y;
  F(T, wsb, x + dt * k3x, y + dt * k3y, l4x, l4y, status);
  m.x = x + (dt * (((k1x + 2.0 * k2x) + 2.0 * k3x) + k4x)) / 6.0;
  m.y = y + (dt * (((k1y + 2.0 * k2y) + 2.0 * k3y) + k4y)) / 6.0;
  m.z = z + (dt * (((k1z + 2.0 * k2z) + 2.0 * k3z) + k4z)) / 6.0;
  m.vx = k1x + (dt * (((l1x + 2.0 * l2x) + 2.0 * l3x) + l4x)) / 6.0;
  m.vy = k1y + (dt * (((l1y + 2.0 * l2y) + 2.0 * l3y) + l4y)) / 6.0;
  m.vz = k1z + (dt * (((0.0 + 2.0 * 0.0) + 2.0 * 0.0) + 0.0)) / 6.0;
  m.t = m.t + dt;
  m.ax = l1x;
  m.ay = l1y;
  m.az = 0.0;
}

// The reference's lens interior (electrostatic_lens.py:86-118): N = rint(L/dz) steps of dt = dz/vz,
// bore test sqrt(x^2 + y^2) > d/2 after each, made as x^2 + y^2 > p[1] with the device copy's
// squared-radius threshold of d/2 (beyond_r2; a reference-mode lens has no other use for p[1]).
// Returns false when the molecule leaves the bore.
template <int FORCE, class Sink>
__device__ __forceinline__ bool lens_ref(const ctraj_elem &E, const Tables &T, int state, double g, MState &m,
                                         uint32_t &status, WalkCounters &wc, Sink &sink) {
  const Force<FORCE> F = make_force<FORCE>(E, g);
  const WsRef &wsb = ws_of(T, state);
  const double thr2 = E.p[1];
  const int N = E.n_steps;
  const double dt = E.p[2] / m.vz;
  for (int i = 0; i < N; ++i) {
    ref_step<FORCE>(F, T, wsb, dt, m, status);
    ++wc.steps;
    if constexpr (Sink::kRows) sink.row(m);
    if (beyond_r2(m.x, m.y, thr2)) return false;
  }
  return true;
}

// h/6 is formed as h * (1/6) (one rounding of the constant, one of the product), the same on host
// and device.
constexpr double kSixth = 1.0 / 6.0;

// Classical RK4 on the transverse state (x, y, vx, vy) with step h; vz is constant in the lens.  The
// force does not depend on the velocity, so the stage velocities enter only through their sums, and
// with slopes s_k = a(x_k) / ks (ks = 1 for the plain step; the folded hexapole step below divides
// the force scale out) the stage points and the result expand to
//   x2 = x + h/2 v,  x3 = x2 + (h/2)(ks h/2) s1,  x4 = (x + h v) + h (ks h/2) s2,
//   x' = (x + h v) + (h/6)(ks h) (s1 + s2 + s3),  v' = v + (ks h/6) (2 (s1 + s2 + s3) + (s4 - s1)),
// the classical stages in exact arithmetic with 20 operations on the state where the stage
// velocities take 28.  HexK holds the constants, formed by hex_k in this order wherever a step size
// is chosen (host or lane); h/6 is formed as h * (1/6).  The CPU oracle mirrors every operation.
struct HexK {
  double h, hh, c3, c4, cx, kh6;  // h, h/2, (h/2)(ks h/2), h (ks h/2), (h/6)(ks h), ks h/6
};

__host__ __device__ __forceinline__ HexK hex_k(double ks, double h) {
  const double hh = 0.5 * h, h6 = h * kSixth, khh = ks * hh, kh = ks * h;
  return HexK{h, hh, hh * khh, h * khh, h6 * kh, ks * h6};
}

// The expanded step; slope(x, y, r2, sx, sy) gives the stage slope at (x, y), r2 = x^2 + y^2 (the
// first stage's r2 is the caller's).
template <class Slope>
__device__ __forceinline__ void rk4_expanded(const HexK &K, const Slope &slope, double &x, double &y, double &vx,
                                             double &vy, double r2) {
  double s1x, s1y, s2x, s2y, s3x, s3y, s4x, s4y;
  slope(x, y, r2, s1x, s1y);
  const double x2 = __builtin_fma(K.hh, vx, x), y2 = __builtin_fma(K.hh, vy, y);
  slope(x2, y2, __builtin_fma(x2, x2, y2 * y2), s2x, s2y);
  const double x3 = __builtin_fma(K.c3, s1x, x2), y3 = __builtin_fma(K.c3, s1y, y2);
  const double xv = __builtin_fma(K.h, vx, x), yv = __builtin_fma(K.h, vy, y);
  slope(x3, y3, __builtin_fma(x3, x3, y3 * y3), s3x, s3y);
  const double x4 = __builtin_fma(K.c4, s2x, xv), y4 = __builtin_fma(K.c4, s2y, yv);
  double sumx = s1x + s2x, sumy = s1y + s2y;
  sumx = sumx + s3x;
  sumy = sumy + s3y;
  slope(x4, y4, __builtin_fma(x4, x4, y4 * y4), s4x, s4y);
  x = __builtin_fma(K.cx, sumx, xv);
  y = __builtin_fma(K.cx, sumy, yv);
  vx = __builtin_fma(K.kh6, __builtin_fma(2.0, sumx, s4x - s1x), vx);
  vy = __builtin_fma(K.kh6, __builtin_fma(2.0, sumy, s4y - s1y), vy);
}

// The plain step (slopes = accelerations): K = hex_k(1.0, h); rk4_step forms it from h.
template <int FORCE>
__device__ __forceinline__ void rk4_step_k(const Force<FORCE> &F, const Tables &T, const WsRef &wsb, const HexK &K,
                                           double &x, double &y, double &vx, double &vy, double r2, uint32_t &status) {
  rk4_expanded(
      K,
      [&](double px, double py, double pr2, double &ax, double &ay) {
        if constexpr (FORCE == CTRAJ_FORCE_WSPLINE_HEX) force_hex_r2(wsb, T.umax, F.ks, F.cu, F.g, px, py, pr2, ax, ay);
        else F(T, wsb, px, py, ax, ay, status);
      },
      x, y, vx, vy, r2);
}

template <int FORCE>
__device__ __forceinline__ void rk4_step(const Force<FORCE> &F, const Tables &T, const WsRef &wsb, double h,
                                         double &x, double &y, double &vx, double &vy, double r2, uint32_t &status) {
  rk4_step_k<FORCE>(F, T, wsb, hex_k(1.0, h), x, y, vx, vy, r2, status);
}

// The hexapole step with the force scale folded into the step constants.  The acceleration is
// a = ks W'(cu r^2) (x, y) - g y_hat = ks s with the slope s = W' (x, y) + (0, gk), gk = -g/ks, so
// the step takes K = hex_k(ks, h): one multiplication fewer per slope than the plain step.
__device__ __forceinline__ void slope_hex_k(const WsRef &wsb, double umax, double cu, double gk, double x, double y,
                                            double r2, double &sx, double &sy) {
  const double w = wspline_eval(wsb, umax, r2 * cu);
  sx = w * x;
  sy = __builtin_fma(w, y, gk);
}

__device__ __forceinline__ void rk4_step_hex_k(const WsRef &wsb, double umax, double cu, double gk, const HexK &K,
                                               double &x, double &y, double &vx, double &vy, double r2) {
  rk4_expanded(
      K, [&](double px, double py, double pr2, double &sx, double &sy) { slope_hex_k(wsb, umax, cu, gk, px, py, pr2, sx, sy); },
      x, y, vx, vy, r2);
}

// One configuration-mode lens step of size h: the folded hexapole step where the lens folds,
// rk4_step otherwise.
template <int FORCE>
__device__ __forceinline__ void lens_step(const Force<FORCE> &F, const Tables &T, const WsRef &wsb, double h,
                                          double &x, double &y, double &vx, double &vy, double r2, uint32_t &status) {
  if constexpr (FORCE == CTRAJ_FORCE_WSPLINE_HEX) {
    if (hex_folds(F.ks)) {
      rk4_step_hex_k(wsb, T.umax, F.cu, F.gk, hex_k(F.ks, h), x, y, vx, vy, r2);
      return;
    }
  }
  rk4_step<FORCE>(F, T, wsb, h, x, y, vx, vy, r2, status);
}

// Number of full RK4 steps in a lens of length L at axial speed vz: floor(L / (vz dt)), 0 for a
// non-finite or negative quotient, clipped to cap (CTRAJ_ST_CAPPED).
__device__ __forceinline__ int rk4_full_steps(double L, double vz, double dt, int cap, uint32_t &status) {
  const double q = __builtin_floor(L / (vz * dt));
  if (!(q >= 0.0)) return 0;
  if (q > (double)cap) {
    status |= CTRAJ_ST_CAPPED;
    return cap;
  }
  return (int)q;
}

// Config-mode lens interior: n = floor(L/(vz dt)) steps of dt, then one step of L/vz - n dt, so
// the molecule leaves at z1.  Bore test x^2 + y^2 > (d/2)^2 after every step.  Rows store the
// force at the row's position; the exit row stores (0, -g, 0).  z and t after k full steps are
// formed from the entry values, z_k = z_in + k (vz dt) and t_k = t_in + k dt (one FMA each, not
// accumulated step by step), and after the partial step z = z1, t = (t_in + n dt) + hp.
template <int FORCE, class Sink>
__device__ __forceinline__ bool lens_rk4(const ctraj_elem &E, const Tables &T, int state, double g, int cap,
                                         MState &m, uint32_t &status, WalkCounters &wc, Sink &sink) {
  const Force<FORCE> F = make_force<FORCE>(E, g);
  const WsRef &wsb = ws_of(T, state);
  const double R2 = E.p[1];
  const double h = E.p[2];
  const double L = E.p[3];
  const double dzs = m.vz * h;
  const int nfull = rk4_full_steps(L, m.vz, h, cap, status);
  double x = m.x, y = m.y, vx = m.vx, vy = m.vy;
  const double z_in = m.z, t_in = m.t;
  double z = z_in, t = t_in;
  double r2 = __builtin_fma(x, x, y * y);
  bool alive = true;
  for (int i = 0; i <= nfull; ++i) {
    if (i < nfull) {
      lens_step<FORCE>(F, T, wsb, h, x, y, vx, vy, r2, status);
      z = __builtin_fma((double)(i + 1), dzs, z_in);
      t = __builtin_fma((double)(i + 1), h, t_in);
    } else {
      const double hp = L / m.vz - (double)nfull * h;
      lens_step<FORCE>(F, T, wsb, hp, x, y, vx, vy, r2, status);
      z = E.z1;
      t = __builtin_fma((double)nfull, h, t_in) + hp;
    }
    ++wc.steps;
    r2 = __builtin_fma(x, x, y * y);
    if constexpr (Sink::kRows) {
      MState r = m;
      r.x = x; r.y = y; r.z = z; r.vx = vx; r.vy = vy; r.t = t;
      if (i < nfull || r2 > R2) {
        F(T, wsb, x, y, r.ax, r.ay, status);
      } else {
        r.ax = 0.0; r.ay = -g;
      }
      r.az = 0.0;
      sink.row(r);
    }
    if (r2 > R2) {
      alive = false;
      break;
    }
  }
  m.x = x; m.y = y; m.vx = vx; m.vy = vy; m.z = z; m.t = t;
  if (alive) {
    m.ax = 0.0; m.ay = -g; m.az = 0.0;
  } else {
    uint32_t st = 0;
    F(T, wsb, x, y, m.ax, m.ay, st);
    m.az = 0.0;
  }
  return alive;
}

// Element kinds a walker instantiation handles (KM bit mask); the host picks the narrowest mask
// that covers the beamline, which keeps unused code (and its registers) out of the kernel.
constexpr int KM_CIRC = 1, KM_RECT = 2, KM_PLATES = 4, KM_ALL = 7;

// The element table is written by the host before a launch and never by a kernel, so a streaming
// walker (one that prefetches its next molecule) reads it through the constant address space:
// wave-uniform reads become scalar loads, which do not wait behind the wavefront's outstanding
// vector loads and stores (vector memory operations complete in order, so a vector load of an
// element would also wait for the prefetched next molecule).  The lane-refill kernels, which
// prefetch nothing and are short of scalar registers, keep vector loads (SE = false).
__device__ __forceinline__ ctraj_elem elem_at(const ctraj_elem *elems, int e) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef const __attribute__((address_space(4))) ctraj_elem const_elem_t;
  const const_elem_t &E = ((const_elem_t *)elems)[e];
  ctraj_elem r;
  r.kind = E.kind;
  r.label_hit = E.label_hit;
  r.label_inside = E.label_inside;
  r.force = E.force;
  r.integrator = E.integrator;
  r.table = E.table;
  r.n_steps = E.n_steps;
  r.reserved = E.reserved;
  r.z0 = E.z0;
  r.z1 = E.z1;
#pragma unroll
  for (int k = 0; k < 8; ++k) r.p[k] = E.p[k];
  return r;
#else
  return elems[e];  // the host pass never runs device code
#endif
}

// The label of each element visited, read in a streaming walker's loop and kept in a vector
// register.  Without the empty asm the compiler sinks the read of the stopping element's label past
// the loop, as a per-lane vector load of the element table, and the wait for that load (vector
// memory operations complete in order) also waits for the prefetched next molecule.
__device__ __forceinline__ int stop_label(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

template <bool SE>
__device__ __forceinline__ ctraj_elem elem_get(const ctraj_elem *elems, int e) {
  if constexpr (SE) return elem_at(elems, e);
  else return elems[e];
}

template <int FORCE, int INTEG, class Sink, int KM = KM_ALL, bool SE = false>
__device__ __forceinline__ void walk(const ctraj_elem *__restrict__ elems, int n_elems, const Tables &T, double g,
                                     int cap, int state, int label_detected, MState &m, int &label, uint32_t &status,
                                     WalkCounters &wc, Sink &sink) {
  if constexpr (Sink::kRows) sink.row(m);
  label = label_detected;
  for (int e = 0; e < n_elems; ++e) {
    const ctraj_elem E = elem_get<SE>(elems, e);
    const int lhit = SE ? stop_label(E.label_hit) : E.label_hit;
    if (E.kind == CTRAJ_CIRC) {
      // CircularAperture.propagate_through (apertures.py:92-115): test at z0 and z1; x0, y0 unused.
      // p[1] holds the squared-radius threshold of d/2 (beyond_r2)
      const double thr2 = E.p[1];
      drift_dt(m, time_to(m, E.z0), g);
      ++wc.faces;
      if constexpr (Sink::kRows) sink.row(m);
      if (beyond_r2(m.x, m.y, thr2)) { label = lhit; return; }
      drift_dt(m, time_to(m, E.z1), g);
      ++wc.faces;
      if constexpr (Sink::kRows) sink.row(m);
      if (beyond_r2(m.x, m.y, thr2)) { label = lhit; return; }
    } else if ((KM & KM_RECT) && E.kind == CTRAJ_RECT) {
      // RectangularAperture.propagate_through (apertures.py:165-189): strict box test
      const double x1 = E.p[0], x2 = E.p[1], y1 = E.p[2], y2 = E.p[3];
      drift_dt(m, time_to(m, E.z0), g);
      ++wc.faces;
      if constexpr (Sink::kRows) sink.row(m);
      if (!((x1 < m.x && m.x < x2) && (y1 < m.y && m.y < y2))) { label = lhit; return; }
      drift_dt(m, time_to(m, E.z1), g);
      ++wc.faces;
      if constexpr (Sink::kRows) sink.row(m);
      if (!((x1 < m.x && m.x < x2) && (y1 < m.y && m.y < y2))) { label = lhit; return; }
    } else if ((KM & KM_PLATES) && E.kind == CTRAJ_PLATES) {
      // FieldPlates.propagate_through (apertures.py:227-270): x-bounded plates with wall-hit time
      const double x1 = E.p[0], x2 = E.p[1];
      drift_dt(m, time_to(m, E.z0), g);
      ++wc.faces;
      if constexpr (Sink::kRows) sink.row(m);
      if (!(x1 < m.x && m.x < x2)) { label = lhit; return; }
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
      if constexpr (Sink::kRows) sink.row(m);
      if (!inside) { label = lhit; return; }
    } else if (INTEG != CTRAJ_INTEG_NONE && E.kind == CTRAJ_LENS) {
      // ElectrostaticLens.propagate_through (electrostatic_lens.py:48-77)
      drift_dt(m, time_to(m, E.z0), g);
      ++wc.faces;
      if constexpr (Sink::kRows) sink.row(m);
      if (rho_of(m.x, m.y) > E.p[0]) { label = lhit; return; }
      if constexpr (INTEG == CTRAJ_INTEG_NONE) {
      } else if constexpr (INTEG == CTRAJ_INTEG_REF) {
        if (!lens_ref<FORCE>(E, T, state, g, m, status, wc, sink)) { label = E.label_inside; return; }
        // exit hop with the stored a = l1 of the last step (electrostatic_lens.py:74-77)
        drift_dt(m, time_to(m, E.z1), g);
        ++wc.faces;
        if constexpr (Sink::kRows) sink.row(m);
      } else {
        if (!lens_rk4<FORCE>(E, T, state, g, cap, m, status, wc, sink)) { label = E.label_inside; return; }
      }
    }
  }
}

}  // namespace ctraj
