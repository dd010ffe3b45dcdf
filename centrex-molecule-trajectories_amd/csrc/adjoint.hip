// adjoint.hip -- sigmoid-smoothed detected flux and its discrete adjoint (gfx950).
//
// k_smooth<GRAD> walks 64 molecules per wavefront work item through the beamline without stopping
// them.  Each circular-aperture face multiplies the molecule's weight by sigmoid((r - rho)/s); the
// lens is integrated with the same classical RK4 (and the same FMA sequence) as the propagation
// kernels.  With GRAD the wavefront then runs the reverse sweep:
//   * aperture faces: the sigmoid factor's derivative and the adjoint of the ballistic drift;
//   * lens: for each checkpoint segment (last first) the K steps are recomputed from the stored
//     checkpoint into a per-wavefront recompute store, then the adjoint of each RK4 step is applied
//     backwards, accumulating d/dV through the field scale and d/d(vz) through the final partial step;
//   * the mean axial speed collects the adjoint of vz.
// Per-parameter sums are reduced over the wavefront (xor butterfly) at wave-uniform points and
// written per work item; k_smooth_reduce adds the items of each parameter set in a fixed order,
// so J and the gradient do not depend on scheduling.
//
// Workspace of one wavefront (doubles, lane-minor so that every access is one 512-byte row):
//   faces [n_face_slots][4][64]   state (x, y, vx, vy) on each aperture face and lens entrance
//   ckpt  [n_ck_slots][4][64]     lens state before steps 0, K, 2K, ...
//   sub   [ceil(K/8)][4][64]      state every 8 steps of the segment being reversed
// and, in LDS, a [8][4][64] window of recomputed step inputs per wavefront.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "ctraj_device.hpp"
#include "ctraj_internal.hpp"
#include "ctraj_wave.hpp"

using namespace ctraj;

namespace {

struct SmoothArgs {
  const double *x, *y, *vx, *vy, *eps;
  const double *params;           // [n_cfg][n_params]
  double *partial;                // [n_items][n_params + 1]: gradient sums, then the J sum
  double *ws;                     // workspace, ws_stride doubles per wavefront
  unsigned long long *n_capped;   // may be null
  int64_t n, items_per_cfg, n_items, ws_stride;
  double z_src, sigma_vz, inv_s;
  int n_params, K, max_steps, ck_per_lens, off_ck, off_sub;  // offsets in rows of 256 doubles
  int rec_lds;                  // LDS offset (doubles) of the per-wavefront recompute windows
};

// Inside a checkpoint segment the reverse sweep works in windows of kSub steps: the segment is
// recomputed once to store the state every kSub steps (HBM rows, cache-resident), then each
// window, last first, is recomputed into the wavefront's LDS slice and reversed.  With K <= kSub
// (the default K = kSub) a segment is one window: every step is recomputed once.
constexpr int kSub = 8;

// sigmoid(w) and 1 - sigmoid(w) without cancellation.
__device__ __forceinline__ void sigmoid2(double w, double &phi, double &omphi) {
  const double e = exp(-__builtin_fabs(w));
  const double q = 1.0 / (1.0 + e);
  const double eq = e * q;
  phi = w >= 0.0 ? q : eq;
  omphi = w >= 0.0 ? eq : q;
}

// Ballistic drift by tau under a = (0, -g): the propagation kernels' drift with the stored
// acceleration (0, -g, 0) (x gains + (0 * tau^2) * 0.5 = +0 there, which changes nothing).
__device__ __forceinline__ void drift_g(double &x, double &y, double vx, double &vy, double tau, double g) {
  x = x + vx * tau;
  y = (y + vy * tau) + ((-g) * (tau * tau)) * 0.5;
  vy = vy + (-g) * tau;
}

// Adjoint of drift_g given the state after the drift (vx, vy_after): updates (vbx, vby) in place
// (pbx, pby pass through) and returns d/dtau.
__device__ __forceinline__ double drift_g_adj(double vx, double vy_after, double g, double tau, double pbx,
                                              double pby, double &vbx, double &vby) {
  const double taubar = (pbx * vx + pby * vy_after) - g * vby;
  vbx = vbx + tau * pbx;
  vby = vby + tau * pby;
  return taubar;
}

// Hexapole force and its derivatives at r2 = x^2 + y^2:
//   f  = ks W'(u), u = min(cu r2, umax)       (a = f (x, y) - g y_hat; identical to force_hex_r2)
//   d  = 2 ks cu dW'/du  (= 2 df/dr2; 0 where u is clamped)
//   gv = df/dV = (f + ks cu dW'/du r2) / V    (ks and cu are both proportional to V)
__device__ __forceinline__ void hex_eval_d(const WsRef &wsb, double umax, double ks, double cu, double kc2,
                                           double invV, double r2, double &f, double &d, double &gv) {
  const double u0 = r2 * cu;
  const double u = clamp_to(u0, umax);
  const unsigned k = (unsigned)(int)u;
  const double s = __builtin_amdgcn_fract(u);
  const dbl2_t ab = wsb.ab[k];
  const double A = ab.x, Bc = ab.y, C = wsb.c[k];
  const double P = __builtin_fma(__builtin_fma(A, s, Bc), s, C);
  double dP = __builtin_fma(2.0 * A, s, Bc);
  if (u0 > umax) dP = 0.0;
  f = P * ks;
  d = kc2 * dP;
  gv = __builtin_fma(0.5 * d, r2, f) * invV;
}

struct LensK {
  WsRef wsb;
  double umax, ks, cu, kc2, invV, g;
};

// The forward sweep's lens step and its recomputation in the reverse sweep: full steps of the
// lens's uniform h take the folded hexapole step with the constants formed once per lens (as the
// propagation's lens kernel does), the final partial step hp takes lens_step; where the force scale
// does not fold every step takes the plain RK4 step.  The forward sweep and both recomputations call
// this one function, so a recomputed state equals the forward state bit for bit.
struct SmoothStep {
  HexK hk;
  bool fold;
};

__device__ __forceinline__ SmoothStep smooth_step_k(const Force<CTRAJ_FORCE_WSPLINE_HEX> &F, double h) {
  return SmoothStep{hex_k(F.ks, h), hex_folds(F.ks)};
}

__device__ __forceinline__ void smooth_step(const Force<CTRAJ_FORCE_WSPLINE_HEX> &F, const Tables &T, const WsRef &wsb,
                                            const SmoothStep &S, bool full, double hp, double &x, double &y,
                                            double &vx, double &vy, double r2, uint32_t &st) {
  if (full && S.fold)
    rk4_step_hex_k(wsb, T.umax, F.cu, F.gk, S.hk, x, y, vx, vy, r2);
  else
    lens_step<CTRAJ_FORCE_WSPLINE_HEX>(F, T, wsb, full ? S.hk.h : hp, x, y, vx, vy, r2, st);
}

// Reverse of one classical RK4 step of size h from input state (x, y, vx, vy), derived from the
// stages with their velocities formed; the forward step (rk4_expanded) forms the same stages in
// expanded form, equal in exact arithmetic, so this is its adjoint up to rounding.
// In: (pbx, pby, vbx, vby) = adjoint of the step's output.  Out: adjoint of its input.
// Vbar accumulates d/dV; with WANT_H the derivative with respect to h is returned.
template <bool WANT_H>
__device__ __forceinline__ double rk4_adjoint(const LensK &L, double h, double x, double y, double vx, double vy,
                                              double &pbx, double &pby, double &vbx, double &vby, double &Vbar) {
  const double hh = 0.5 * h, h6 = h * kSixth, h3 = 2.0 * h6;
  const double g = L.g;
  double f1, d1, g1, f2, d2, g2, f3, d3, g3, f4, d4, g4;
  // forward stages with their velocities (the classical form: rk4_step's expanded stages up to rounding)
  hex_eval_d(L.wsb, L.umax, L.ks, L.cu, L.kc2, L.invV, __builtin_fma(x, x, y * y), f1, d1, g1);
  const double a1x = f1 * x, a1y = __builtin_fma(f1, y, -g);
  const double p2x = __builtin_fma(hh, vx, x), p2y = __builtin_fma(hh, vy, y);
  const double v2x = __builtin_fma(hh, a1x, vx), v2y = __builtin_fma(hh, a1y, vy);
  hex_eval_d(L.wsb, L.umax, L.ks, L.cu, L.kc2, L.invV, __builtin_fma(p2x, p2x, p2y * p2y), f2, d2, g2);
  const double a2x = f2 * p2x, a2y = __builtin_fma(f2, p2y, -g);
  const double p3x = __builtin_fma(hh, v2x, x), p3y = __builtin_fma(hh, v2y, y);
  const double v3x = __builtin_fma(hh, a2x, vx), v3y = __builtin_fma(hh, a2y, vy);
  hex_eval_d(L.wsb, L.umax, L.ks, L.cu, L.kc2, L.invV, __builtin_fma(p3x, p3x, p3y * p3y), f3, d3, g3);
  const double a3x = f3 * p3x, a3y = __builtin_fma(f3, p3y, -g);
  const double p4x = __builtin_fma(h, v3x, x), p4y = __builtin_fma(h, v3y, y);
  hex_eval_d(L.wsb, L.umax, L.ks, L.cu, L.kc2, L.invV, __builtin_fma(p4x, p4x, p4y * p4y), f4, d4, g4);

  // The reverse sweep below forms each sum of products with fused multiply-adds (it is checked
  // against the CPU adjoint to a tolerance, not bit for bit).
  // stage 4: a4 = A(p4), p4 = p1 + h v3, v4 = v1 + h a3
  const double ab4x = h6 * vbx, ab4y = h6 * vby;
  double q = __builtin_fma(p4x, ab4x, p4y * ab4y);
  double dq = d4 * q;
  const double pb4x = __builtin_fma(f4, ab4x, dq * p4x), pb4y = __builtin_fma(f4, ab4y, dq * p4y);
  Vbar = __builtin_fma(q, g4, Vbar);
  const double vb4x = h6 * pbx, vb4y = h6 * pby;
  double Px = pbx + pb4x, Py = pby + pb4y;
  const double vb3x = __builtin_fma(h3, pbx, h * pb4x), vb3y = __builtin_fma(h3, pby, h * pb4y);
  double Vx = vbx + vb4x, Vy = vby + vb4y;
  const double ab3x = __builtin_fma(h3, vbx, h * vb4x), ab3y = __builtin_fma(h3, vby, h * vb4y);
  // stage 3: a3 = A(p3), p3 = p1 + hh v2, v3 = v1 + hh a2
  q = __builtin_fma(p3x, ab3x, p3y * ab3y);
  dq = d3 * q;
  const double pb3x = __builtin_fma(f3, ab3x, dq * p3x), pb3y = __builtin_fma(f3, ab3y, dq * p3y);
  Vbar = __builtin_fma(q, g3, Vbar);
  Px = Px + pb3x;
  Py = Py + pb3y;
  const double vb2x = __builtin_fma(h3, pbx, hh * pb3x), vb2y = __builtin_fma(h3, pby, hh * pb3y);
  Vx = Vx + vb3x;
  Vy = Vy + vb3y;
  const double ab2x = __builtin_fma(h3, vbx, hh * vb3x), ab2y = __builtin_fma(h3, vby, hh * vb3y);
  // stage 2: a2 = A(p2), p2 = p1 + hh v1, v2 = v1 + hh a1
  q = __builtin_fma(p2x, ab2x, p2y * ab2y);
  dq = d2 * q;
  const double pb2x = __builtin_fma(f2, ab2x, dq * p2x), pb2y = __builtin_fma(f2, ab2y, dq * p2y);
  Vbar = __builtin_fma(q, g2, Vbar);
  Px = Px + pb2x;
  Py = Py + pb2y;
  Vx = __builtin_fma(h6, pbx, __builtin_fma(hh, pb2x, Vx) + vb2x);
  Vy = __builtin_fma(h6, pby, __builtin_fma(hh, pb2y, Vy) + vb2y);
  const double ab1x = __builtin_fma(h6, vbx, hh * vb2x), ab1y = __builtin_fma(h6, vby, hh * vb2y);
  // stage 1: a1 = A(p1)
  q = __builtin_fma(x, ab1x, y * ab1y);
  dq = d1 * q;
  Px = Px + __builtin_fma(f1, ab1x, dq * x);
  Py = Py + __builtin_fma(f1, ab1y, dq * y);
  Vbar = __builtin_fma(q, g1, Vbar);
  double hbar = 0.0;
  if constexpr (WANT_H) {
    const double a4x = f4 * p4x, a4y = __builtin_fma(f4, p4y, -g);
    const double v4x = __builtin_fma(h, a3x, vx), v4y = __builtin_fma(h, a3y, vy);
    const double Svx = ((vx + 2.0 * v2x) + 2.0 * v3x) + v4x, Svy = ((vy + 2.0 * v2y) + 2.0 * v3y) + v4y;
    const double Sax = ((a1x + 2.0 * a2x) + 2.0 * a3x) + a4x, Say = ((a1y + 2.0 * a2y) + 2.0 * a3y) + a4y;
    hbar = (((pbx * Svx + pby * Svy) + (vbx * Sax + vby * Say)) / 6.0 +
            0.5 * (((pb2x * vx + pb2y * vy) + (vb2x * a1x + vb2y * a1y)) +
                   ((pb3x * v2x + pb3y * v2y) + (vb3x * a2x + vb3y * a2y)))) +
           ((pb4x * v3x + pb4y * v3y) + (vb4x * a3x + vb4y * a3y));
  }
  pbx = Px;
  pby = Py;
  vbx = Vx;
  vby = Vy;
  return hbar;
}

// Lane j of the wavefront owns accumulators for parameters j, j + 64, j + 128.
__device__ __forceinline__ void acc_add(double &a0, double &a1, double &a2, int p, int lane, double v) {
  if ((p & 63) == lane) {
    const int b = p >> 6;
    if (b == 0) a0 += v;
    else if (b == 1) a1 += v;
    else a2 += v;
  }
}

__device__ __forceinline__ double *row(double *W, int slot, int lane) { return W + (size_t)slot * 256 + lane; }

__device__ __forceinline__ void put4(double *W, int slot, int lane, double a, double b, double c, double d) {
  double *r = row(W, slot, lane);
  r[0] = a;
  r[64] = b;
  r[128] = c;
  r[192] = d;
}

__device__ __forceinline__ void get4(const double *W, int slot, int lane, double &a, double &b, double &c, double &d) {
  const double *r = W + (size_t)slot * 256 + lane;
  a = r[0];
  b = r[64];
  c = r[128];
  d = r[192];
}

// z of the exit face of element e (CIRC: z1; LENS: its z0 parameter + L); z_src before the first.
__device__ __forceinline__ double exit_z(const KBeam &B, const double *P, int e, int pidx_e, double z_src) {
  if (e < 0) return z_src;
  const ctraj_elem &E = B.elems[e];
  if (E.kind == CTRAJ_LENS) return P[pidx_e + 1] + E.p[3];
  return E.z1;
}

__device__ __forceinline__ int n_params_of(const ctraj_elem &E) { return E.kind == CTRAJ_LENS ? 2 : 1; }

template <bool GRAD>
__global__ void __launch_bounds__(512) k_smooth(KBeam B, KTab KT, SmoothArgs A) {
  extern __shared__ __attribute__((aligned(16))) double lds[];
  const Tables T = stage_tables(lds, KT);
  __syncthreads();
  const WsRef wsb = ws_of(T, 0);
  const int lane = lane_id();
  const int64_t wpb = blockDim.x >> 6;
  const int64_t gwave = (int64_t)blockIdx.x * wpb + (threadIdx.x >> 6);
  const int64_t nwaves = (int64_t)gridDim.x * wpb;
  double *W = GRAD ? A.ws + gwave * A.ws_stride : nullptr;
  const double g = B.g;
  const int np = A.n_params;
  const int K = A.K;
  unsigned long long capped = 0;

  for (int64_t item = gwave; item < A.n_items; item += nwaves) {
    const int64_t cfg = item / A.items_per_cfg;
    const int64_t i = (item - cfg * A.items_per_cfg) * 64 + lane;
    const double *P = A.params + cfg * np;
    const double mu = P[np - 1];
    double x = 0.0, y = 0.0, vx = 0.0, vy = 0.0, vz = 1.0;
    bool ok = false;
    if (i < A.n) {
      x = A.x[i];
      y = A.y[i];
      vx = A.vx[i];
      vy = A.vy[i];
      vz = mu + A.sigma_vz * A.eps[i];
      ok = vz > 0.0;
      if (!ok) vz = 1.0;
    }
    double Jw = ok ? 1.0 : 0.0;  // product of the sigmoid factors
    double zc = A.z_src;
    int fslot = 0, cslot = 0, pidx = 0;

    // ---------------- forward ----------------
    for (int e = 0; e < B.n_elems; ++e) {
      const ctraj_elem &E = B.elems[e];
      if (E.kind == CTRAJ_CIRC) {
        const double r = P[pidx];
        const int nf = E.z1 > E.z0 ? 2 : 1;
        for (int f = 0; f < nf; ++f) {
          const double zf = f ? E.z1 : E.z0;
          if (ok) {
            drift_g(x, y, vx, vy, (zf - zc) / vz, g);
            if constexpr (GRAD) put4(W, fslot, lane, x, y, vx, vy);
            double phi, om;
            sigmoid2((r - __builtin_sqrt(x * x + y * y)) * A.inv_s, phi, om);
            Jw = Jw * phi;
          }
          zc = zf;
          ++fslot;
        }
        pidx += 1;
      } else {  // CTRAJ_LENS
        const double V = P[pidx], z0 = P[pidx + 1];
        const double scale = V / E.p[7];
        const double ks0 = E.p[4] * scale;
        const Force<CTRAJ_FORCE_WSPLINE_HEX> F{ks0, E.p[5] * scale, g, hex_folds(ks0) ? -g / ks0 : 0.0};
        const double h = E.p[2], Ll = E.p[3];
        const SmoothStep SK = smooth_step_k(F, h);
        if (!(scale >= 0.0)) {  // a negative or NaN field scale has no table: the set contributes 0
          ok = false;
          Jw = 0.0;
        }
        if (ok) {
          drift_g(x, y, vx, vy, (z0 - zc) / vz, g);
          if constexpr (GRAD) put4(W, fslot, lane, x, y, vx, vy);
          uint32_t st = 0;
          const int nfull = rk4_full_steps(Ll, vz, h, A.max_steps, st);
          capped += (st & CTRAJ_ST_CAPPED) ? 1u : 0u;
          const double hp = Ll / vz - (double)nfull * h;
          double r2 = __builtin_fma(x, x, y * y);
          for (int s = 0; s <= nfull; ++s) {
            if constexpr (GRAD) {
              if (s % K == 0) put4(W, A.off_ck + cslot + s / K, lane, x, y, vx, vy);
            }
            smooth_step(F, T, wsb, SK, s < nfull, hp, x, y, vx, vy, r2, st);
            r2 = __builtin_fma(x, x, y * y);
          }
        }
        zc = z0 + Ll;
        ++fslot;
        cslot += A.ck_per_lens;
        pidx += 2;
      }
    }

    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    if constexpr (GRAD) {
      // ---------------- reverse ----------------
      double pbx = 0.0, pby = 0.0, vbx = 0.0, vby = 0.0, vzbar = 0.0;
      double zexit_bar = 0.0;  // adjoint of the exit z of the element before the current one
      for (int e = B.n_elems - 1; e >= 0; --e) {
        const ctraj_elem &E = B.elems[e];
        pidx -= n_params_of(E);
        const int pprev = e > 0 ? pidx - n_params_of(B.elems[e - 1]) : 0;
        const double zprev = exit_z(B, P, e - 1, pprev, A.z_src);
        if (E.kind == CTRAJ_CIRC) {
          const double r = P[pidx];
          const int nf = E.z1 > E.z0 ? 2 : 1;
          double rbar = 0.0, zfrom_bar = 0.0;
          for (int f = nf - 1; f >= 0; --f) {
            --fslot;
            const double zf = f ? E.z1 : E.z0;
            const double zfrom = f ? E.z0 : zprev;
            if (ok) {
              double fx, fy, fvx, fvy;
              get4(W, fslot, lane, fx, fy, fvx, fvy);
              const double rho = __builtin_sqrt(fx * fx + fy * fy);
              double phi, om;
              sigmoid2((r - rho) * A.inv_s, phi, om);
              const double wbar = Jw * om;
              rbar = rbar + wbar * A.inv_s;
              if (rho > 0.0) {
                const double rb = -(wbar * A.inv_s) / rho;
                pbx = pbx + rb * fx;
                pby = pby + rb * fy;
              }
              const double tau = (zf - zfrom) / vz;
              const double taubar = drift_g_adj(fvx, fvy, g, tau, pbx, pby, vbx, vby);
              vzbar = vzbar - (taubar * tau) / vz;
              if (f == 0) zfrom_bar = -taubar / vz;
            }
          }
          acc_add(a0, a1, a2, pidx, lane, wave_sum(rbar));
          zexit_bar = zfrom_bar;
        } else {  // CTRAJ_LENS
          const double V = P[pidx], z0 = P[pidx + 1];
          const double scale = V / E.p[7];
          const double ks = E.p[4] * scale, cu = E.p[5] * scale;
          const LensK LK{wsb, T.umax, ks, cu, (2.0 * ks) * cu, 1.0 / V, g};
          const Force<CTRAJ_FORCE_WSPLINE_HEX> F{ks, cu, g, hex_folds(ks) ? -g / ks : 0.0};
          const double h = E.p[2], Ll = E.p[3];
          const SmoothStep SK = smooth_step_k(F, h);
          double Vbar = 0.0, z0bar = zexit_bar, zfrom_bar = 0.0;
          cslot -= A.ck_per_lens;
          --fslot;
          if (ok) {
            uint32_t st = 0;
            const int nfull = rk4_full_steps(Ll, vz, h, A.max_steps, st);
            const double hp = Ll / vz - (double)nfull * h;
            const int ns = nfull + 1;
            const int nseg = (ns + K - 1) / K;
            lds_double *rec = (lds_double *)(lds + A.rec_lds) + (threadIdx.x >> 6) * (kSub * 256) + lane;
            for (int sg = nseg - 1; sg >= 0; --sg) {
              const int s0 = sg * K, s1 = min(s0 + K, ns);
              const int nsub = (s1 - s0 + kSub - 1) / kSub;
              double x, y, vx, vy;
              get4(W, A.off_ck + cslot + sg, lane, x, y, vx, vy);
              double r2 = __builtin_fma(x, x, y * y);
              // the state every kSub steps of the segment (a one-window segment starts from the
              // checkpoint itself, still in registers)
              for (int q = 0; q < nsub; ++q) {
                if (nsub > 1) put4(W, A.off_sub + q, lane, x, y, vx, vy);
                if (q + 1 < nsub) {
                  for (int s = s0 + q * kSub; s < s0 + (q + 1) * kSub; ++s) {
                    smooth_step(F, T, wsb, SK, s < nfull, hp, x, y, vx, vy, r2, st);
                    r2 = __builtin_fma(x, x, y * y);
                  }
                }
              }
              // windows, last first: recompute into LDS, then reverse
              for (int q = nsub - 1; q >= 0; --q) {
                const int t0 = s0 + q * kSub, t1 = min(t0 + kSub, s1);
                if (nsub > 1) get4(W, A.off_sub + q, lane, x, y, vx, vy);
                r2 = __builtin_fma(x, x, y * y);
                for (int s = t0; s < t1 - 1; ++s) {
                  lds_double *r = rec + (s - t0) * 256;
                  r[0] = x;
                  r[64] = y;
                  r[128] = vx;
                  r[192] = vy;
                  smooth_step(F, T, wsb, SK, s < nfull, hp, x, y, vx, vy, r2, st);
                  r2 = __builtin_fma(x, x, y * y);
                }
                for (int s = t1 - 1; s >= t0; --s) {
                  if (s < t1 - 1) {
                    const lds_double *r = rec + (s - t0) * 256;
                    x = r[0];
                    y = r[64];
                    vx = r[128];
                    vy = r[192];
                  }
                  if (s == nfull) {
                    const double hbar = rk4_adjoint<true>(LK, hp, x, y, vx, vy, pbx, pby, vbx, vby, Vbar);
                    vzbar = vzbar - hbar * ((Ll / vz) / vz);  // hp = L/vz - nfull h
                  } else {
                    (void)rk4_adjoint<false>(LK, h, x, y, vx, vy, pbx, pby, vbx, vby, Vbar);
                  }
                }
              }
            }
            // drift from the previous exit to the lens entrance z0
            double ex, ey, evx, evy;
            get4(W, fslot, lane, ex, ey, evx, evy);
            const double tau = (z0 - zprev) / vz;
            const double taubar = drift_g_adj(evx, evy, g, tau, pbx, pby, vbx, vby);
            vzbar = vzbar - (taubar * tau) / vz;
            z0bar = z0bar + taubar / vz;
            zfrom_bar = -taubar / vz;
          }
          acc_add(a0, a1, a2, pidx, lane, wave_sum(Vbar));
          acc_add(a0, a1, a2, pidx + 1, lane, wave_sum(z0bar));
          zexit_bar = zfrom_bar;
        }
      }
      acc_add(a0, a1, a2, np - 1, lane, wave_sum(vzbar));  // d vz / d mu = 1
    }
    acc_add(a0, a1, a2, np, lane, wave_sum(Jw));
    double *out = A.partial + item * (int64_t)(np + 1);
    if (lane <= np) out[lane] = a0;
    if (64 + lane <= np) out[64 + lane] = a1;
    if (128 + lane <= np) out[128 + lane] = a2;
  }
  if (A.n_capped) {
    const double c = wave_sum((double)capped);
    if (lane == 0 && c > 0.0) atomicAdd(A.n_capped, (unsigned long long)c);
  }
}

// J[c] = (sum of item J sums) / n and grad[c][j] = (sum of item sums) / n, each over the items of
// parameter set c in a fixed order (strided per thread, then a fixed LDS tree).
__global__ void __launch_bounds__(256) k_smooth_reduce(const double *partial, int64_t items_per_cfg, int np, double inv_n,
                                                       double *J, double *grad) {
  __shared__ double red[256];
  const int64_t c = blockIdx.x;
  const double *base = partial + c * items_per_cfg * (int64_t)(np + 1);
  for (int j = 0; j <= np; ++j) {
    if (j < np && !grad) continue;
    double s = 0.0;
    for (int64_t it = threadIdx.x; it < items_per_cfg; it += blockDim.x) s += base[it * (np + 1) + j];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x >> 1; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      if (j == np) J[c] = red[0] * inv_n;
      else grad[c * np + j] = red[0] * inv_n;
    }
    __syncthreads();
  }
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------------
static int smooth_layout(ctraj_ctx *c, int *np, int *n_faces, int *n_lens, double *theta) {
  if (!c->have_beamline) return fail(c, CTRAJ_ESTATE, "no beamline set");
  int p = 0, nf = 0, nl = 0;
  for (const ctraj_elem &E : c->h_elems) {
    if (E.kind == CTRAJ_CIRC) {
      if (theta) theta[p] = E.p[0];
      p += 1;
      nf += E.z1 > E.z0 ? 2 : 1;
    } else if (E.kind == CTRAJ_LENS) {
      if (E.force != CTRAJ_FORCE_WSPLINE_HEX || E.integrator != CTRAJ_INTEG_RK4)
        return fail(c, CTRAJ_EINVAL, "smooth objective: lenses must use the hexapole W' spline and classical RK4");
      if (!(E.p[7] != 0.0)) return fail(c, CTRAJ_EINVAL, "smooth objective: lens V (p[7]) must be nonzero");
      if (theta) {
        theta[p] = E.p[7];
        theta[p + 1] = E.z0;
      }
      p += 2;
      nf += 1;
      nl += 1;
    } else {
      return fail(c, CTRAJ_EINVAL, "smooth objective: only circular apertures and lenses are supported");
    }
  }
  *np = p + 1;
  *n_faces = nf;
  *n_lens = nl;
  return CTRAJ_OK;
}

extern "C" {

int32_t ctraj_smooth_params(ctraj_ctx *c, int32_t *n_params, double *theta) {
  if (!c || !n_params) return fail(c, CTRAJ_EINVAL, "null argument");
  int np, nf, nl;
  const int rc = smooth_layout(c, &np, &nf, &nl, theta);
  if (rc != CTRAJ_OK) return rc;
  *n_params = np;
  return CTRAJ_OK;
}

int32_t ctraj_smooth_objective(ctraj_ctx *c, const ctraj_smooth *s) {
  if (!c || !s) return fail(c, CTRAJ_EINVAL, "null argument");
  HIPCHK(c, hipSetDevice(c->device));
  int np, nf, nl;
  int rc = smooth_layout(c, &np, &nf, &nl, nullptr);
  if (rc != CTRAJ_OK) return rc;
  if (s->n_params != np)
    return fail(c, CTRAJ_EINVAL, "n_params is " + std::to_string(s->n_params) + ", the beamline has " + std::to_string(np));
  if (np + 1 > 192) return fail(c, CTRAJ_EINVAL, "smooth objective: at most 191 parameters");
  if (s->n <= 0) return fail(c, CTRAJ_EINVAL, "n must be > 0");
  if (s->n_cfg < 1) return fail(c, CTRAJ_EINVAL, "n_cfg must be >= 1");
  if (!s->x || !s->y || !s->vx || !s->vy || !s->eps || !s->params || !s->J)
    return fail(c, CTRAJ_EINVAL, "source arrays, params and J are required");
  if (!(s->softness > 0.0)) return fail(c, CTRAJ_EINVAL, "softness must be > 0");
  if (s->ckpt_every < 1) return fail(c, CTRAJ_EINVAL, "ckpt_every must be >= 1");
  if (s->max_steps < 1) return fail(c, CTRAJ_EINVAL, "max_steps must be >= 1");
  const bool grad = s->grad != nullptr;
  // with gradients, 8 wavefronts per workgroup share the tables and keep 16 KB of LDS each
  const int block = s->block ? s->block : (grad ? 512 : 256);
  if (block % 64 || block < 64 || block > 512) return fail(c, CTRAJ_EINVAL, "block must be a multiple of 64 up to 512");
  if (nl > 0 && (!c->d_ws || c->ws_states < 1)) return fail(c, CTRAJ_ESTATE, "no W' spline set");

  KTab kt = ctx_ktab(c);
  kt.ws_states = nl > 0 ? 1 : 0;  // state 0 only
  kt.rt_total = 0;
  const size_t lds_tab = ktab_lds_bytes(kt);
  const size_t lds = lds_tab + (grad ? (size_t)(block / 64) * kSub * 256 * sizeof(double) : 0);
  const void *kern = grad ? (const void *)k_smooth<true> : (const void *)k_smooth<false>;
  if ((rc = lds_limit(c, kern, lds)) != CTRAJ_OK) return rc;
  if ((rc = block_ok(c, kern, block)) != CTRAJ_OK) return rc;

  SmoothArgs A{};
  A.x = s->x; A.y = s->y; A.vx = s->vx; A.vy = s->vy; A.eps = s->eps;
  A.params = s->params;
  A.n_capped = reinterpret_cast<unsigned long long *>(s->n_capped);
  A.n = s->n;
  A.items_per_cfg = (s->n + 63) / 64;
  A.n_items = A.items_per_cfg * s->n_cfg;
  A.z_src = s->z_src;
  A.sigma_vz = s->sigma_vz;
  A.inv_s = 1.0 / s->softness;
  A.n_params = np;
  A.K = s->ckpt_every;
  A.max_steps = s->max_steps;
  A.ck_per_lens = (s->max_steps + 1 + s->ckpt_every - 1) / s->ckpt_every;
  A.off_ck = nf;
  A.off_sub = nf + nl * A.ck_per_lens;
  A.ws_stride = grad ? (int64_t)(A.off_sub + (s->ckpt_every + kSub - 1) / kSub) * 256 : 0;
  A.rec_lds = (int)(lds_tab / sizeof(double));

  const int wpb = block / 64;
  int grid = grid_for(c, kern, block, lds, A.n_items * 64, 0);
  if (grad) {
    // keep the checkpoint/recompute store within 4 GiB
    const int64_t per_block = A.ws_stride * 8 * wpb;
    const int64_t max_blocks = std::max<int64_t>(1, ((int64_t)4 << 30) / per_block);
    grid = (int)std::min<int64_t>(grid, max_blocks);
    if ((rc = ensure_buf(c, c->sm_ws, (size_t)per_block * grid)) != CTRAJ_OK) return rc;
    A.ws = (double *)c->sm_ws.p;
  }
  if ((rc = ensure_buf(c, c->sm_part, (size_t)A.n_items * (np + 1) * sizeof(double))) != CTRAJ_OK) return rc;
  A.partial = (double *)c->sm_part.p;

  const KBeam B = make_kbeam(c->d_elems, c->n_elems, c->phys);
  if (grad)
    hipLaunchKernelGGL(k_smooth<true>, dim3(grid), dim3(block), lds, c->stream, B, kt, A);
  else
    hipLaunchKernelGGL(k_smooth<false>, dim3(grid), dim3(block), lds, c->stream, B, kt, A);
  HIPCHK(c, hipGetLastError());
  hipLaunchKernelGGL(k_smooth_reduce, dim3(s->n_cfg), dim3(256), 0, c->stream, (const double *)A.partial,
                     A.items_per_cfg, np, 1.0 / (double)s->n, s->J, s->grad);
  HIPCHK(c, hipGetLastError());
  return CTRAJ_OK;
}

}  // extern "C"
