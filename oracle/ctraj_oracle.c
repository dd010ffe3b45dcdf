/*
 * ctraj_oracle.c -- CPU FP64 restatement of the reference's hot path.  TEST INFRASTRUCTURE ONLY:
 * it is loaded by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, always as
 * the checker, never as the thing measured or shipped.  The product (libctraj.so) never links it.
 *
 * What it restates (reference = otimgren/centrex-molecule-trajectories, src/trajectories/):
 *   molecule.py:26-68       Molecule.x(dt), v(dt), update_trajectory            -> orc_drift
 *   molecule.py:111-144     Trajectory rows                                     -> Rows
 *   beamline.py:20-35       Beamline.propagate_through (walk, break, "Detected") -> walk_one
 *   apertures.py:92-115     CircularAperture.propagate_through                   -> walk_one CIRC
 *   apertures.py:165-189    RectangularAperture.propagate_through                -> walk_one RECT
 *   apertures.py:227-270    FieldPlates.propagate_through                        -> walk_one PLATES
 *   electrostatic_lens.py:48-118  lens entrance / RK4 variant / exit hop         -> lens_ref
 *   electrostatic_lens.py:168-228 lens_acceleration with interp1d (np.interp)   -> rtable_eval
 *   post_processing.py:9-75 take_timestep + find_radial_pos_dist resampling     -> orc_zgrid_from_rows
 * plus two things the reference does not have: the configuration-mode lens (hexapole |E| = cE r^2,
 * cubic spline of W(|E|), classical RK4 at fixed dt with a final partial step), whose operation
 * order, including every fused multiply-add, is the one libctraj.so documents; and orc_smooth, the
 * sigmoid-smoothed detected flux with its reverse-mode gradient (the CPU FP64 adjoint).
 *
 * Arithmetic: compiled with -ffp-contract=off; fma() is used exactly where the device code uses
 * __builtin_fma, so results are comparable bit for bit.  One deliberate difference from numpy:
 * the reference's `delta_t**2` on a numpy scalar goes through libm pow(); this file (like the
 * device) uses dt*dt, the correctly rounded square (they differ by 1 ulp in ~0.1% of cases).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/ctraj.h"

typedef struct {
  double x, y, z, vx, vy, vz, t, ax, ay, az;
} Mol;

/* ---- trajectory rows (Trajectory.update, molecule.py:133-144) ---- */
typedef struct {
  double *buf; /* [cap][10] or NULL */
  int cap, count;
} Rows;

static void put_row(Rows *R, const Mol *m) {
  if (!R) return;
  if (R->buf && R->count < R->cap) {
    double *r = R->buf + (size_t)R->count * 10;
    r[0] = m->x; r[1] = m->y; r[2] = m->z;
    r[3] = m->vx; r[4] = m->vy; r[5] = m->vz;
    r[6] = m->ax; r[7] = m->ay; r[8] = m->az;
    r[9] = m->t;
  }
  R->count++;
}

/* ---- tables ---- */
typedef struct {
  const double *ws; /* [state][nseg][4] */
  int nseg;
  const double *rt_r[8], *rt_a[8];
  int rt_n[8];
} OTab;

/* np.interp at r on knots (rr, aa) -- the interp1d call of electrostatic_lens.py:217.  Outside the
 * table the reference raises ValueError (interp1d bounds_error); here the status bit is set and
 * the end value returned. */
static double rtable_eval(const OTab *T, int id, double r, uint32_t *status) {
  const double *rr = T->rt_r[id], *aa = T->rt_a[id];
  const int n = T->rt_n[id];
  if (r != r) return r;
  if (r < rr[0]) { *status |= CTRAJ_ST_OOT; return aa[0]; }
  if (r > rr[n - 1]) { *status |= CTRAJ_ST_OOT; return aa[n - 1]; }
  if (r == rr[n - 1]) return aa[n - 1];
  int lo = 0, hi = n - 1; /* invariant rr[lo] <= r < rr[hi] */
  while (hi - lo > 1) {
    int mid = (lo + hi) / 2;
    if (rr[mid] <= r) lo = mid; else hi = mid;
  }
  if (rr[lo] == r) return aa[lo];
  const double slope = (aa[lo + 1] - aa[lo]) / (rr[lo + 1] - rr[lo]);
  return slope * (r - rr[lo]) + aa[lo];
}

/* lens_acceleration (electrostatic_lens.py:215-228) with tabulated a_r(r) */
static void force_rtable(const OTab *T, int id, double g, double x, double y, double *ax, double *ay,
                         uint32_t *st) {
  const double r = sqrt(x * x + y * y);
  const double ar = rtable_eval(T, id, r, st);
  double fx = 0.0, fy = 0.0;
  if (r != 0.0) {
    fx = (ar * x) / r;
    fy = (ar * y) / r;
  }
  *ax = fx;
  *ay = fy - g;
}

/* W'(u) on the spline of one state: u clamped to the largest double below nseg, segment k = trunc(u),
 * s = u - k (exact), W' = (A s + B) s + C. */
static double wspline_eval(const OTab *T, int state, double u) {
  u = fmin(u, nextafter((double)T->nseg, 0.0));
  const int k = (int)u;
  const double s = u - (double)k;
  const double *c = T->ws + 4 * ((size_t)state * T->nseg + k);
  return fma(fma(c[0], s, c[1]), s, c[2]);
}

typedef struct {
  int kind; /* CTRAJ_FORCE_* */
  int id;   /* r table */
  double ks, cu, g;
} OForce;

static void force_eval(const OForce *F, const OTab *T, int state, double x, double y, double *ax, double *ay,
                       uint32_t *st) {
  if (F->kind == CTRAJ_FORCE_RTABLE) {
    force_rtable(T, F->id, F->g, x, y, ax, ay, st);
  } else if (F->kind == CTRAJ_FORCE_WSPLINE_HEX) {
    /* |E| = cE r^2: a = ks W'(cu r^2) (x, y) - g y_hat, ks = -2 cE/m */
    const double r2 = fma(x, x, y * y);
    const double f = wspline_eval(T, state, r2 * F->cu) * F->ks;
    *ax = f * x;
    *ay = fma(f, y, -F->g);
  } else {
    /* |E| = cE r: a = ks W'(cu r) (x, y)/r - g y_hat, ks = -cE/m */
    const double r = sqrt(x * x + y * y);
    double fx = 0.0, fy = 0.0;
    if (r != 0.0) {
      const double f = (wspline_eval(T, state, r * F->cu) * F->ks) / r;
      fx = f * x;
      fy = f * y;
    }
    *ax = fx;
    *ay = fy - F->g;
  }
}

/* Molecule.update_trajectory(delta_t) (molecule.py:58-68) */
static void orc_drift(Mol *m, double dt, double g) {
  if (dt != 0.0) { /* `if not delta_t` keeps x and v (molecule.py:31,41) */
    const double d2 = dt * dt;
    const double nx = (m->x + m->vx * dt) + (m->ax * d2) / 2.0;
    const double ny = (m->y + m->vy * dt) + (m->ay * d2) / 2.0;
    const double nz = (m->z + m->vz * dt) + (m->az * d2) / 2.0;
    m->vx = m->vx + m->ax * dt;
    m->vy = m->vy + m->ay * dt;
    m->vz = m->vz + m->az * dt;
    m->x = nx;
    m->y = ny;
    m->z = nz;
  }
  m->t = m->t + dt;
  m->ax = 0.0;
  m->ay = -g;
  m->az = 0.0;
}

static void edge_upd(double *e, double d) {
  d = fabs(d);
  if (d < *e) *e = d;
}

/* electrostatic_lens.py:86-118: N = rint(L/dz) steps of dt = dz/vz; RK4 variant */
static int lens_ref(const ctraj_elem *E, const OTab *T, int state, double g, Mol *m, uint32_t *st, uint64_t *steps,
                    Rows *R, double *edge) {
  OForce F = {E->force, E->table, E->p[4], E->p[5], g};
  const double half_d = E->p[0];
  const double dt = E->p[2] / m->vz;
  for (int i = 0; i < E->n_steps; ++i) {
    const double x = m->x, y = m->y, z = m->z;
    const double k1x = m->vx, k1y = m->vy, k1z = m->vz;
    double l1x, l1y, l2x, l2y, l3x, l3y, l4x, l4y;
    force_eval(&F, T, state, x, y, &l1x, &l1y, st);
    const double k2x = k1x + (dt * l1x) / 2.0, k2y = k1y + (dt * l1y) / 2.0, k2z = k1z + (dt * 0.0) / 2.0;
    force_eval(&F, T, state, x + dt * k1x, y + dt * k1y, &l2x, &l2y, st);
    const double k3x = k1x + (dt * l2x) / 2.0, k3y = k1y + (dt * l2y) / 2.0, k3z = k1z + (dt * 0.0) / 2.0;
    force_eval(&F, T, state, x + (dt * k2x) / 2.0, y + (dt * k2y) / 2.0, &l3x, &l3y, st);
    const double k4x = k1x + dt * l3x, k4y = k1y + dt * l3y, k4z = k1z + dt * 0.0;
    force_eval(&F, T, state, x + dt * k3x, y + dt * k3y, &l4x, &l4y, st);
    m->x = x + (dt * (((k1x + 2.0 * k2x) + 2.0 * k3x) + k4x)) / 6.0;
    m->y = y + (dt * (((k1y + 2.0 * k2y) + 2.0 * k3y) + k4y)) / 6.0;
    m->z = z + (dt * (((k1z + 2.0 * k2z) + 2.0 * k3z) + k4z)) / 6.0;
    m->vx = k1x + (dt * (((l1x + 2.0 * l2x) + 2.0 * l3x) + l4x)) / 6.0;
    m->vy = k1y + (dt * (((l1y + 2.0 * l2y) + 2.0 * l3y) + l4y)) / 6.0;
    m->vz = k1z + (dt * (((0.0 + 2.0 * 0.0) + 2.0 * 0.0) + 0.0)) / 6.0;
    m->t = m->t + dt;
    m->ax = l1x;
    m->ay = l1y;
    m->az = 0.0;
    (*steps)++;
    put_row(R, m);
    const double rho = sqrt(m->x * m->x + m->y * m->y);
    edge_upd(edge, rho - half_d);
    if (rho > half_d) return 0;
  }
  return 1;
}

/* classical RK4 on (x, y, vx, vy); the force does not depend on the velocity, so the stages expand
 * to x2 = x + h/2 v, x3 = x2 + (h/2)(h/2) a1, x4 = (x + h v) + h (h/2) a2,
 * x' = (x + h v) + (h/6) h (a1 + a2 + a3), v' = v + (h/6)(2 (a1 + a2 + a3) + (a4 - a1)); h/6 formed as
 * h*(1/6); the device's rk4_step (rk4_expanded with hex_k(1, h)), same operations in the same order */
static void rk4_step(const OForce *F, const OTab *T, int state, double h, double *x, double *y, double *vx,
                     double *vy, uint32_t *st) {
  const double hh = 0.5 * h, h6 = h * (1.0 / 6.0);
  const double c3 = hh * hh, c4 = h * hh, cx = h6 * h;
  double a1x, a1y, a2x, a2y, a3x, a3y, a4x, a4y;
  force_eval(F, T, state, *x, *y, &a1x, &a1y, st);
  const double x2 = fma(hh, *vx, *x), y2 = fma(hh, *vy, *y);
  force_eval(F, T, state, x2, y2, &a2x, &a2y, st);
  const double x3 = fma(c3, a1x, x2), y3 = fma(c3, a1y, y2);
  const double xv = fma(h, *vx, *x), yv = fma(h, *vy, *y);
  force_eval(F, T, state, x3, y3, &a3x, &a3y, st);
  const double x4 = fma(c4, a2x, xv), y4 = fma(c4, a2y, yv);
  double sumx = a1x + a2x, sumy = a1y + a2y;
  sumx = sumx + a3x;
  sumy = sumy + a3y;
  force_eval(F, T, state, x4, y4, &a4x, &a4y, st);
  *x = fma(cx, sumx, xv);
  *y = fma(cx, sumy, yv);
  *vx = fma(h6, fma(2.0, sumx, a4x - a1x), *vx);
  *vy = fma(h6, fma(2.0, sumy, a4y - a1y), *vy);
}

/* A hexapole lens folds its force scale into the step constants when |ks| lies in [2^-500, 2^500]:
 * a = ks W'(cu r^2) (x, y) - g y_hat = ks s with s = W' (x, y) + (0, gk), gk = -g/ks.  The force does
 * not depend on the velocity, so the classical RK4 stages on (x, v) expand to
 *   x2 = x + h/2 v,  x3 = x2 + (h/2)(ks h/2) s1,  x4 = (x + h v) + h (ks h/2) s2,
 *   x' = (x + h v) + (h/6)(ks h)(s1 + s2 + s3),  v' = v + (ks h/6)(2 (s1 + s2 + s3) + (s4 - s1))
 * (the device's hex_k and rk4_step_hex_k, same operations in the same order). */
static int hex_folds(double ks) {
  const double a = fabs(ks);
  return a >= 0x1p-500 && a <= 0x1p500;
}

static void slope_hex_k(const OTab *T, int state, double cu, double gk, double x, double y, double *sx, double *sy) {
  const double w = wspline_eval(T, state, fma(x, x, y * y) * cu);
  *sx = w * x;
  *sy = fma(w, y, gk);
}

static void rk4_step_hex_k(const OForce *F, const OTab *T, int state, double h, double *x, double *y, double *vx,
                           double *vy) {
  const double ks = F->ks;
  const double hh = 0.5 * h, h6 = h * (1.0 / 6.0), khh = ks * hh, kh = ks * h;
  const double c3 = hh * khh, c4 = h * khh, cx = h6 * kh, kh6 = ks * h6;
  const double gk = -F->g / ks, cu = F->cu;
  double s1x, s1y, s2x, s2y, s3x, s3y, s4x, s4y;
  slope_hex_k(T, state, cu, gk, *x, *y, &s1x, &s1y);
  const double x2 = fma(hh, *vx, *x), y2 = fma(hh, *vy, *y);
  slope_hex_k(T, state, cu, gk, x2, y2, &s2x, &s2y);
  const double x3 = fma(c3, s1x, x2), y3 = fma(c3, s1y, y2);
  const double xv = fma(h, *vx, *x), yv = fma(h, *vy, *y);
  slope_hex_k(T, state, cu, gk, x3, y3, &s3x, &s3y);
  const double x4 = fma(c4, s2x, xv), y4 = fma(c4, s2y, yv);
  double sumx = s1x + s2x, sumy = s1y + s2y;
  sumx = sumx + s3x;
  sumy = sumy + s3y;
  slope_hex_k(T, state, cu, gk, x4, y4, &s4x, &s4y);
  *x = fma(cx, sumx, xv);
  *y = fma(cx, sumy, yv);
  *vx = fma(kh6, fma(2.0, sumx, s4x - s1x), *vx);
  *vy = fma(kh6, fma(2.0, sumy, s4y - s1y), *vy);
}

/* one configuration-mode lens step: the folded hexapole step where the lens folds, else rk4_step */
static void lens_step(const OForce *F, const OTab *T, int state, double h, double *x, double *y, double *vx,
                      double *vy, uint32_t *st) {
  if (F->kind == CTRAJ_FORCE_WSPLINE_HEX && hex_folds(F->ks))
    rk4_step_hex_k(F, T, state, h, x, y, vx, vy);
  else
    rk4_step(F, T, state, h, x, y, vx, vy, st);
}

/* configuration-mode lens interior: floor(L/(vz dt)) steps of dt, then L/vz - n dt; bore test
 * x^2 + y^2 > (d/2)^2 after each step; rows store the force at the row (exit row: (0,-g,0));
 * after k full steps z = z_in + k (vz dt), t = t_in + k dt (one fma each), after the partial step
 * z = z1, t = (t_in + n dt) + hp */
static int lens_rk4(const ctraj_elem *E, const OTab *T, int state, double g, int cap, Mol *m, uint32_t *st,
                    uint64_t *steps, Rows *R, double *edge) {
  OForce F = {E->force, E->table, E->p[4], E->p[5], g};
  const double R2 = E->p[1], h = E->p[2], L = E->p[3];
  const double dzs = m->vz * h;
  int nfull;
  {
    const double q = floor(L / (m->vz * h));
    if (!(q >= 0.0)) nfull = 0;
    else if (q > (double)cap) { *st |= CTRAJ_ST_CAPPED; nfull = cap; }
    else nfull = (int)q;
  }
  const double z_in = m->z, t_in = m->t;
  double x = m->x, y = m->y, vx = m->vx, vy = m->vy, z = z_in, t = t_in;
  int alive = 1;
  for (int i = 0; i <= nfull; ++i) {
    if (i < nfull) {
      lens_step(&F, T, state, h, &x, &y, &vx, &vy, st);
      z = fma((double)(i + 1), dzs, z_in);
      t = fma((double)(i + 1), h, t_in);
    } else {
      const double hp = L / m->vz - (double)nfull * h;
      lens_step(&F, T, state, hp, &x, &y, &vx, &vy, st);
      z = E->z1;
      t = fma((double)nfull, h, t_in) + hp;
    }
    (*steps)++;
    const double r2 = fma(x, x, y * y);
    edge_upd(edge, sqrt(r2) - E->p[0]);
    if (R) {
      Mol r = *m;
      r.x = x; r.y = y; r.z = z; r.vx = vx; r.vy = vy; r.t = t; r.az = 0.0;
      if (i < nfull || r2 > R2) {
        uint32_t s2 = 0;
        force_eval(&F, T, state, x, y, &r.ax, &r.ay, &s2);
      } else {
        r.ax = 0.0;
        r.ay = -g;
      }
      put_row(R, &r);
    }
    if (r2 > R2) { alive = 0; break; }
  }
  m->x = x; m->y = y; m->vx = vx; m->vy = vy; m->z = z; m->t = t;
  if (alive) {
    m->ax = 0.0; m->ay = -g; m->az = 0.0;
  } else {
    uint32_t s2 = 0;
    force_eval(&F, T, state, x, y, &m->ax, &m->ay, &s2);
    m->az = 0.0;
  }
  return alive;
}

/* Beamline.propagate_through for one molecule; returns the label */
static int walk_one(const ctraj_elem *el, int n_el, const ctraj_phys *P, const OTab *T, int cap, int state, Mol *m,
                    uint32_t *st, uint64_t *faces, uint64_t *steps, Rows *R, double *edge) {
  const double g = P->g;
  put_row(R, m);
  for (int e = 0; e < n_el; ++e) {
    const ctraj_elem *E = &el[e];
    if (E->kind == CTRAJ_CIRC) {
      const double zz[2] = {E->z0, E->z1};
      for (int f = 0; f < 2; ++f) {
        orc_drift(m, (zz[f] - m->z) / m->vz, g);
        (*faces)++;
        put_row(R, m);
        const double rho = sqrt(m->x * m->x + m->y * m->y);
        edge_upd(edge, rho - E->p[0]);
        if (rho > E->p[0]) return E->label_hit;
      }
    } else if (E->kind == CTRAJ_RECT) {
      const double zz[2] = {E->z0, E->z1};
      for (int f = 0; f < 2; ++f) {
        orc_drift(m, (zz[f] - m->z) / m->vz, g);
        (*faces)++;
        put_row(R, m);
        edge_upd(edge, m->x - E->p[0]); edge_upd(edge, m->x - E->p[1]);
        edge_upd(edge, m->y - E->p[2]); edge_upd(edge, m->y - E->p[3]);
        if (!((E->p[0] < m->x && m->x < E->p[1]) && (E->p[2] < m->y && m->y < E->p[3]))) return E->label_hit;
      }
    } else if (E->kind == CTRAJ_PLATES) {
      const double x1 = E->p[0], x2 = E->p[1];
      orc_drift(m, (E->z0 - m->z) / m->vz, g);
      (*faces)++;
      put_row(R, m);
      edge_upd(edge, m->x - x1); edge_upd(edge, m->x - x2);
      if (!(x1 < m->x && m->x < x2)) return E->label_hit;
      double dt = (E->z1 - m->z) / m->vz;
      double xe = m->x;
      if (dt != 0.0) xe = (m->x + m->vx * dt) + (m->ax * (dt * dt)) / 2.0;
      edge_upd(edge, xe - x1); edge_upd(edge, xe - x2);
      const int inside = (x1 < xe && xe < x2);
      if (!inside) {
        if (m->vx < 0.0) dt = (x1 - m->x) / m->vx;
        else if (m->vx > 0.0) dt = (x2 - m->x) / m->vx;
      }
      orc_drift(m, dt, g);
      (*faces)++;
      put_row(R, m);
      if (!inside) return E->label_hit;
    } else if (E->kind == CTRAJ_LENS) {
      orc_drift(m, (E->z0 - m->z) / m->vz, g);
      (*faces)++;
      put_row(R, m);
      const double rho = sqrt(m->x * m->x + m->y * m->y);
      edge_upd(edge, rho - E->p[0]);
      if (rho > E->p[0]) return E->label_hit;
      if (E->integrator == CTRAJ_INTEG_REF) {
        if (!lens_ref(E, T, state, g, m, st, steps, R, edge)) return E->label_inside;
        orc_drift(m, (E->z1 - m->z) / m->vz, g); /* exit hop with the stored a = l1 */
        (*faces)++;
        put_row(R, m);
      } else {
        if (!lens_rk4(E, T, state, g, cap, m, st, steps, R, edge)) return E->label_inside;
      }
    }
  }
  return P->label_detected;
}

/*
 * orc_propagate: propagate n molecules.  in[7] = x,y,z,vx,vy,vz,t (z/t may be NULL: z_const,
 * t_const); a0[3] optional; outputs optional.  rows: [n][cap][10] with n_rows[n] (rows beyond the
 * molecule's count or the cap are left untouched).  edge: min |distance to a tested edge| in m.
 * work[0] += faces, work[1] += lens steps.  n_threads > 1 uses OpenMP.
 */
int orc_propagate(const ctraj_elem *el, int n_el, const ctraj_phys *P, const double *ws, int nseg,
                  const double *const *rt_r, const double *const *rt_a, const int32_t *rt_n, int64_t n,
                  const double *const *in, double z_const, double t_const, const double *const *a0,
                  const uint8_t *state, double *const *out, double *const *out_a, int32_t *label, uint8_t *status,
                  double *rows, int32_t cap, int32_t *n_rows, double *edge, uint64_t *work, int n_threads) {
  OTab T;
  memset(&T, 0, sizeof T);
  T.ws = ws;
  T.nseg = nseg;
  for (int j = 0; j < 8; ++j) {
    T.rt_r[j] = rt_r ? rt_r[j] : NULL;
    T.rt_a[j] = rt_a ? rt_a[j] : NULL;
    T.rt_n[j] = rt_n ? rt_n[j] : 0;
  }
  const int lens_cap = P->max_lens_steps > 0 ? P->max_lens_steps : (1 << 20);
  uint64_t tf = 0, ts = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : tf, ts) num_threads(n_threads > 0 ? n_threads : 1)
  for (int64_t i = 0; i < n; ++i) {
    Mol m;
    m.x = in[0][i];
    m.y = in[1][i];
    m.z = in[2] ? in[2][i] : z_const;
    m.vx = in[3][i];
    m.vy = in[4][i];
    m.vz = in[5][i];
    m.t = in[6] ? in[6][i] : t_const;
    m.ax = (a0 && a0[0]) ? a0[0][i] : 0.0;
    m.ay = (a0 && a0[1]) ? a0[1][i] : -P->g;
    m.az = (a0 && a0[2]) ? a0[2][i] : 0.0;
    const int st = state ? state[i] : 0;
    uint32_t s = 0;
    uint64_t f = 0, k = 0;
    double ed = INFINITY;
    Rows R = {rows ? rows + (size_t)i * cap * 10 : NULL, cap, 0};
    const int lab = walk_one(el, n_el, P, &T, lens_cap, st, &m, &s, &f, &k, (rows || n_rows) ? &R : NULL, &ed);
    if (out) {
      const double v[7] = {m.x, m.y, m.z, m.vx, m.vy, m.vz, m.t};
      for (int q = 0; q < 7; ++q)
        if (out[q]) out[q][i] = v[q];
    }
    if (out_a) {
      if (out_a[0]) out_a[0][i] = m.ax;
      if (out_a[1]) out_a[1][i] = m.ay;
      if (out_a[2]) out_a[2][i] = m.az;
    }
    if (label) label[i] = lab;
    if (status) status[i] = (uint8_t)s;
    if (n_rows) n_rows[i] = R.count;
    if (edge) edge[i] = ed;
    tf += f;
    ts += k;
  }
  if (work) {
    work[0] += tf;
    work[1] += ts;
  }
  return 0;
}

/*
 * find_radial_pos_dist / find_vel_dist (post_processing.py:20-140) for one molecule's rows, at
 * z_k = z_start + k*z_step, k < cap.  out[6][cap] = x, y, vx, vy, vz, t (t = t_n + dt); samples
 * past the last row (:42-43) are NaN.
 */
void orc_zgrid_from_rows(const double *rows, int n_rows, double z_start, double z_step, int cap, double *out) {
  const double *last = rows + (size_t)(n_rows - 1) * 10;
  for (int k = 0; k < cap; ++k) {
    const double zk = z_start + (double)k * z_step;
    for (int f = 0; f < 6; ++f) out[(size_t)f * cap + k] = NAN;
    if (last[2] < zk) continue;
    int exact = -1;
    for (int r = 0; r < n_rows; ++r)
      if (rows[(size_t)r * 10 + 2] == zk) { exact = r; break; }
    if (exact >= 0) {
      const double *p = rows + (size_t)exact * 10;
      out[0 * cap + k] = p[0]; out[1 * cap + k] = p[1];
      out[2 * cap + k] = p[3]; out[3 * cap + k] = p[4]; out[4 * cap + k] = p[5];
      out[5 * cap + k] = p[9];
      continue;
    }
    int first_ge = n_rows; /* (x[:,2] < z).tolist().index(False) */
    for (int r = 0; r < n_rows; ++r)
      if (!(rows[(size_t)r * 10 + 2] < zk)) { first_ge = r; break; }
    int nidx = first_ge - 1;
    if (nidx < 0) nidx += n_rows; /* python index -1 */
    const double *p = rows + (size_t)nidx * 10;
    const double dt = (zk - p[2]) / p[5];
    const double d2 = dt * dt;
    out[0 * cap + k] = (p[0] + p[3] * dt) + (p[6] * d2) / 2.0;
    out[1 * cap + k] = (p[1] + p[4] * dt) + (p[7] * d2) / 2.0;
    out[2 * cap + k] = p[3] + p[6] * dt;
    out[3 * cap + k] = p[4] + p[7] * dt;
    out[4 * cap + k] = p[5] + p[8] * dt;
    out[5 * cap + k] = p[9] + dt;
  }
}

/* ---- Philox4x32-10 and the Gaussian source ---- */
void orc_philox4x32_10(uint32_t *c, uint32_t k0, uint32_t k1) {
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
    c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
  }
}

/* out[d][i] = mean[d] + sigma[d] * z, z from the pair (i + offset, d/2) as libctraj draws it */
void orc_draw_gaussian(uint64_t seed, uint64_t offset, int64_t n, int nd, const double *mean, const double *sigma,
                       double *const *out, uint32_t *raw /* optional [n][ceil(nd/2)][4] */) {
  const int np = (nd + 1) / 2;
  for (int64_t i = 0; i < n; ++i) {
    for (int p = 0; p < np; ++p) {
      const uint64_t ii = (uint64_t)i + offset;
      uint32_t c[4] = {(uint32_t)ii, (uint32_t)(ii >> 32), (uint32_t)p, 0x43545241u};
      orc_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
      if (raw) memcpy(raw + ((size_t)i * np + p) * 4, c, 16);
      const uint64_t m1 = ((uint64_t)c[0] << 21) | (c[1] >> 11);
      const uint64_t m2 = ((uint64_t)c[2] << 21) | (c[3] >> 11);
      const double u1 = ((double)m1 + 0.5) * 0x1p-53;
      const double u2 = ((double)m2 + 0.5) * 0x1p-53;
      const double r = sqrt(-2.0 * log(u1));
      const double th = 6.283185307179586 * u2;
      const double z0 = r * cos(th), z1 = r * sin(th);
      if (out[2 * p]) out[2 * p][i] = mean[2 * p] + sigma[2 * p] * z0;
      if (2 * p + 1 < nd && out[2 * p + 1]) out[2 * p + 1][i] = mean[2 * p + 1] + sigma[2 * p + 1] * z1;
    }
  }
}

/* ctraj_draw_uniform_u8: state id floor(n_values * c0 / 2^32), counter (i_lo, i_hi, 'STAT', 'CTRA') */
void orc_draw_uniform_u8(uint64_t seed, uint64_t offset, int64_t n, int n_values, uint8_t *out) {
  for (int64_t i = 0; i < n; ++i) {
    const uint64_t k = (uint64_t)i + offset;
    uint32_t c[4] = {(uint32_t)k, (uint32_t)(k >> 32), 0x53544154u, 0x43545241u};
    orc_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    out[i] = (uint8_t)(((uint64_t)c[0] * (uint64_t)n_values) >> 32);
  }
}

/* ctraj_draw_disc: CeNTREXPositionDistribution.draw (distributions.py:112-119) from counter
 * (i_lo, i_hi, 'DISC', 'CTRA'): theta = 2 pi u1, r = radius sqrt(u2), uniforms as in orc_draw_gaussian */
void orc_draw_disc(uint64_t seed, uint64_t offset, int64_t n, double radius, double z, double *x, double *y,
                   double *zo) {
  for (int64_t i = 0; i < n; ++i) {
    const uint64_t k = (uint64_t)i + offset;
    uint32_t c[4] = {(uint32_t)k, (uint32_t)(k >> 32), 0x43534944u, 0x43545241u};
    orc_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const uint64_t m1 = ((uint64_t)c[0] << 21) | (c[1] >> 11);
    const uint64_t m2 = ((uint64_t)c[2] << 21) | (c[3] >> 11);
    const double u1 = ((double)m1 + 0.5) * 0x1p-53;
    const double u2 = ((double)m2 + 0.5) * 0x1p-53;
    const double th = 6.283185307179586 * u1;
    const double r = sqrt(u2) * radius;
    x[i] = r * cos(th);
    y[i] = r * sin(th);
    if (zo) zo[i] = z;
  }
}

int orc_abi_version(void) { return CTRAJ_ABI_VERSION; }

/*
 * Sigmoid-smoothed detected flux and its gradient (the differentiable objective of libctraj.so's
 * ctraj_smooth_objective; the reference has no counterpart).  For parameter set c and molecule i:
 * vz = mu + sigma_vz * eps_i (vz <= 0 contributes 0); walk the beamline without stopping; every
 * CIRC face (z0, and z1 when z1 > z0) multiplies by sigmoid((r - rho)/s); lenses: classical RK4
 * as rk4_step with ks, cu scaled by V / p[7] and z1 = z0 + L.  Parameters in beamline order: CIRC r;
 * LENS (V, z0); then mu.  J[c] = mean of the products; grad[c][j] = dJ/dtheta_j by reverse-mode
 * differentiation of exactly these operations, with every lens step input stored (the device
 * recomputes them from checkpoints; recomputation reproduces them bit for bit).
 * per_mol (optional, [n_cfg][n][n_params + 1]) receives each molecule's gradient terms and product.
 * Sums run over molecules in index order.
 */
static void sig2(double w, double *phi, double *om) {
  const double e = exp(-fabs(w));
  const double q = 1.0 / (1.0 + e);
  const double eq = e * q;
  *phi = w >= 0.0 ? q : eq;
  *om = w >= 0.0 ? eq : q;
}

static void drift_g(double *x, double *y, double vx, double *vy, double tau, double g) {
  *x = *x + vx * tau;
  *y = (*y + *vy * tau) + ((-g) * (tau * tau)) * 0.5;
  *vy = *vy + (-g) * tau;
}

typedef struct {
  const double *ws; /* state-0 spline [nseg][4] */
  double umax, ks, cu, kc2, invV, g;
} SLens;

static void hex_d(const SLens *L, double r2, double *f, double *d, double *gv) {
  const double u0 = r2 * L->cu;
  const double u = fmin(u0, L->umax);
  const int k = (int)u;
  const double s = u - (double)k;
  const double *c = L->ws + 4 * (size_t)k;
  const double P = fma(fma(c[0], s, c[1]), s, c[2]);
  double dP = fma(2.0 * c[0], s, c[1]);
  if (u0 > L->umax) dP = 0.0;
  *f = P * L->ks;
  *d = L->kc2 * dP;
  *gv = (*f + (0.5 * *d) * r2) * L->invV;
}

static double rk4_adj(const SLens *L, double h, double x, double y, double vx, double vy, double *pbx, double *pby,
                      double *vbx, double *vby, double *Vbar, int want_h) {
  const double hh = 0.5 * h, h6 = h * (1.0 / 6.0), h3 = 2.0 * h6, g = L->g;
  double f1, d1, g1, f2, d2, g2, f3, d3, g3, f4, d4, g4;
  hex_d(L, fma(x, x, y * y), &f1, &d1, &g1);
  const double a1x = f1 * x, a1y = fma(f1, y, -g);
  const double p2x = fma(hh, vx, x), p2y = fma(hh, vy, y);
  const double v2x = fma(hh, a1x, vx), v2y = fma(hh, a1y, vy);
  hex_d(L, fma(p2x, p2x, p2y * p2y), &f2, &d2, &g2);
  const double a2x = f2 * p2x, a2y = fma(f2, p2y, -g);
  const double p3x = fma(hh, v2x, x), p3y = fma(hh, v2y, y);
  const double v3x = fma(hh, a2x, vx), v3y = fma(hh, a2y, vy);
  hex_d(L, fma(p3x, p3x, p3y * p3y), &f3, &d3, &g3);
  const double a3x = f3 * p3x, a3y = fma(f3, p3y, -g);
  const double p4x = fma(h, v3x, x), p4y = fma(h, v3y, y);
  hex_d(L, fma(p4x, p4x, p4y * p4y), &f4, &d4, &g4);
  const double pbx0 = *pbx, pby0 = *pby, vbx0 = *vbx, vby0 = *vby;

  const double ab4x = h6 * vbx0, ab4y = h6 * vby0;
  double q = p4x * ab4x + p4y * ab4y;
  double dq = d4 * q;
  const double pb4x = f4 * ab4x + dq * p4x, pb4y = f4 * ab4y + dq * p4y;
  *Vbar = *Vbar + q * g4;
  const double vb4x = h6 * pbx0, vb4y = h6 * pby0;
  double Px = pbx0 + pb4x, Py = pby0 + pb4y;
  const double vb3x = h3 * pbx0 + h * pb4x, vb3y = h3 * pby0 + h * pb4y;
  double Vx = vbx0 + vb4x, Vy = vby0 + vb4y;
  const double ab3x = h3 * vbx0 + h * vb4x, ab3y = h3 * vby0 + h * vb4y;
  q = p3x * ab3x + p3y * ab3y;
  dq = d3 * q;
  const double pb3x = f3 * ab3x + dq * p3x, pb3y = f3 * ab3y + dq * p3y;
  *Vbar = *Vbar + q * g3;
  Px = Px + pb3x;
  Py = Py + pb3y;
  const double vb2x = h3 * pbx0 + hh * pb3x, vb2y = h3 * pby0 + hh * pb3y;
  Vx = Vx + vb3x;
  Vy = Vy + vb3y;
  const double ab2x = h3 * vbx0 + hh * vb3x, ab2y = h3 * vby0 + hh * vb3y;
  q = p2x * ab2x + p2y * ab2y;
  dq = d2 * q;
  const double pb2x = f2 * ab2x + dq * p2x, pb2y = f2 * ab2y + dq * p2y;
  *Vbar = *Vbar + q * g2;
  Px = Px + pb2x;
  Py = Py + pb2y;
  Vx = ((Vx + hh * pb2x) + vb2x) + h6 * pbx0;
  Vy = ((Vy + hh * pb2y) + vb2y) + h6 * pby0;
  const double ab1x = h6 * vbx0 + hh * vb2x, ab1y = h6 * vby0 + hh * vb2y;
  q = x * ab1x + y * ab1y;
  dq = d1 * q;
  Px = Px + (f1 * ab1x + dq * x);
  Py = Py + (f1 * ab1y + dq * y);
  *Vbar = *Vbar + q * g1;
  double hbar = 0.0;
  if (want_h) {
    const double a4x = f4 * p4x, a4y = fma(f4, p4y, -g);
    const double v4x = fma(h, a3x, vx), v4y = fma(h, a3y, vy);
    const double Svx = ((vx + 2.0 * v2x) + 2.0 * v3x) + v4x, Svy = ((vy + 2.0 * v2y) + 2.0 * v3y) + v4y;
    const double Sax = ((a1x + 2.0 * a2x) + 2.0 * a3x) + a4x, Say = ((a1y + 2.0 * a2y) + 2.0 * a3y) + a4y;
    hbar = (((pbx0 * Svx + pby0 * Svy) + (vbx0 * Sax + vby0 * Say)) / 6.0 +
            0.5 * (((pb2x * vx + pb2y * vy) + (vb2x * a1x + vb2y * a1y)) +
                   ((pb3x * v2x + pb3y * v2y) + (vb3x * a2x + vb3y * a2y)))) +
           ((pb4x * v3x + pb4y * v3y) + (vb4x * a3x + vb4y * a3y));
  }
  *pbx = Px;
  *pby = Py;
  *vbx = Vx;
  *vby = Vy;
  return hbar;
}

static double exit_z_of(const ctraj_elem *el, const double *P, int e, int pidx_e, double z_src) {
  if (e < 0) return z_src;
  if (el[e].kind == CTRAJ_LENS) return P[pidx_e + 1] + el[e].p[3];
  return el[e].z1;
}

/* one molecule; out[0..np-1] = gradient terms (if grad), out[np] = product.  Returns lens steps
 * capped (0/1 per lens). */
static int smooth_one(const ctraj_elem *el, int n_el, double g, const double *ws, int nseg, const double *P, int np,
                      double x, double y, double vx, double vy, double vz, double z_src, double inv_s, int max_steps,
                      int grad, double *fs /* [n_el*2][4] */, double *steps /* [max_steps+1][4] per lens */,
                      double *out) {
  for (int j = 0; j <= np; ++j) out[j] = 0.0;
  if (!(vz > 0.0)) return 0;
  const double umax = nextafter((double)nseg, 0.0);
  double Jw = 1.0, zc = z_src;
  int fslot = 0, pidx = 0, capped = 0, lens_i = 0;
  const int per_lens = 4 * (max_steps + 1);
  for (int e = 0; e < n_el; ++e) {
    const ctraj_elem *E = &el[e];
    if (E->kind == CTRAJ_CIRC) {
      const double r = P[pidx];
      const int nf = E->z1 > E->z0 ? 2 : 1;
      for (int f = 0; f < nf; ++f) {
        const double zf = f ? E->z1 : E->z0;
        drift_g(&x, &y, vx, &vy, (zf - zc) / vz, g);
        double *S = fs + 4 * fslot;
        S[0] = x; S[1] = y; S[2] = vx; S[3] = vy;
        double phi, om;
        sig2((r - sqrt(x * x + y * y)) * inv_s, &phi, &om);
        Jw = Jw * phi;
        zc = zf;
        ++fslot;
      }
      pidx += 1;
    } else {
      const double V = P[pidx], z0 = P[pidx + 1];
      const double scale = V / E->p[7];
      if (!(scale >= 0.0)) return capped; /* no table for a negative or NaN field scale: contributes 0 */
      OForce F = {CTRAJ_FORCE_WSPLINE_HEX, 0, E->p[4] * scale, E->p[5] * scale, g};
      OTab T;
      memset(&T, 0, sizeof T);
      T.ws = ws;
      T.nseg = nseg;
      const double h = E->p[2], L = E->p[3];
      drift_g(&x, &y, vx, &vy, (z0 - zc) / vz, g);
      double *S = fs + 4 * fslot;
      S[0] = x; S[1] = y; S[2] = vx; S[3] = vy;
      double q = floor(L / (vz * h));
      int nfull;
      if (!(q >= 0.0)) nfull = 0;
      else if (q > (double)max_steps) { nfull = max_steps; capped++; }
      else nfull = (int)q;
      const double hp = L / vz - (double)nfull * h;
      double *ST = steps + (size_t)lens_i * per_lens;
      uint32_t st = 0;
      for (int s = 0; s <= nfull; ++s) {
        ST[4 * s] = x; ST[4 * s + 1] = y; ST[4 * s + 2] = vx; ST[4 * s + 3] = vy;
        rk4_step(&F, &T, 0, s < nfull ? h : hp, &x, &y, &vx, &vy, &st);
      }
      zc = z0 + L;
      ++fslot;
      ++lens_i;
      pidx += 2;
    }
  }
  out[np] = Jw;
  if (!grad) return capped;

  double pbx = 0, pby = 0, vbx = 0, vby = 0, vzbar = 0, zexit_bar = 0;
  for (int e = n_el - 1; e >= 0; --e) {
    const ctraj_elem *E = &el[e];
    pidx -= E->kind == CTRAJ_LENS ? 2 : 1;
    const int pprev = e > 0 ? pidx - (el[e - 1].kind == CTRAJ_LENS ? 2 : 1) : 0;
    const double zprev = exit_z_of(el, P, e - 1, pprev, z_src);
    if (E->kind == CTRAJ_CIRC) {
      const double r = P[pidx];
      const int nf = E->z1 > E->z0 ? 2 : 1;
      double rbar = 0.0, zfrom_bar = 0.0;
      for (int f = nf - 1; f >= 0; --f) {
        --fslot;
        const double zf = f ? E->z1 : E->z0;
        const double zfrom = f ? E->z0 : zprev;
        const double *S = fs + 4 * fslot;
        const double rho = sqrt(S[0] * S[0] + S[1] * S[1]);
        double phi, om;
        sig2((r - rho) * inv_s, &phi, &om);
        const double wbar = Jw * om;
        rbar = rbar + wbar * inv_s;
        if (rho > 0.0) {
          const double rb = -(wbar * inv_s) / rho;
          pbx = pbx + rb * S[0];
          pby = pby + rb * S[1];
        }
        const double tau = (zf - zfrom) / vz;
        const double taubar = (pbx * S[2] + pby * S[3]) - g * vby;
        vbx = vbx + tau * pbx;
        vby = vby + tau * pby;
        vzbar = vzbar - (taubar * tau) / vz;
        if (f == 0) zfrom_bar = -taubar / vz;
      }
      out[pidx] += rbar;
      zexit_bar = zfrom_bar;
    } else {
      --lens_i;
      const double V = P[pidx], z0 = P[pidx + 1];
      const double scale = V / E->p[7];
      SLens LK = {ws, umax, E->p[4] * scale, E->p[5] * scale, 0, 1.0 / V, g};
      LK.kc2 = (2.0 * LK.ks) * LK.cu;
      const double h = E->p[2], L = E->p[3];
      double Vbar = 0.0, z0bar = zexit_bar;
      --fslot;
      double q = floor(L / (vz * h));
      int nfull;
      if (!(q >= 0.0)) nfull = 0;
      else if (q > (double)max_steps) nfull = max_steps;
      else nfull = (int)q;
      const double hp = L / vz - (double)nfull * h;
      const double *ST = steps + (size_t)lens_i * per_lens;
      for (int s = nfull; s >= 0; --s) {
        const double *I = ST + 4 * s;
        if (s == nfull) {
          const double hbar = rk4_adj(&LK, hp, I[0], I[1], I[2], I[3], &pbx, &pby, &vbx, &vby, &Vbar, 1);
          vzbar = vzbar - hbar * ((L / vz) / vz);
        } else {
          (void)rk4_adj(&LK, h, I[0], I[1], I[2], I[3], &pbx, &pby, &vbx, &vby, &Vbar, 0);
        }
      }
      const double *S = fs + 4 * fslot;
      const double tau = (z0 - zprev) / vz;
      const double taubar = (pbx * S[2] + pby * S[3]) - g * vby;
      vbx = vbx + tau * pbx;
      vby = vby + tau * pby;
      vzbar = vzbar - (taubar * tau) / vz;
      z0bar = z0bar + taubar / vz;
      out[pidx] += Vbar;
      out[pidx + 1] += z0bar;
      zexit_bar = -taubar / vz;
    }
  }
  out[np - 1] += vzbar;
  return capped;
}

int orc_smooth(const ctraj_elem *el, int n_el, double g, const double *ws, int nseg, int64_t n, const double *x,
               const double *y, const double *vx, const double *vy, const double *eps, double z_src, double sigma_vz,
               double softness, int n_cfg, int np, const double *params, int max_steps, double *J, double *grad,
               double *per_mol, int64_t *n_capped, int n_threads) {
  int n_lens = 0;
  for (int e = 0; e < n_el; ++e) {
    if (el[e].kind == CTRAJ_LENS) n_lens++;
    else if (el[e].kind != CTRAJ_CIRC) return -1;
  }
  const double inv_s = 1.0 / softness;
  int64_t capped_total = 0;
  for (int c = 0; c < n_cfg; ++c) {
    const double *P = params + (size_t)c * np;
    double *pm = per_mol ? per_mol + (size_t)c * n * (np + 1) : (double *)malloc((size_t)n * (np + 1) * sizeof(double));
    int64_t cap_c = 0;
#pragma omp parallel num_threads(n_threads > 0 ? n_threads : 1) reduction(+ : cap_c)
    {
      double *fs = (double *)malloc((size_t)(2 * n_el + 2) * 4 * sizeof(double));
      double *steps = (double *)malloc((size_t)(n_lens > 0 ? n_lens : 1) * 4 * (max_steps + 1) * sizeof(double));
#pragma omp for schedule(dynamic, 64)
      for (int64_t i = 0; i < n; ++i) {
        const double vz = P[np - 1] + sigma_vz * eps[i];
        cap_c += smooth_one(el, n_el, g, ws, nseg, P, np, x[i], y[i], vx[i], vy[i], vz, z_src, inv_s, max_steps,
                            grad != NULL, fs, steps, pm + (size_t)i * (np + 1));
      }
      free(fs);
      free(steps);
    }
    capped_total += cap_c;
    for (int j = 0; j <= np; ++j) {
      if (j < np && !grad) continue;
      double s = 0.0;
      for (int64_t i = 0; i < n; ++i) s += pm[(size_t)i * (np + 1) + j];
      if (j == np) J[c] = s / (double)n;
      else grad[(size_t)c * np + j] = s / (double)n;
    }
    if (!per_mol) free(pm);
  }
  if (n_capped) *n_capped = capped_total;
  return 0;
}
