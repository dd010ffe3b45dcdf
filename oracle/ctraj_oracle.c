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
 * plus the configuration-mode lens the reference does not have (hexapole |E| = cE r^2, cubic
 * spline of W(|E|), classical RK4 at fixed dt with a final partial step), whose operation order,
 * including every fused multiply-add, is the one libctraj.so documents.
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

/* classical RK4 on (x, y, vx, vy), running slope sums, h/6 formed as h*(1/6); same operation order
 * as the device */
static void rk4_step(const OForce *F, const OTab *T, int state, double h, double *x, double *y, double *vx,
                     double *vy, uint32_t *st) {
  const double hh = 0.5 * h, h6 = h * (1.0 / 6.0);
  double ax, ay;
  force_eval(F, T, state, *x, *y, &ax, &ay, st);
  double xs = fma(hh, *vx, *x), ys = fma(hh, *vy, *y);
  double ux = fma(hh, ax, *vx), uy = fma(hh, ay, *vy);
  double svx = ax, svy = ay;
  force_eval(F, T, state, xs, ys, &ax, &ay, st);
  double sx = fma(2.0, ux, *vx), sy = fma(2.0, uy, *vy);
  svx = fma(2.0, ax, svx);
  svy = fma(2.0, ay, svy);
  xs = fma(hh, ux, *x);
  ys = fma(hh, uy, *y);
  ux = fma(hh, ax, *vx);
  uy = fma(hh, ay, *vy);
  force_eval(F, T, state, xs, ys, &ax, &ay, st);
  sx = fma(2.0, ux, sx);
  sy = fma(2.0, uy, sy);
  svx = fma(2.0, ax, svx);
  svy = fma(2.0, ay, svy);
  xs = fma(h, ux, *x);
  ys = fma(h, uy, *y);
  ux = fma(h, ax, *vx);
  uy = fma(h, ay, *vy);
  force_eval(F, T, state, xs, ys, &ax, &ay, st);
  sx = sx + ux;
  sy = sy + uy;
  svx = svx + ax;
  svy = svy + ay;
  *x = fma(h6, sx, *x);
  *y = fma(h6, sy, *y);
  *vx = fma(h6, svx, *vx);
  *vy = fma(h6, svy, *vy);
}

/* configuration-mode lens interior: floor(L/(vz dt)) steps of dt, then L/vz - n dt; bore test
 * x^2 + y^2 > (d/2)^2 after each step; rows store the force at the row (exit row: (0,-g,0)) */
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
  double x = m->x, y = m->y, vx = m->vx, vy = m->vy, z = m->z, t = m->t;
  int alive = 1;
  for (int i = 0; i <= nfull; ++i) {
    if (i < nfull) {
      rk4_step(&F, T, state, h, &x, &y, &vx, &vy, st);
      z = z + dzs;
      t = t + h;
    } else {
      const double hp = L / m->vz - (double)nfull * h;
      rk4_step(&F, T, state, hp, &x, &y, &vx, &vy, st);
      z = E->z1;
      t = t + hp;
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

int orc_abi_version(void) { return CTRAJ_ABI_VERSION; }
