/*
 * ctraj.h -- C ABI of libctraj.so, the MI355X (gfx950) FP64 molecular-trajectory engine.
 *
 * The reference (otimgren/centrex-molecule-trajectories) is pure Python with no FFI; its hot path is
 * TrajectorySimulator.run_simulation -> Beamline.propagate_through -> BeamlineElement.propagate_through
 * -> Molecule.update_trajectory / ElectrostaticLens.propagate_inside_lens.  Every entry point below
 * replaces one piece of that path and is bound from Python with ctypes (see the `trajectories`
 * package next to csrc/).  Conventions:
 *   - every function returns CTRAJ_OK (0) or a negative CTRAJ_E* code and never aborts;
 *     ctraj_last_error() gives the text of the last failure on that context;
 *   - every array argument is a DEVICE pointer unless its comment says "host";
 *   - all work is queued on the context's one HIP stream; ctraj_sync() waits for it;
 *   - one context per host thread; nothing is retained after a call returns except the
 *     beamline/tables set on the context and the batch objects.
 */
#ifndef CTRAJ_H
#define CTRAJ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CTRAJ_ABI_VERSION 1

/* return codes */
#define CTRAJ_OK 0
#define CTRAJ_EINVAL (-1) /* bad argument (null pointer, size, unsupported element combination) */
#define CTRAJ_EHIP (-2)   /* HIP runtime error; text in ctraj_last_error */
#define CTRAJ_ENOMEM (-3) /* device or pinned allocation failed */
#define CTRAJ_ESTATE (-4) /* call out of order (no beamline / tables set) */

#define CTRAJ_MAX_ELEMS 64
#define CTRAJ_MAX_LABELS 64
#define CTRAJ_MAX_STATES 4

/* ---- beamline elements (apertures.py:22-286, electrostatic_lens.py:23-125) ---- */
#define CTRAJ_CIRC 1   /* CircularAperture   apertures.py:84-122  p[0] = d/2                      */
#define CTRAJ_RECT 2   /* RectangularAperture apertures.py:148-196 p[0..3] = x1, x2, y1, y2        */
#define CTRAJ_PLATES 3 /* FieldPlates        apertures.py:214-277 p[0..1] = x1, x2                */
#define CTRAJ_LENS 4   /* ElectrostaticLens  electrostatic_lens.py:24-125, p[] below              */

/* lens force models */
#define CTRAJ_FORCE_RTABLE 0      /* radial acceleration a_r(r) from a linear table (electrostatic_lens.py:168-228) */
#define CTRAJ_FORCE_WSPLINE_HEX 1 /* a = -grad W(|E|)/m with |E| = cE r^2 (hexapole); W'(|E|) from a cubic spline */
#define CTRAJ_FORCE_WSPLINE_LIN 2 /* same, with |E| = cE r (the reference's linear field model) */

/* lens integrators */
#define CTRAJ_INTEG_REF 0 /* N = rint(L/dz) steps of dt = dz/vz, the reference's RK4 variant (electrostatic_lens.py:79-118) */
#define CTRAJ_INTEG_RK4 1 /* classical RK4 at fixed dt, then one partial step of L/vz - n*dt so the lens ends at z1 */

/*
 * LENS parameters:
 *   p[0] = d/2 (bore radius)      p[1] = (d/2)^2             p[2] = dz (REF) or dt (RK4)
 *   p[3] = L                      p[4] = ks force scale      p[5] = cu = cE/dE (|E| -> spline coordinate)
 *   p[6] = unused                 p[7] = V (electrode voltage; used by the adjoint for d/dV)
 * WSPLINE_HEX: a = ks * W'(E) * (x, y) with ks = -2 cE / m;  u = r^2 * cu
 * WSPLINE_LIN: a = ks * W'(E) * (x, y) / r with ks = -cE / m; u = r * cu
 */
typedef struct ctraj_elem {
  int32_t kind;         /* CTRAJ_CIRC .. CTRAJ_LENS */
  int32_t label_hit;    /* label of a molecule stopped here: element name, or "Lens entrance" for a lens */
  int32_t label_inside; /* LENS: label "Inside lens" */
  int32_t force;        /* LENS: CTRAJ_FORCE_* */
  int32_t integrator;   /* LENS: CTRAJ_INTEG_* */
  int32_t table;        /* LENS with FORCE_RTABLE: r-table id */
  int32_t n_steps;      /* LENS with INTEG_REF: int(rint(L/dz)) */
  int32_t reserved;
  double z0, z1;
  double p[8];
} ctraj_elem;

typedef struct ctraj_phys {
  double g;               /* gravitational acceleration along -y (reference: scipy.constants.g) */
  int32_t n_labels;       /* distinct labels, <= CTRAJ_MAX_LABELS */
  int32_t label_detected; /* label of survivors (beamline.py:34-35) */
  int32_t n_states;       /* W-spline state tables (1..CTRAJ_MAX_STATES); molecule state ids must be < n_states */
  int32_t max_lens_steps; /* safety cap on RK4 steps per lens per molecule (0 = 1<<20; at most 1<<30) */
} ctraj_phys;

/* structure-of-arrays phase space (device pointers) */
typedef struct ctraj_soa {
  double *x, *y, *z, *vx, *vy, *vz, *t;
} ctraj_soa;

/* per-molecule status bits */
#define CTRAJ_ST_OOT 1u    /* a lens force evaluation fell outside its table (reference: interp1d ValueError) */
#define CTRAJ_ST_CAPPED 2u /* RK4 step count clipped to max_lens_steps */

/* work counters (uint64, accumulated) */
#define CTRAJ_WORK_FACES 0  /* element faces reached (drift + test) */
#define CTRAJ_WORK_STEPS 1  /* lens integration steps taken */
#define CTRAJ_WORK_OOT 2    /* molecules with CTRAJ_ST_OOT */
#define CTRAJ_WORK_CAPPED 3 /* molecules with CTRAJ_ST_CAPPED */
#define CTRAJ_WORK_FACES_POST 4 /* of CTRAJ_WORK_FACES, the faces of the elements after the lens when a run is
                                 * split into kernels (0 for a single-kernel run) */
#define CTRAJ_WORK_N 8

typedef struct ctraj_ctx ctraj_ctx;
typedef struct ctraj_batch ctraj_batch;

/* ---- context, memory, timing ---- */
int32_t ctraj_abi_version(void);
int32_t ctraj_ctx_create(int32_t device, ctraj_ctx **out);
int32_t ctraj_ctx_destroy(ctraj_ctx *ctx);
const char *ctraj_last_error(const ctraj_ctx *ctx);
/* n_cu, max LDS bytes per block, warp size, gcn arch number (950 for gfx950) -- host outputs */
int32_t ctraj_device_info(ctraj_ctx *ctx, int32_t *n_cu, int32_t *lds_per_block, int32_t *wave, int32_t *arch);
int32_t ctraj_dev_alloc(ctraj_ctx *ctx, uint64_t bytes, void **ptr);
int32_t ctraj_dev_free(ctraj_ctx *ctx, void *ptr);
int32_t ctraj_host_alloc(ctraj_ctx *ctx, uint64_t bytes, void **ptr); /* pinned host memory */
int32_t ctraj_host_free(ctraj_ctx *ctx, void *ptr);
#define CTRAJ_H2D 1
#define CTRAJ_D2H 2
#define CTRAJ_D2D 3
int32_t ctraj_memcpy(ctraj_ctx *ctx, void *dst, const void *src, uint64_t bytes, int32_t kind); /* async on the stream */
int32_t ctraj_fill_u8(ctraj_ctx *ctx, void *dst, uint8_t value, uint64_t bytes);               /* fill kernel */
int32_t ctraj_fill_i64(ctraj_ctx *ctx, int64_t *dst, int64_t value, int64_t n);                 /* fill kernel */
int32_t ctraj_sync(ctraj_ctx *ctx);
/* HIP events on the context stream: record into slot (0..63), elapsed ms between two recorded slots */
int32_t ctraj_event_record(ctraj_ctx *ctx, int32_t slot);
int32_t ctraj_event_elapsed(ctraj_ctx *ctx, int32_t slot_a, int32_t slot_b, float *ms);

/* ---- beamline and lens tables ---- */
/* Beamline(elements) with its stable z0 sort already applied by the caller (beamline.py:17-18,44-45). host arrays */
int32_t ctraj_set_beamline(ctraj_ctx *ctx, const ctraj_elem *elems, int32_t n_elems, const ctraj_phys *phys);
/* linear a_r(r) table with np.interp semantics (the interp1d built at electrostatic_lens.py:194-209). host arrays;
 * r ascending, n >= 2; slopes are formed here as (a[j+1]-a[j])/(r[j+1]-r[j]) */
int32_t ctraj_set_rtable(ctraj_ctx *ctx, int32_t id, const double *r, const double *a, int32_t n);
/* W'(|E|) spline for each state: coef[state][seg][4] = {A, B, C, 0}, W'(E) = (A s + B) s + C with
 * s = u - seg, u = E/dE (host array) */
int32_t ctraj_set_wspline(ctraj_ctx *ctx, int32_t n_states, int32_t n_seg, const double *coef);

/* ---- forward propagation: replaces run_simulation's per-molecule loop (trajectory_simulator.py:52-78)
 *      and Beamline.propagate_through (beamline.py:20-38).  A beamline with exactly one lens runs as
 *      three kernels queued on the stream, per slice of at most 2^26 molecules: a streaming walk to
 *      and through the lens entrance that queues the molecules entering the lens (a context-owned
 *      queue of 64 bytes per molecule of the slice), a persistent lane-refill kernel that integrates
 *      the lens and writes each molecule's state where the lens left it back into its queue entry,
 *      and a streaming walk of the elements after the lens that records the lens molecules.  For a
 *      hexapole lens two small kernels between the first two list the queue longest predicted lens
 *      time first (5 more bytes per molecule), the order the lens kernel takes it in.
 *      CTRAJ_RUN_LANES, a WSPLINE_HEX lens whose force scale p[4] lies outside [2^-500, 2^500] in
 *      magnitude (V0 = 0, say: its RK4 step cannot fold ks into the step constants), or a
 *      reference-mode lens with stored accelerations requested use the single-kernel lane-refill
 *      walker; other beamlines, or CTRAJ_RUN_SIMPLE, run one molecule per thread.  All give
 *      bit-identical results. ---- */
#define CTRAJ_RUN_STATUS_OR 1 /* OR status bits into run->status instead of overwriting */
#define CTRAJ_RUN_SIMPLE 2    /* one molecule per thread start to finish (no lane refill) */
#define CTRAJ_RUN_LANES 4     /* one lens: the single-kernel lane-refill walker instead of the split kernels */
#define CTRAJ_RUN_PHASE_TIMES 8 /* time the launch's kernels with HIP events on the stream (ctraj_phase_times) */
#define CTRAJ_RUN_QUEUE_ORDER 16 /* split path: the lens kernel takes the queue in the order the walk filled it
                                  * instead of longest predicted lens time first (results are the same) */
typedef struct ctraj_run {
  int64_t n;             /* molecules */
  int64_t index_base;    /* global index of molecule 0 (for first_index / aoi_index in chunked runs) */
  ctraj_soa in;          /* initial phase space; in.z / in.t may be NULL, then z_const / t_const */
  double z_const, t_const;
  const double *a0[3];   /* initial stored acceleration; NULL -> (0, -g, 0) (molecule.py:15-16) */
  const uint8_t *state;  /* per-molecule state id, NULL -> 0 */
  ctraj_soa out;         /* final phase space; NULL fields are not written (may alias in) */
  double *out_a[3];      /* final stored acceleration (optional) */
  int32_t *label;        /* per-molecule label (optional) */
  uint8_t *status;       /* per-molecule CTRAJ_ST_* bits (optional) */
  int64_t *tally;        /* [n_states][n_labels] counts, accumulated (optional) */
  int64_t *first_index;  /* [n_labels] atomic min of the global index reaching each label (optional) */
  uint64_t aoi_mask;     /* labels whose molecules are compacted (apertures_of_interest) */
  int64_t aoi_cap;
  int64_t *aoi_count;    /* [1] accumulated; may exceed aoi_cap (then entries past the cap are dropped) */
  int64_t *aoi_index;    /* [aoi_cap] global indices of compacted molecules */
  int32_t *aoi_label;    /* [aoi_cap] their labels (optional) */
  ctraj_soa aoi_out;     /* [aoi_cap] their final phase space (optional fields) */
  uint64_t *work;        /* [CTRAJ_WORK_N] counters, accumulated (optional) */
  int32_t block;         /* threads per block, multiple of 64: up to 1024 for the split path's lens kernel and 768
                          * for the lane walker (one lens), else 256; 0 = 1024 for the split path's lens kernel,
                          * 768 for the lane walker when the lens tables exceed 40 KB of LDS, else 256 */
  int32_t grid;          /* blocks (0 = enough for n, capped at 8 per CU x n_cu x 8) */
  int32_t flags;         /* CTRAJ_RUN_* */
  int32_t refill;        /* lane refill: idle lanes per wavefront that trigger a refill (0 = 10 split path, 24 single kernel) */
} ctraj_run;
int32_t ctraj_propagate(ctraj_ctx *ctx, const ctraj_run *run);
/* ctraj_propagate for callers that cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid allocation).  No counterpart in the reference.  Same run, same results, bit for bit.
 * Every device buffer its kernels use beyond the caller's inputs and outputs -- the lens-entry queue, the queue's
 * claim and count words, the lens tally words and list of molecules leaving the lens, the lane walker's molecule
 * counter and its copy of the run -- lies in a library-owned
 * array of CTRAJ_NOALLOC_SCRATCH_BYTES per device, reserved when the library is loaded on a device and held until
 * the process ends; a beamline with a lens is processed in slices of at most as many molecules as that queue holds
 * (about a million).  No hipMalloc, hipFree or event creation happens in the call: CTRAJ_RUN_PHASE_TIMES, and a
 * run->grid larger than the scratch array's queue counts can take, are refused with CTRAJ_EINVAL.  The beamline and
 * tables must be set beforehand (those calls allocate).  Calls on one device share its array, which keeps what the
 * last call left there: calls that overlap on one device (several host threads or streams) race. */
#define CTRAJ_NOALLOC_SCRATCH_BYTES (64ull << 20)
int32_t ctraj_propagate_noalloc(ctraj_ctx *ctx, const ctraj_run *run);
/* Device times (ms) of the kernels of the launches made with CTRAJ_RUN_PHASE_TIMES since the last call, oldest
 * first: walk_ms[i] is the walk to and through the lens entrance of a run split into kernels, lens_ms[i] its lens
 * kernel and post_ms[i] its walk of the elements after the lens and recording of the lens molecules (walk and post
 * are 0 for a single-kernel run, and for a run of several slices -- more than 2^26 molecules --, whose whole time
 * is lens_ms[i]).  No counterpart in the reference (it has no device).  Waits for those launches;
 * writes at most max entries (the most recent ones; at most 64 are kept) to the host arrays, sets *count, and
 * forgets every pending measurement. */
int32_t ctraj_phase_times(ctraj_ctx *ctx, int32_t max, float *walk_ms, float *lens_ms, float *post_ms,
                          int32_t *count);

/* ---- saved trajectories (molecule.py:111-167 rows; post_processing.py:20-140 resampling) ---- */
#define CTRAJ_REC_ROWS 1  /* every trajectory row; buf = [10][cap][n]: x,y,z,vx,vy,vz,ax,ay,az,t */
#define CTRAJ_REC_ZGRID 2 /* samples at z_k = z_start + k*z_step; buf = [6][cap][n]: x,y,vx,vy,vz,t */
#define CTRAJ_REC_ZGRID_MOL 3 /* the same samples molecule-major: buf = [n][cap][6], each molecule contiguous */
typedef struct ctraj_record {
  int64_t n;              /* molecules to record */
  const int64_t *index;   /* indices into `in` (NULL -> 0..n-1) */
  ctraj_soa in;
  double z_const, t_const;
  const double *a0[3];
  const uint8_t *state;
  int32_t mode;           /* CTRAJ_REC_* */
  int32_t cap;            /* rows (ROWS) or samples (ZGRID) per molecule */
  double z_start, z_step; /* ZGRID */
  double *buf;            /* see mode; unwritten entries are NaN */
  int32_t *n_rows;        /* [n] rows produced (ROWS; may exceed cap) or samples written (ZGRID) */
  int32_t *label;         /* [n] final label (optional) */
  int32_t block;          /* threads per block, multiple of 64: up to 512 for beamlines of circular apertures and
                           * lenses, else 256 (0 = 256 for ZGRID_MOL, else 128) */
  int32_t reserved;
} ctraj_record;
int32_t ctraj_record_trajectories(ctraj_ctx *ctx, const ctraj_record *rec);

/* ---- stored trajectories at one z: find_radial_pos_dist / find_vel_dist (post_processing.py:20-140) ----
 * Molecule j owns rows [row_off[j], row_off[j+1]) of `rows`, each row 10 doubles x, y, z, vx, vy, vz, ax, ay, az,
 * t, in the order the reference appends them (molecule.py:133-144) and with its NaN rows dropped
 * (molecule.py:160-167).  out[f][j] for f = x, y, vx, vy, vz, t (a NULL out[f] is not written) is the molecule's
 * state at z by the reference's rule: valid[j] = 0 and NaN outputs when its last row lies before z (:42-43);
 * otherwise the first row with that z exactly (:52-56), else the row before the first row with !(z_row < z)
 * advanced to z by take_timestep with its stored acceleration (:61-72, :9-17), the last row when that is the
 * first row (Python's index -1).  row_off is checked on the host (one copy) before the launch. */
typedef struct ctraj_sample {
  int64_t n;              /* molecules */
  const int64_t *row_off; /* device [n + 1], non-decreasing, row_off[0] >= 0, row_off[n] <= n_rows */
  int64_t n_rows;         /* rows in `rows` */
  const double *rows;     /* device [n_rows][10] */
  double z;
  double *out[6];         /* device [n] each */
  uint8_t *valid;         /* device [n] (optional) */
} ctraj_sample;
int32_t ctraj_sample_rows(ctraj_ctx *ctx, const ctraj_sample *s);

/* ---- device source sampling: Distribution.draw (distributions.py:32-162) as Philox4x32-10 + Box-Muller ---- */
/* out[d][i] = mean[d] + sigma[d] * N(0,1) drawn from counter (i + offset, pair(d)) with key seed; host mean/sigma,
 * device out[d] (a NULL out[d] skips that dimension) */
int32_t ctraj_draw_gaussian(ctraj_ctx *ctx, uint64_t seed, uint64_t offset, int64_t n, int32_t n_dims,
                            const double *mean, const double *sigma, double *const *out);

/* out[i] = floor(n_values * u32 / 2^32) for molecule i (+ offset), u32 the first word of Philox4x32-10 with
 * counter (i_lo, i_hi, 'STAT' = 0x53544154, 'CTRA') and key (seed_lo, seed_hi): a uniform lens-state id per
 * molecule (exactly uniform when n_values divides 2^32).  1 <= n_values <= 256; out is device [n]. */
int32_t ctraj_draw_uniform_u8(ctraj_ctx *ctx, uint64_t seed, uint64_t offset, int64_t n, int32_t n_values,
                              uint8_t *out);

/* CeNTREXPositionDistribution.draw (distributions.py:112-119) on the device: molecule i (+ offset) at
 * (r cos theta, r sin theta, z) with theta = 2 pi u1 and r = radius sqrt(u2), u1 and u2 the 53-bit uniforms of
 * Philox4x32-10 counter (i_lo, i_hi, 'DISC' = 0x43534944, 'CTRA') and key (seed_lo, seed_hi), formed as in
 * ctraj_draw_gaussian.  x, y, z_out are device [n]; a NULL z_out skips z.  radius must be finite and >= 0. */
int32_t ctraj_draw_disc(ctraj_ctx *ctx, uint64_t seed, uint64_t offset, int64_t n, double radius, double z,
                        double *x, double *y, double *z_out);

/* index-list forms of the two draws: output slot j holds molecule index[j] (instead of j + offset) of the same
 * streams, so any subset of a large run can be regenerated on the device (device index[n], out as above) */
int32_t ctraj_draw_gaussian_at(ctraj_ctx *ctx, uint64_t seed, const int64_t *index, int64_t n, int32_t n_dims,
                               const double *mean, const double *sigma, double *const *out);
int32_t ctraj_draw_uniform_u8_at(ctraj_ctx *ctx, uint64_t seed, const int64_t *index, int64_t n, int32_t n_values,
                                 uint8_t *out);

/* ascending in-place radix sort of device keys [n] */
int32_t ctraj_sort_i64(ctraj_ctx *ctx, int64_t *keys, int64_t n);

/* ---- gather (device arrays): dst[f][j] = src[f][index[j] - index_base], j < n, f < n_fields (<= 8);
 * NULL dst[f] skips field f.  Every index[j] - index_base must lie in [0, src_len). ---- */
int32_t ctraj_gather_f64(ctraj_ctx *ctx, int64_t n, const int64_t *index, int64_t index_base, int64_t src_len,
                         int32_t n_fields, const double *const *src, double *const *dst);
int32_t ctraj_gather_u8(ctraj_ctx *ctx, int64_t n, const int64_t *index, int64_t index_base, int64_t src_len,
                        const uint8_t *src, uint8_t *dst);

/* ---- batched scan: many beamline configurations in one persistent, graph-replayed launch ----
 * Replaces running TrajectorySimulator.run_simulation (trajectory_simulator.py:34-103) once per
 * configuration.  Every configuration is a z0-sorted slice of one element array (at most one lens,
 * hexapole W' spline + classical RK4, lens state 0) and draws its own molecules on the device:
 * molecule j of configuration c is ctraj_draw_gaussian(seed_c, offset 0)'s molecule j over the six
 * dimensions (x, y, z, vx, vy, vz) with t = 0 and a = (0, -g, 0), so a batch equals per-configuration
 * ctraj_propagate runs over those draws, molecule for molecule.  Lanes of one wavefront may hold
 * molecules of different configurations; each molecule runs exactly the arithmetic of ctraj_propagate. */
typedef struct ctraj_batch_cfg {
  int32_t elem_begin, n_elems; /* slice of the batch element array (z0-sorted) */
  int64_t n;                   /* molecules in this configuration */
  uint64_t seed;               /* device source stream of this configuration */
  double mean[6], sigma[6];    /* Gaussian source for x, y, z, vx, vy, vz */
} ctraj_batch_cfg;
/* A batch with a lens runs as split kernels per slice of at most 2^27 molecules, like ctraj_propagate:
 * every molecule of the slice is drawn and walked to its lens entrance, the molecules entering a lens
 * are queued (64 B each, a batch-owned queue reused by every slice), persistent wavefronts integrate
 * the queue with lane refill and write each molecule back into its entry, and a streaming kernel walks
 * the molecules that left their lens through the rest of their configuration and tallies.  A lens
 * whose force scale does not fold (see ctraj_propagate) takes the plain RK4 step for its own molecules
 * only.  fused = 1, or a batch whose queue could not be allocated, runs one lane-refill kernel that
 * draws, walks and integrates. */
typedef struct ctraj_batch_tune {
  int32_t block;     /* threads per block of the lens kernel: a multiple of 64 up to 1024 (0 = 1024); with
                      * fused: 64, 128, 192 or 256 (0 = 256) */
  int32_t chunk;     /* fused: molecules a wavefront claims at a time from the global queue (0 = 128) */
  int32_t grid;      /* persistent blocks (0 = as many as are resident) */
  int32_t refill;    /* refill a wavefront when at least this many lanes are idle, 1..64 (0 = 16) */
  int32_t use_graph; /* 1: capture the launches (counter and tally resets, kernels) once as a hipGraph and replay */
  int32_t fused;     /* 1: the single-kernel walker instead of the split kernels */
} ctraj_batch_tune;
/* host arrays; the context's W' spline (ctraj_set_wspline, state 0) at creation is copied into the batch and
 * shared by all configurations */
int32_t ctraj_batch_create(ctraj_ctx *ctx, const ctraj_elem *elems, int32_t n_elems_total, const ctraj_batch_cfg *cfgs,
                           int32_t n_cfgs, const ctraj_phys *phys, ctraj_batch **out);
/* runs every configuration once; tally_host [n_cfgs][n_labels] and work_host [CTRAJ_WORK_N] are host outputs
 * (may be NULL); elapsed_ms (host, may be NULL) is the device time of the launch measured with HIP events */
int32_t ctraj_batch_run(ctraj_batch *b, const ctraj_batch_tune *tune, int64_t *tally_host, uint64_t *work_host,
                        float *elapsed_ms);
/* 1 if the batch runs the split kernels by default (fused = 0), 0 if it runs the single kernel */
int32_t ctraj_batch_split_path(const ctraj_batch *b);
int32_t ctraj_batch_destroy(ctraj_batch *b);

/* ---- differentiable detected flux and its adjoint (no counterpart in the reference: the objective
 * and gradients of the sigmoid-smoothed detected flux over the beamline of ctraj_set_beamline) ----
 * For parameter set c and molecule i: vz_i = mu_c + sigma_vz * eps_i (molecules with vz_i <= 0
 * contribute 0); the molecule starts at (x_i, y_i, z_src) with (vx_i, vy_i, vz_i) and is walked through
 * every element without being stopped.  Each CIRC face (z0, and z1 when z1 > z0) contributes the factor
 * sigmoid((r - rho) / softness), rho = sqrt(x^2 + y^2) on the face;  J_c = (1/n) sum_i prod factors.
 * Lens bores and entrances are not factors.  Parameters of set c, in beamline order: each CIRC its
 * radius r (its d/2 is ignored); each LENS (V, z0): field scaled by V / V_table (p[7], nonzero) and
 * z1 = z0 + L; last, mu.  Only CIRC and LENS (WSPLINE_HEX, INTEG_RK4, lens state 0) elements are
 * accepted.  The gradient is the exact derivative of the computed J (discrete adjoint of the drifts
 * and of every RK4 step, including the final partial step), obtained by reverse sweeps over lens
 * checkpoints taken every ckpt_every steps.  Results are deterministic: per-64-molecule partial sums
 * are combined in a fixed order. */
typedef struct ctraj_smooth {
  int64_t n;                            /* molecules */
  const double *x, *y, *vx, *vy, *eps;  /* device [n] */
  double z_src, sigma_vz, softness;     /* softness: sigmoid width in m (> 0) */
  int32_t n_cfg, n_params;              /* parameter sets in this launch; entries per set */
  const double *params;                 /* device [n_cfg][n_params] */
  double *J;                            /* device [n_cfg] */
  double *grad;                         /* device [n_cfg][n_params], or NULL: objective only */
  int64_t *n_capped;                    /* device [1] or NULL: += molecules whose lens step count was cut to max_steps */
  int32_t ckpt_every;                   /* K >= 1 */
  int32_t max_steps;                    /* full RK4 steps per lens at most (>= 1); bounds the checkpoint store */
  int32_t block;                        /* threads per block, multiple of 64 (0 = 256) */
  int32_t reserved;
} ctraj_smooth;
/* number of parameters of the current beamline (CIRC: 1, LENS: 2, plus mu) and, when theta is not NULL,
 * its first n_params - 1 entries as set in the element table (host array of at least n_params) */
int32_t ctraj_smooth_params(ctraj_ctx *ctx, int32_t *n_params, double *theta);
int32_t ctraj_smooth_objective(ctraj_ctx *ctx, const ctraj_smooth *s);

#ifdef __cplusplus
}
#endif
#endif /* CTRAJ_H */
