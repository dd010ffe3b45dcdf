// ctraj.hip -- kernels and C ABI of libctraj.so (gfx950).
//
// Kernels:
//   k_propagate        streaming beamline walker, one thread per molecule, grid-stride over a SoA
//                      store: the whole beamline (FULL), or the split path's walk to and through the
//                      lens entrance that queues lens molecules (PRE) and its walk after the lens
//                      (POST); wave-aggregated tallies and kept-molecule compaction.
//   k_ord_scan/scatter the split path's claim list, longest predicted lens time first (ctraj_order.hpp).
//   k_lens             the split path's lens integration: persistent wavefronts with lane refill.
//   k_propagate_lanes  the single-kernel lane-refill walker (one lens).
//   k_record           the walker re-run for selected molecules with a row / z-grid sink.
//   k_draw*            Philox4x32-10 device sources (Gaussian, uniform ids, uniform disc).
//   k_gather_*         kept-molecule bookkeeping.
// Elsewhere: the batched scan (batch.hip), the smoothed flux and its adjoint (adjoint.hip), the
// radix sort (sort.hip), the resampling of stored rows (resample.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "ctraj_device.hpp"
#include "ctraj_internal.hpp"
#include "ctraj_lanes.hpp"
#include "ctraj_order.hpp"
#include "ctraj_sinks.hpp"
#include "ctraj_wave.hpp"

using namespace ctraj;

// ---------------------------------------------------------------------------------------------
// k_propagate: one molecule per thread, grid-stride over the SoA store.  Each wavefront stages its
// kept (aoi) molecules in its own LDS slice and flushes them to the dense output with one global
// atomic per ~AOI_WAVE/2 molecules (no workgroup barrier in the loop).  The next molecule's inputs
// are loaded before the current one is walked, so HBM latency overlaps the walk.  A flush runs at
// the start of the iteration after the one that filled the slice, before that iteration's input
// loads: vector memory operations complete in order, so the wait for those loads at the end of the
// iteration would otherwise also wait for the flush's stores to reach memory.
// INTEG = CTRAJ_INTEG_NONE instantiates the walker without lens code (beamlines with no lens).
// ---------------------------------------------------------------------------------------------
constexpr int AOI_WAVE = 128;          // staged entries per wavefront
constexpr int AOI_ENTRY_BYTES = 8 + 4 + 7 * 8;

struct AoiStage {
  long long *idx;  // [AOI_WAVE]
  int *label;      // [AOI_WAVE]
  double *val;     // [7][AOI_WAVE]
};

// Copy a wavefront's staged entries [0, cnt) to the output (called by the whole wavefront).
__device__ __forceinline__ void aoi_flush_wave(const ctraj_run &R, const AoiStage &A, int cnt, int lane) {
  unsigned long long base = 0;
  if (lane == 0) base = atomicAdd((unsigned long long *)R.aoi_count, (unsigned long long)cnt);
  base = __shfl(base, 0);
  double *outs[7] = {R.aoi_out.x, R.aoi_out.y, R.aoi_out.z, R.aoi_out.vx, R.aoi_out.vy, R.aoi_out.vz, R.aoi_out.t};
  for (int j = lane; j < cnt; j += 64) {
    const int64_t slot = (int64_t)base + j;
    if (slot < R.aoi_cap) {
      R.aoi_index[slot] = A.idx[j];
      if (R.aoi_label) R.aoi_label[slot] = A.label[j];
#pragma unroll
      for (int f = 0; f < 7; ++f)
        if (outs[f]) outs[f][slot] = A.val[f * AOI_WAVE + j];
    }
  }
}

struct InMol {
  double x, y, z, vx, vy, vz, t, ax, ay, az;
  int st;
  int status;
  int64_t idx;
  int outcome;  // POST: LENS_* of the queue entry (LENS_QUEUED: an empty slot)
};

// ---- propagation of beamlines with one lens in three kernels ----
// PRE walks every molecule up to and through the lens entrance and queues the molecules that enter
// the lens, each wavefront into its own region of the queue; k_lens integrates the queue with lane
// refill and writes each molecule's state where the lens left it back into its queue entry, with
// the outcome (stopped by the bore, or left the lens at z1); POST walks the molecules that left the
// lens through the remaining elements and records every lens molecule.  A molecule is recorded
// (outputs, tallies, kept-molecule compaction) by PRE if it stops before the lens, else by POST --
// except, with lens tallies (PipeArgs::lt), a molecule stopped in a hexapole lens, which k_lens
// counts itself; POST then reads only the molecules k_lens listed as leaving the lens.
enum : int { PHASE_FULL = 0, PHASE_PRE = 1, PHASE_POST = 2 };

// Lens step constants: the half and sixth of the step (as rk4_step forms them) and, for a folding
// hexapole lens, ks times the half step, the step and the sixth, and gk = -g/ks (rk4_step_hex_k).
struct StepK {
  double gk;
  HexK hk;  // the folded hexapole step's constants, hex_k(ks, h)
  HexK h1;  // the plain step's, hex_k(1, h)
};

__host__ __device__ inline StepK step_k(double h, double ks, double g) {
  return StepK{hex_folds(ks) ? -g / ks : 0.0, hex_k(ks, h), hex_k(1.0, h)};
}

struct PipeArgs {
  longlong2 *q;                // lens-entry queue, 4 vectors per entry
  uint32_t *wcount;            // PRE: entries queued by each wavefront (a prefix of its region)
  int64_t region;              // queue entries owned by each PRE wavefront, a multiple of 64
  int n_pre_waves;
  unsigned long long *claim;   // lens kernel: next unclaimed 64-entry chunk of each region group, [8][32]
  int lens_e;
  // claim order (all null: the lens kernel claims the regions in order)
  uint8_t *bkt;                // PRE: lifetime bucket of each queue entry
  uint32_t *ord_hist;          // PRE: bucket counts of each PRE workgroup [block][ORD_BUCKETS]; k_ord_scan turns
                               // each bucket's column into exclusive prefixes over the workgroups
  uint32_t *ord_tot;           // [ORD_BUCKETS] entries per bucket, then the queue's total entry count
  uint32_t *perm;              // queue slots in claim order
  float ord_k2[CTRAJ_MAX_STATES];  // per state: the acceleration per unit radius at the W' table's end
  int ord_ref;                 // the lens integrates in the reference's steps of dz (else of dt)
  // lens tallies (null: the lens kernel writes every molecule back and POST reads every slot): the
  // lens kernel counts the molecules stopped in the lens into this block, which POST adds to the run's
  // tallies under the lens's "inside" label, and lists the slots of those that leave the lens, the
  // only ones POST then walks.  Words (zeroed before the lens kernel): [LT_N] listed slots,
  // [LT_STOP + s] stopped molecules of state s, [LT_OOT] / [LT_CAP] out-of-table / capped ones among
  // them, [LT_FIRST] ~(slice index) of the first of them (0: none); the list from word LT_LIST.
  uint32_t *lt;
};
enum : int { LT_N = 0, LT_STOP = 1, LT_OOT = LT_STOP + CTRAJ_MAX_STATES, LT_CAP, LT_FIRST, LT_LIST = 64 };
constexpr int POST_ITEMS = 4;  // POST with lens tallies: listed slots per thread

// Outcome of a lens molecule, written by k_lens into its queue entry's key (queue_put_done).
enum : int { LENS_QUEUED = 0, LENS_STOPPED = 1, LENS_LEFT = 2 };

__device__ __forceinline__ void load_mol(const ctraj_run &R, const KBeam &B, int64_t i, InMol &q) {
  q.x = R.in.x[i];
  q.y = R.in.y[i];
  q.z = R.in.z ? R.in.z[i] : R.z_const;
  q.vx = R.in.vx[i];
  q.vy = R.in.vy[i];
  q.vz = R.in.vz[i];
  q.t = R.in.t ? R.in.t[i] : R.t_const;
  q.ax = R.a0[0] ? R.a0[0][i] : 0.0;
  q.ay = R.a0[1] ? R.a0[1][i] : -B.g;
  q.az = R.a0[2] ? R.a0[2][i] : 0.0;
  q.st = R.state ? (int)R.state[i] : 0;  // clamped to the states present where it is used, not here:
                                         // a use right after the load would wait for the whole prefetch
  q.status = 0;
  q.idx = i;
}

// POST: the molecule of queue slot e as the lens kernel left it.
__device__ __forceinline__ void load_entry(const PipeArgs &P, int64_t e, InMol &q) {
  MState m;
  int64_t key;
  queue_get_done(P.q, e, m, key);
  q.x = m.x; q.y = m.y; q.z = m.z; q.vx = m.vx; q.vy = m.vy; q.vz = m.vz; q.t = m.t;
  q.idx = key & QKEY_INDEX;
  q.status = (int)((key >> 40) & 0xff);
  q.outcome = (int)((key >> 48) & 0xff);
  q.st = (int)((uint64_t)key >> 56);
}

// POST's item j: with lens tallies the j-th listed slot, else queue slot j (entry k of PRE wavefront
// w's region; slots past the region's count are empty).
__device__ __forceinline__ void load_slot(const PipeArgs &P, int64_t j, InMol &q) {
  q.outcome = LENS_QUEUED;
  if (P.lt) {
    load_entry(P, (int64_t)P.lt[LT_LIST + j], q);
    return;
  }
  const int64_t w = j / P.region;
  if (j - w * P.region < (int64_t)P.wcount[w]) load_entry(P, j, q);
}

// Inputs of a run that passes positions and velocities only (no time, stored acceleration or state
// arrays): every load is issued unconditionally.
__device__ __forceinline__ void load_min(const ctraj_run &R, int64_t i, InMol &q) {
  q.x = R.in.x[i];
  q.y = R.in.y[i];
  q.z = R.in.z[i];
  q.vx = R.in.vx[i];
  q.vy = R.in.vy[i];
  q.vz = R.in.vz[i];
  q.idx = i;
}

// MIN: the run passes positions and velocities only; the inputs are then prefetched two iterations
// ahead (and the constant time, acceleration and state are not carried per molecule).
template <int FORCE, int INTEG, int PHASE, int KM, bool OUT, bool MIN = false>
__global__ void __launch_bounds__(256) k_propagate(KBeam B, KTab KT, ctraj_run R, int aoi_lds_off, PipeArgs P) {
  // POST: the listed slots (lens tallies), else every slot of the queue
  const int64_t n_items =
      PHASE == PHASE_POST ? (P.lt ? (int64_t)P.lt[LT_N] : (int64_t)P.n_pre_waves * P.region) : R.n;
  // POST with lens tallies: the grid is sized for every slot of the queue, but the listed slots are
  // usually a few per cent of them; only as many workgroups as give each thread about POST_ITEMS
  // slots take part (the others leave at once; the first also merges the lens kernel's tallies
  // below), which keeps the same-address atomics of the workgroups' final flushes few
  int64_t n_wg = gridDim.x;
  if constexpr (PHASE == PHASE_POST) {
    if (P.lt) {
      const int64_t per_wg = (int64_t)blockDim.x * POST_ITEMS;
      n_wg = std::min<int64_t>(gridDim.x, std::max<int64_t>(1, (n_items + per_wg - 1) / per_wg));
      if (blockIdx.x >= n_wg) return;
    }
  }
  extern __shared__ __attribute__((aligned(16))) double lds[];
  __shared__ uint32_t s_hist[CTRAJ_MAX_STATES * CTRAJ_MAX_LABELS];
  __shared__ unsigned long long s_first[CTRAJ_MAX_LABELS];
  __shared__ unsigned long long s_work[5];
  __shared__ uint32_t s_ord[PHASE == PHASE_PRE ? ORD_BUCKETS : 1];  // PRE: this workgroup's lifetime buckets
  const int nbins = B.n_states * B.n_labels;
  for (int b = threadIdx.x; b < nbins; b += blockDim.x) s_hist[b] = 0;
  if constexpr (PHASE == PHASE_PRE)
    for (int b = threadIdx.x; b < ORD_BUCKETS; b += blockDim.x) s_ord[b] = 0;
  for (int b = threadIdx.x; b < B.n_labels; b += blockDim.x) s_first[b] = ~0ull >> 1;
  if (threadIdx.x < 5) s_work[threadIdx.x] = 0;
  Tables T{};
  if constexpr (INTEG != CTRAJ_INTEG_NONE && PHASE == PHASE_FULL) T = stage_tables(lds, KT);
  // POST stages the lens tables only when stored accelerations are written (the host passes empty
  // tables otherwise): a molecule stopped by the bore stores the force where it stopped
  if constexpr (PHASE == PHASE_POST && OUT) T = stage_tables(lds, KT);
  const int lane = lane_id();
  const int wave = threadIdx.x >> 6;
  AoiStage A;
  {
    char *base = reinterpret_cast<char *>(lds + aoi_lds_off) + (size_t)wave * AOI_WAVE * AOI_ENTRY_BYTES;
    A.idx = reinterpret_cast<long long *>(base);
    A.val = reinterpret_cast<double *>(base + AOI_WAVE * 8);
    A.label = reinterpret_cast<int *>(base + AOI_WAVE * 64);
  }
  int aoi_n = 0;  // entries staged by this wavefront (wave-uniform)
  bool aoi_full = false;  // the slice holds more than AOI_WAVE - 64 entries: flush before staging more
  __syncthreads();

  const int64_t stride = n_wg * blockDim.x;
  WalkCounters wc{0u, 0u};
  uint32_t n_oot = 0, n_cap = 0, faces_post = 0;
  const int64_t gwave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  int64_t qn = 0;  // PRE: entries this wavefront has queued
  // PRE: the lens entrance, read once with scalar loads (a vector load of it in the loop would wait
  // for the prefetched molecules and the queue stores in flight)
  double lens_z0 = 0.0, lens_r = 0.0;
  int lens_hit = 0;
  float ord_bore2 = 0.0f, ord_len = 0.0f, ord_h = 0.0f, ord_cap = 0.0f;  // lifetime prediction (P.bkt)
  if constexpr (PHASE == PHASE_PRE) {
    const ctraj_elem L = elem_at(B.elems, P.lens_e);
    lens_z0 = L.z0;
    lens_r = L.p[0];
    lens_hit = L.label_hit;
    ord_bore2 = (float)L.p[1];
    ord_h = (float)L.p[2];
    ord_len = (float)L.p[3];
    ord_cap = P.ord_ref ? (float)L.n_steps : (float)B.cap;
  }
  int lens_inside = 0;
  if constexpr (PHASE == PHASE_POST) {
    lens_inside = elem_at(B.elems, P.lens_e).label_inside;
    if (P.lt && blockIdx.x == 0 && threadIdx.x == 0) {
      // lens tallies: the molecules the lens kernel counted (stopped in the lens, not walked here)
      for (int s = 0; s < B.n_states; ++s)
        if (R.tally && P.lt[LT_STOP + s])
          atomicAdd((unsigned long long *)&R.tally[s * B.n_labels + lens_inside], (unsigned long long)P.lt[LT_STOP + s]);
      if (R.work && P.lt[LT_OOT]) atomicAdd((unsigned long long *)&R.work[2], (unsigned long long)P.lt[LT_OOT]);
      if (R.work && P.lt[LT_CAP]) atomicAdd((unsigned long long *)&R.work[3], (unsigned long long)P.lt[LT_CAP]);
      if (R.first_index && P.lt[LT_FIRST])
        atomicMin((unsigned long long *)&R.first_index[lens_inside],
                  (unsigned long long)(R.index_base + (int64_t)~P.lt[LT_FIRST]));
    }
  }
  const int64_t j0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t w0 = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63);
  // one grid-stride iteration: molecule j (inputs in q, loaded earlier), then q is refilled
  auto body = [&](InMol &q, const int64_t j, const int64_t wbase) {
    const bool active = j < n_items && (PHASE != PHASE_POST || q.outcome != LENS_QUEUED);
    MState m;
    m.x = q.x; m.y = q.y; m.z = q.z; m.vx = q.vx; m.vy = q.vy; m.vz = q.vz;
    int st = 0;
    uint32_t status = 0;
    if constexpr (MIN) {  // the values load_mol gives absent arrays
      m.t = R.t_const;
      m.ax = 0.0; m.ay = -B.g; m.az = 0.0;
    } else {
      m.t = q.t;
      m.ax = q.ax; m.ay = q.ay; m.az = q.az;
      st = min(q.st, B.n_states - 1);
      status = (uint32_t)q.status;
    }
    const int64_t i = q.idx;
    if (aoi_full) {  // wave-uniform
      aoi_flush_wave(R, A, aoi_n, lane);
      aoi_n = 0;
      aoi_full = false;
    }
    const int outcome = q.outcome;
    if constexpr (MIN) {  // the molecule two iterations ahead; unconditional (the last index repeats at the end)
      load_min(R, j + 2 * stride < n_items ? j + 2 * stride : n_items - 1, q);
    } else if (PHASE == PHASE_POST) {
      if (j + stride < n_items) load_slot(P, j + stride, q);
      else q.outcome = LENS_QUEUED;
    } else if (j + stride < n_items) {
      load_mol(R, B, j + stride, q);  // prefetch the next molecule
    }
    int label = 0;
    bool enq = false;
    if (active) {
      if constexpr (PHASE == PHASE_FULL) {
        NullSink sk;
        walk<FORCE, INTEG, NullSink, KM, true>(B.elems, B.n_elems, T, B.g, B.cap, st, B.label_detected, m, label, status,
                                         wc, sk);
      } else if constexpr (PHASE == PHASE_PRE) {
        label = B.label_detected;
        if (walk_plain<KM, true>(B.elems, 0, P.lens_e, B.g, m, label, wc)) {
          drift_dt(m, time_to(m, lens_z0), B.g);  // lens entrance (electrostatic_lens.py:59-64)
          ++wc.faces;
          if (rho_of(m.x, m.y) > lens_r) label = lens_hit;
          else enq = true;
        }
      } else if constexpr (PHASE == PHASE_POST) {
        if (outcome == LENS_LEFT) {
          // the row at z1 stores (0, -g, 0) (configuration mode); the reference lens's exit hop was made
          // by k_lens with its stored a = l1 and ended with the same row
          m.ax = 0.0; m.ay = -B.g; m.az = 0.0;
          label = B.label_detected;
          const uint32_t f0 = wc.faces;
          walk_plain<KM, true>(B.elems, P.lens_e + 1, B.n_elems, B.g, m, label, wc);
          faces_post += wc.faces - f0;
        } else {
          label = lens_inside;
          if constexpr (OUT && INTEG == CTRAJ_INTEG_RK4) {
            // stopped by the bore: the row stores the force at the stopping point
            if (R.out_a[0] || R.out_a[1] || R.out_a[2]) {
              const Force<FORCE> F = make_force<FORCE>(elem_at(B.elems, P.lens_e), B.g);
              uint32_t st2 = 0;
              F(T, ws_of(T, st), m.x, m.y, m.ax, m.ay, st2);
              m.az = 0.0;
            }
          }
        }
      }
    }
    if constexpr (PHASE == PHASE_PRE) {
      // wavefront-private region of the queue: no atomics
      const uint64_t qm = __ballot(enq);
      if (enq) {
        const int64_t e = gwave * P.region + qn + lanes_below(qm);
        queue_put(P.q, e, m, i, st);
        if (P.bkt) {
          const int b = ord_bucket(P.ord_k2[st], (float)m.x, (float)m.y, (float)m.vx, (float)m.vy, (float)m.vz,
                                   ord_bore2, ord_len, ord_h, ord_cap, P.ord_ref);
          P.bkt[e] = (uint8_t)b;
          atomicAdd(&s_ord[b], 1u);
        }
      }
      qn += __popcll(qm);
    }
    if (active && !enq) {
      if constexpr (OUT) {  // per-molecule outputs (an instantiation without them keeps fewer pointers live)
        if (R.out.x) R.out.x[i] = m.x;
        if (R.out.y) R.out.y[i] = m.y;
        if (R.out.z) R.out.z[i] = m.z;
        if (R.out.vx) R.out.vx[i] = m.vx;
        if (R.out.vy) R.out.vy[i] = m.vy;
        if (R.out.vz) R.out.vz[i] = m.vz;
        if (R.out.t) R.out.t[i] = m.t;
        if (R.out_a[0]) R.out_a[0][i] = m.ax;
        if (R.out_a[1]) R.out_a[1][i] = m.ay;
        if (R.out_a[2]) R.out_a[2][i] = m.az;
        if (R.label) R.label[i] = label;
        if (R.status)
          R.status[i] = (R.flags & CTRAJ_RUN_STATUS_OR) ? (uint8_t)(R.status[i] | status) : (uint8_t)status;
      }
      n_oot += (status & CTRAJ_ST_OOT) ? 1u : 0u;
      n_cap += (status & CTRAJ_ST_CAPPED) ? 1u : 0u;
      wave_hist(s_hist, R.first_index ? s_first : nullptr, st * B.n_labels + label, label, R.index_base + i);
    }
    if (R.aoi_mask) {
      const bool keep = active && !enq && ((R.aoi_mask >> label) & 1ull);
      const uint64_t mask = __ballot(keep);
      if (mask) {
        if (keep) {
          const int slot = aoi_n + lanes_below(mask);
          A.idx[slot] = R.index_base + i;
          A.label[slot] = label;
          A.val[0 * AOI_WAVE + slot] = m.x;
          A.val[1 * AOI_WAVE + slot] = m.y;
          A.val[2 * AOI_WAVE + slot] = m.z;
          A.val[3 * AOI_WAVE + slot] = m.vx;
          A.val[4 * AOI_WAVE + slot] = m.vy;
          A.val[5 * AOI_WAVE + slot] = m.vz;
          A.val[6 * AOI_WAVE + slot] = m.t;
        }
        aoi_n += __popcll(mask);
        aoi_full = aoi_n > AOI_WAVE - 64;
      }
    }
    };
  // wave-uniform trip count: every lane of a wavefront runs the same iterations
  if constexpr (MIN) {
    // two input buffers, each consumed two iterations after its loads were issued; the loads are
    // unconditional, so the compiler can count them and wait for the older buffer only
    InMol qa, qb;
    load_min(R, j0 < n_items ? j0 : n_items - 1, qa);
    load_min(R, j0 + stride < n_items ? j0 + stride : n_items - 1, qb);
    for (int64_t wbase = w0, j = j0; wbase < n_items; wbase += 2 * stride, j += 2 * stride) {
      body(qa, j, wbase);
      if (wbase + stride >= n_items) break;
      body(qb, j + stride, wbase + stride);
    }
  } else {
    InMol q;
    q.outcome = LENS_QUEUED;
    if constexpr (PHASE == PHASE_POST) {
      if (j0 < n_items) load_slot(P, j0, q);
    } else {
      if (j0 < n_items) load_mol(R, B, j0, q);
    }
    for (int64_t wbase = w0, j = j0; wbase < n_items; wbase += stride, j += stride) body(q, j, wbase);
  }
  if (R.aoi_mask && aoi_n > 0) aoi_flush_wave(R, A, aoi_n, lane);
  if constexpr (PHASE == PHASE_PRE) {
    if (lane == 0) P.wcount[gwave] = (uint32_t)qn;
  }
  if (R.work) {
    if (wc.faces) atomicAdd(&s_work[0], (unsigned long long)wc.faces);
    if (wc.steps) atomicAdd(&s_work[1], (unsigned long long)wc.steps);
    if (n_oot) atomicAdd(&s_work[2], (unsigned long long)n_oot);
    if (n_cap) atomicAdd(&s_work[3], (unsigned long long)n_cap);
    if (faces_post) atomicAdd(&s_work[4], (unsigned long long)faces_post);
  }
  __syncthreads();
  if (R.tally)
    for (int b = threadIdx.x; b < nbins; b += blockDim.x)
      if (s_hist[b]) atomicAdd((unsigned long long *)&R.tally[b], (unsigned long long)s_hist[b]);
  if (R.first_index)  // the stored minimum only decreases: a read of it that is not above ours spares the atomic
    for (int b = threadIdx.x; b < B.n_labels; b += blockDim.x)
      if (s_first[b] != (~0ull >> 1) && (int64_t)s_first[b] < __atomic_load_n(&R.first_index[b], __ATOMIC_RELAXED))
        atomicMin((unsigned long long *)&R.first_index[b], s_first[b]);
  if (R.work && threadIdx.x < 5 && s_work[threadIdx.x])
    atomicAdd((unsigned long long *)&R.work[threadIdx.x], s_work[threadIdx.x]);
  if constexpr (PHASE == PHASE_PRE) {
    if (P.bkt)  // every workgroup writes its whole row
      for (int b = threadIdx.x; b < ORD_BUCKETS; b += blockDim.x)
        P.ord_hist[(size_t)blockIdx.x * ORD_BUCKETS + b] = s_ord[b];
  }
}

// ---------------------------------------------------------------------------------------------
// k_propagate_lanes: persistent lane-refill walker (beamlines with exactly one lens)
// ---------------------------------------------------------------------------------------------
struct LaneArgs {
  unsigned long long *counter;  // next unclaimed molecule (zeroed before launch)
  int lens_e;                   // index of the lens in the element table
  int refill_min;               // idle lanes that trigger a refill section
  int claim;                    // molecules claimed per atomic
  double gk;                    // -g/ks of a hexapole lens that folds its force scale (FOLD)
};

// Record the molecules of lanes with s.fin set (convergent call).  A molecule stopped by the bore
// gets its stored acceleration (the force at the stopping point) here, only if it is asked for.
template <int FORCE>
__device__ __forceinline__ void record_finished(const KBeam &B, const ctraj_run &R, LaneMol &s, uint32_t *s_hist,
                                                unsigned long long *s_first, int lane, uint32_t &n_oot,
                                                uint32_t &n_cap, const Force<FORCE> &F, const Tables &T) {
  if (!__ballot(s.fin)) return;
  if (s.fin) {
    const int64_t i = s.gidx;
    MState &m = s.m;
    if (s.fin == 2 && R.out_a[0]) {
      uint32_t st2 = 0;
      F(T, ws_of(T, s.st), m.x, m.y, m.ax, m.ay, st2);
      m.az = 0.0;
    }
    if (R.out.x) R.out.x[i] = m.x;
    if (R.out.y) R.out.y[i] = m.y;
    if (R.out.z) R.out.z[i] = m.z;
    if (R.out.vx) R.out.vx[i] = m.vx;
    if (R.out.vy) R.out.vy[i] = m.vy;
    if (R.out.vz) R.out.vz[i] = m.vz;
    if (R.out.t) R.out.t[i] = m.t;
    if (R.out_a[0]) R.out_a[0][i] = m.ax;
    if (R.out_a[1]) R.out_a[1][i] = m.ay;
    if (R.out_a[2]) R.out_a[2][i] = m.az;
    if (R.label) R.label[i] = s.label;
    if (R.status) R.status[i] = (R.flags & CTRAJ_RUN_STATUS_OR) ? (uint8_t)(R.status[i] | s.status) : (uint8_t)s.status;
    n_oot += (s.status & CTRAJ_ST_OOT) ? 1u : 0u;
    n_cap += (s.status & CTRAJ_ST_CAPPED) ? 1u : 0u;
    const int64_t gidx = R.index_base + i;
    wave_hist(s_hist, R.first_index ? s_first : nullptr, s.st * B.n_labels + s.label, s.label, gidx);
    if (R.aoi_mask) {
      const bool keep = (R.aoi_mask >> s.label) & 1ull;
      const uint64_t mask = __ballot(keep);
      if (mask) {
        const int leader = __ffsll((unsigned long long)mask) - 1;
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd((unsigned long long *)R.aoi_count, (unsigned long long)__popcll(mask));
        base = __shfl(base, leader);
        if (keep) {
          const int64_t slot = (int64_t)base + lanes_below(mask);
          if (slot < R.aoi_cap) {
            R.aoi_index[slot] = gidx;
            if (R.aoi_label) R.aoi_label[slot] = s.label;
            if (R.aoi_out.x) R.aoi_out.x[slot] = m.x;
            if (R.aoi_out.y) R.aoi_out.y[slot] = m.y;
            if (R.aoi_out.z) R.aoi_out.z[slot] = m.z;
            if (R.aoi_out.vx) R.aoi_out.vx[slot] = m.vx;
            if (R.aoi_out.vy) R.aoi_out.vy[slot] = m.vy;
            if (R.aoi_out.vz) R.aoi_out.vz[slot] = m.vz;
            if (R.aoi_out.t) R.aoi_out.t[slot] = m.t;
          }
        }
      }
    }
    s.fin = 0;
  }
}


// FOLD: a hexapole lens whose force scale folds into the step constants (rk4_step_hex_k); the host
// launches FOLD = false for other lenses, which take rk4_step.
template <int FORCE, int INTEG, bool FOLD = false>
__global__ void __launch_bounds__(768) k_propagate_lanes(KBeam B, KTab KT, const ctraj_run *__restrict__ Rg,
                                                         LaneArgs LA) {
  extern __shared__ __attribute__((aligned(16))) double lds[];
  __shared__ uint32_t s_hist[CTRAJ_MAX_STATES * CTRAJ_MAX_LABELS];
  __shared__ unsigned long long s_first[CTRAJ_MAX_LABELS];
  __shared__ unsigned long long s_work[4];
  const int nbins = B.n_states * B.n_labels;
  for (int b = threadIdx.x; b < nbins; b += blockDim.x) s_hist[b] = 0;
  for (int b = threadIdx.x; b < B.n_labels; b += blockDim.x) s_first[b] = ~0ull >> 1;
  if (threadIdx.x < 4) s_work[threadIdx.x] = 0;
  const Tables T = stage_tables(lds, KT);
  __syncthreads();

  const int lane = lane_id();
  // hot lens parameters (live across the walker loop); the rest is re-read where it is used
  const ctraj_elem &L0 = B.elems[LA.lens_e];
  const Force<FORCE> F = make_force<FORCE>(L0, B.g);
  const double R2 = L0.p[1], z1 = L0.z1;
  const int label_inside = L0.label_inside;
  WalkCounters wc{0u, 0u};
  uint32_t n_oot = 0, n_cap = 0;
  unsigned long long steps_w = 0;  // lens steps of this wavefront (counted with the exec mask)
  int64_t next = 0, end = 0;  // this wavefront's claimed molecules [next, end)
  bool exhausted = false;
  LaneMol s;
  s.phase = PH_NEED;
  s.fin = 0;
  s.gidx = 0;
  s.st = 0;
  s.label = 0;
  s.status = 0;

  for (;;) {
    const uint64_t lensm = __ballot(s.phase == PH_LENS);
    const uint64_t idlem = __ballot(s.phase == PH_NEED || s.phase == PH_POST);
    if (lensm == 0 || __popcll(idlem) >= LA.refill_min) {
      for (;;) {
        // lanes that left the lens walk the rest of the beamline
        if (s.phase == PH_POST) {
          if constexpr (INTEG == CTRAJ_INTEG_REF) {
            drift_dt(s.m, time_to(s.m, z1), B.g);  // exit hop with the stored a = l1
            ++wc.faces;
          } else {
            s.m.ax = 0.0;  // the exit row at z1 stores (0, -g, 0)
            s.m.ay = -B.g;
            s.m.az = 0.0;
          }
          int lab = B.label_detected;
          walk_plain(B.elems, LA.lens_e + 1, B.n_elems, B.g, s.m, lab, wc);
          s.label = lab;
          s.fin = 1;
          s.phase = PH_NEED;
        }
        {
          const ctraj_run &R = *opaque(Rg);
          record_finished<FORCE>(B, R, s, s_hist, R.first_index ? s_first : nullptr, lane, n_oot, n_cap, F, T);
        }
        const uint64_t need = __ballot(s.phase == PH_NEED);
        if (!need || exhausted) {
          if (s.phase == PH_NEED && exhausted) s.phase = PH_DONE;
          break;
        }
        if (next >= end) {
          unsigned long long c = 0;
          if (lane == 0) c = atomicAdd(LA.counter, (unsigned long long)LA.claim);
          c = __shfl(c, 0);
          next = (int64_t)c;
          const int64_t n_all = opaque(Rg)->n;
          end = next + LA.claim < n_all ? next + LA.claim : n_all;
          if (next >= end) {
            exhausted = true;
            continue;
          }
        }
        const int64_t avail = end - next;
        if (s.phase == PH_NEED) {
          const int rank = lanes_below(need);
          if (rank < avail) {
            const ctraj_run &R = *opaque(Rg);
            const ctraj_elem &L = *opaque(B.elems + LA.lens_e);
            const double half_d = L.p[0], Hs = L.p[2], Ll = L.p[3];
            const int64_t i = next + rank;
            s.gidx = i;
            MState &m = s.m;
            m.x = R.in.x[i];
            m.y = R.in.y[i];
            m.z = R.in.z ? R.in.z[i] : R.z_const;
            m.vx = R.in.vx[i];
            m.vy = R.in.vy[i];
            m.vz = R.in.vz[i];
            m.t = R.in.t ? R.in.t[i] : R.t_const;
            m.ax = R.a0[0] ? R.a0[0][i] : 0.0;
            m.ay = R.a0[1] ? R.a0[1][i] : -B.g;
            m.az = R.a0[2] ? R.a0[2][i] : 0.0;
            s.st = R.state ? min((int)R.state[i], B.n_states - 1) : 0;
            s.wsb = ws_of(T, s.st);
            s.status = 0;
            int lab = B.label_detected;
            bool ok = walk_plain(B.elems, 0, LA.lens_e, B.g, m, lab, wc);
            if (ok) {
              drift_dt(m, time_to(m, L.z0), B.g);
              ++wc.faces;
              if (rho_of(m.x, m.y) > half_d) {
                ok = false;
                lab = L.label_hit;
              }
            }
            if (!ok) {
              s.label = lab;
              s.fin = 1;
            } else {
              s.i = 0;
              if constexpr (INTEG == CTRAJ_INTEG_RK4) {
                s.nfull = rk4_full_steps(Ll, m.vz, Hs, B.cap, s.status);
                s.dzs = m.vz * Hs;
                s.hp = Ll / m.vz - (double)s.nfull * Hs;
                s.h = s.nfull == 0 ? s.hp : Hs;
                s.r2 = __builtin_fma(m.x, m.x, m.y * m.y);
                s.phase = PH_LENS;
              } else {
                s.nfull = L.n_steps;
                s.h = Hs / m.vz;
                s.phase = s.nfull > 0 ? PH_LENS : PH_POST;
              }
            }
          }
        }
        next += avail < (int64_t)__popcll(need) ? avail : (int64_t)__popcll(need);
      }
      if (!__ballot(s.phase == PH_LENS)) break;
    }

    steps_w += (unsigned long long)__popcll(__ballot(s.phase == PH_LENS));
    if (s.phase == PH_LENS) {
      MState &m = s.m;
      if constexpr (INTEG == CTRAJ_INTEG_RK4) {
        // branch-free step tail (every lane in the lens runs the same instructions), except where
        // a molecule finishes: z and t stay at their entry values until then
        const bool last = (s.i == s.nfull);
        if constexpr (FOLD) {
          rk4_step_hex_k(s.wsb, T.umax, F.cu, LA.gk, hex_k(F.ks, s.h), m.x, m.y, m.vx, m.vy, s.r2);
        } else {
          rk4_step<FORCE>(F, T, s.wsb, s.h, m.x, m.y, m.vx, m.vy, s.r2, s.status);
        }
        s.r2 = __builtin_fma(m.x, m.x, m.y * m.y);
        ++s.i;
        s.h = (s.i == s.nfull) ? s.hp : s.h;
        const bool dead = s.r2 > R2;
        if (last) {  // after the partial step
          m.z = z1;
          m.t = __builtin_fma((double)s.nfull, L0.p[2], m.t) + s.hp;
        } else if (dead) {  // stopped by the bore after s.i full steps of the lens's dt
          m.z = __builtin_fma((double)s.i, s.dzs, m.z);
          m.t = __builtin_fma((double)s.i, L0.p[2], m.t);
        }
        s.label = dead ? label_inside : s.label;
        s.fin = dead ? 2 : 0;
        s.phase = dead ? (int)PH_NEED : (last ? (int)PH_POST : (int)PH_LENS);
      } else {
        ref_step<FORCE>(F, T, s.wsb, s.h, m, s.status);
        ++s.i;
        if (beyond_r2(m.x, m.y, L0.p[1])) {  // sqrt(x^2 + y^2) > d/2 (lens_ref)
          s.label = label_inside;
          s.fin = 1;
          s.phase = PH_NEED;
        } else if (s.i == s.nfull) {
          s.phase = PH_POST;
        }
      }
    }
  }

  const ctraj_run &R = *Rg;
  if (R.work) {
    if (wc.faces) atomicAdd(&s_work[0], (unsigned long long)wc.faces);
    if (lane == 0 && steps_w) atomicAdd(&s_work[1], steps_w);
    if (n_oot) atomicAdd(&s_work[2], (unsigned long long)n_oot);
    if (n_cap) atomicAdd(&s_work[3], (unsigned long long)n_cap);
  }
  __syncthreads();
  if (R.tally)
    for (int b = threadIdx.x; b < nbins; b += blockDim.x)
      if (s_hist[b]) atomicAdd((unsigned long long *)&R.tally[b], (unsigned long long)s_hist[b]);
  if (R.first_index)  // the stored minimum only decreases: a read of it that is not above ours spares the atomic
    for (int b = threadIdx.x; b < B.n_labels; b += blockDim.x)
      if (s_first[b] != (~0ull >> 1) && (int64_t)s_first[b] < __atomic_load_n(&R.first_index[b], __ATOMIC_RELAXED))
        atomicMin((unsigned long long *)&R.first_index[b], s_first[b]);
  if (R.work && threadIdx.x < 4 && s_work[threadIdx.x])
    atomicAdd((unsigned long long *)&R.work[threadIdx.x], s_work[threadIdx.x]);
}

// ---------------------------------------------------------------------------------------------
// k_lens: integration of the lens-entry queue.  Persistent wavefronts keep their lanes busy by
// lane refill: every iteration advances the lanes that take full steps by one step; when at least
// `refill_min` lanes do not, or none does, a convergent section finishes the molecules that stopped
// taking full steps -- those that reached the end of the lens take the final partial step
// (configuration mode) or the exit hop with the stored a = l1 (reference mode) -- writes each
// finished molecule's state back into its queue entry with its outcome (queue_put_done), and gives
// the empty lanes the next queued molecules.  A lane's rem is its full steps left (> 0: it takes
// steps; 0: it reached the end of the lens), REM_EMPTY when it holds no molecule, and at most
// REM_STOPPED once the bore stopped it: -2 - k after its k-th full step in configuration mode (the
// lane forms z and t from k when it finishes, not step by step), REM_STOPPED in reference mode; so
// one compare per iteration gives both the step's lane mask and the refill count, and a step updates
// rem with one compare, one select and one subtraction.  Finished molecules are walked
// through the remaining elements and recorded by POST, so the refill section carries no element
// walk and no recording.  Wavefronts claim the queue in 64-entry chunks, region by region, from
// eight counters: counter g serves the regions of the PRE workgroups dispatched to XCD g
// (workgroup b runs on XCD b mod 8), and a wavefront starts with its own XCD's counter and moves to
// the next one when it runs dry, which spreads the claim atomics over eight addresses.  Chunk c of
// group g is entries [64 (c % cpr), +64) of the group's region c / cpr (cpr = region / 64), of
// which the first wcount are filled, so a claim needs no search and a chunk past a region's count
// is simply claimed over.  With a claim list (longest-lived first, see ORD_BUCKETS) counter g hands
// out the list's chunks g, g + 8, g + 16, ... instead.  The per-molecule arithmetic is the one of
// walk().
// ---------------------------------------------------------------------------------------------

// The hexapole RK4 instantiations are held to 64 VGPRs (eight wavefronts per SIMD), so the host's
// 1024-thread workgroups run two per CU beside two copies of one state's 24 KB table.  They fit
// without spilling vector registers; the two scalar registers they spill are reloaded after the
// lens loop only.  Unheld, the scheduler spends up to 72 on the lens end's partial step (the step
// loop itself needs 64), which would leave one workgroup per CU.  The others need more registers
// than that.
constexpr int REM_EMPTY = -1, REM_STOPPED = -2;

// ORD: the wavefronts claim the list P.perm (longest-lived first) instead of the regions in order.
template <int FORCE, int INTEG, bool ORD>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(
    FORCE == CTRAJ_FORCE_WSPLINE_HEX && INTEG == CTRAJ_INTEG_RK4 ? 8 : 1)))
k_lens(KBeam B, KTab KT, PipeArgs P, int refill_min, StepK K, unsigned long long *work) {
  extern __shared__ __attribute__((aligned(16))) double lds[];
  // lens tallies (P.lt) are built into the hexapole RK4 instantiations only; the others would need
  // more registers for them
  constexpr bool LT = FORCE == CTRAJ_FORCE_WSPLINE_HEX && INTEG == CTRAJ_INTEG_RK4;
  __shared__ unsigned long long s_work[2];
  __shared__ uint32_t s_lt[LT ? LT_FIRST + 1 : 1];  // lens tallies: this workgroup's words of P.lt
  if (threadIdx.x < 2) s_work[threadIdx.x] = 0;
  if (LT && threadIdx.x <= LT_FIRST) s_lt[threadIdx.x] = 0;
  const Tables T = stage_tables(lds, KT);
  __syncthreads();
  // eight claim counters: counter g serves the regions of one XCD's PRE workgroups (n_chunks over
  // cpr chunks per region) or, with the claim list, the list's chunks g, g + 8, g + 16, ...
  const uint32_t cpr = ORD ? 0u : (uint32_t)(P.region >> 6);
  const uint32_t n_q = ORD ? P.ord_tot[ORD_BUCKETS] : 0u;
  // chunks: per counter, its group's regions x cpr; with the list, the list's ceil(n_q / 64) (counter
  // g's claim c is chunk c * 8 + g)
  const unsigned long long n_chunks =
      ORD ? (unsigned long long)((n_q + 63u) >> 6) : (unsigned long long)(P.n_pre_waves >> 3) * cpr;
  int grp = blockIdx.x & 7, tried = 0;

  const int lane = lane_id();
  const ctraj_elem &L0 = B.elems[P.lens_e];
  const Force<FORCE> F = make_force<FORCE>(L0, B.g);
  // the lens: z1, bore (x^2 + y^2 beyond R2 stops a molecule: (d/2)^2 for RK4, the squared-radius
  // threshold of d/2 for the reference), step (dt for RK4, dz for the reference), length; for RK4
  // the uniform step's half and sixth and, for a folding hexapole, the velocity constants ks Hs/2,
  // ks Hs, ks Hs/6 come formed by the host in K, so all stay in scalar registers
  const double R2 = L0.p[1], z1 = L0.z1, Hs = L0.p[2], Ll = L0.p[3];
  const int n_ref = L0.n_steps;
  uint32_t steps_w = 0, faces_w = 0;  // wave-uniform counts
  int64_t next = 0, end = 0;
  bool exhausted = false;

  // lane state: the molecule (configuration mode: z and t stay at their entry values until the lane
  // finishes; reference mode: they advance with the steps; vz is constant in the lens), its queue
  // entry and key, rem (see above), its full-step count plus 3 (configuration mode) or the step dt
  // (reference mode), and its state table
  MState m;
  m.x = m.y = m.z = m.vx = m.vy = m.vz = m.t = 0.0;
  m.ax = 0.0; m.ay = -B.g; m.az = 0.0;
  double r2 = 0.0, h = 0.0;
  int rem = REM_EMPTY, nfull3 = 0;
  uint32_t status = 0;
  int64_t slot = 0, key = 0;
  WsRef wsb = ws_of(T, 0);

  for (;;) {
    bool act = rem > 0;
    uint64_t am = __builtin_amdgcn_ballot_w64(act);
    if (64 - __popcll(am) >= refill_min) {
      for (;;) {
        // finish the molecules that hold a lane but take no more full steps: stopped by the bore, or
        // at the end of the lens, which takes one more step (configuration mode) or the exit hop
        // (reference mode)
        const bool fin = rem == 0 || rem <= REM_STOPPED;
        const bool dead = rem <= REM_STOPPED;
        const uint64_t ends = __builtin_amdgcn_ballot_w64(rem == 0);
        if constexpr (INTEG == CTRAJ_INTEG_RK4) steps_w += (uint32_t)__popcll(ends);
        else faces_w += (uint32_t)__popcll(ends);
        if (fin) {
          int outcome = LENS_STOPPED;
          if constexpr (INTEG == CTRAJ_INTEG_RK4) {
            if (!dead) {  // the end of the lens: one partial step of L/vz - n dt, so the lens ends at z1
              const int nfull = nfull3 - 3;
              const double hp = Ll / m.vz - (double)nfull * Hs;
              if constexpr (FORCE == CTRAJ_FORCE_WSPLINE_HEX) {  // the host launches only folding lenses here
                rk4_step_hex_k(wsb, T.umax, F.cu, K.gk, hex_k(F.ks, hp), m.x, m.y, m.vx, m.vy, r2);
              } else {
                rk4_step<FORCE>(F, T, wsb, hp, m.x, m.y, m.vx, m.vy, r2, status);
              }
              m.z = z1;
              m.t = __builtin_fma((double)nfull, Hs, m.t) + hp;
              r2 = __builtin_fma(m.x, m.x, m.y * m.y);
              outcome = r2 > R2 ? LENS_STOPPED : LENS_LEFT;
            } else {  // stopped by the bore after k full steps: z and t from their entry values
              const double k = (double)(REM_STOPPED - rem);
              m.z = __builtin_fma(k, m.vz * Hs, m.z);
              m.t = __builtin_fma(k, Hs, m.t);
            }
          } else {
            if (!dead) {  // exit hop with the stored a = l1 of the last step (electrostatic_lens.py:74-77)
              drift_dt(m, time_to(m, z1), B.g);
              outcome = LENS_LEFT;
            }
          }
          if (LT && P.lt && outcome == LENS_STOPPED) {
            // lens tallies: the molecule is counted here and needs no further walk
            atomicAdd(&s_lt[LT_STOP + (int)((uint64_t)key >> 56)], 1u);
            atomicMax(&s_lt[LT_FIRST], ~(uint32_t)key);  // slice index < 2^32
            if (status) {
              if (status & CTRAJ_ST_OOT) atomicAdd(&s_lt[LT_OOT], 1u);
              if (status & CTRAJ_ST_CAPPED) atomicAdd(&s_lt[LT_CAP], 1u);
            }
          } else {
            queue_put_done(P.q, slot, m.x, m.y, m.z, m.vx, m.vy, m.vz, m.t,
                           key | ((int64_t)status << 40) | ((int64_t)outcome << 48));
            if (LT && P.lt) {  // list the slot for POST: one atomic per wavefront, by its first lane here
              const uint64_t lm = __builtin_amdgcn_ballot_w64(true);
              const int rank = lanes_below(lm);
              uint32_t base = 0;
              if (rank == 0) base = atomicAdd(&P.lt[LT_N], (uint32_t)__popcll(lm));
              base = __builtin_amdgcn_readfirstlane(base);
              P.lt[LT_LIST + base + (uint32_t)rank] = (uint32_t)slot;
            }
          }
          rem = REM_EMPTY;
        }
        const uint64_t need = __builtin_amdgcn_ballot_w64(rem == REM_EMPTY);
        if (!need || exhausted) break;
        if (next >= end) {
          unsigned long long c = 0;
          if (lane == 0) c = atomicAdd(P.claim + 32 * grp, 1ull);
          c = uniform_u64(c);  // every lane is active here, lane 0 first
          if (ORD ? c * 8 + (unsigned)grp >= n_chunks : c >= n_chunks) {
            if (++tried == 8) exhausted = true;
            grp = (grp + 1) & 7;
            continue;
          }
          if constexpr (ORD) {
            next = (int64_t)(c * 8 + (unsigned)grp) * 64;
            end = next + 64 < (int64_t)n_q ? next + 64 : (int64_t)n_q;
          } else {
            const uint32_t r = (uint32_t)c / cpr;
            const uint32_t w = ((r >> 2) << 5) | ((uint32_t)grp << 2) | (r & 3u);
            const int64_t base = (int64_t)((uint32_t)c - r * cpr) * 64;
            const int64_t filled = (int64_t)P.wcount[w] - base;
            next = (int64_t)w * P.region + base;
            end = next + (filled < 64 ? filled : 64);  // next >= end: an empty chunk, claim again
          }
          continue;
        }
        const int64_t avail = end - next;
        if (rem == REM_EMPTY) {
          const int rank = lanes_below(need);
          if (rank < avail) {
            slot = next + rank;
            if constexpr (ORD) slot = P.perm[slot];
            queue_get_done(P.q, slot, m, key);
            m.ax = 0.0;  // the lens entrance row stores (0, -g, 0)
            m.ay = -B.g;
            m.az = 0.0;
            wsb = ws_of(T, (int)((uint64_t)key >> 56));
            status = 0;
            if constexpr (INTEG == CTRAJ_INTEG_RK4) {
              rem = rk4_full_steps(Ll, m.vz, Hs, B.cap, status);
              nfull3 = rem + 3;
              r2 = __builtin_fma(m.x, m.x, m.y * m.y);
            } else {
              rem = n_ref;
              h = Hs / m.vz;  // dt = dz / v_z (electrostatic_lens.py:88)
            }
          }
        }
        next += avail < (int64_t)__popcll(need) ? avail : (int64_t)__popcll(need);
      }
      act = rem > 0;
      am = __builtin_amdgcn_ballot_w64(act);
      if (!am) break;  // the queue is exhausted and every molecule has been finished
    }

    steps_w += (uint32_t)__popcll(am);
    if (act) {
      if constexpr (INTEG == CTRAJ_INTEG_RK4) {
        if constexpr (FORCE == CTRAJ_FORCE_WSPLINE_HEX)  // the host launches only folding lenses here
          rk4_step_hex_k(wsb, T.umax, F.cu, K.gk, K.hk, m.x, m.y, m.vx, m.vy, r2);
        else
          rk4_step_k<FORCE>(F, T, wsb, K.h1, m.x, m.y, m.vx, m.vy, r2, status);
        r2 = __builtin_fma(m.x, m.x, m.y * m.y);
        rem = rem - (r2 > R2 ? nfull3 : 1);  // stopped after its k-th step: -2 - k
      } else {
        ref_step<FORCE>(F, T, wsb, h, m, status);  // the row stores a = l1
        rem = beyond_r2(m.x, m.y, R2) ? REM_STOPPED : rem - 1;  // R2: the threshold of d/2 (lens_ref)
      }
    }
  }

  if (work || (LT && P.lt)) {  // kernel arguments: every thread takes the same branch
    if (work && lane == 0) {
      if (steps_w) atomicAdd(&s_work[1], (unsigned long long)steps_w);
      if (faces_w) atomicAdd(&s_work[0], (unsigned long long)faces_w);
    }
    __syncthreads();
    if (work && threadIdx.x < 2 && s_work[threadIdx.x]) atomicAdd(&work[threadIdx.x], s_work[threadIdx.x]);
    if (LT && P.lt && threadIdx.x >= LT_STOP && threadIdx.x <= LT_FIRST && s_lt[threadIdx.x]) {
      if (threadIdx.x == LT_FIRST) atomicMax(&P.lt[LT_FIRST], s_lt[LT_FIRST]);
      else atomicAdd(&P.lt[threadIdx.x], s_lt[threadIdx.x]);
    }
  }
}

// ---------------------------------------------------------------------------------------------
// k_record: re-run selected molecules and store their rows (or z-grid samples)
// ---------------------------------------------------------------------------------------------
// Workgroups of up to 512 threads for the circular-aperture-only instantiations (two wavefronts per
// SIMD fit their registers, and one workgroup per CU is all that several lens-state tables leave
// room for in LDS); 256 for the general ones, whose z-grid form needs more registers.
template <int FORCE, int INTEG, int MODE, int KM>
__global__ void __launch_bounds__(KM == KM_CIRC ? 512 : 256) k_record(KBeam B, KTab KT, ctraj_record RC) {
  extern __shared__ __attribute__((aligned(16))) double lds[];
  const Tables T = stage_tables(lds, KT);
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < RC.n; s += stride) {
    const int64_t i = RC.index ? RC.index[s] : s;
    MState m;
    m.x = RC.in.x[i];
    m.y = RC.in.y[i];
    m.z = RC.in.z ? RC.in.z[i] : RC.z_const;
    m.vx = RC.in.vx[i];
    m.vy = RC.in.vy[i];
    m.vz = RC.in.vz[i];
    m.t = RC.in.t ? RC.in.t[i] : RC.t_const;
    m.ax = RC.a0[0] ? RC.a0[0][i] : 0.0;
    m.ay = RC.a0[1] ? RC.a0[1][i] : -B.g;
    m.az = RC.a0[2] ? RC.a0[2][i] : 0.0;
    const int st = RC.state ? min((int)RC.state[i], B.n_states - 1) : 0;
    int label;
    uint32_t status = 0;
    WalkCounters wc{0u, 0u};
    if constexpr (MODE == CTRAJ_REC_ROWS) {
      RowSink sk{RC.buf, RC.n, s, RC.cap, 0};
      walk<FORCE, INTEG, decltype(sk), KM>(B.elems, B.n_elems, T, B.g, B.cap, st, B.label_detected, m, label, status, wc, sk);
      sk.finish(m);
      if (RC.n_rows) RC.n_rows[s] = sk.count;
    } else {
      ZGridSinkT<MODE == CTRAJ_REC_ZGRID_MOL> sk;
      sk.buf = RC.buf; sk.n = RC.n; sk.slot = s; sk.cap = RC.cap;
      sk.z_start = RC.z_start; sk.z_step = RC.z_step;
      sk.init();
      walk<FORCE, INTEG, decltype(sk), KM>(B.elems, B.n_elems, T, B.g, B.cap, st, B.label_detected, m, label, status, wc, sk);
      sk.finish(m);
      if (RC.n_rows) RC.n_rows[s] = sk.written;
    }
    if (RC.label) RC.label[s] = label;
  }
}

// ---------------------------------------------------------------------------------------------
// Philox4x32-10 Gaussian source
// ---------------------------------------------------------------------------------------------
struct DrawArgs {
  double mean[8], sigma[8];
  double *out[8];
};

__global__ void __launch_bounds__(256) k_draw(uint64_t seed, uint64_t offset, int64_t n, int nd, DrawArgs A) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    for (int p = 0; 2 * p < nd; ++p) {
      double z0, z1;
      gaussian_pair(seed, (uint64_t)i + offset, (uint32_t)p, z0, z1);
      if (A.out[2 * p]) A.out[2 * p][i] = A.mean[2 * p] + A.sigma[2 * p] * z0;
      if (2 * p + 1 < nd && A.out[2 * p + 1]) A.out[2 * p + 1][i] = A.mean[2 * p + 1] + A.sigma[2 * p + 1] * z1;
    }
  }
}

// index-list forms: slot j draws molecule index[j]
__global__ void __launch_bounds__(256) k_draw_at(uint64_t seed, const int64_t *index, int64_t n, int nd, DrawArgs A) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
    const uint64_t i = (uint64_t)index[j];
    for (int p = 0; 2 * p < nd; ++p) {
      double z0, z1;
      gaussian_pair(seed, i, (uint32_t)p, z0, z1);
      if (A.out[2 * p]) A.out[2 * p][j] = A.mean[2 * p] + A.sigma[2 * p] * z0;
      if (2 * p + 1 < nd && A.out[2 * p + 1]) A.out[2 * p + 1][j] = A.mean[2 * p + 1] + A.sigma[2 * p + 1] * z1;
    }
  }
}

__device__ __forceinline__ uint8_t uniform_u8(uint64_t seed, uint64_t k, uint32_t n_values) {
  uint32_t c[4] = {(uint32_t)k, (uint32_t)(k >> 32), 0x53544154u, 0x43545241u};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  return (uint8_t)(((uint64_t)c[0] * n_values) >> 32);
}

__global__ void __launch_bounds__(256) k_draw_u8_at(uint64_t seed, const int64_t *index, int64_t n, uint32_t n_values,
                                                    uint8_t *out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride)
    out[j] = uniform_u8(seed, (uint64_t)index[j], n_values);
}

__global__ void __launch_bounds__(256) k_draw_u8(uint64_t seed, uint64_t offset, int64_t n, uint32_t n_values,
                                                  uint8_t *out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    out[i] = uniform_u8(seed, (uint64_t)i + offset, n_values);
  }
}

__global__ void __launch_bounds__(256) k_draw_disc(uint64_t seed, uint64_t offset, int64_t n, double radius,
                                                    double z, double *x, double *y, double *zo) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double px, py;
    disc_point(seed, (uint64_t)i + offset, radius, px, py);
    x[i] = px;
    y[i] = py;
    if (zo) zo[i] = z;
  }
}

struct GatherArgs {
  const double *src[8];
  double *dst[8];
};

// Reads are scattered (kept molecules are sparse), writes are dense.
__global__ void __launch_bounds__(256) k_gather_f64(int64_t n, const int64_t *index, int64_t base, int nf,
                                                    GatherArgs G) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
    const int64_t i = index[j] - base;
    for (int f = 0; f < nf; ++f)
      if (G.dst[f]) G.dst[f][j] = G.src[f][i];
  }
}

__global__ void __launch_bounds__(256) k_gather_u8(int64_t n, const int64_t *index, int64_t base, const uint8_t *src,
                                                   uint8_t *dst) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) dst[j] = src[index[j] - base];
}

__global__ void __launch_bounds__(256) k_fill_i64(int64_t *dst, int64_t v, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = v;
}

__global__ void __launch_bounds__(256) k_fill_u8(uint8_t *dst, uint8_t v, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = v;
}

// ---------------------------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------------------------
namespace ctraj {

static thread_local std::string g_err;

int fail(ctraj_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->err = msg;
  g_err = msg;
  return code;
}

int hipfail(ctraj_ctx *ctx, hipError_t e, const char *what) {
  return fail(ctx, CTRAJ_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

KTab ctx_ktab(const ctraj_ctx *ctx) {
  KTab k{};
  k.ws = ctx->d_ws;
  k.ws_nseg = ctx->ws_nseg;
  k.ws_states = ctx->d_ws ? ctx->ws_states : 0;
  k.umax = ctx->ws_nseg > 0 ? std::nextafter((double)ctx->ws_nseg, 0.0) : 0.0;
  k.rt = ctx->d_rt;
  k.rt_total = ctx->d_rt ? ctx->rt_total : 0;
  return k;
}

size_t ktab_lds_bytes(const KTab &k) {
  const size_t nws = (size_t)k.ws_states * k.ws_nseg * 3;
  return (nws + (nws & 1) + (size_t)k.rt_total * 4) * sizeof(double);
}

// Validate an element table and find the (force, integrator) pair shared by its lenses.
int check_elems(ctraj_ctx *ctx, const ctraj_elem *e, int n, const ctraj_phys *p, int *force, int *integ) {
  *force = CTRAJ_FORCE_WSPLINE_HEX;
  *integ = CTRAJ_INTEG_RK4;
  bool have_lens = false;
  for (int i = 0; i < n; ++i) {
    const ctraj_elem &E = e[i];
    if (E.kind < CTRAJ_CIRC || E.kind > CTRAJ_LENS) return fail(ctx, CTRAJ_EINVAL, "unknown element kind");
    if (E.label_hit < 0 || E.label_hit >= p->n_labels) return fail(ctx, CTRAJ_EINVAL, "element label out of range");
    if (i > 0 && E.z0 < e[i - 1].z0) return fail(ctx, CTRAJ_EINVAL, "elements must be sorted by z0");
    if (E.kind == CTRAJ_LENS) {
      if (E.label_inside < 0 || E.label_inside >= p->n_labels)
        return fail(ctx, CTRAJ_EINVAL, "lens label_inside out of range");
      if (E.force < 0 || E.force > 2 || E.integrator < 0 || E.integrator > 1)
        return fail(ctx, CTRAJ_EINVAL, "bad lens force/integrator");
      if (have_lens && (E.force != *force || E.integrator != *integ))
        return fail(ctx, CTRAJ_EINVAL, "all lenses of a beamline must share force model and integrator");
      if (E.integrator == CTRAJ_INTEG_RK4 && !(E.p[2] > 0.0)) return fail(ctx, CTRAJ_EINVAL, "RK4 lens needs dt > 0");
      if (E.force != CTRAJ_FORCE_RTABLE && !(E.p[5] >= 0.0))
        return fail(ctx, CTRAJ_EINVAL, "lens spline scale p[5] must be >= 0 (use |V|)");
      if (!(E.p[0] >= 0.0)) return fail(ctx, CTRAJ_EINVAL, "lens bore radius must be >= 0");
      if (E.integrator == CTRAJ_INTEG_REF && E.n_steps < 0) return fail(ctx, CTRAJ_EINVAL, "negative lens step count");
      if (E.force == CTRAJ_FORCE_RTABLE && (E.table < 0 || E.table >= 8))
        return fail(ctx, CTRAJ_EINVAL, "lens r-table id out of range (0..7)");
      *force = E.force;
      *integ = E.integrator;
      have_lens = true;
    }
  }
  if (p->label_detected < 0 || p->label_detected >= p->n_labels)
    return fail(ctx, CTRAJ_EINVAL, "label_detected out of range");
  return CTRAJ_OK;
}

// Tables a launch needs must be present: r tables of RTABLE lenses, the W spline (with enough
// states) for WSPLINE lenses.
int check_tables(ctraj_ctx *ctx, const ctraj_elem *e, int n, int n_states) {
  for (int i = 0; i < n; ++i) {
    if (e[i].kind != CTRAJ_LENS) continue;
    if (e[i].force == CTRAJ_FORCE_RTABLE) {
      if (ctx->rt_n[e[i].table] < 2) return fail(ctx, CTRAJ_ESTATE, "lens r-table not set");
    } else {
      if (!ctx->d_ws) return fail(ctx, CTRAJ_ESTATE, "lens W spline not set");
      if (n_states > ctx->ws_states) return fail(ctx, CTRAJ_ESTATE, "n_states exceeds the W spline states set");
    }
  }
  return CTRAJ_OK;
}

// The squared-radius threshold of a circular aperture of radius d: the largest double whose correctly
// rounded square root is <= d (see beyond_r2).
double circ_r2_threshold(double d) {
  if (std::isnan(d) || std::isinf(d)) return d;  // NaN: never beyond; +inf: never; -inf: always
  if (d < 0.0) return -1.0;                        // x^2 + y^2 >= 0 > -1: always beyond (NaN r2 never)
  double t = d * d;
  while (std::sqrt(t) > d) t = std::nextafter(t, -HUGE_VAL);
  for (;;) {
    const double nx = std::nextafter(t, HUGE_VAL);
    if (std::isinf(nx) || std::sqrt(nx) > d) break;
    t = nx;
  }
  return t;
}

// Device copy of an element table: a circular aperture carries its squared-radius threshold (p[1]);
// so does a reference-mode lens for its bore test (its p[1] has no other use; configuration mode
// compares x^2 + y^2 with (d/2)^2 itself); an RTABLE lens carries its table's knot offset (table),
// knot count (reserved), inverse knot spacing (p[6]) and first and last knots (p[4], p[5], which
// only the spline forces use) instead of the table id.
void resolve_elems(const ctraj_ctx *ctx, const ctraj_elem *e, int n, ctraj_elem *out) {
  for (int i = 0; i < n; ++i) {
    out[i] = e[i];
    if (e[i].kind == CTRAJ_CIRC) out[i].p[1] = circ_r2_threshold(e[i].p[0]);
    if (e[i].kind == CTRAJ_LENS && e[i].integrator == CTRAJ_INTEG_REF) out[i].p[1] = circ_r2_threshold(e[i].p[0]);
    if (e[i].kind == CTRAJ_LENS && e[i].force == CTRAJ_FORCE_RTABLE) {
      const int id = e[i].table;
      out[i].table = ctx->rt_off[id];
      out[i].reserved = ctx->rt_n[id];
      out[i].p[6] = ctx->rt_invh[id];
      const std::vector<double> &tb = ctx->rt_host[id];
      out[i].p[4] = tb.empty() ? 0.0 : tb[0];
      out[i].p[5] = tb.empty() ? 0.0 : tb[tb.size() - 4];
    }
  }
}

static int upload_elems(ctraj_ctx *c) {
  if (!c->have_beamline) return CTRAJ_OK;
  std::vector<ctraj_elem> res(c->h_elems.size());
  resolve_elems(c, c->h_elems.data(), (int)c->h_elems.size(), res.data());
  HIPCHK(c, hipSetDevice(c->device));
  if (!c->d_elems) HIPCHK(c, hipMalloc((void **)&c->d_elems, sizeof(ctraj_elem) * CTRAJ_MAX_ELEMS));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (!res.empty()) HIPCHK(c, hipMemcpy(c->d_elems, res.data(), sizeof(ctraj_elem) * res.size(), hipMemcpyHostToDevice));
  return CTRAJ_OK;
}

int check_phys(ctraj_ctx *ctx, const ctraj_phys *p) {
  if (p->n_labels < 1 || p->n_labels > CTRAJ_MAX_LABELS) return fail(ctx, CTRAJ_EINVAL, "n_labels out of range");
  if (p->n_states < 1 || p->n_states > CTRAJ_MAX_STATES) return fail(ctx, CTRAJ_EINVAL, "n_states out of range");
  if (p->max_lens_steps < 0) return fail(ctx, CTRAJ_EINVAL, "max_lens_steps < 0");
  // the lens kernels keep a stopped molecule's step count in its negative lane state (-2 - steps)
  if (p->max_lens_steps > (1 << 30)) return fail(ctx, CTRAJ_EINVAL, "max_lens_steps > 2^30");
  return CTRAJ_OK;
}

KBeam make_kbeam(const ctraj_elem *d_elems, int n, const ctraj_phys &p) {
  KBeam b;
  b.elems = d_elems;
  b.n_elems = n;
  b.g = p.g;
  b.n_labels = p.n_labels;
  b.label_detected = p.label_detected;
  b.n_states = p.n_states;
  b.cap = p.max_lens_steps > 0 ? p.max_lens_steps : (1 << 20);
  return b;
}

int lds_limit(ctraj_ctx *ctx, const void *kern, size_t bytes) {
  if (bytes > (size_t)ctx->lds_max) return fail(ctx, CTRAJ_EINVAL, "lens tables exceed the LDS of one workgroup");
  if (bytes > 65536) HIPCHK(ctx, hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  return CTRAJ_OK;
}

int block_ok(ctraj_ctx *ctx, const void *kern, int block) {
  hipFuncAttributes fa{};
  HIPCHK(ctx, hipFuncGetAttributes(&fa, kern));
  if (block < 1 || block > fa.maxThreadsPerBlock)
    return fail(ctx, CTRAJ_EINVAL, "block of " + std::to_string(block) + " threads exceeds the kernel's limit of " +
                                       std::to_string(fa.maxThreadsPerBlock));
  return CTRAJ_OK;
}

int grid_for(ctraj_ctx *ctx, const void *kern, int block, size_t lds, int64_t n, int req) {
  if (req > 0) return req;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, block, lds) != hipSuccess || per_cu < 1) per_cu = 1;
  const int64_t need = (n + block - 1) / block;
  const int64_t full = (int64_t)per_cu * ctx->n_cu;
  return (int)std::max<int64_t>(1, std::min(need, full));
}

int ensure_buf(ctraj_ctx *ctx, ctraj_devbuf &b, size_t bytes) {
  if (b.bytes >= bytes && b.p) return CTRAJ_OK;
  if (b.p) {
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
  }
  if (hipMalloc(&b.p, bytes) != hipSuccess) {
    b.p = nullptr;
    (void)hipGetLastError();
    return fail(ctx, CTRAJ_ENOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
  }
  b.bytes = bytes;
  return CTRAJ_OK;
}

}  // namespace ctraj

// kernel tables over (force, integrator)
#define CT_FI_LIST(X)                                 \
  X(CTRAJ_FORCE_RTABLE, CTRAJ_INTEG_REF)             \
  X(CTRAJ_FORCE_RTABLE, CTRAJ_INTEG_RK4)             \
  X(CTRAJ_FORCE_WSPLINE_HEX, CTRAJ_INTEG_REF)        \
  X(CTRAJ_FORCE_WSPLINE_HEX, CTRAJ_INTEG_RK4)        \
  X(CTRAJ_FORCE_WSPLINE_LIN, CTRAJ_INTEG_REF)        \
  X(CTRAJ_FORCE_WSPLINE_LIN, CTRAJ_INTEG_RK4)

using PropFn = void (*)(KBeam, KTab, ctraj_run, int, PipeArgs);
using LensFn = void (*)(KBeam, KTab, PipeArgs, int, StepK, unsigned long long *);
using LaneFn = void (*)(KBeam, KTab, const ctraj_run *, LaneArgs);
using RecFn = void (*)(KBeam, KTab, ctraj_record);

// KM_CIRC when every element in [e0, e1) is a circular aperture, else KM_ALL
static int kinds_mask(const std::vector<ctraj_elem> &el, int e0, int e1) {
  for (int e = e0; e < e1; ++e)
    if (el[e].kind != CTRAJ_CIRC && el[e].kind != CTRAJ_LENS) return KM_ALL;
  return KM_CIRC;
}

// true if the run asks for any per-molecule output array
static bool wants_outputs(const ctraj_run *R) {
  return R->out.x || R->out.y || R->out.z || R->out.vx || R->out.vy || R->out.vz || R->out.t || R->out_a[0] ||
         R->out_a[1] || R->out_a[2] || R->label || R->status;
}

template <int PHASE>
// min_in: the run passes positions and velocities only (k_propagate's MIN form, instantiated for the
// circular-aperture walk without per-molecule outputs)
static PropFn walker_kernel(int km, bool out, bool min_in) {
  constexpr int F = CTRAJ_FORCE_WSPLINE_HEX, I = CTRAJ_INTEG_NONE;
  if (km == KM_CIRC && !out && min_in) return k_propagate<F, I, PHASE, KM_CIRC, false, true>;
  if (km == KM_CIRC)
    return out ? k_propagate<F, I, PHASE, KM_CIRC, true> : k_propagate<F, I, PHASE, KM_CIRC, false>;
  return out ? k_propagate<F, I, PHASE, KM_ALL, true> : k_propagate<F, I, PHASE, KM_ALL, false>;
}

static bool min_inputs(const ctraj_run *R) {
  return R->in.z && !R->in.t && !R->a0[0] && !R->a0[1] && !R->a0[2] && !R->state;
}

static PropFn prop_kernel(int f, int i, bool has_lens, int km, bool out, bool min_in) {
  if (!has_lens) return walker_kernel<PHASE_FULL>(km, out, min_in);
#define X(F, I) \
  if (f == F && i == I) return k_propagate<F, I, PHASE_FULL, KM_ALL, true>;
  CT_FI_LIST(X)
#undef X
  return nullptr;
}

static LensFn lens_kernel(int f, int i, bool ord) {
  if (ord) {  // the claim list is built for hexapole lenses only
    if (f == CTRAJ_FORCE_WSPLINE_HEX && i == CTRAJ_INTEG_RK4) return k_lens<CTRAJ_FORCE_WSPLINE_HEX, CTRAJ_INTEG_RK4, true>;
    if (f == CTRAJ_FORCE_WSPLINE_HEX && i == CTRAJ_INTEG_REF) return k_lens<CTRAJ_FORCE_WSPLINE_HEX, CTRAJ_INTEG_REF, true>;
    return nullptr;
  }
#define X(F, I) \
  if (f == F && i == I) return k_lens<F, I, false>;
  CT_FI_LIST(X)
#undef X
  return nullptr;
}

// POST: the walk after the lens and the recording of every lens molecule.  Only a run that asks for
// stored accelerations needs the lens force (configuration mode: the force where the bore stopped a
// molecule); the other instantiations do not depend on the lens model.
static PropFn post_kernel(int f, int km, bool out, bool accel) {
  constexpr int H = CTRAJ_FORCE_WSPLINE_HEX, I = CTRAJ_INTEG_RK4;
  if (accel) {
    if (f == CTRAJ_FORCE_RTABLE) return k_propagate<CTRAJ_FORCE_RTABLE, I, PHASE_POST, KM_ALL, true>;
    if (f == CTRAJ_FORCE_WSPLINE_LIN) return k_propagate<CTRAJ_FORCE_WSPLINE_LIN, I, PHASE_POST, KM_ALL, true>;
    return k_propagate<H, I, PHASE_POST, KM_ALL, true>;
  }
  if (km == KM_CIRC) return out ? k_propagate<H, I, PHASE_POST, KM_CIRC, true> : k_propagate<H, I, PHASE_POST, KM_CIRC, false>;
  return out ? k_propagate<H, I, PHASE_POST, KM_ALL, true> : k_propagate<H, I, PHASE_POST, KM_ALL, false>;
}

static LaneFn lane_kernel(int f, int i, bool fold) {
  if (fold && f == CTRAJ_FORCE_WSPLINE_HEX && i == CTRAJ_INTEG_RK4)
    return k_propagate_lanes<CTRAJ_FORCE_WSPLINE_HEX, CTRAJ_INTEG_RK4, true>;
#define X(F, I) \
  if (f == F && i == I) return k_propagate_lanes<F, I>;
  CT_FI_LIST(X)
#undef X
  return nullptr;
}

template <int F, int I, int KM>
static RecFn rec_kernel_km(int mode) {
  if (mode == CTRAJ_REC_ROWS) return k_record<F, I, CTRAJ_REC_ROWS, KM>;
  if (mode == CTRAJ_REC_ZGRID) return k_record<F, I, CTRAJ_REC_ZGRID, KM>;
  return k_record<F, I, CTRAJ_REC_ZGRID_MOL, KM>;
}

static RecFn rec_kernel(int f, int i, int mode, int km) {
#define X(F, I)                                                                                       \
  if (f == F && i == I) return km == KM_CIRC ? rec_kernel_km<F, I, KM_CIRC>(mode) : rec_kernel_km<F, I, KM_ALL>(mode);
  CT_FI_LIST(X)
#undef X
  return nullptr;
}

// Device memory of ctraj_propagate_noalloc: one array per device (each device that loads the library
// holds its own copy until the process ends).  In host code `scratch` names a host-side stand-in, not
// the device array: the array's address on each device is resolved once with hipGetSymbolAddress
// (scratch_on_device) and only that address is handed to kernels.
static __device__ __align__(256) char scratch[CTRAJ_NOALLOC_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static char *g_scratch_dev[256];  // per device ordinal, resolved by ctraj_ctx_create

// The calling thread's current device must be `device`.
static hipError_t scratch_on_device(int device, char **out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0 || device >= 256) return hipErrorInvalidDevice;
  if (!g_scratch_dev[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_dev[device] = (char *)p;
  }
  *out = g_scratch_dev[device];
  return hipSuccess;
}

// Scratch layout of ctraj_propagate_noalloc (byte offsets into the device's scratch array).
constexpr size_t SCR_CLAIM = 0;            // [8][32] claim counters
constexpr size_t SCR_COUNTER = 2048;       // lane walker's molecule counter
constexpr size_t SCR_RUN = 4096;           // device copy of the run (lane walker)
constexpr size_t SCR_WCOUNT = 8192;        // per-PRE-wavefront queue counts
constexpr size_t SCR_WCOUNT_MAX = 16384;   // PRE wavefronts at most
constexpr size_t SCR_QUEUE = SCR_WCOUNT + 4 * SCR_WCOUNT_MAX;
static_assert(sizeof(ctraj_run) <= SCR_WCOUNT - SCR_RUN, "run copy must fit its scratch slot");

// Molecules per slice of a run split into kernels: the lens-entry queue holds one slice.
constexpr int64_t SLICE_MAX = (int64_t)1 << 26;

// The run restricted to molecules [s0, s0 + len): per-molecule arrays offset by s0, global indices
// (first_index, kept molecules) by index_base + s0; the accumulated outputs are shared.
static ctraj_run slice_of(const ctraj_run &R, int64_t s0, int64_t len) {
  ctraj_run S = R;
  auto off = [s0](auto *p) { return p ? p + s0 : p; };
  S.n = len;
  S.index_base = R.index_base + s0;
  S.in.x = off(R.in.x); S.in.y = off(R.in.y); S.in.z = off(R.in.z); S.in.vx = off(R.in.vx);
  S.in.vy = off(R.in.vy); S.in.vz = off(R.in.vz); S.in.t = off(R.in.t);
  for (int k = 0; k < 3; ++k) {
    S.a0[k] = off(R.a0[k]);
    S.out_a[k] = off(R.out_a[k]);
  }
  S.state = off(R.state);
  S.out.x = off(R.out.x); S.out.y = off(R.out.y); S.out.z = off(R.out.z); S.out.vx = off(R.out.vx);
  S.out.vy = off(R.out.vy); S.out.vz = off(R.out.vz); S.out.t = off(R.out.t);
  S.label = off(R.label);
  S.status = off(R.status);
  return S;
}

__global__ void k_put_run(ctraj_run R, ctraj_run *dst) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = R;
}

// ctraj_propagate and ctraj_propagate_noalloc.  sdev == nullptr: the context's buffers (grown as
// needed); otherwise sdev is this device's scratch array (its resolved device address), every device
// buffer the launches need lies in it, and nothing is allocated.
static int32_t propagate_impl(ctraj_ctx *c, const ctraj_run *R, char *sdev) {
  if (!c || !R) return fail(c, CTRAJ_EINVAL, "null argument");
  if (!c->have_beamline) return fail(c, CTRAJ_ESTATE, "no beamline set");
  if (R->n < 0) return fail(c, CTRAJ_EINVAL, "n < 0");
  if (R->n == 0) return CTRAJ_OK;
  if (!R->in.x || !R->in.y || !R->in.vx || !R->in.vy || !R->in.vz) return fail(c, CTRAJ_EINVAL, "missing inputs");
  if (R->aoi_mask && (!R->aoi_count || !R->aoi_index || R->aoi_cap < 0))
    return fail(c, CTRAJ_EINVAL, "aoi compaction needs aoi_count, aoi_index and aoi_cap");
  if (c->phys.n_labels < 64 && (R->aoi_mask >> c->phys.n_labels))
    return fail(c, CTRAJ_EINVAL, "aoi_mask names labels beyond n_labels");
  if (sdev && (R->flags & CTRAJ_RUN_PHASE_TIMES))
    return fail(c, CTRAJ_EINVAL, "CTRAJ_RUN_PHASE_TIMES creates events: not available without allocation");
  int rc0 = check_tables(c, c->h_elems.data(), c->n_elems, c->phys.n_states);
  if (rc0) return rc0;
  KTab kt = ctx_ktab(c);
  if (c->force == CTRAJ_FORCE_RTABLE) kt.ws_states = 0;
  const size_t lds = ktab_lds_bytes(kt);
  KBeam B = make_kbeam(c->d_elems, c->n_elems, c->phys);
  int n_lens = 0, lens_e = -1;
  for (int e = 0; e < c->n_elems; ++e)
    if (c->h_elems[e].kind == CTRAJ_LENS) {
      ++n_lens;
      lens_e = e;
    }
  // A hexapole lens whose force scale does not fold into the step constants (hex_folds) takes the
  // single-kernel walker, which tests that per molecule; so does a reference-mode lens when stored
  // accelerations are asked for (the row of a molecule stopped by the bore stores the l1 of its last
  // step, which the lens kernel does not pass on).
  const bool accel = R->out_a[0] || R->out_a[1] || R->out_a[2];
  const bool folds = c->force != CTRAJ_FORCE_WSPLINE_HEX || c->integ != CTRAJ_INTEG_RK4 ||
                     (lens_e >= 0 && hex_folds(c->h_elems[lens_e].p[4]));
  const bool pipeline = n_lens == 1 && !(R->flags & (CTRAJ_RUN_SIMPLE | CTRAJ_RUN_LANES)) && folds &&
                        !(accel && c->integ == CTRAJ_INTEG_REF);
  const bool lanes = n_lens == 1 && !(R->flags & CTRAJ_RUN_SIMPLE) && !pipeline;
  // The lane-refill kernels keep their tables once per workgroup: with several lens-state tables
  // (more than ~40 KB) one large workgroup per CU holds as many wavefronts as its registers allow,
  // where 256-thread workgroups would be limited by LDS to one per CU.  The lens kernel of the split
  // path takes 1024-thread workgroups always: at 64 VGPRs two of them hold 8 wavefronts per SIMD
  // beside two copies of one state's table (256-thread workgroups: 6, limited by LDS).
  const int block = R->block ? R->block : (pipeline ? 1024 : (lanes && lds > 40960 ? 768 : 256));
  const int max_block = pipeline ? 1024 : (lanes ? 768 : 256);
  if (block < 64 || block > max_block || block % 64)
    return fail(c, CTRAJ_EINVAL, "block must be a multiple of 64 up to " + std::to_string(max_block));
  HIPCHK(c, hipSetDevice(c->device));
  // CTRAJ_RUN_PHASE_TIMES: four events of the context's ring bracket the kernels of this launch
  hipEvent_t *pev = nullptr;
  if (R->flags & CTRAJ_RUN_PHASE_TIMES) {
    const int slot = (int)(c->pt_next % 64);
    for (int k = 0; k < 4; ++k)
      if (!c->pt_ev[slot][k]) HIPCHK(c, hipEventCreate(&c->pt_ev[slot][k]));
    pev = c->pt_ev[slot];
  }
  if (pipeline) {
    // claim order, longest-lived first: hexapole lenses (whose W' table gives the prediction's force
    // scale), unless the run asks for queue order; the scratch array of the allocation-free entry
    // point holds no room for it
    const bool ordered = !sdev && c->force == CTRAJ_FORCE_WSPLINE_HEX && !(R->flags & CTRAJ_RUN_QUEUE_ORDER);
    LensFn kl = lens_kernel(c->force, c->integ, ordered);
    if (!kl) return fail(c, CTRAJ_EINVAL, "no kernel for this force/integrator");
    const bool out = wants_outputs(R);
    // lens tallies: for a hexapole lens in RK4 steps, without per-molecule outputs, and unless the
    // molecules stopped in the lens are kept (apertures of interest), the lens kernel tallies those
    // itself and POST walks only the molecules that leave the lens
    const int lab_in = c->h_elems[lens_e].label_inside;
    const bool lens_tally = c->force == CTRAJ_FORCE_WSPLINE_HEX && c->integ == CTRAJ_INTEG_RK4 && !out &&
                            !(lab_in >= 0 && lab_in < 64 && ((R->aoi_mask >> lab_in) & 1ull));
    // PRE: walk to the lens and through its entrance; each wavefront queues the molecules that enter
    // the lens into its own region (no atomics)
    const int blk = 256;
    const size_t lds_aoi = R->aoi_mask ? (size_t)(blk / 64) * AOI_WAVE * AOI_ENTRY_BYTES : 0;
    PropFn kpre = walker_kernel<PHASE_PRE>(kinds_mask(c->h_elems, 0, lens_e), out, min_inputs(R));
    // POST: the elements after the lens, and the recording of the lens molecules; it needs the lens
    // tables only for stored accelerations
    PropFn kpost = post_kernel(c->force, kinds_mask(c->h_elems, lens_e + 1, c->n_elems), out, accel);
    KTab kt_post = kt;
    if (!accel) kt_post.ws_states = kt_post.rt_total = 0;
    const size_t lds_post_tab = ktab_lds_bytes(kt_post);
    const size_t lds_post = lds_post_tab + lds_aoi;
    int rc = lds_limit(c, (const void *)kpre, lds_aoi);
    if (rc) return rc;
    if ((rc = lds_limit(c, (const void *)kpost, lds_post)) != CTRAJ_OK) return rc;
    if ((rc = lds_limit(c, (const void *)kl, lds)) != CTRAJ_OK) return rc;
    if ((rc = block_ok(c, (const void *)kl, block)) != CTRAJ_OK) return rc;
    // a slice's queue holds threads * ceil(len / threads) entries (region = a multiple of 64 per
    // PRE wavefront), so the PRE grid and the slice length follow from the queue's capacity
    const int64_t n_first = std::min<int64_t>(R->n, SLICE_MAX);
    int gpre = (grid_for(c, (const void *)kpre, blk, lds_aoi, n_first, R->grid) + 7) & ~7;  // whole XCD groups
    int64_t slice = n_first;
    PipeArgs P{};
    if (sdev) {
      // 64 B of queue entry and 4 B of listed slot per entry, and the lens tally words
      const size_t q_entries = (CTRAJ_NOALLOC_SCRATCH_BYTES - SCR_QUEUE - 4 * LT_LIST) / 68 & ~(size_t)63;
      const int g_cap = (int)std::min<size_t>(q_entries / blk, SCR_WCOUNT_MAX / (blk / 64)) & ~7;
      if (R->grid > g_cap)
        return fail(c, CTRAJ_EINVAL, "grid of " + std::to_string(R->grid) + " blocks exceeds the " +
                                         std::to_string(g_cap) + " the scratch array's queue counts can take");
      gpre = std::min(gpre, g_cap);
      const int64_t threads = (int64_t)gpre * blk;
      slice = std::min<int64_t>(R->n, (int64_t)(q_entries / threads) * threads);
      P.q = (longlong2 *)(sdev + SCR_QUEUE);
      P.claim = (unsigned long long *)(sdev + SCR_CLAIM);
      P.wcount = (uint32_t *)(sdev + SCR_WCOUNT);
      if (lens_tally) P.lt = (uint32_t *)(sdev + SCR_QUEUE + q_entries * 64);
    } else {
      const int64_t threads = (int64_t)gpre * blk;
      const size_t cap = (size_t)threads * (size_t)((slice + threads - 1) / threads);
      // the queue, then (lens tallies) the tally words and the listed slots
      if ((rc = ensure_buf(c, c->pipe, cap * 64 + 256 + (lens_tally ? 4 * (LT_LIST + cap) : 0))) != CTRAJ_OK)
        return rc;
      if ((rc = ensure_buf(c, c->pipe_cnt, 8 * 256 + (size_t)gpre * (blk / 64) * sizeof(uint32_t))) != CTRAJ_OK)
        return rc;
      P.q = (longlong2 *)c->pipe.p;
      P.claim = (unsigned long long *)c->pipe_cnt.p;
      P.wcount = (uint32_t *)((char *)c->pipe_cnt.p + 8 * 256);
      if (lens_tally) P.lt = (uint32_t *)((char *)c->pipe.p + cap * 64 + 256);
    }
    const int64_t threads = (int64_t)gpre * blk;
    P.n_pre_waves = gpre * (blk / 64);
    P.lens_e = lens_e;
    if (ordered) {
      const size_t cap = (size_t)threads * (size_t)((slice + threads - 1) / threads);
      const size_t b_bkt = (cap + 255) & ~(size_t)255, b_perm = cap * 4, b_tot = 2048;
      const size_t b_hist = (size_t)gpre * ORD_BUCKETS * 4;
      if ((rc = ensure_buf(c, c->ord, b_bkt + b_perm + b_tot + b_hist)) != CTRAJ_OK) return rc;
      char *o = (char *)c->ord.p;
      P.ord_tot = (uint32_t *)o;
      P.ord_hist = (uint32_t *)(o + b_tot);
      P.perm = (uint32_t *)(o + b_tot + b_hist);
      P.bkt = (uint8_t *)(o + b_tot + b_hist + b_perm);
      const double ks = c->h_elems[lens_e].p[4];
      for (int st = 0; st < c->ws_states && st < CTRAJ_MAX_STATES; ++st) P.ord_k2[st] = (float)(ks * c->ws_end[st]);
      P.ord_ref = c->integ == CTRAJ_INTEG_REF;
    }
    // refill at 10 idle lanes by default (a sweep of 6-16 on the 10^7-molecule ES-lens run: 10 and
    // 12 fastest, 3 % ahead of 8)
    const int refill = R->refill > 0 ? std::min(R->refill, 64) : 10;
    const int grid = grid_for(c, (const void *)kl, block, lds, (int64_t)1 << 40, 0);
    const StepK K = step_k(c->h_elems[lens_e].p[2], c->h_elems[lens_e].p[4], c->phys.g);
    const bool one_slice = slice >= R->n;
    for (int64_t s0 = 0; s0 < R->n; s0 += slice) {
      const int64_t len = std::min(slice, R->n - s0);
      const ctraj_run S = slice_of(*R, s0, len);
      P.region = (len + threads - 1) / threads * 64;
      if ((rc = ctraj_fill_i64(c, (int64_t *)P.claim, 0, 8 * 32)) != CTRAJ_OK) return rc;
      if (P.lt && (rc = ctraj_fill_i64(c, (int64_t *)P.lt, 0, LT_LIST / 2)) != CTRAJ_OK) return rc;
      if (pev && s0 == 0) HIPCHK(c, hipEventRecord(pev[0], c->stream));
      hipLaunchKernelGGL(kpre, dim3(gpre), dim3(blk), lds_aoi, c->stream, B, kt, S, 0, P);
      HIPCHK(c, hipGetLastError());
      if (ordered) {
        hipLaunchKernelGGL(k_ord_scan, dim3(ORD_BUCKETS), dim3(256), 0, c->stream, P.ord_hist, gpre, P.ord_tot);
        hipLaunchKernelGGL(k_ord_scatter, dim3(gpre), dim3(256), 0, c->stream,
                           OrdList{P.bkt, P.ord_hist, P.ord_tot, P.perm, P.wcount, P.region});
        HIPCHK(c, hipGetLastError());
      }
      if (pev && one_slice) HIPCHK(c, hipEventRecord(pev[1], c->stream));
      hipLaunchKernelGGL(kl, dim3(grid), dim3(block), lds, c->stream, B, kt, P, refill, K,
                         (unsigned long long *)R->work);
      HIPCHK(c, hipGetLastError());
      if (pev && one_slice) HIPCHK(c, hipEventRecord(pev[2], c->stream));
      const int gpost = grid_for(c, (const void *)kpost, blk, lds_post, (int64_t)P.n_pre_waves * P.region, 0);
      hipLaunchKernelGGL(kpost, dim3(gpost), dim3(blk), lds_post, c->stream, B, kt_post, S,
                         (int)(lds_post_tab / sizeof(double)), P);
      HIPCHK(c, hipGetLastError());
    }
    if (pev) {
      HIPCHK(c, hipEventRecord(one_slice ? pev[3] : pev[1], c->stream));
      c->pt_split[c->pt_next % 64] = one_slice;
    }
  } else if (lanes) {
    const double ks = c->h_elems[lens_e].p[4];
    LaneFn k = lane_kernel(c->force, c->integ, hex_folds(ks));
    if (!k) return fail(c, CTRAJ_EINVAL, "no kernel for this force/integrator");
    int rc = lds_limit(c, (const void *)k, lds);
    if (rc) return rc;
    rc = block_ok(c, (const void *)k, block);
    if (rc) return rc;
    void *counter = nullptr;
    ctraj_run *d_run = nullptr;
    if (sdev) {
      counter = sdev + SCR_COUNTER;
      d_run = (ctraj_run *)(sdev + SCR_RUN);
      hipLaunchKernelGGL(k_put_run, dim3(1), dim3(64), 0, c->stream, *R, d_run);
      HIPCHK(c, hipGetLastError());
    } else {
      if (!c->d_counter) HIPCHK(c, hipMalloc((void **)&c->d_counter, 256));
      counter = c->d_counter;
      if (!c->d_run) HIPCHK(c, hipMalloc(&c->d_run, sizeof(ctraj_run)));
      if (!c->have_last_run || std::memcmp(&c->last_run, R, sizeof(ctraj_run)) != 0) {
        HIPCHK(c, hipMemcpyAsync(c->d_run, R, sizeof(ctraj_run), hipMemcpyHostToDevice, c->stream));
        c->last_run = *R;
        c->have_last_run = true;
      }
      d_run = (ctraj_run *)c->d_run;
    }
    rc = ctraj_fill_i64(c, (int64_t *)counter, 0, 1);
    if (rc) return rc;
    LaneArgs LA;
    LA.counter = (unsigned long long *)counter;
    LA.lens_e = lens_e;
    LA.refill_min = R->refill > 0 ? std::min(R->refill, 64) : 24;
    LA.claim = 128;
    LA.gk = hex_folds(ks) ? -c->phys.g / ks : 0.0;
    const int grid = grid_for(c, (const void *)k, block, lds, R->n, R->grid);
    if (pev) HIPCHK(c, hipEventRecord(pev[0], c->stream));
    hipLaunchKernelGGL(k, dim3(grid), dim3(block), lds, c->stream, B, kt, (const ctraj_run *)d_run, LA);
    if (pev) {
      HIPCHK(c, hipEventRecord(pev[1], c->stream));
      c->pt_split[c->pt_next % 64] = false;
    }
  } else {
    PropFn k = prop_kernel(c->force, c->integ, n_lens > 0, kinds_mask(c->h_elems, 0, c->n_elems), wants_outputs(R),
                           min_inputs(R));
    if (!k) return fail(c, CTRAJ_EINVAL, "no kernel for this force/integrator");
    if (n_lens == 0) kt.ws_states = kt.rt_total = 0;
    size_t lds_tab = n_lens > 0 ? lds : 0;
    // kept-molecule staging area after the tables: AOI_WAVE entries per wavefront
    const int aoi_off = (int)(lds_tab / sizeof(double));
    const size_t lds_all = lds_tab + (R->aoi_mask ? (size_t)(block / 64) * AOI_WAVE * AOI_ENTRY_BYTES : 0);
    int rc = lds_limit(c, (const void *)k, lds_all);
    if (rc) return rc;
    rc = block_ok(c, (const void *)k, block);
    if (rc) return rc;
    const int grid = grid_for(c, (const void *)k, block, lds_all, R->n, R->grid);
    if (pev) HIPCHK(c, hipEventRecord(pev[0], c->stream));
    hipLaunchKernelGGL(k, dim3(grid), dim3(block), lds_all, c->stream, B, kt, *R, aoi_off, PipeArgs{});
    if (pev) {
      HIPCHK(c, hipEventRecord(pev[1], c->stream));
      c->pt_split[c->pt_next % 64] = false;
    }
  }
  HIPCHK(c, hipGetLastError());
  if (pev && ++c->pt_next - c->pt_read > 64) c->pt_read = c->pt_next - 64;
  return CTRAJ_OK;
}

extern "C" {

int32_t ctraj_abi_version(void) { return CTRAJ_ABI_VERSION; }

int32_t ctraj_ctx_create(int32_t device, ctraj_ctx **out) {
  if (!out) return fail(nullptr, CTRAJ_EINVAL, "out is NULL");
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess) return hipfail(nullptr, e, "hipGetDeviceCount");
  if (device < 0 || device >= ndev) return fail(nullptr, CTRAJ_EINVAL, "no such device");
  e = hipSetDevice(device);
  if (e != hipSuccess) return hipfail(nullptr, e, "hipSetDevice");
  ctraj_ctx *c = new ctraj_ctx();
  c->device = device;
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return hipfail(nullptr, e, "hipStreamCreate");
  }
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return hipfail(nullptr, e, "hipGetDeviceProperties");
  }
  c->n_cu = prop.multiProcessorCount;
  c->lds_max = (int)prop.sharedMemPerBlock;
  int lds_optin = 0;
  if (hipDeviceGetAttribute(&lds_optin, hipDeviceAttributeSharedMemPerBlockOptin, device) == hipSuccess &&
      lds_optin > c->lds_max)
    c->lds_max = lds_optin;
  c->arch = 0;
  if (std::strncmp(prop.gcnArchName, "gfx", 3) == 0) c->arch = (int)std::strtol(prop.gcnArchName + 3, nullptr, 10);
  c->wave = prop.warpSize;
  for (int i = 0; i < 64; ++i) c->ev[i] = nullptr;
  // the device's scratch array (ctraj_propagate_noalloc), resolved once per device with it current;
  // a failure here only disables ctraj_propagate_noalloc
  if (scratch_on_device(device, &c->scratch) != hipSuccess) {
    c->scratch = nullptr;
    (void)hipGetLastError();
  }
  *out = c;
  return CTRAJ_OK;
}

int32_t ctraj_ctx_destroy(ctraj_ctx *c) {
  if (!c) return CTRAJ_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  if (c->d_elems) (void)hipFree(c->d_elems);
  if (c->d_ws) (void)hipFree(c->d_ws);
  if (c->d_rt) (void)hipFree(c->d_rt);
  if (c->d_counter) (void)hipFree(c->d_counter);
  if (c->d_run) (void)hipFree(c->d_run);
  if (c->sm_ws.p) (void)hipFree(c->sm_ws.p);
  if (c->sm_part.p) (void)hipFree(c->sm_part.p);
  if (c->sort_tmp.p) (void)hipFree(c->sort_tmp.p);
  if (c->pipe.p) (void)hipFree(c->pipe.p);
  if (c->pipe_cnt.p) (void)hipFree(c->pipe_cnt.p);
  if (c->ord.p) (void)hipFree(c->ord.p);
  if (c->sort_alt.p) (void)hipFree(c->sort_alt.p);
  for (int i = 0; i < 64; ++i) {
    if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
    for (int k = 0; k < 4; ++k)
      if (c->pt_ev[i][k]) (void)hipEventDestroy(c->pt_ev[i][k]);
  }
  (void)hipStreamDestroy(c->stream);
  delete c;
  return CTRAJ_OK;
}

const char *ctraj_last_error(const ctraj_ctx *c) { return c ? c->err.c_str() : g_err.c_str(); }

int32_t ctraj_device_info(ctraj_ctx *c, int32_t *n_cu, int32_t *lds, int32_t *wave, int32_t *arch) {
  if (!c) return fail(nullptr, CTRAJ_EINVAL, "ctx is NULL");
  if (n_cu) *n_cu = c->n_cu;
  if (lds) *lds = c->lds_max;
  if (wave) *wave = c->wave;
  if (arch) *arch = c->arch;
  return CTRAJ_OK;
}

int32_t ctraj_dev_alloc(ctraj_ctx *c, uint64_t bytes, void **ptr) {
  if (!c || !ptr) return fail(c, CTRAJ_EINVAL, "null argument");
  HIPCHK(c, hipSetDevice(c->device));
  *ptr = nullptr;
  hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
  if (e != hipSuccess) return fail(c, CTRAJ_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  return CTRAJ_OK;
}

int32_t ctraj_dev_free(ctraj_ctx *c, void *ptr) {
  if (!c) return fail(nullptr, CTRAJ_EINVAL, "ctx is NULL");
  if (ptr) HIPCHK(c, hipFree(ptr));
  return CTRAJ_OK;
}

int32_t ctraj_host_alloc(ctraj_ctx *c, uint64_t bytes, void **ptr) {
  if (!c || !ptr) return fail(c, CTRAJ_EINVAL, "null argument");
  *ptr = nullptr;
  hipError_t e = hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault);
  if (e != hipSuccess) return fail(c, CTRAJ_ENOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
  return CTRAJ_OK;
}

int32_t ctraj_host_free(ctraj_ctx *c, void *ptr) {
  if (!c) return fail(nullptr, CTRAJ_EINVAL, "ctx is NULL");
  if (ptr) HIPCHK(c, hipHostFree(ptr));
  return CTRAJ_OK;
}

int32_t ctraj_memcpy(ctraj_ctx *c, void *dst, const void *src, uint64_t bytes, int32_t kind) {
  if (!c) return fail(nullptr, CTRAJ_EINVAL, "ctx is NULL");
  if (bytes == 0) return CTRAJ_OK;
  if (!dst || !src) return fail(c, CTRAJ_EINVAL, "null pointer");
  hipMemcpyKind k = kind == CTRAJ_H2D ? hipMemcpyHostToDevice
                    : kind == CTRAJ_D2H ? hipMemcpyDeviceToHost
                    : kind == CTRAJ_D2D ? hipMemcpyDeviceToDevice
                                        : hipMemcpyDefault;
  if (kind < 1 || kind > 3) return fail(c, CTRAJ_EINVAL, "bad copy kind");
  HIPCHK(c, hipMemcpyAsync(dst, src, bytes, k, c->stream));
  return CTRAJ_OK;
}

int32_t ctraj_fill_u8(ctraj_ctx *c, void *dst, uint8_t v, uint64_t bytes) {
  if (!c) return fail(nullptr, CTRAJ_EINVAL, "ctx is NULL");
  if (bytes == 0) return CTRAJ_OK;
  if (!dst) return fail(c, CTRAJ_EINVAL, "null pointer");
  const int grid = (int)std::min<uint64_t>((bytes + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_u8, dim3(grid), dim3(256), 0, c->stream, (uint8_t *)dst, v, (int64_t)bytes);
  HIPCHK(c, hipGetLastError());
  return CTRAJ_OK;
}

int32_t ctraj_fill_i64(ctraj_ctx *c, int64_t *dst, int64_t v, int64_t n) {
  if (!c) return fail(nullptr, CTRAJ_EINVAL, "ctx is NULL");
  if (n <= 0) return CTRAJ_OK;
  if (!dst) return fail(c, CTRAJ_EINVAL, "null pointer");
  const int grid = (int)std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_i64, dim3(grid), dim3(256), 0, c->stream, dst, v, n);
  HIPCHK(c, hipGetLastError());
  return CTRAJ_OK;
}

int32_t ctraj_sync(ctraj_ctx *c) {
  if (!c) return fail(nullptr, CTRAJ_EINVAL, "ctx is NULL");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CTRAJ_OK;
}

int32_t ctraj_event_record(ctraj_ctx *c, int32_t slot) {
  if (!c || slot < 0 || slot >= 64) return fail(c, CTRAJ_EINVAL, "bad event slot");
  if (!c->ev[slot]) HIPCHK(c, hipEventCreate(&c->ev[slot]));
  HIPCHK(c, hipEventRecord(c->ev[slot], c->stream));
  return CTRAJ_OK;
}

int32_t ctraj_event_elapsed(ctraj_ctx *c, int32_t a, int32_t b, float *ms) {
  if (!c || !ms || a < 0 || a >= 64 || b < 0 || b >= 64 || !c->ev[a] || !c->ev[b])
    return fail(c, CTRAJ_EINVAL, "bad event slots");
  HIPCHK(c, hipEventSynchronize(c->ev[b]));
  HIPCHK(c, hipEventElapsedTime(ms, c->ev[a], c->ev[b]));
  return CTRAJ_OK;
}

int32_t ctraj_set_beamline(ctraj_ctx *c, const ctraj_elem *elems, int32_t n, const ctraj_phys *phys) {
  if (!c || !phys || (n > 0 && !elems)) return fail(c, CTRAJ_EINVAL, "null argument");
  if (n < 0 || n > CTRAJ_MAX_ELEMS) return fail(c, CTRAJ_EINVAL, "n_elems out of range");
  int rc = check_phys(c, phys);
  if (rc) return rc;
  int f, in;
  rc = check_elems(c, elems, n, phys, &f, &in);
  if (rc) return rc;
  c->h_elems.assign(elems, elems + n);
  c->n_elems = n;
  c->phys = *phys;
  c->force = f;
  c->integ = in;
  c->have_beamline = true;
  return upload_elems(c);
}

int32_t ctraj_set_rtable(ctraj_ctx *c, int32_t id, const double *r, const double *a, int32_t n) {
  if (!c || !r || !a) return fail(c, CTRAJ_EINVAL, "null argument");
  if (id < 0 || id >= 8) return fail(c, CTRAJ_EINVAL, "r-table id out of range (0..7)");
  if (n < 2) return fail(c, CTRAJ_EINVAL, "r-table needs at least 2 knots");
  for (int j = 0; j < n; ++j) {
    if (!std::isfinite(r[j]) || !std::isfinite(a[j])) return fail(c, CTRAJ_EINVAL, "r-table has non-finite entries");
    if (j > 0 && !(r[j] > r[j - 1])) return fail(c, CTRAJ_EINVAL, "r-table knots must be strictly increasing");
  }
  std::vector<double> tb((size_t)n * 4, 0.0);
  for (int j = 0; j < n; ++j) {
    tb[4 * j] = r[j];
    tb[4 * j + 1] = a[j];
    tb[4 * j + 2] = j + 1 < n ? (a[j + 1] - a[j]) / (r[j + 1] - r[j]) : 0.0;
  }
  c->rt_host[id] = tb;
  // rebuild the concatenated device table
  int total = 0;
  for (int k = 0; k < 8; ++k) {
    c->rt_off[k] = total;
    c->rt_n[k] = (int)(c->rt_host[k].size() / 4);
    c->rt_invh[k] = c->rt_n[k] >= 2 ? 1.0 / (c->rt_host[k][4] - c->rt_host[k][0]) : 0.0;
    total += c->rt_n[k];
  }
  std::vector<double> all((size_t)total * 4);
  for (int k = 0; k < 8; ++k)
    std::copy(c->rt_host[k].begin(), c->rt_host[k].end(), all.begin() + (size_t)c->rt_off[k] * 4);
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->d_rt) HIPCHK(c, hipFree(c->d_rt));
  c->d_rt = nullptr;
  HIPCHK(c, hipMalloc((void **)&c->d_rt, all.size() * sizeof(double)));
  HIPCHK(c, hipMemcpy(c->d_rt, all.data(), all.size() * sizeof(double), hipMemcpyHostToDevice));
  c->rt_total = total;
  return upload_elems(c);
}

int32_t ctraj_set_wspline(ctraj_ctx *c, int32_t n_states, int32_t n_seg, const double *coef) {
  if (!c || !coef) return fail(c, CTRAJ_EINVAL, "null argument");
  if (n_states < 1 || n_states > CTRAJ_MAX_STATES) return fail(c, CTRAJ_EINVAL, "n_states out of range");
  if (n_seg < 1) return fail(c, CTRAJ_EINVAL, "n_seg < 1");
  const size_t cnt = (size_t)n_states * n_seg * 4;
  for (size_t i = 0; i < cnt; ++i)
    if (!std::isfinite(coef[i])) return fail(c, CTRAJ_EINVAL, "W spline has non-finite coefficients");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->d_ws) HIPCHK(c, hipFree(c->d_ws));
  c->d_ws = nullptr;
  HIPCHK(c, hipMalloc((void **)&c->d_ws, cnt * sizeof(double)));
  HIPCHK(c, hipMemcpy(c->d_ws, coef, cnt * sizeof(double), hipMemcpyHostToDevice));
  c->ws_states = n_states;
  c->ws_nseg = n_seg;
  for (int st = 0; st < n_states; ++st) {
    const double *last = coef + ((size_t)st * n_seg + n_seg - 1) * 4;
    c->ws_end[st] = (last[0] + last[1]) + last[2];
  }
  return CTRAJ_OK;
}

int32_t ctraj_propagate(ctraj_ctx *c, const ctraj_run *R) { return propagate_impl(c, R, nullptr); }

int32_t ctraj_propagate_noalloc(ctraj_ctx *c, const ctraj_run *R) {
  if (!c || !R) return fail(c, CTRAJ_EINVAL, "null argument");
  if (!c->scratch) return fail(c, CTRAJ_EHIP, "the scratch array's device address could not be resolved");
  return propagate_impl(c, R, c->scratch);
}

int32_t ctraj_phase_times(ctraj_ctx *c, int32_t max, float *walk_ms, float *lens_ms, float *post_ms, int32_t *count) {
  if (!c || !count || max < 0 || (max > 0 && (!walk_ms || !lens_ms || !post_ms)))
    return fail(c, CTRAJ_EINVAL, "null argument");
  *count = 0;
  const int64_t first = std::max(c->pt_read, c->pt_next - (int64_t)max);
  int32_t k = 0;
  for (int64_t j = first; j < c->pt_next; ++j, ++k) {
    hipEvent_t *e = c->pt_ev[j % 64];
    if (c->pt_split[j % 64]) {
      HIPCHK(c, hipEventSynchronize(e[3]));
      HIPCHK(c, hipEventElapsedTime(&walk_ms[k], e[0], e[1]));
      HIPCHK(c, hipEventElapsedTime(&lens_ms[k], e[1], e[2]));
      HIPCHK(c, hipEventElapsedTime(&post_ms[k], e[2], e[3]));
    } else {
      HIPCHK(c, hipEventSynchronize(e[1]));
      HIPCHK(c, hipEventElapsedTime(&lens_ms[k], e[0], e[1]));
      walk_ms[k] = post_ms[k] = 0.0f;
    }
  }
  c->pt_read = c->pt_next;
  *count = k;
  return CTRAJ_OK;
}

int32_t ctraj_record_trajectories(ctraj_ctx *c, const ctraj_record *RC) {
  if (!c || !RC) return fail(c, CTRAJ_EINVAL, "null argument");
  if (!c->have_beamline) return fail(c, CTRAJ_ESTATE, "no beamline set");
  if (RC->n < 0 || RC->cap < 0) return fail(c, CTRAJ_EINVAL, "negative size");
  if (RC->n == 0) return CTRAJ_OK;
  if (!RC->in.x || !RC->in.y || !RC->in.vx || !RC->in.vy || !RC->in.vz || !RC->buf)
    return fail(c, CTRAJ_EINVAL, "missing inputs or buffer");
  if (RC->mode != CTRAJ_REC_ROWS && RC->mode != CTRAJ_REC_ZGRID && RC->mode != CTRAJ_REC_ZGRID_MOL)
    return fail(c, CTRAJ_EINVAL, "bad record mode");
  if (RC->mode != CTRAJ_REC_ROWS && !(RC->z_step > 0.0)) return fail(c, CTRAJ_EINVAL, "z_step must be > 0");
  int rc0 = check_tables(c, c->h_elems.data(), c->n_elems, c->phys.n_states);
  if (rc0) return rc0;
  const int km = kinds_mask(c->h_elems, 0, c->n_elems);
  RecFn k = rec_kernel(c->force, c->integ, RC->mode, km);
  KTab kt = ctx_ktab(c);
  if (c->force == CTRAJ_FORCE_RTABLE) kt.ws_states = 0;
  const size_t lds = ktab_lds_bytes(kt);
  // one large workgroup per CU when the tables fill LDS (several lens states), else 128 threads
  const int block = RC->block ? RC->block : (RC->mode == CTRAJ_REC_ZGRID_MOL ? 256 : 128);
  if (block < 64 || block % 64) return fail(c, CTRAJ_EINVAL, "block must be a positive multiple of 64");
  int rc = lds_limit(c, (const void *)k, lds);
  if (rc) return rc;
  rc = block_ok(c, (const void *)k, block);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  const int grid = grid_for(c, (const void *)k, block, lds, RC->n, 0);
  KBeam B = make_kbeam(c->d_elems, c->n_elems, c->phys);
  hipLaunchKernelGGL(k, dim3(grid), dim3(block), lds, c->stream, B, kt, *RC);
  HIPCHK(c, hipGetLastError());
  return CTRAJ_OK;
}

int32_t ctraj_draw_gaussian(ctraj_ctx *c, uint64_t seed, uint64_t offset, int64_t n, int32_t nd, const double *mean,
                            const double *sigma, double *const *out) {
  if (!c || !mean || !sigma || !out) return fail(c, CTRAJ_EINVAL, "null argument");
  if (nd < 1 || nd > 8) return fail(c, CTRAJ_EINVAL, "n_dims must be 1..8");
  if (n < 0) return fail(c, CTRAJ_EINVAL, "n < 0");
  if (n == 0) return CTRAJ_OK;
  DrawArgs A{};
  for (int d = 0; d < nd; ++d) {
    A.mean[d] = mean[d];
    A.sigma[d] = sigma[d];
    A.out[d] = out[d];
  }
  HIPCHK(c, hipSetDevice(c->device));
  const int grid = (int)std::min<int64_t>((n + 255) / 256, (int64_t)c->n_cu * 16);
  hipLaunchKernelGGL(k_draw, dim3(grid), dim3(256), 0, c->stream, seed, offset, n, nd, A);
  HIPCHK(c, hipGetLastError());
  return CTRAJ_OK;
}

int32_t ctraj_draw_gaussian_at(ctraj_ctx *c, uint64_t seed, const int64_t *index, int64_t n, int32_t nd,
                               const double *mean, const double *sigma, double *const *out) {
  if (!c || !mean || !sigma || !out || (n > 0 && !index)) return fail(c, CTRAJ_EINVAL, "null argument");
  if (nd < 1 || nd > 8) return fail(c, CTRAJ_EINVAL, "n_dims must be 1..8");
  if (n < 0) return fail(c, CTRAJ_EINVAL, "n < 0");
  if (n == 0) return CTRAJ_OK;
  DrawArgs A{};
  for (int d = 0; d < nd; ++d) {
    A.mean[d] = mean[d];
    A.sigma[d] = sigma[d];
    A.out[d] = out[d];
  }
  HIPCHK(c, hipSetDevice(c->device));
  const int grid = (int)std::min<int64_t>((n + 255) / 256, (int64_t)c->n_cu * 16);
  hipLaunchKernelGGL(k_draw_at, dim3(grid), dim3(256), 0, c->stream, seed, index, n, nd, A);
  HIPCHK(c, hipGetLastError());
  return CTRAJ_OK;
}

int32_t ctraj_draw_uniform_u8_at(ctraj_ctx *c, uint64_t seed, const int64_t *index, int64_t n, int32_t n_values,
                                 uint8_t *out) {
  if (!c || !out || (n > 0 && !index)) return fail(c, CTRAJ_EINVAL, "null argument");
  if (n_values < 1 || n_values > 256) return fail(c, CTRAJ_EINVAL, "n_values must be 1..256");
  if (n < 0) return fail(c, CTRAJ_EINVAL, "n < 0");
  if (n == 0) return CTRAJ_OK;
  HIPCHK(c, hipSetDevice(c->device));
  const int grid = (int)std::min<int64_t>((n + 255) / 256, (int64_t)c->n_cu * 16);
  hipLaunchKernelGGL(k_draw_u8_at, dim3(grid), dim3(256), 0, c->stream, seed, index, n, (uint32_t)n_values, out);
  HIPCHK(c, hipGetLastError());
  return CTRAJ_OK;
}

int32_t ctraj_draw_uniform_u8(ctraj_ctx *c, uint64_t seed, uint64_t offset, int64_t n, int32_t n_values,
                              uint8_t *out) {
  if (!c || !out) return fail(c, CTRAJ_EINVAL, "null argument");
  if (n_values < 1 || n_values > 256) return fail(c, CTRAJ_EINVAL, "n_values must be 1..256");
  if (n < 0) return fail(c, CTRAJ_EINVAL, "n < 0");
  if (n == 0) return CTRAJ_OK;
  HIPCHK(c, hipSetDevice(c->device));
  const int grid = (int)std::min<int64_t>((n + 255) / 256, (int64_t)c->n_cu * 16);
  hipLaunchKernelGGL(k_draw_u8, dim3(grid), dim3(256), 0, c->stream, seed, offset, n, (uint32_t)n_values, out);
  HIPCHK(c, hipGetLastError());
  return CTRAJ_OK;
}

int32_t ctraj_draw_disc(ctraj_ctx *c, uint64_t seed, uint64_t offset, int64_t n, double radius, double z, double *x,
                        double *y, double *z_out) {
  if (!c || ((!x || !y) && n > 0)) return fail(c, CTRAJ_EINVAL, "null argument");
  if (n < 0) return fail(c, CTRAJ_EINVAL, "n < 0");
  if (!(radius >= 0.0) || !std::isfinite(radius)) return fail(c, CTRAJ_EINVAL, "radius must be finite and >= 0");
  if (n == 0) return CTRAJ_OK;
  HIPCHK(c, hipSetDevice(c->device));
  const int grid = (int)std::min<int64_t>((n + 255) / 256, (int64_t)c->n_cu * 16);
  hipLaunchKernelGGL(k_draw_disc, dim3(grid), dim3(256), 0, c->stream, seed, offset, n, radius, z, x, y, z_out);
  HIPCHK(c, hipGetLastError());
  return CTRAJ_OK;
}

// The index range is checked on the host before launch (one copy of the index array), so a bad
// index is an EINVAL, never an out-of-bounds read.
static int check_gather_index(ctraj_ctx *c, int64_t n, const int64_t *index, int64_t base, int64_t src_len) {
  std::vector<int64_t> h((size_t)n);
  HIPCHK(c, hipMemcpyAsync(h.data(), index, sizeof(int64_t) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int64_t j = 0; j < n; ++j)
    if (h[j] - base < 0 || h[j] - base >= src_len)
      return fail(c, CTRAJ_EINVAL, "gather index " + std::to_string(h[j]) + " outside [index_base, index_base + src_len)");
  return CTRAJ_OK;
}

int32_t ctraj_gather_f64(ctraj_ctx *c, int64_t n, const int64_t *index, int64_t index_base, int64_t src_len,
                         int32_t n_fields, const double *const *src, double *const *dst) {
  if (!c || !src || !dst || (n > 0 && !index)) return fail(c, CTRAJ_EINVAL, "null argument");
  if (n_fields < 1 || n_fields > 8) return fail(c, CTRAJ_EINVAL, "n_fields must be 1..8");
  if (n < 0 || src_len < 0) return fail(c, CTRAJ_EINVAL, "negative size");
  if (n == 0) return CTRAJ_OK;
  GatherArgs G{};
  for (int f = 0; f < n_fields; ++f) {
    if (dst[f] && !src[f]) return fail(c, CTRAJ_EINVAL, "gather field with a destination but no source");
    G.src[f] = src[f];
    G.dst[f] = dst[f];
  }
  HIPCHK(c, hipSetDevice(c->device));
  int rc = check_gather_index(c, n, index, index_base, src_len);
  if (rc != CTRAJ_OK) return rc;
  const int grid = (int)std::min<int64_t>((n + 255) / 256, (int64_t)c->n_cu * 16);
  hipLaunchKernelGGL(k_gather_f64, dim3(grid), dim3(256), 0, c->stream, n, index, index_base, n_fields, G);
  HIPCHK(c, hipGetLastError());
  return CTRAJ_OK;
}

int32_t ctraj_gather_u8(ctraj_ctx *c, int64_t n, const int64_t *index, int64_t index_base, int64_t src_len,
                        const uint8_t *src, uint8_t *dst) {
  if (!c || !src || !dst || (n > 0 && !index)) return fail(c, CTRAJ_EINVAL, "null argument");
  if (n < 0 || src_len < 0) return fail(c, CTRAJ_EINVAL, "negative size");
  if (n == 0) return CTRAJ_OK;
  HIPCHK(c, hipSetDevice(c->device));
  int rc = check_gather_index(c, n, index, index_base, src_len);
  if (rc != CTRAJ_OK) return rc;
  const int grid = (int)std::min<int64_t>((n + 255) / 256, (int64_t)c->n_cu * 16);
  hipLaunchKernelGGL(k_gather_u8, dim3(grid), dim3(256), 0, c->stream, n, index, index_base, src, dst);
  HIPCHK(c, hipGetLastError());
  return CTRAJ_OK;
}

}  // extern "C"
