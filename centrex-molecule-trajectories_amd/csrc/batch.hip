// batch.hip -- many beamline configurations in one persistent, graph-replayed launch (gfx950).
//
// The molecules of all configurations form one index space [0, total): configuration c owns
// [begin_c, begin_c + n_c); a molecule finds its configuration by a binary search of the begin
// offsets (LDS) and draws itself from that configuration's Philox stream (the ctraj_draw_gaussian
// stream).  By default the scan runs as split kernels per slice of at most 2^27 molecules, like the
// single-run path: k_batch_pre draws, walks each molecule to and through its configuration's lens
// entrance and queues the lens molecules; k_batch_integ integrates the queue with lane refill and
// the configuration's lens constants per lane; k_batch_post walks the molecules that left their
// lens through the rest of their configuration.  k_batch_lanes does all of it in one lane-refill
// kernel, wavefronts claiming `chunk` indices at a time from one global counter (reset by a fill
// kernel inside the captured graph).  Tallies are (configuration, label) histograms in LDS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "ctraj_device.hpp"
#include "ctraj_internal.hpp"
#include "ctraj_lanes.hpp"
#include "ctraj_wave.hpp"

using namespace ctraj;

namespace {

constexpr int kMaxBins = 4096;  // configurations x labels tallied in LDS
constexpr int64_t kPipeMax = (int64_t)1 << 27;  // molecules per slice of the split path (64 B of queue each)


struct BCfg {
  int64_t begin, n;
  uint64_t seed;
  double mean[6], sigma[6];
  int32_t e0, ne, lens, pad;  // element slice; lens index within the slice (-1: none)
};

// Lens constants of one configuration, staged in the split path's lens kernel's LDS.
struct BLensC {
  double ks, cu, R2, z1, Hs, Ll, gk;  // gk = -g/ks where the lens folds ks into its step (hex_folds), else 0
  int32_t label_inside, e0, ne, lens;  // element slice and the lens's index in it
};

struct BArgs {
  const BCfg *cfgs;
  const BLensC *lensk;
  const ctraj_elem *elems;
  unsigned long long *counter;
  int64_t *tally;               // [n_cfg][n_labels]
  unsigned long long *work;     // [CTRAJ_WORK_N]
  int64_t total;
  double g;
  int n_cfg, n_labels, label_detected, cap, chunk, refill_min;
  double Hs;                    // the lens dt shared by every configuration (split path's lens kernel, UH)
  int off_begin, off_hist;      // in doubles from the LDS base (off_begin: the lens constants in the lens kernel)
};

// Split scan: the PRE kernel queues the molecules that enter their configuration's lens into
// per-wavefront regions (the single-run queue layout, key = configuration index); the lens kernel
// claims 64-entry chunks from eight per-XCD counters.
struct BPipe {
  longlong2 *q;
  uint32_t *wcount;             // entries queued by each PRE wavefront
  unsigned long long *claim;    // [8][32]
  int64_t region;
  int n_pre_waves;
  int64_t s0, s1;               // the slice of batch molecule indices [s0, s1) this PRE launch walks
  // the lens kernel counts the molecules the bore stops into the tallies itself and lists the queue
  // slots of those that leave their lens, the only ones phase 3 reads: [0] their count (zeroed
  // before the lens kernel), the list from word BLT_LIST
  uint32_t *lt;
};
constexpr int BLT_LIST = 64;
constexpr int BPOST_ITEMS = 4;  // phase 3: listed slots per thread

struct BLane {
  int c, e0, ne, lens, label_inside;
  double ks, cu, R2, z1, gk, hs;  // hs: the lens's full step dt
};

__device__ __forceinline__ int find_cfg(const int64_t *s_begin, int n_cfg, int64_t gi) {
  int lo = 0, hi = n_cfg;  // largest c with begin[c] <= gi
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (s_begin[mid] <= gi) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Molecule j of configuration C from its Philox stream, walked through the configuration's
// elements `el` before its lens and through the lens entrance.  Returns true if it enters the lens;
// otherwise it has finished with label `lab`.
template <int KM>
__device__ __forceinline__ bool batch_start(const BCfg &C, const ctraj_elem *el, uint64_t j, double g,
                                            int label_detected, MState &m, int &lab, WalkCounters &wc) {
  double z[6];
  gaussian_pair(C.seed, j, 0u, z[0], z[1]);
  gaussian_pair(C.seed, j, 1u, z[2], z[3]);
  gaussian_pair(C.seed, j, 2u, z[4], z[5]);
  m.x = C.mean[0] + C.sigma[0] * z[0];
  m.y = C.mean[1] + C.sigma[1] * z[1];
  m.z = C.mean[2] + C.sigma[2] * z[2];
  m.vx = C.mean[3] + C.sigma[3] * z[3];
  m.vy = C.mean[4] + C.sigma[4] * z[4];
  m.vz = C.mean[5] + C.sigma[5] * z[5];
  m.t = 0.0;
  m.ax = 0.0;
  m.ay = -g;
  m.az = 0.0;
  lab = label_detected;
  const bool ok = walk_plain<KM, true>(el, 0, C.lens >= 0 ? C.lens : C.ne, g, m, lab, wc);
  if (!ok || C.lens < 0) return false;
  const ctraj_elem &L = el[C.lens];
  drift_dt(m, time_to(m, L.z0), g);
  ++wc.faces;
  if (rho_of(m.x, m.y) > L.p[0]) {
    lab = L.label_hit;
    return false;
  }
  return true;
}

// (configuration, label) tally of the lanes with `fin` set into the workgroup's LDS histogram:
// one atomic per distinct key in the wavefront.  Every lane takes part.
__device__ __forceinline__ void batch_hist(uint32_t *s_hist, bool fin, int key, int lane) {
  uint64_t todo = __ballot(fin);
  while (todo) {
    const int leader = __ffsll((unsigned long long)todo) - 1;
    const int lk = __shfl(key, leader);
    const uint64_t same = __ballot(key == lk) & todo;
    if (lane == leader) atomicAdd(&s_hist[lk], (uint32_t)__popcll(same));
    todo &= ~same;
  }
}

__global__ void __launch_bounds__(256) k_batch_lanes(BArgs A, KTab KT) {
  extern __shared__ __attribute__((aligned(16))) double lds[];
  __shared__ unsigned long long s_work[4];
  const Tables T = stage_tables(lds, KT);
  int64_t *s_begin = reinterpret_cast<int64_t *>(lds + A.off_begin);
  uint32_t *s_hist = reinterpret_cast<uint32_t *>(lds + A.off_hist);
  const int nbins = A.n_cfg * A.n_labels;
  for (int i = threadIdx.x; i < A.n_cfg; i += blockDim.x) s_begin[i] = A.cfgs[i].begin;
  for (int b = threadIdx.x; b < nbins; b += blockDim.x) s_hist[b] = 0;
  if (threadIdx.x < 4) s_work[threadIdx.x] = 0;
  __syncthreads();

  const WsRef wsb = ws_of(T, 0);
  const int lane = lane_id();
  const double g = A.g;
  WalkCounters wc{0u, 0u};
  uint32_t n_oot = 0, n_cap = 0;
  unsigned long long steps_w = 0;
  int64_t next = 0, end = 0;
  bool exhausted = false;
  LaneMol s;
  s.phase = PH_NEED;
  s.fin = 0;
  s.label = 0;
  s.status = 0;
  s.wsb = wsb;
  BLane b{0, 0, 0, -1, 0, 0.0, 0.0, 0.0, 0.0};

  for (;;) {
    const uint64_t lensm = __ballot(s.phase == PH_LENS);
    const uint64_t idlem = __ballot(s.phase == PH_NEED || s.phase == PH_POST);
    if (lensm == 0 || __popcll(idlem) >= A.refill_min) {
      for (;;) {
        if (s.phase == PH_POST) {
          s.m.ax = 0.0;  // the lens exit row stores (0, -g, 0)
          s.m.ay = -g;
          s.m.az = 0.0;
          int lab = A.label_detected;
          walk_plain(opaque(A.elems) + b.e0, b.lens + 1, b.ne, g, s.m, lab, wc);
          s.label = lab;
          s.fin = 1;
          s.phase = PH_NEED;
        }
        if (__ballot(s.fin)) {
          if (s.fin) {
            n_oot += (s.status & CTRAJ_ST_OOT) ? 1u : 0u;
            n_cap += (s.status & CTRAJ_ST_CAPPED) ? 1u : 0u;
          }
          // lanes without a finished molecule carry a key no one else uses
          batch_hist(s_hist, s.fin != 0, s.fin ? b.c * A.n_labels + s.label : nbins + lane, lane);
          s.fin = 0;
        }
        const uint64_t need = __ballot(s.phase == PH_NEED);
        if (!need || exhausted) {
          if (s.phase == PH_NEED && exhausted) s.phase = PH_DONE;
          break;
        }
        if (next >= end) {
          unsigned long long c = 0;
          if (lane == 0) c = atomicAdd(A.counter, (unsigned long long)A.chunk);
          c = __shfl(c, 0);
          next = (int64_t)c;
          end = next + A.chunk < A.total ? next + A.chunk : A.total;
          if (next >= end) {
            exhausted = true;
            continue;
          }
        }
        const int64_t avail = end - next;
        if (s.phase == PH_NEED) {
          const int rank = lanes_below(need);
          if (rank < avail) {
            const int64_t gi = next + rank;
            const int c = find_cfg(s_begin, A.n_cfg, gi);
            const BCfg &C = opaque(A.cfgs)[c];
            const uint64_t j = (uint64_t)(gi - C.begin);
            double z[6];  // the arithmetic of batch_start, written out: inlining the helper here costs registers
            gaussian_pair(C.seed, j, 0u, z[0], z[1]);
            gaussian_pair(C.seed, j, 1u, z[2], z[3]);
            gaussian_pair(C.seed, j, 2u, z[4], z[5]);
            MState &m = s.m;
            m.x = C.mean[0] + C.sigma[0] * z[0];
            m.y = C.mean[1] + C.sigma[1] * z[1];
            m.z = C.mean[2] + C.sigma[2] * z[2];
            m.vx = C.mean[3] + C.sigma[3] * z[3];
            m.vy = C.mean[4] + C.sigma[4] * z[4];
            m.vz = C.mean[5] + C.sigma[5] * z[5];
            m.t = 0.0;
            m.ax = 0.0;
            m.ay = -g;
            m.az = 0.0;
            b.c = c;
            b.e0 = C.e0;
            b.ne = C.ne;
            b.lens = C.lens;
            s.status = 0;
            const ctraj_elem *el = opaque(A.elems) + b.e0;
            int lab = A.label_detected;
            bool ok = walk_plain(el, 0, b.lens >= 0 ? b.lens : b.ne, g, m, lab, wc);
            if (ok && b.lens >= 0) {
              const ctraj_elem &L = el[b.lens];
              drift_dt(m, time_to(m, L.z0), g);
              ++wc.faces;
              if (rho_of(m.x, m.y) > L.p[0]) {
                ok = false;
                lab = L.label_hit;
              }
            }
            if (ok && b.lens >= 0) {
              const ctraj_elem &L = el[b.lens];
              const double Hs = L.p[2], Ll = L.p[3];
              b.ks = L.p[4];
              b.cu = L.p[5];
              b.gk = hex_folds(b.ks) ? -g / b.ks : 0.0;
              b.R2 = L.p[1];
              b.z1 = L.z1;
              b.label_inside = L.label_inside;
              b.hs = Hs;
              s.i = 0;
              s.nfull = rk4_full_steps(Ll, m.vz, Hs, A.cap, s.status);
              s.dzs = m.vz * Hs;
              s.hp = Ll / m.vz - (double)s.nfull * Hs;
              s.h = s.nfull == 0 ? s.hp : Hs;
              s.r2 = __builtin_fma(m.x, m.x, m.y * m.y);
              s.phase = PH_LENS;
            }
            if (s.phase != PH_LENS) {
              s.label = lab;
              s.fin = 1;
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
      const Force<CTRAJ_FORCE_WSPLINE_HEX> F{b.ks, b.cu, g, b.gk};
      const bool last = (s.i == s.nfull);
      lens_step<CTRAJ_FORCE_WSPLINE_HEX>(F, T, wsb, s.h, m.x, m.y, m.vx, m.vy, s.r2, s.status);
      s.r2 = __builtin_fma(m.x, m.x, m.y * m.y);
      ++s.i;
      s.h = (s.i == s.nfull) ? s.hp : s.h;
      const bool dead = s.r2 > b.R2;
      // z and t stay at their entry values until the molecule finishes (the walker's definition)
      if (last) {
        m.z = b.z1;
        m.t = __builtin_fma((double)s.nfull, b.hs, m.t) + s.hp;
      } else if (dead) {
        m.z = __builtin_fma((double)s.i, s.dzs, m.z);
        m.t = __builtin_fma((double)s.i, b.hs, m.t);
      }
      s.label = dead ? b.label_inside : s.label;
      s.fin = dead ? 2 : 0;
      s.phase = dead ? (int)PH_NEED : (last ? (int)PH_POST : (int)PH_LENS);
    }
  }

  if (wc.faces) atomicAdd(&s_work[0], (unsigned long long)wc.faces);
  if (lane == 0 && steps_w) atomicAdd(&s_work[1], steps_w);
  if (n_oot) atomicAdd(&s_work[2], (unsigned long long)n_oot);
  if (n_cap) atomicAdd(&s_work[3], (unsigned long long)n_cap);
  __syncthreads();
  for (int k = threadIdx.x; k < nbins; k += blockDim.x)
    if (s_hist[k]) atomicAdd((unsigned long long *)&A.tally[k], (unsigned long long)s_hist[k]);
  if (threadIdx.x < 4 && s_work[threadIdx.x]) atomicAdd(&A.work[threadIdx.x], s_work[threadIdx.x]);
}

// Split scan, phase 1: every molecule of the slice [s0, s1) is drawn and walked to and through
// its configuration's lens entrance (grid-stride, wave-uniform trip count); the molecules that enter
// a lens are queued in the wavefront's region, the others are tallied.
template <int KM>
__global__ void __launch_bounds__(256) k_batch_pre(BArgs A, BPipe P) {
  extern __shared__ __attribute__((aligned(16))) double lds[];
  __shared__ unsigned long long s_work[4];
  int64_t *s_begin = reinterpret_cast<int64_t *>(lds);
  uint32_t *s_hist = reinterpret_cast<uint32_t *>(lds + A.n_cfg);
  const int nbins = A.n_cfg * A.n_labels;
  for (int i = threadIdx.x; i < A.n_cfg; i += blockDim.x) s_begin[i] = A.cfgs[i].begin;
  for (int b = threadIdx.x; b < nbins; b += blockDim.x) s_hist[b] = 0;
  if (threadIdx.x < 4) s_work[threadIdx.x] = 0;
  __syncthreads();

  const int lane = lane_id();
  const double g = A.g;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t gwave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  WalkCounters wc{0u, 0u};
  int64_t qn = 0;
  for (int64_t wbase = P.s0 + (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63); wbase < P.s1; wbase += stride) {
    const int64_t gi = wbase + lane;
    const bool active = gi < P.s1;
    MState m;
    int lab = 0, c = 0;
    bool enq = false;
    if (active) {
      c = find_cfg(s_begin, A.n_cfg, gi);
      const BCfg &C = opaque(A.cfgs)[c];
      enq = batch_start<KM>(C, opaque(A.elems) + C.e0, (uint64_t)(gi - C.begin), g, A.label_detected, m, lab, wc);
    }
    const uint64_t qm = __ballot(enq);
    if (enq) queue_put(P.q, gwave * P.region + qn + lanes_below(qm), m, c, 0);
    qn += __popcll(qm);
    const bool fin = active && !enq;
    if (__ballot(fin)) batch_hist(s_hist, fin, fin ? c * A.n_labels + lab : nbins + lane, lane);
  }
  if (lane == 0) P.wcount[gwave] = (uint32_t)qn;
  if (wc.faces) atomicAdd(&s_work[0], (unsigned long long)wc.faces);
  __syncthreads();
  for (int k = threadIdx.x; k < nbins; k += blockDim.x)
    if (s_hist[k]) atomicAdd((unsigned long long *)&A.tally[k], (unsigned long long)s_hist[k]);
  if (threadIdx.x < 4 && s_work[threadIdx.x]) atomicAdd(&A.work[threadIdx.x], s_work[threadIdx.x]);
}

// Split scan, phase 2: persistent wavefronts integrate the queued molecules through their
// configuration's lens with lane refill and write each molecule's state where the lens left it back
// into its queue entry, with its outcome (the single-run lens kernel's scheme, with the lens
// constants per lane: the configuration's lens constants staged in LDS, its field constants and
// step in the lane's registers).  A lane's rem is its full steps left (> 0), 0 at the lens end (one
// partial step of L/vz - n dt to go), REM_STOPPED once the bore stopped it, REM_EMPTY when it holds
// no molecule.  UH: every configuration's lens has the same dt and folds its force scale (the folded
// step's constants, hex_k of the lens's ks and dt, are formed per molecule); otherwise each step
// chooses the folded or the plain RK4 step for its molecule's lens (lens_step).  The arithmetic per
// molecule is the one of k_batch_lanes.
constexpr int BREM_EMPTY = -1, BREM_STOPPED = -2;
enum : int { BLENS_STOPPED = 1, BLENS_LEFT = 2 };

template <bool UH>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(UH ? 6 : 1)))
k_batch_integ(BArgs A, BPipe P, KTab KT) {
  extern __shared__ __attribute__((aligned(16))) double lds[];
  __shared__ unsigned long long s_steps;
  if (threadIdx.x == 0) s_steps = 0;
  const Tables T = stage_tables(lds, KT);
  BLensC *s_cfg = reinterpret_cast<BLensC *>(lds + A.off_begin);
  // per configuration, the molecules the bore stopped; then the out-of-table and capped ones among them
  uint32_t *s_stop = reinterpret_cast<uint32_t *>(lds + A.off_hist);
  for (int i = threadIdx.x; i < A.n_cfg; i += blockDim.x) s_cfg[i] = A.lensk[i];
  for (int i = threadIdx.x; i < A.n_cfg + 2; i += blockDim.x) s_stop[i] = 0;
  __syncthreads();

  const WsRef wsb = ws_of(T, 0);
  const int lane = lane_id();
  const double g = A.g;
  const uint32_t cpr = (uint32_t)(P.region >> 6);
  const unsigned long long n_chunks = (unsigned long long)(P.n_pre_waves >> 3) * cpr;
  int grp = blockIdx.x & 7, tried = 0;
  uint32_t steps_w = 0;  // wave-uniform
  int64_t next = 0, end = 0;
  bool exhausted = false;

  MState m;
  m.x = m.y = m.z = m.vx = m.vy = m.vz = m.t = 0.0;
  // z and t stay at their entry values while the molecule steps: a stopped one is only counted, and
  // one that leaves takes z1 and t_in + n dt + hp (the walker's definition)
  double r2 = 0.0, ks = 0.0, cu = 0.0, gk = 0.0, R2 = 0.0, Hs = A.Hs;
  HexK hk{};  // UH: the folded step's constants of the molecule's lens
  int rem = BREM_EMPTY, nfull = 0, c = 0;
  uint32_t status = 0;
  int64_t slot = 0, key = 0;

  for (;;) {
    bool act = rem > 0;
    uint64_t am = __builtin_amdgcn_ballot_w64(act);
    if (64 - __popcll(am) >= A.refill_min) {
      for (;;) {
        // finish the molecules that take no more full steps: stopped by the bore, or at the lens end
        // with its partial step to take
        const bool fin = rem == 0 || rem == BREM_STOPPED;
        steps_w += (uint32_t)__popcll(__builtin_amdgcn_ballot_w64(rem == 0));
        if (fin) {
          int outcome = BLENS_STOPPED;
          if (rem == 0) {
            const BLensC &K = s_cfg[c];
            const double hp = K.Ll / m.vz - (double)nfull * Hs;
            const Force<CTRAJ_FORCE_WSPLINE_HEX> F{ks, cu, g, gk};
            lens_step<CTRAJ_FORCE_WSPLINE_HEX>(F, T, wsb, hp, m.x, m.y, m.vx, m.vy, r2, status);
            m.z = K.z1;
            m.t = __builtin_fma((double)nfull, Hs, m.t) + hp;
            r2 = __builtin_fma(m.x, m.x, m.y * m.y);
            outcome = r2 > R2 ? BLENS_STOPPED : BLENS_LEFT;
          }
          if (outcome == BLENS_STOPPED) {  // counted here: the molecule needs no further walk
            atomicAdd(&s_stop[c], 1u);
            if (status) {
              if (status & CTRAJ_ST_OOT) atomicAdd(&s_stop[A.n_cfg], 1u);
              if (status & CTRAJ_ST_CAPPED) atomicAdd(&s_stop[A.n_cfg + 1], 1u);
            }
          } else {
            queue_put_done(P.q, slot, m.x, m.y, m.z, m.vx, m.vy, m.vz, m.t,
                           key | ((int64_t)status << 40) | ((int64_t)outcome << 48));
            // list the slot for phase 3: one atomic per wavefront, by its first lane here
            const uint64_t lm = __builtin_amdgcn_ballot_w64(true);
            const int rank = lanes_below(lm);
            uint32_t base = 0;
            if (rank == 0) base = atomicAdd(&P.lt[0], (uint32_t)__popcll(lm));
            base = __builtin_amdgcn_readfirstlane(base);
            P.lt[BLT_LIST + base + (uint32_t)rank] = (uint32_t)slot;
          }
          rem = BREM_EMPTY;
        }
        const uint64_t need = __builtin_amdgcn_ballot_w64(rem == BREM_EMPTY);
        if (!need || exhausted) break;
        if (next >= end) {
          unsigned long long cl = 0;
          if (lane == 0) cl = atomicAdd(P.claim + 32 * grp, 1ull);
          cl = uniform_u64(cl);  // every lane is active here, lane 0 first
          if (cl >= n_chunks) {
            if (++tried == 8) exhausted = true;
            grp = (grp + 1) & 7;
            continue;
          }
          const uint32_t r = (uint32_t)cl / cpr;
          const uint32_t w = ((r >> 2) << 5) | ((uint32_t)grp << 2) | (r & 3u);
          const int64_t base = (int64_t)((uint32_t)cl - r * cpr) * 64;
          const int64_t filled = (int64_t)P.wcount[w] - base;
          next = (int64_t)w * P.region + base;
          end = next + (filled < 64 ? filled : 64);  // next >= end: an empty chunk, claim again
          continue;
        }
        const int64_t avail = end - next;
        if (rem == BREM_EMPTY) {
          const int rank = lanes_below(need);
          if (rank < avail) {
            slot = next + rank;
            queue_get_done(P.q, slot, m, key);
            c = (int)(key & QKEY_INDEX);
            const BLensC &K = s_cfg[c];
            ks = K.ks;
            cu = K.cu;
            gk = K.gk;
            R2 = K.R2;
            if constexpr (!UH) Hs = K.Hs;
            if constexpr (UH) {
              hk = hex_k(ks, Hs);
            }
            status = 0;
            nfull = rk4_full_steps(K.Ll, m.vz, Hs, A.cap, status);
            r2 = __builtin_fma(m.x, m.x, m.y * m.y);
            rem = nfull;
          }
        }
        next += avail < (int64_t)__popcll(need) ? avail : (int64_t)__popcll(need);
      }
      act = rem > 0;
      am = __builtin_amdgcn_ballot_w64(act);
      if (!am) break;  // the queue is exhausted and every molecule has been written back
    }

    steps_w += (uint32_t)__popcll(am);
    if (act) {
      if constexpr (UH) {
        rk4_step_hex_k(wsb, T.umax, cu, gk, hk, m.x, m.y, m.vx, m.vy, r2);
      } else {
        const Force<CTRAJ_FORCE_WSPLINE_HEX> F{ks, cu, g, gk};
        lens_step<CTRAJ_FORCE_WSPLINE_HEX>(F, T, wsb, Hs, m.x, m.y, m.vx, m.vy, r2, status);
      }
      r2 = __builtin_fma(m.x, m.x, m.y * m.y);
      rem = r2 > R2 ? BREM_STOPPED : rem - 1;
    }
  }

  if (lane == 0 && steps_w) atomicAdd(&s_steps, (unsigned long long)steps_w);
  __syncthreads();
  if (threadIdx.x == 0 && s_steps) atomicAdd(&A.work[1], s_steps);
  for (int i = threadIdx.x; i < A.n_cfg; i += blockDim.x)
    if (s_stop[i])
      atomicAdd((unsigned long long *)&A.tally[(int64_t)i * A.n_labels + s_cfg[i].label_inside],
                (unsigned long long)s_stop[i]);
  if (threadIdx.x < 2 && s_stop[A.n_cfg + threadIdx.x])
    atomicAdd(&A.work[2 + threadIdx.x], (unsigned long long)s_stop[A.n_cfg + threadIdx.x]);
}

// Split scan, phase 3: the molecules the lens kernel listed as leaving their lens (grid-stride,
// wave-uniform trip count) walk their configuration's elements after the lens (the lens exit row
// stores (0, -g, 0)); (configuration, label) tallied in LDS.  Those the bore stopped were tallied by
// the lens kernel.
template <int KM>
__global__ void __launch_bounds__(256) k_batch_post(BArgs A, BPipe P) {
  // as many workgroups as give each thread about BPOST_ITEMS listed slots take part (the grid is
  // sized for every slot of the queue; the listed ones are usually a few per cent of them), which
  // keeps the workgroups' final tally atomics few
  const int64_t n_items = (int64_t)P.lt[0];
  const int64_t per_wg = (int64_t)blockDim.x * BPOST_ITEMS;
  const int64_t n_wg = std::min<int64_t>(gridDim.x, std::max<int64_t>(1, (n_items + per_wg - 1) / per_wg));
  if (blockIdx.x >= n_wg) return;
  extern __shared__ __attribute__((aligned(16))) double lds[];
  __shared__ unsigned long long s_work[4];
  uint32_t *s_hist = reinterpret_cast<uint32_t *>(lds);
  const int nbins = A.n_cfg * A.n_labels;
  for (int b = threadIdx.x; b < nbins; b += blockDim.x) s_hist[b] = 0;
  if (threadIdx.x < 4) s_work[threadIdx.x] = 0;
  __syncthreads();

  const int lane = lane_id();
  const double g = A.g;
  const int64_t stride = n_wg * blockDim.x;
  WalkCounters wc{0u, 0u};
  uint32_t n_oot = 0, n_cap = 0;
  for (int64_t wbase = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63); wbase < n_items; wbase += stride) {
    const int64_t j = wbase + lane;
    const bool fin = j < n_items;
    int c = 0, label = 0;
    if (fin) {
      MState m;
      int64_t key;
      queue_get_done(P.q, (int64_t)P.lt[BLT_LIST + j], m, key);
      c = (int)(key & QKEY_INDEX);
      const uint32_t status = (uint32_t)((key >> 40) & 0xff);
      const BLensC &K = A.lensk[c];
      m.ax = 0.0;
      m.ay = -g;
      m.az = 0.0;
      label = A.label_detected;
      walk_plain<KM>(opaque(A.elems) + K.e0, K.lens + 1, K.ne, g, m, label, wc);
      n_oot += (status & CTRAJ_ST_OOT) ? 1u : 0u;
      n_cap += (status & CTRAJ_ST_CAPPED) ? 1u : 0u;
    }
    if (__ballot(fin)) batch_hist(s_hist, fin, fin ? c * A.n_labels + label : nbins + lane, lane);
  }
  if (wc.faces) atomicAdd(&s_work[0], (unsigned long long)wc.faces);
  if (n_oot) atomicAdd(&s_work[2], (unsigned long long)n_oot);
  if (n_cap) atomicAdd(&s_work[3], (unsigned long long)n_cap);
  __syncthreads();
  for (int k = threadIdx.x; k < nbins; k += blockDim.x)
    if (s_hist[k]) atomicAdd((unsigned long long *)&A.tally[k], (unsigned long long)s_hist[k]);
  if (threadIdx.x < 4 && s_work[threadIdx.x]) atomicAdd(&A.work[threadIdx.x], s_work[threadIdx.x]);
}

__global__ void __launch_bounds__(256) k_zero_u64(unsigned long long *p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0;
}

}  // namespace

struct ctraj_batch {
  ctraj_ctx *ctx = nullptr;
  int n_cfg = 0, n_labels = 0, label_detected = 0, cap = 0;
  int64_t total = 0;
  double g = 0.0;
  BCfg *d_cfgs = nullptr;
  ctraj_elem *d_elems = nullptr;
  unsigned long long *d_counter = nullptr;  // [1] queue, then [CTRAJ_WORK_N] work counters
  int64_t *d_tally = nullptr;
  double *d_ws = nullptr;  // the batch's own copy of the W' spline (state 0), [nseg][4]
  int ws_nseg = 0;
  // split path: per-configuration lens constants, the lens-entry queue and its counters
  BLensC *d_lensk = nullptr;
  void *d_pipe = nullptr, *d_pipe_cnt = nullptr;
  int pre_grid = 0, n_pre_waves = 0;
  int64_t region = 0;
  bool uniform_dt = false;  // every lens has the same dt and folds its force scale (hex_folds)
  double dt = 0.0;
  int64_t slice = 0;        // molecules per slice of the split path
  bool circ_only = false;   // only circular apertures (and lenses) in every configuration
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // captured launches by (block, chunk, grid, refill, fused)
  std::map<std::tuple<int, int, int, int, int>, hipGraphExec_t> graphs;
};

static int batch_launch(ctraj_batch *b, int block, int chunk, int grid, int refill, bool fused) {
  ctraj_ctx *c = b->ctx;
  KTab kt{};
  kt.ws = b->d_ws;
  kt.ws_nseg = b->ws_nseg;
  kt.ws_states = b->d_ws ? 1 : 0;
  kt.umax = b->ws_nseg > 0 ? std::nextafter((double)b->ws_nseg, 0.0) : 0.0;
  kt.rt = nullptr;
  kt.rt_total = 0;
  const size_t tab = ktab_lds_bytes(kt) / sizeof(double);
  BArgs A{};
  A.cfgs = b->d_cfgs;
  A.lensk = b->d_lensk;
  A.elems = b->d_elems;
  A.counter = b->d_counter;
  A.work = b->d_counter + 1;
  A.tally = b->d_tally;
  A.total = b->total;
  A.g = b->g;
  A.n_cfg = b->n_cfg;
  A.n_labels = b->n_labels;
  A.label_detected = b->label_detected;
  A.cap = b->cap;
  A.chunk = chunk;
  A.refill_min = refill;
  const size_t hist_bytes = (size_t)b->n_cfg * b->n_labels * sizeof(uint32_t);
  const int64_t nz = 1 + CTRAJ_WORK_N;
  const int64_t nt = (int64_t)b->n_cfg * b->n_labels;
  int rc = CTRAJ_OK;
  if (fused || !b->d_pipe) {
    A.off_begin = (int)tab;
    A.off_hist = (int)(tab + b->n_cfg);
    const size_t lds = (tab + b->n_cfg) * sizeof(double) + hist_bytes;
    if ((rc = lds_limit(c, (const void *)k_batch_lanes, lds)) != CTRAJ_OK) return rc;
    if ((rc = block_ok(c, (const void *)k_batch_lanes, block)) != CTRAJ_OK) return rc;
    if (grid <= 0) grid = grid_for(c, (const void *)k_batch_lanes, block, lds, (int64_t)1 << 40, 0);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, c->stream, b->d_counter, nz);
    hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)std::min<int64_t>((nt + 255) / 256, 1024)), dim3(256), 0,
                       c->stream, (unsigned long long *)b->d_tally, nt);
    hipLaunchKernelGGL(k_batch_lanes, dim3(grid), dim3(block), lds, c->stream, A, kt);
    HIPCHK(c, hipGetLastError());
    return CTRAJ_OK;
  }
  BPipe P{};
  P.q = (longlong2 *)b->d_pipe;
  P.claim = (unsigned long long *)b->d_pipe_cnt;
  P.wcount = (uint32_t *)((char *)b->d_pipe_cnt + 8 * 256);
  P.region = b->region;
  P.n_pre_waves = b->n_pre_waves;
  P.lt = (uint32_t *)((char *)b->d_pipe + (size_t)b->n_pre_waves * (size_t)b->region * 64);
  // phase 1: LDS = configuration offsets + histogram
  const size_t lds_pre = (size_t)b->n_cfg * sizeof(int64_t) + hist_bytes;
  const void *kp = b->circ_only ? (const void *)k_batch_pre<KM_CIRC> : (const void *)k_batch_pre<KM_ALL>;
  if ((rc = lds_limit(c, kp, lds_pre)) != CTRAJ_OK) return rc;
  // phase 2: LDS = tables + lens constants
  const size_t cfg_dbl = (size_t)b->n_cfg * (sizeof(BLensC) / sizeof(double));
  A.off_begin = (int)tab;
  A.off_hist = (int)(tab + cfg_dbl);  // the lens kernel's stopped-molecule counts
  const size_t lds_lens = (tab + cfg_dbl) * sizeof(double) + (size_t)(b->n_cfg + 2) * sizeof(uint32_t);
  const void *kl = b->uniform_dt ? (const void *)k_batch_integ<true> : (const void *)k_batch_integ<false>;
  A.Hs = b->dt;
  if ((rc = lds_limit(c, kl, lds_lens)) != CTRAJ_OK) return rc;
  if ((rc = block_ok(c, kl, block)) != CTRAJ_OK) return rc;
  if (grid <= 0) grid = grid_for(c, kl, block, lds_lens, (int64_t)1 << 40, 0);
  // phase 3: LDS = histogram
  const void *kq = b->circ_only ? (const void *)k_batch_post<KM_CIRC> : (const void *)k_batch_post<KM_ALL>;
  if ((rc = lds_limit(c, kq, hist_bytes)) != CTRAJ_OK) return rc;
  const int gpost = grid_for(c, kq, 256, hist_bytes, (int64_t)b->n_pre_waves * b->region, 0);
  // queue counters and tallies are reset by fill kernels (no memset nodes in the captured graph)
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(256), 0, c->stream, (unsigned long long *)b->d_pipe_cnt,
                     (int64_t)8 * 32);
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, c->stream, b->d_counter, nz);
  hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)std::min<int64_t>((nt + 255) / 256, 1024)), dim3(256), 0, c->stream,
                     (unsigned long long *)b->d_tally, nt);
  using PreK = void (*)(BArgs, BPipe);
  using LensK = void (*)(BArgs, BPipe, KTab);
  // slices of at most b->slice molecules reuse the queue; the tallies and work counters accumulate
  for (int64_t s0 = 0; s0 < b->total; s0 += b->slice) {
    P.s0 = s0;
    P.s1 = std::min(b->total, s0 + b->slice);
    if (s0 > 0)
      hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(256), 0, c->stream, (unsigned long long *)b->d_pipe_cnt,
                         (int64_t)8 * 32);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, c->stream, (unsigned long long *)P.lt, (int64_t)BLT_LIST / 2);
    hipLaunchKernelGGL((PreK)kp, dim3(b->pre_grid), dim3(256), lds_pre, c->stream, A, P);
    HIPCHK(c, hipGetLastError());
    hipLaunchKernelGGL((LensK)kl, dim3(grid), dim3(block), lds_lens, c->stream, A, P, kt);
    HIPCHK(c, hipGetLastError());
    hipLaunchKernelGGL((PreK)kq, dim3(gpost), dim3(256), hist_bytes, c->stream, A, P);
    HIPCHK(c, hipGetLastError());
  }
  return CTRAJ_OK;
}

extern "C" {

int32_t ctraj_batch_create(ctraj_ctx *c, const ctraj_elem *elems, int32_t n_elems_total, const ctraj_batch_cfg *cfgs,
                           int32_t n_cfgs, const ctraj_phys *phys, ctraj_batch **out) {
  if (out) *out = nullptr;
  if (!c || !elems || !cfgs || !phys || !out) return fail(c, CTRAJ_EINVAL, "null argument");
  if (n_cfgs < 1 || n_elems_total < 1) return fail(c, CTRAJ_EINVAL, "need at least one configuration and element");
  int rc = check_phys(c, phys);
  if (rc != CTRAJ_OK) return rc;
  if ((int64_t)n_cfgs * phys->n_labels > kMaxBins)
    return fail(c, CTRAJ_EINVAL, "configurations x labels exceeds " + std::to_string(kMaxBins) + "; split the batch");
  std::vector<BCfg> hc(n_cfgs);
  int64_t total = 0;
  bool any_lens = false;
  for (int k = 0; k < n_cfgs; ++k) {
    const ctraj_batch_cfg &C = cfgs[k];
    if (C.elem_begin < 0 || C.n_elems < 0 || C.elem_begin + C.n_elems > n_elems_total)
      return fail(c, CTRAJ_EINVAL, "configuration " + std::to_string(k) + ": element slice out of range");
    if (C.n < 0) return fail(c, CTRAJ_EINVAL, "configuration " + std::to_string(k) + ": n < 0");
    int force, integ;
    if ((rc = check_elems(c, elems + C.elem_begin, C.n_elems, phys, &force, &integ)) != CTRAJ_OK) return rc;
    int lens = -1;
    for (int e = 0; e < C.n_elems; ++e) {
      const ctraj_elem &E = elems[C.elem_begin + e];
      if (E.kind != CTRAJ_LENS) continue;
      if (lens >= 0) return fail(c, CTRAJ_EINVAL, "configuration " + std::to_string(k) + ": more than one lens");
      if (E.force != CTRAJ_FORCE_WSPLINE_HEX || E.integrator != CTRAJ_INTEG_RK4)
        return fail(c, CTRAJ_EINVAL, "batched lenses use the hexapole W' spline and classical RK4");
      lens = e;
    }
    any_lens |= lens >= 0;
    BCfg &B = hc[k];
    B.begin = total;
    B.n = C.n;
    B.seed = C.seed;
    for (int d = 0; d < 6; ++d) {
      B.mean[d] = C.mean[d];
      B.sigma[d] = C.sigma[d];
    }
    B.e0 = C.elem_begin;
    B.ne = C.n_elems;
    B.lens = lens;
    B.pad = 0;
    total += C.n;
  }
  if (any_lens && !c->d_ws) return fail(c, CTRAJ_ESTATE, "lens W spline not set");
  HIPCHK(c, hipSetDevice(c->device));
  ctraj_batch *b = new ctraj_batch();
  b->ctx = c;
  b->n_cfg = n_cfgs;
  b->n_labels = phys->n_labels;
  b->label_detected = phys->label_detected;
  b->cap = phys->max_lens_steps > 0 ? phys->max_lens_steps : (1 << 20);
  b->g = phys->g;
  b->total = total;
  const size_t nt = (size_t)n_cfgs * phys->n_labels;
  if (hipMalloc((void **)&b->d_cfgs, sizeof(BCfg) * n_cfgs) != hipSuccess ||
      hipMalloc((void **)&b->d_elems, sizeof(ctraj_elem) * n_elems_total) != hipSuccess ||
      hipMalloc((void **)&b->d_counter, sizeof(unsigned long long) * (1 + CTRAJ_WORK_N)) != hipSuccess ||
      hipMalloc((void **)&b->d_tally, sizeof(int64_t) * nt) != hipSuccess ||
      hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
    (void)hipGetLastError();
    ctraj_batch_destroy(b);
    return fail(c, CTRAJ_ENOMEM, "batch allocation failed");
  }
  if (any_lens) {
    b->ws_nseg = c->ws_nseg;
    if (hipMalloc((void **)&b->d_ws, sizeof(double) * 4 * b->ws_nseg) != hipSuccess) {
      (void)hipGetLastError();
      ctraj_batch_destroy(b);
      return fail(c, CTRAJ_ENOMEM, "batch allocation failed");
    }
  }
  // split path: lens constants per configuration and a lens-entry queue sized for the batch
  std::vector<BLensC> hl(n_cfgs);
  b->uniform_dt = true;
  bool have_dt = false;
  bool all_fold = true;
  b->circ_only = true;
  for (int e = 0; e < n_elems_total; ++e)
    if (elems[e].kind != CTRAJ_CIRC && elems[e].kind != CTRAJ_LENS) b->circ_only = false;
  for (int k = 0; k < n_cfgs; ++k) {
    BLensC &K = hl[k];
    K = BLensC{};
    K.e0 = hc[k].e0;
    K.ne = hc[k].ne;
    K.lens = hc[k].lens;
    if (hc[k].lens >= 0) {
      const ctraj_elem &L = elems[hc[k].e0 + hc[k].lens];
      K.ks = L.p[4];
      K.cu = L.p[5];
      K.R2 = L.p[1];
      K.z1 = L.z1;
      K.Hs = L.p[2];
      K.Ll = L.p[3];
      K.gk = hex_folds(K.ks) ? -b->g / K.ks : 0.0;
      all_fold = all_fold && hex_folds(K.ks);
      K.label_inside = L.label_inside;
      if (!have_dt) {
        b->dt = K.Hs;
        have_dt = true;
      } else if (K.Hs != b->dt) {
        b->uniform_dt = false;
      }
    }
  }
  b->uniform_dt = b->uniform_dt && all_fold;
  if (any_lens && total > 0) {
    const int64_t slice = std::min(total, kPipeMax);
    const size_t lds_pre = (size_t)n_cfgs * sizeof(int64_t) + nt * sizeof(uint32_t);
    const void *kp = b->circ_only ? (const void *)k_batch_pre<KM_CIRC> : (const void *)k_batch_pre<KM_ALL>;
    int gpre = grid_for(c, kp, 256, lds_pre, slice, 0);
    gpre = (gpre + 7) & ~7;  // whole groups of eight workgroups: one queue-region group per XCD
    const int64_t threads = (int64_t)gpre * 256;
    b->pre_grid = gpre;
    b->n_pre_waves = gpre * 4;
    b->slice = slice;
    b->region = (slice + threads - 1) / threads * 64;
    // the queue, then the lens kernel's count and list of the molecules that leave their lens
    const size_t qbytes = (size_t)b->n_pre_waves * (size_t)b->region * (64 + 4) + 4 * BLT_LIST;
    if (hipMalloc((void **)&b->d_lensk, sizeof(BLensC) * n_cfgs) != hipSuccess ||
        hipMalloc(&b->d_pipe, qbytes) != hipSuccess ||
        hipMalloc(&b->d_pipe_cnt, 8 * 256 + (size_t)b->n_pre_waves * sizeof(uint32_t)) != hipSuccess) {
      // not enough memory for the queue: the single-kernel walker needs none
      (void)hipGetLastError();
      if (b->d_pipe) (void)hipFree(b->d_pipe);
      b->d_pipe = nullptr;
    }
  }
  // the device copy of the elements: circular apertures carry their squared-radius threshold
  std::vector<ctraj_elem> dev_elems(elems, elems + n_elems_total);
  for (auto &E : dev_elems)
    if (E.kind == CTRAJ_CIRC) E.p[1] = circ_r2_threshold(E.p[0]);
  if (hipStreamSynchronize(c->stream) != hipSuccess ||
      (b->d_lensk && hipMemcpy(b->d_lensk, hl.data(), sizeof(BLensC) * n_cfgs, hipMemcpyHostToDevice) != hipSuccess) ||
      (b->d_ws && hipMemcpy(b->d_ws, c->d_ws, sizeof(double) * 4 * b->ws_nseg, hipMemcpyDeviceToDevice) != hipSuccess) ||
      hipMemcpy(b->d_cfgs, hc.data(), sizeof(BCfg) * n_cfgs, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_elems, dev_elems.data(), sizeof(ctraj_elem) * n_elems_total, hipMemcpyHostToDevice) != hipSuccess) {
    const hipError_t e = hipGetLastError();
    ctraj_batch_destroy(b);
    return hipfail(c, e, "batch upload");
  }
  *out = b;
  return CTRAJ_OK;
}

int32_t ctraj_batch_run(ctraj_batch *b, const ctraj_batch_tune *tune, int64_t *tally_host, uint64_t *work_host,
                        float *elapsed_ms) {
  if (!b) return fail(nullptr, CTRAJ_EINVAL, "null batch");
  ctraj_ctx *c = b->ctx;
  HIPCHK(c, hipSetDevice(c->device));
  const bool fused = (tune && tune->fused) || !b->d_pipe;
  const int block = tune && tune->block ? tune->block : (fused ? 256 : 1024);
  const int chunk = tune && tune->chunk ? tune->chunk : 128;
  const int grid = tune ? tune->grid : 0;
  const int refill = tune && tune->refill ? tune->refill : 16;  // the tuner's usual pick on either path
  const bool use_graph = tune && tune->use_graph;
  if (block % 64 || block < 64 || block > (fused ? 256 : 1024))
    return fail(c, CTRAJ_EINVAL, fused ? "block must be 64, 128, 192 or 256" : "block must be a multiple of 64 up to 1024");
  if (chunk < 1 || chunk > (1 << 20)) return fail(c, CTRAJ_EINVAL, "chunk must be 1..2^20");
  if (refill < 1 || refill > 64) return fail(c, CTRAJ_EINVAL, "refill must be 1..64");
  if (grid < 0) return fail(c, CTRAJ_EINVAL, "grid < 0");
  HIPCHK(c, hipEventRecord(b->ev0, c->stream));
  int rc = CTRAJ_OK;
  if (use_graph) {
    const auto key = std::make_tuple(block, chunk, grid, refill, fused ? 1 : 0);
    auto it = b->graphs.find(key);
    if (it == b->graphs.end()) {
      hipGraph_t graph = nullptr;
      HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      rc = batch_launch(b, block, chunk, grid, refill, fused);
      const hipError_t e = hipStreamEndCapture(c->stream, &graph);
      if (rc != CTRAJ_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
      }
      if (e != hipSuccess) return hipfail(c, e, "hipStreamEndCapture");
      hipGraphExec_t exec = nullptr;
      const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ei != hipSuccess) return hipfail(c, ei, "hipGraphInstantiate");
      it = b->graphs.emplace(key, exec).first;
    }
    HIPCHK(c, hipGraphLaunch(it->second, c->stream));
  } else {
    rc = batch_launch(b, block, chunk, grid, refill, fused);
    if (rc != CTRAJ_OK) return rc;
  }
  HIPCHK(c, hipEventRecord(b->ev1, c->stream));
  if (tally_host)
    HIPCHK(c, hipMemcpyAsync(tally_host, b->d_tally, sizeof(int64_t) * b->n_cfg * b->n_labels, hipMemcpyDeviceToHost,
                             c->stream));
  if (work_host) {
    uint64_t w[1 + CTRAJ_WORK_N];
    HIPCHK(c, hipMemcpyAsync(w, b->d_counter, sizeof(w), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < CTRAJ_WORK_N; ++k) work_host[k] = w[1 + k];
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (elapsed_ms) HIPCHK(c, hipEventElapsedTime(elapsed_ms, b->ev0, b->ev1));
  return CTRAJ_OK;
}

int32_t ctraj_batch_split_path(const ctraj_batch *b) { return b && b->d_pipe ? 1 : 0; }

int32_t ctraj_batch_destroy(ctraj_batch *b) {
  if (!b) return CTRAJ_OK;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  for (auto &kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
  if (b->d_cfgs) (void)hipFree(b->d_cfgs);
  if (b->d_elems) (void)hipFree(b->d_elems);
  if (b->d_counter) (void)hipFree(b->d_counter);
  if (b->d_tally) (void)hipFree(b->d_tally);
  if (b->d_ws) (void)hipFree(b->d_ws);
  if (b->d_lensk) (void)hipFree(b->d_lensk);
  if (b->d_pipe) (void)hipFree(b->d_pipe);
  if (b->d_pipe_cnt) (void)hipFree(b->d_pipe_cnt);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  delete b;
  return CTRAJ_OK;
}

}  // extern "C"
