// ctraj_order.hpp -- the lens kernels' claim list: longest predicted lens time first.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

// Claim order of the lens kernel.  A wavefront's lanes stay busy only while its molecules end
// together, and the lens kernel ends when the last-claimed molecules do, so the queue is claimed
// longest-lived first: the walk kernel that fills the queue (PRE) files each queued molecule under a
// bucket of its predicted number of lens steps (ORD_WIDTH steps per bucket, the last bucket
// open-ended) and counts its workgroup's buckets, k_ord_scan and k_ord_scatter list the queue's
// slots by bucket, highest first, and the lens kernel claims that list in 64-slot chunks.  The
// order changes which lane integrates a molecule when, never its arithmetic: every molecule's result
// is the same in any order.
constexpr int ORD_BUCKETS = 256;
constexpr float ORD_WIDTH = 2.0f;

// The claim list of one queue: the walk kernel's per-entry buckets and per-workgroup bucket counts,
// the bucket totals and the list itself (queue slots); the queue's regions as the walk filled them.
struct OrdList {
  uint8_t *bkt;            // lifetime bucket of each queue entry
  uint32_t *hist;          // [walk workgroup][ORD_BUCKETS] counts; k_ord_scan makes them column prefixes
  uint32_t *tot;           // [ORD_BUCKETS] totals, then the queue's total entry count
  uint32_t *perm;          // queue slots in claim order
  const uint32_t *wcount;  // entries queued by each walk wavefront
  int64_t region;          // entries per walk wavefront's region
};

// Fast single-precision helpers for the prediction (its error only moves a molecule between
// neighbouring buckets): the hardware reciprocal and square root (1 ulp), and polynomial arc
// functions.  acos: Abramowitz & Stegun 4.4.45, |error| <= 6.7e-5 on [0, 1], with
// acos(-x) = pi - acos(x).  atan on [0, 1]: the odd degree-7 polynomial
// a - 0.327622764 a^3 + 0.15931422 a^5 - 0.0464964749 a^7 (|error| <= 2.1e-4 rad, largest at a = 1),
// extended to atan2 by the octant symmetries.  A bucket is 2 lens steps, some 5e-4 of a focusing
// period at the lens's field.
__device__ __forceinline__ float ord_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float ord_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

__device__ __forceinline__ float ord_acos(float x) {
  const float ax = __builtin_fabsf(x);
  const float p = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(-0.0187293f, ax, 0.0742610f), ax, -0.2121144f), ax,
                                 1.5707288f);
  const float r = ord_sqrt(1.0f - ax) * p;
  return x < 0.0f ? 3.14159265f - r : r;
}

__device__ __forceinline__ float ord_atan2(float y, float x) {
  const float ax = __builtin_fabsf(x), ay = __builtin_fabsf(y);
  const float mx = __builtin_fmaxf(ax, ay), mn = __builtin_fminf(ax, ay);
  const float a = mx > 0.0f ? mn * ord_rcp(mx) : 0.0f, s = a * a;
  float r = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(-0.0464964749f, s, 0.15931422f), s, -0.327622764f), s * a, a);
  r = ay > ax ? 1.57079633f - r : r;
  r = x < 0.0f ? 3.14159265f - r : r;
  return y < 0.0f ? -r : r;
}

// Predicted lens steps of a molecule entering a hexapole lens at (x, y) with transverse velocity
// (vx, vy) and axial speed vz, as the ordering bucket: the time until its radius reaches the bore
// under the linear acceleration k2 r (the force where the W' table saturates; k2 > 0 pushes out,
// r^2(t) = P cosh 2kt + Q sinh 2kt + S; k2 < 0 pulls in, r^2(t) = A + D cos(2wt - phi)), in steps,
// at most the lens's full steps.  Single precision with fast approximations: the prediction orders
// work and is not a result.
__device__ __forceinline__ int ord_bucket(float k2, float x, float y, float vx, float vy, float vz, float bore2,
                                          float len, float h, float cap, int ref) {
  const float r0s = x * x + y * y, v0s = vx * vx + vy * vy, rv = x * vx + y * vy;
  float t = __builtin_inff();  // the molecule stays inside the bore
  if (k2 > 0.0f) {
    const float k = ord_sqrt(k2), ik = ord_rcp(k), iv = v0s * ord_rcp(k2);
    const float P = 0.5f * (r0s + iv), Q = rv * ik, S = 0.5f * (r0s - iv);
    // u = e^(2kt) solves (P + Q) u^2 + 2 (S - bore2) u + (P - Q) = 0 (larger root)
    const float a = P + Q, b = 2.0f * (S - bore2), c = P - Q;
    const float u = (ord_sqrt(__builtin_fmaxf(b * b - 4.0f * a * c, 0.0f)) - b) * ord_rcp(2.0f * a);
    t = __logf(__builtin_fmaxf(u, 1.0f)) * (0.5f * ik);
  } else if (k2 < 0.0f) {
    const float w = ord_sqrt(-k2), iw = ord_rcp(w), iv = v0s * ord_rcp(-k2);
    const float A = 0.5f * (r0s + iv), Bc = 0.5f * (r0s - iv), C = rv * iw;
    const float D = ord_sqrt(Bc * Bc + C * C);
    if (A + D > bore2) {  // first upward crossing of bore2: 2wt = (phi - acos((bore2 - A) / D)) mod 2 pi
      const float al = ord_acos(__builtin_fminf(__builtin_fmaxf((bore2 - A) * ord_rcp(D), -1.0f), 1.0f));
      float th = ord_atan2(C, Bc) - al;
      th -= 6.28318531f * __builtin_floorf(th * 0.159154943f);
      t = th * (0.5f * iw);
    }
  }
  // full steps: floor(L / (vz dt)) capped (configuration mode), n = rint(L/dz) (reference mode)
  const float ih = ord_rcp(h);
  const float steps_max = ref ? cap : __builtin_fminf(__builtin_floorf(len * ord_rcp(vz * h)), cap);
  float s = t * (ref ? vz * ih : ih);
  s = __builtin_fminf(s, steps_max);  // a NaN time gives the full steps
  s = s >= 0.0f ? s : 0.0f;           // and a NaN speed bucket 0
  return (int)__builtin_fminf(s * (1.0f / ORD_WIDTH), (float)(ORD_BUCKETS - 1));
}

// Block-wide exclusive scan of one value per thread (256 threads); returns the sum of the values of
// the threads below, and the total in *total.
__device__ __forceinline__ uint32_t block_excl_scan256(uint32_t v, uint32_t *s, uint32_t *total) {
  const int t = threadIdx.x;
  s[t] = v;
  __syncthreads();
  for (int d = 1; d < 256; d <<= 1) {
    const uint32_t add = t >= d ? s[t - d] : 0u;
    __syncthreads();
    s[t] += add;
    __syncthreads();
  }
  const uint32_t incl = s[t];
  *total = s[255];
  __syncthreads();
  return incl - v;
}

// Claim order, step 1: block b (one per bucket, 256 threads) replaces column b of the walk
// workgroups' bucket counts by its exclusive prefix over the workgroups and writes the bucket's
// total; block 0 also leaves the queue's total in ord_tot[ORD_BUCKETS] (below, after step 2).
static __global__ void __launch_bounds__(256) k_ord_scan(uint32_t *hist, int n_rows, uint32_t *tot) {
  __shared__ uint32_t s[256];
  const int b = blockIdx.x, t = threadIdx.x;
  const int per = (n_rows + 255) / 256;
  const int r0 = min(t * per, n_rows), r1 = min(r0 + per, n_rows);
  uint32_t sum = 0;
  for (int r = r0; r < r1; ++r) sum += hist[(size_t)r * ORD_BUCKETS + b];
  uint32_t total;
  uint32_t run = block_excl_scan256(sum, s, &total);
  for (int r = r0; r < r1; ++r) {
    const size_t k = (size_t)r * ORD_BUCKETS + b;
    const uint32_t v = hist[k];
    hist[k] = run;
    run += v;
  }
  if (t == 0) tot[b] = total;
}

// Claim order, step 2: walk workgroup g's entries of bucket b go to positions base(b) + prefix(g, b)
// on, where base(b) counts the entries of the buckets above b.  The workgroup goes through its four
// wavefronts' regions in chunks of ORD_CHUNK filled slots: each slot takes the next position of its
// bucket (LDS atomics on the workgroup's cursors; the order within a bucket is immaterial), then the
// chunk's (position, slot) pairs are put in position order in LDS and written out, so that the list
// is written in runs of consecutive positions rather than in scattered 4-byte stores.
constexpr int ORD_CHUNK = 4096;

static __global__ void __launch_bounds__(256) k_ord_scatter(OrdList P) {
  __shared__ uint32_t s[256];
  __shared__ uint32_t cur[ORD_BUCKETS], cur0[ORD_BUCKETS], loc[ORD_BUCKETS];
  __shared__ uint32_t s_pos[ORD_CHUNK], s_slot[ORD_CHUNK];
  constexpr int PER = ORD_CHUNK / 256;
  const int t = threadIdx.x;
  // base(b) = sum of the totals of buckets b+1 .. ORD_BUCKETS-1: an exclusive scan in reversed order
  const uint32_t tb = P.tot[ORD_BUCKETS - 1 - t];
  uint32_t total;
  const uint32_t above = block_excl_scan256(tb, s, &total);
  cur[ORD_BUCKETS - 1 - t] = above + P.hist[(size_t)blockIdx.x * ORD_BUCKETS + (ORD_BUCKETS - 1 - t)];
  if (blockIdx.x == 0 && t == 0) P.tot[ORD_BUCKETS] = total;
  // the workgroup's four regions, filled prefixes of c0..c3 slots, as one index range [0, n)
  const int64_t w0 = (int64_t)blockIdx.x * 4;
  const uint32_t p1 = P.wcount[w0], p2 = p1 + P.wcount[w0 + 1], p3 = p2 + P.wcount[w0 + 2];
  const uint32_t n = p3 + P.wcount[w0 + 3];
  for (uint32_t base = 0; base < n; base += ORD_CHUNK) {  // the same trip count for every thread
    const uint32_t m = n - base < (uint32_t)ORD_CHUNK ? n - base : (uint32_t)ORD_CHUNK;
    __syncthreads();
    cur0[t] = cur[t];
    __syncthreads();
    uint32_t pos[PER], slot[PER], bk[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const uint32_t k = (uint32_t)(t + 256 * j);
      bk[j] = 0;
      if (k < m) {
        const uint32_t i = base + k;
        const uint32_t w = (uint32_t)(i >= p1) + (uint32_t)(i >= p2) + (uint32_t)(i >= p3);
        const uint32_t pw = w == 0 ? 0u : (w == 1 ? p1 : (w == 2 ? p2 : p3));
        const int64_t e = (w0 + w) * P.region + (i - pw);
        bk[j] = P.bkt[e];
        slot[j] = (uint32_t)e;
        pos[j] = atomicAdd(&cur[bk[j]], 1u);
      }
    }
    __syncthreads();
    // the chunk's offset of each bucket: its entries before it in claim order (buckets above first)
    uint32_t m_tot;
    loc[ORD_BUCKETS - 1 - t] = block_excl_scan256(cur[ORD_BUCKETS - 1 - t] - cur0[ORD_BUCKETS - 1 - t], s, &m_tot);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      if ((uint32_t)(t + 256 * j) < m) {
        const uint32_t l = loc[bk[j]] + (pos[j] - cur0[bk[j]]);
        s_pos[l] = pos[j];
        s_slot[l] = slot[j];
      }
    }
    __syncthreads();
    for (uint32_t k = t; k < m; k += 256) P.perm[s_pos[k]] = s_slot[k];
  }
}
