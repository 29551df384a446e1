// FP64 replay of Kimera-RPGO's PCM arithmetic on gfx950.
//
// This translation unit is compiled with -ffp-contract=off: every product and
// sum below rounds exactly as written, in the same order as the reference's
// dense Eigen expressions are evaluated by the restatement (dot products
// summed k = 0..n-1).  It provides
//   - the odometry fold        (Pcm.h:516-557, GeometryUtils.h:119-129)
//   - the odometry check       (Pcm.h:604-629)
//   - the pairwise check       (Pcm.h:670-718, GeometryUtils.h:133-186)
//   - adjacency bit packing    (Pcm.h:725-768, row-packed and symmetrized)
//   - the PcmSimple checks     (Pcm.h:581-596, 654-662, PoseWithNode norms)
// and the device functions reused by the FP32 path's near-threshold recheck.
#include "replay64.h"
#include "rpgo_internal.h"

namespace rpgo {

// Robot trajectories of a problem: chain c covers poses [off[c], off[c+1]).
// One chain when n <= 1 (off unused).  landmark: the loop closures are the
// observations of one landmark from poses cf[] (no odometry check, the
// landmark pair check).
struct Chains {
  const int* off;
  int n;
  int landmark = 0;
  __device__ __forceinline__ int of(int k) const {
    if (n <= 1) return 0;
    int lo = 0, hi = n - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (off[mid] <= k) lo = mid;
      else hi = mid - 1;
    }
    return lo;
  }
};

// ---------------------------------------------------------------------------
// per-measurement precompute: PoseWithCovariance(BetweenFactor) records and
// the fold Jacobians H_k = Ad(delta_k^-1)          (GeometryUtils.h:91-129)
// ---------------------------------------------------------------------------
template <int DIM>
__global__ void k_measurements(int n, const double* __restrict__ pose,
                               const double* __restrict__ cov, double* __restrict__ pwc,
                               double* __restrict__ hinv) {
  constexpr int PL = R64<DIM>::PL, D = R64<DIM>::D, PWC = R64<DIM>::PWC;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double rec[PWC];
  R64<DIM>::from_measurement(pose + (size_t)k * PL, cov + (size_t)k * D * D, rec);
  for (int i = 0; i < PWC; ++i) pwc[(size_t)k * PWC + i] = rec[i];
  if (hinv) {
    double inv[PL], H[D * D];
    R64<DIM>::pose_inverse(rec, inv);
    R64<DIM>::adjoint(inv, H);
    for (int i = 0; i < D * D; ++i) hinv[(size_t)k * D * D + i] = H[i];
  }
}

// ---------------------------------------------------------------------------
// sequential fold, one wavefront:  T_{k+1} = T_k * delta_k,
// Sigma_{k+1} = (H_k Sigma_k) H_k^T + Sigma_delta_k   (same order as the oracle)
// lanes 0..D*D-1 own the covariance entries, the next lanes the pose entries.
// Records start+1 .. n_poses-1 are written; the state at `start` comes from
// pose0 (chain start, record `start` written too) or from the stored record.
// ---------------------------------------------------------------------------
template <int DIM>
__global__ void __launch_bounds__(64) k_fold(int start, int n_poses,
                                            const double* __restrict__ pose0,
                                            const double* __restrict__ delta,
                                            const double* __restrict__ hinv,
                                            double* __restrict__ traj) {
  constexpr int PL = R64<DIM>::PL, D = R64<DIM>::D, PWC = R64<DIM>::PWC, DD = D * D;
  __shared__ double S[DD], T[DD], P[PL], Hs[DD], Dl[PWC], ob[kFoldAhead * PWC];
  __shared__ int ri;
  const int l = threadIdx.x;
  if (pose0) {
    // chain start: pose from the values, zero covariance (Pcm.h:534-542)
    if (l < DD) S[l] = 0.0;
    if (l < PL) P[l] = pose0[l];
    if (l == 0) ri = 1;
    for (int e = l; e < PWC; e += 64)
      traj[(size_t)start * PWC + e] = e < PL ? pose0[e] : (e < PL + DD ? 0.0 : 1.0);
  } else {
    // continue from a stored record: the same values the fold held there
    const double* rec = traj + (size_t)start * PWC;
    if (l < DD) S[l] = rec[PL + l];
    if (l < PL) P[l] = rec[l];
    if (l == 0) ri = rec[PL + DD] != 0.0 ? 1 : 0;
  }
  __syncthreads();
  if (start + 1 < n_poses)
    fold_run<DIM>(n_poses - 1 - start, hinv + (size_t)start * DD, delta + (size_t)start * PWC,
                  traj + (size_t)(start + 1) * PWC, S, T, P, Hs, Dl, &ri, ob, l);
}

// Fold of several chains, one wavefront per chain (Pcm.h:516-557 per robot
// prefix): chain c starts at pose0s[c] with zero covariance; its odometry rows
// are [off[c] - c, off[c+1] - c - 1).  Same arithmetic as k_fold.
template <int DIM>
__global__ void __launch_bounds__(64) k_fold_chains(Chains ch, const double* __restrict__ pose0s,
                                                   const double* __restrict__ delta,
                                                   const double* __restrict__ hinv,
                                                   double* __restrict__ traj) {
  constexpr int PL = R64<DIM>::PL, D = R64<DIM>::D, PWC = R64<DIM>::PWC, DD = D * D;
  __shared__ double S[DD], T[DD], P[PL], Hs[DD], Dl[PWC], ob[kFoldAhead * PWC];
  __shared__ int ri;
  const int c = blockIdx.x;
  if (c >= ch.n) return;
  const int p0 = ch.off[c], n_poses = ch.off[c + 1] - p0, o0 = p0 - c;
  double* tr = traj + (size_t)p0 * PWC;
  const int l = threadIdx.x;
  if (l < DD) S[l] = 0.0;
  if (l < PL) P[l] = pose0s[(size_t)c * PL + l];
  if (l == 0) ri = 1;
  __syncthreads();
  for (int e = l; e < PWC; e += 64) tr[e] = e < PL ? P[e] : (e < PL + DD ? 0.0 : 1.0);
  if (n_poses > 1)
    fold_run<DIM>(n_poses - 1, hinv + (size_t)o0 * DD, delta + (size_t)o0 * PWC, tr + PWC, S, T, P, Hs, Dl, &ri,
                  ob, l);
}

// ---------------------------------------------------------------------------
// odometry check (Pcm.h:604-629) : one thread per loop closure
// ---------------------------------------------------------------------------
template <int DIM>
__global__ void __launch_bounds__(256) k_odom_check(int n_lc, const double* __restrict__ traj,
                             const int* __restrict__ from, const int* __restrict__ to,
                             const double* __restrict__ lcm, double thr,
                             double* __restrict__ dist, int* __restrict__ flag, Chains ch) {
  constexpr int PWC = R64<DIM>::PWC;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_lc) return;
  // inter-robot loop closures and landmark observations: no odometry check
  // (Pcm.h:449-465)
  if (ch.landmark || ch.of(from[k]) != ch.of(to[k])) {
    dist[k] = 0.0;
    flag[k] = 1;
    return;
  }
  // the distance is computed also when the check is off (thr < 0: every loop
  // closure passes, Pcm.h:606), as a per-loop-closure feature
  const double d = R64<DIM>::odom_dist(traj + (size_t)from[k] * PWC, traj + (size_t)to[k] * PWC,
                                       lcm + (size_t)k * PWC);
  dist[k] = d;
  flag[k] = (thr < 0.0 || d < thr) ? 1 : 0;
}

// PcmSimple odometry check (Pcm.h:581-596, 604-629): node |to - from| + 1;
// dist receives the translation norm.  check == false passes every loop closure.
template <int DIM>
__global__ void k_odom_simple(int n_lc, const double* __restrict__ traj,
                              const int* __restrict__ from, const int* __restrict__ to,
                              const double* __restrict__ lcm, bool check, double thr_t,
                              double thr_r, double* __restrict__ dist, int* __restrict__ flag, Chains ch) {
  constexpr int PWC = R64<DIM>::PWC;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_lc) return;
  const int a = from[k], b = to[k];
  if (ch.landmark || ch.of(a) != ch.of(b)) {  // no odometry check (Pcm.h:449-465)
    dist[k] = 0.0;
    flag[k] = 1;
    return;
  }
  double tn, rn;
  R64<DIM>::simple_odom(traj + (size_t)a * PWC, traj + (size_t)b * PWC, lcm + (size_t)k * PWC,
                        abs(b - a) + 1, &tn, &rn);
  dist[k] = tn;
  flag[k] = (!check || (tn < thr_t && rn < thr_r)) ? 1 : 0;
}

// Stable compaction of accepted loop closures (single workgroup scan).
// accepted[0..m-1] = indices with flag set, in input order; count at accepted[n].
__global__ void __launch_bounds__(1024) k_compact(int n, const int* __restrict__ flag,
                                                 int* __restrict__ accepted) {
  __shared__ int warp_sums[16];
  __shared__ int base;
  if (threadIdx.x == 0) base = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int start = 0; start < n; start += 1024) {
    const int k = start + threadIdx.x;
    const int f = (k < n) ? flag[k] : 0;
    const unsigned long long b = __ballot(f);
    const int prefix = __popcll(b & ((1ull << lane) - 1ull));
    if (lane == 0) warp_sums[wid] = __popcll(b);
    __syncthreads();
    int off = base;
    for (int w = 0; w < wid; ++w) off += warp_sums[w];
    if (f) accepted[off + prefix] = k;
    __syncthreads();
    if (threadIdx.x == 0) {
      int tot = 0;
      for (int w = 0; w < 16; ++w) tot += warp_sums[w];
      base += tot;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) accepted[n] = base;
}

template <int DIM>
__global__ void k_gather(int m, const int* __restrict__ accepted, const int* __restrict__ from,
                         const int* __restrict__ to, const double* __restrict__ lcm,
                         int* __restrict__ cf, int* __restrict__ ct, double* __restrict__ clcm) {
  constexpr int PWC = R64<DIM>::PWC;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  const int k = accepted[t];
  cf[t] = from[k];
  ct[t] = to[k];
  for (int e = 0; e < PWC; ++e) clcm[(size_t)t * PWC + e] = lcm[(size_t)k * PWC + e];
}

// ---------------------------------------------------------------------------
// pairwise FP64 replay.  One wavefront per (row j, word w) tile of the lower
// triangle: lane = i - 64w, older LC i = (a, b), newer LC j = (c, d).
// Tiles are enumerated by 64-row groups g: group g has 64 rows of g+1 words.
// ---------------------------------------------------------------------------
__device__ inline void tile_of(long long t, int* j, int* w) {
  // largest g with 32 g (g+1) <= t
  long long g = (long long)((sqrt(1.0 + (double)t / 8.0) - 1.0) * 0.5);
  while (32ll * (g + 1) * (g + 2) <= t) ++g;
  while (g > 0 && 32ll * g * (g + 1) > t) --g;
  const long long rem = t - 32ll * g * (g + 1);
  *j = (int)(64 * g + rem / (g + 1));
  *w = (int)(rem % (g + 1));
}

template <int DIM>
__global__ void __launch_bounds__(256) k_pairs64(int m, long long n_tiles,
                                                 const double* __restrict__ traj,
                                                 const int* __restrict__ cf,
                                                 const int* __restrict__ ct,
                                                 const double* __restrict__ clcm, double thr,
                                                 int W, unsigned long long* __restrict__ adj,
                                                 double* __restrict__ dist,
                                                 unsigned long long* __restrict__ n_consistent,
                                                 Chains ch) {
  constexpr int PWC = R64<DIM>::PWC;
  const long long tile = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (tile >= n_tiles) return;
  const int lane = threadIdx.x & 63;
  int j, w;
  tile_of(tile, &j, &w);
  if (j >= m) return;
  const int i = w * 64 + lane;
  const bool active = i < j;
  double d = 0.0;
  if (active) {
    int c = cf[j], dd = ct[j];
    if (ch.of(cf[i]) != ch.of(c)) {  // Pcm.h:691-698: swap, measurement kept
      const int t = c;
      c = dd;
      dd = t;
    }
    d = ch.landmark ? R64<DIM>::landmark_pair_dist(traj, cf[i], clcm + (size_t)i * PWC, cf[j],
                                                   clcm + (size_t)j * PWC)
                    : R64<DIM>::pair_dist(traj, cf[i], ct[i], clcm + (size_t)i * PWC, c, dd,
                                          clcm + (size_t)j * PWC);
    if (dist) {
      dist[(size_t)j * m + i] = d;
      dist[(size_t)i * m + j] = d;
    }
  }
  const bool cons = active && (d < thr);
  const unsigned long long bits = __ballot(cons);
  if (lane == 0) {
    adj[(size_t)j * W + w] = bits;
    if (bits) atomicAdd(n_consistent, (unsigned long long)__popcll(bits));
  }
}

// PcmSimple pairwise check (Pcm.h:654-662, 670-718), same tiling as
// k_pairs64; consistent iff translation and rotation norms are both below
// their thresholds; dist receives the translation norm.
template <int DIM>
__global__ void __launch_bounds__(256) k_pairs_simple(int m, long long n_tiles,
                                                      const double* __restrict__ traj,
                                                      const int* __restrict__ cf,
                                                      const int* __restrict__ ct,
                                                      const double* __restrict__ clcm,
                                                      double thr_t, double thr_r, int W,
                                                      unsigned long long* __restrict__ adj,
                                                      double* __restrict__ dist,
                                                      unsigned long long* __restrict__ n_consistent,
                                                      Chains ch) {
  constexpr int PWC = R64<DIM>::PWC;
  const long long tile = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (tile >= n_tiles) return;
  const int lane = threadIdx.x & 63;
  int j, w;
  tile_of(tile, &j, &w);
  if (j >= m) return;
  const int i = w * 64 + lane;
  const bool active = i < j;
  bool cons = false;
  if (active) {
    double tn, rn;
    int c = cf[j], dd = ct[j];
    if (ch.of(cf[i]) != ch.of(c)) {  // Pcm.h:691-698
      const int t = c;
      c = dd;
      dd = t;
    }
    if (ch.landmark)
      R64<DIM>::simple_landmark_pair(traj, cf[i], clcm + (size_t)i * PWC, cf[j], clcm + (size_t)j * PWC, &tn,
                                     &rn);
    else
      R64<DIM>::simple_pair(traj, cf[i], ct[i], clcm + (size_t)i * PWC, c, dd, clcm + (size_t)j * PWC, &tn,
                            &rn);
    cons = tn < thr_t && rn < thr_r;
    if (dist) {
      dist[(size_t)j * m + i] = tn;
      dist[(size_t)i * m + j] = tn;
    }
  }
  const unsigned long long bits = __ballot(cons);
  if (lane == 0) {
    adj[(size_t)j * W + w] = bits;
    if (bits) atomicAdd(n_consistent, (unsigned long long)__popcll(bits));
  }
}

// Mirror the lower-triangle bits into the upper triangle: one wavefront per
// 64x64 block (row block J, word W <= J); lane r holds row 64J + r.
__global__ void __launch_bounds__(64) k_symmetrize(int m, int W, unsigned long long* adj) {
  const int J = blockIdx.y, Wd = blockIdx.x;
  if (Wd > J) return;
  const int lane = threadIdx.x;
  const int row = 64 * J + lane;
  const unsigned long long word = row < m ? adj[(size_t)row * W + Wd] : 0ull;
  unsigned long long mine = 0ull;
  for (int c = 0; c < 64; ++c) {
    const unsigned long long b = __ballot((word >> c) & 1ull);
    if (lane == c) mine = b;
  }
  const int orow = 64 * Wd + lane;
  if (orow < m) {
    if (Wd == J) adj[(size_t)orow * W + J] |= mine;
    else adj[(size_t)orow * W + J] = mine;
  }
}

// Queues up to this length are replayed one wavefront per pair (about half
// the latency of a lone lane-per-pair replay, at 1/64 of the throughput);
// longer ones one lane per pair, by the same kernel.
constexpr int kRecheckWaveMax = 2048;

// FP64 replay of pairs queued by the FP32 kernel: the reference decision
// d < thr replaces the FP32 one; both triangle bits are patched.
// One queued pair (i < j): FP64 distance, patch both bits when the FP32
// decision differs, count the flip and move the consistent-pair count.
template <int DIM>
__device__ __forceinline__ void recheck_one(int2 q, const double* __restrict__ traj,
                                            const int* __restrict__ cf, const int* __restrict__ ct,
                                            const double* __restrict__ clcm, double thr, int m, int W,
                                            unsigned long long* __restrict__ adj, double* __restrict__ dist,
                                            unsigned long long* __restrict__ flips,
                                            unsigned long long* __restrict__ n_consistent, Chains ch) {
  constexpr int PWC = R64<DIM>::PWC;
  const int i = q.x, j = q.y;
  if (i < 0 || j <= i || j >= m) return;
  int c = cf[j], dd = ct[j];
  if (ch.of(cf[i]) != ch.of(c)) {  // Pcm.h:691-698
    const int t = c;
    c = dd;
    dd = t;
  }
  const double d = ch.landmark ? R64<DIM>::landmark_pair_dist(traj, cf[i], clcm + (size_t)i * PWC, cf[j],
                                                              clcm + (size_t)j * PWC)
                               : R64<DIM>::pair_dist(traj, cf[i], ct[i], clcm + (size_t)i * PWC, c, dd,
                                                     clcm + (size_t)j * PWC);
  const bool cons = d < thr;
  unsigned long long* wj = adj + (size_t)j * W + (i >> 6);
  unsigned long long* wi = adj + (size_t)i * W + (j >> 6);
  const unsigned long long bj = 1ull << (i & 63), bi = 1ull << (j & 63);
  const bool was = ((*wj) & bj) != 0ull;
  if (cons != was) {
    if (cons) {
      atomicOr(wj, bj);
      atomicOr(wi, bi);
      if (n_consistent) atomicAdd(n_consistent, 1ull);
    } else {
      atomicAnd(wj, ~bj);
      atomicAnd(wi, ~bi);
      if (n_consistent) atomicAdd(n_consistent, ~0ull);  // -1 modulo 2^64
    }
    atomicAdd(flips, 1ull);
  }
  if (dist) {
    dist[(size_t)j * m + i] = d;
    dist[(size_t)i * m + j] = d;
  }
}

template <int DIM>
__global__ void __launch_bounds__(64) k_recheck64(int n, const int2* __restrict__ queue, const double* __restrict__ traj,
                            const int* __restrict__ cf, const int* __restrict__ ct,
                            const double* __restrict__ clcm, double thr, int m, int W,
                            unsigned long long* __restrict__ adj, double* __restrict__ dist,
                            unsigned long long* __restrict__ flips, Chains ch) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  recheck_one<DIM>(queue[t], traj, cf, ct, clcm, thr, m, W, adj, dist, flips, nullptr, ch);
}

// The same with the queue length read on the device (no host round trip):
// grid-stride over min(*qcount, cap) entries.  A longer queue is left to the
// host, which sees *qcount > cap after the stage and replays every pair.
template <int DIM>
__global__ void __launch_bounds__(64) k_recheck64_dev(const unsigned int* __restrict__ qcount, int cap,
                                                      const int2* __restrict__ queue,
                                                      const double* __restrict__ traj,
                                                      const int* __restrict__ cf, const int* __restrict__ ct,
                                                      const double* __restrict__ clcm, double thr, int m, int W,
                                                      unsigned long long* __restrict__ adj,
                                                      double* __restrict__ dist,
                                                      unsigned long long* __restrict__ flips,
                                                      unsigned long long* __restrict__ n_consistent,
                                                      Chains ch) {
  const unsigned int nq = *qcount;
  if (nq > (unsigned)cap) return;
  const int n = (int)nq;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x)
    recheck_one<DIM>(queue[t], traj, cf, ct, clcm, thr, m, W, adj, dist, flips, n_consistent, ch);
}

// ---------------------------------------------------------------------------
// Wavefront-parallel FP64 replay of one queued SE(3) pair.  The sequential
// replay is one long dependent chain per lane (about 40 us for a lone pair,
// which sets the stage's length at 5k poses).  Here one wavefront takes one
// pair: the 12 pose entries, the 36 adjoint entries and each 6x6 product
// entry go to their own lanes, each summed in the same order as mm64 /
// mmt64 / compose_entry, so every value is the sequential one to the bit;
// the LLT tests and the Mahalanobis norm (logmap, LU inverse) stay
// sequential on lane 0 with the R64<3> functions themselves.  Workgroups are
// one wavefront, so __syncthreads() orders the LDS exchanges.
// ---------------------------------------------------------------------------
namespace wp {

constexpr int PL = 12, D = 6, PWC = R64<3>::PWC;

struct Scratch {
  double Pa[PWC], Pb[PWC], Pc[PWC], Pd[PWC], Zab[PWC], Zcd[PWC];
  double A[PWC], B[PWC], X[PWC], Y[PWC], W[PWC], L[PWC];
  double t0[PL], t1[PL], t2[PL], H[D * D], T[D * D], HSH[D * D], C[D * D];
  int ok;
};

// entry e < 12 of a^-1 (R64<3>::pose_inverse)
__device__ __forceinline__ double inverse_entry(const double* a, int e) {
  if (e < 9) return a[(e % 3) * 3 + e / 3];
  const int i = e - 9;
  double s = a[0 * 3 + i] * (-a[9]);
  s += a[1 * 3 + i] * (-a[10]);
  s += a[2 * 3 + i] * (-a[11]);
  return s;
}

// entry (i, j) of Ad(g) (R64<3>::adjoint)
__device__ __forceinline__ double adjoint_entry(const double* g, int i, int j) {
  const double* R = g;
  const double tx = g[9], ty = g[10], tz = g[11];
  if (i < 3) return j < 3 ? R[i * 3 + j] : 0.0;
  if (j >= 3) return R[(i - 3) * 3 + (j - 3)];
  if (i == 3) return -tz * R[3 + j] + ty * R[6 + j];
  if (i == 4) return tz * R[j] - tx * R[6 + j];
  return -ty * R[j] + tx * R[3 + j];
}

__device__ __forceinline__ void pose_inverse(const double* a, double* o, int lane) {
  if (lane < PL) o[lane] = inverse_entry(a, lane);
  __syncthreads();
}

__device__ __forceinline__ void pose_compose(const double* a, const double* b, double* o, int lane) {
  if (lane < PL) o[lane] = R64<3>::compose_entry(a, b, lane);
  __syncthreads();
}

__device__ __forceinline__ void adjoint(const double* g, double* H, int lane) {
  if (lane < D * D) H[lane] = adjoint_entry(g, lane / D, lane % D);
  __syncthreads();
}

// out = (H S) H^T (congruence64<6>: mm64 then mmt64)
__device__ __forceinline__ void congruence(const double* H, const double* S, double* T, double* out, int lane) {
  const int i = lane / D, j = lane % D;
  if (lane < D * D) {
    double s = H[i * D] * S[j];
#pragma unroll
    for (int k = 1; k < D; ++k) s += H[i * D + k] * S[k * D + j];
    T[lane] = s;
  }
  __syncthreads();
  if (lane < D * D) {
    double s = T[i * D] * H[j * D];
#pragma unroll
    for (int k = 1; k < D; ++k) s += T[i * D + k] * H[j * D + k];
    out[lane] = s;
  }
  __syncthreads();
}

__device__ __forceinline__ double flag_and(const double* a, const double* b) {
  return (a[PL + D * D] != 0.0 && b[PL + D * D] != 0.0) ? 1.0 : 0.0;
}

// pwc_between with the LLT fallback
__device__ void between(Scratch& w, const double* a, const double* b, double* o, int lane) {
  pose_inverse(a, w.t0, lane);            // ainv
  pose_compose(w.t0, b, w.t1, lane);      // pose
  pose_inverse(w.t1, w.t2, lane);         // rinv
  adjoint(w.t2, w.H, lane);
  congruence(w.H, a + PL, w.T, w.HSH, lane);
  if (lane < D * D) w.C[lane] = b[PL + lane] - w.HSH[lane];
  __syncthreads();
  if (lane == 0) w.ok = llt_ok64<D>(w.C) ? 1 : 0;
  __syncthreads();
  if (!w.ok) {
    pose_inverse(b, w.t0, lane);          // binv
    pose_compose(w.t0, a, w.t2, lane);    // r2
    pose_inverse(w.t2, w.t0, lane);       // r2inv
    adjoint(w.t0, w.H, lane);
    congruence(w.H, b + PL, w.T, w.HSH, lane);
    if (lane < D * D) w.C[lane] = a[PL + lane] - w.HSH[lane];
    __syncthreads();
  }
  if (lane < PL) o[lane] = w.t1[lane];
  if (lane < D * D) o[PL + lane] = w.C[lane];
  if (lane == 0) o[PL + D * D] = flag_and(a, b);
  __syncthreads();
}

// pwc_compose
__device__ void compose(Scratch& w, const double* a, const double* b, double* o, int lane) {
  pose_compose(a, b, w.t1, lane);         // pose
  pose_inverse(b, w.t0, lane);            // binv
  adjoint(w.t0, w.H, lane);
  congruence(w.H, a + PL, w.T, w.HSH, lane);
  if (lane < PL) o[lane] = w.t1[lane];
  if (lane < D * D) o[PL + lane] = w.HSH[lane] + b[PL + lane];
  if (lane == 0) o[PL + D * D] = flag_and(a, b);
  __syncthreads();
}

// pwc_inverse (covariance kept)
__device__ void inverse(const double* a, double* o, int lane) {
  if (lane < PL) o[lane] = inverse_entry(a, lane);
  if (lane < D * D + 1) o[PL + lane] = a[PL + lane];
  __syncthreads();
}

__device__ void load(const double* __restrict__ src, double* dst, int lane) {
  if (lane < PWC) dst[lane] = src[lane];
}

}  // namespace wp

// One wavefront per queued SE(3) pair (i < j), grid-stride over the queue:
// the pair_dist of R64<3> computed lane-parallel, then the same patch of both
// bits, count and flip as the sequential replay.
__global__ void __launch_bounds__(64) k_recheck64_wave(const unsigned int* __restrict__ qcount, int cap,
                                                       const int2* __restrict__ queue,
                                                       const double* __restrict__ traj,
                                                       const int* __restrict__ cf, const int* __restrict__ ct,
                                                       const double* __restrict__ clcm, double thr, int m, int W,
                                                       unsigned long long* __restrict__ adj,
                                                       double* __restrict__ dist,
                                                       unsigned long long* __restrict__ flips,
                                                       unsigned long long* __restrict__ n_consistent, Chains ch) {
  using namespace wp;
  __shared__ Scratch w;
  __shared__ double dsh;
  const unsigned int nq = *qcount;
  if (nq > (unsigned)cap) return;
  const int lane = threadIdx.x;
  if (nq > (unsigned)kRecheckWaveMax) {  // long queue: one lane per pair
    for (int t = blockIdx.x * 64 + lane; t < (int)nq; t += gridDim.x * 64)
      recheck_one<3>(queue[t], traj, cf, ct, clcm, thr, m, W, adj, dist, flips, n_consistent, ch);
    return;
  }
  for (int t = blockIdx.x; t < (int)nq; t += gridDim.x) {
    const int2 q = queue[t];
    const int i = q.x, j = q.y;
    if (i < 0 || j <= i || j >= m) continue;  // uniform over the wavefront
    int c = cf[j], dd = ct[j];
    if (ch.of(cf[i]) != ch.of(c)) {  // Pcm.h:691-698
      const int tt = c;
      c = dd;
      dd = tt;
    }
    load(traj + (size_t)cf[i] * PWC, w.Pa, lane);
    load(traj + (size_t)ct[i] * PWC, w.Pb, lane);
    load(traj + (size_t)c * PWC, w.Pc, lane);
    load(traj + (size_t)dd * PWC, w.Pd, lane);
    load(clcm + (size_t)i * PWC, w.Zab, lane);
    load(clcm + (size_t)j * PWC, w.Zcd, lane);
    __syncthreads();
    between(w, w.Pa, w.Pc, w.A, lane);    // a_odom_c
    compose(w, w.A, w.Zcd, w.X, lane);    // a_path_d
    inverse(w.X, w.Y, lane);
    compose(w, w.Y, w.Zab, w.W, lane);    // d_path_b
    between(w, w.Pb, w.Pd, w.B, lane);    // b_odom_d
    compose(w, w.W, w.B, w.L, lane);      // loop
    if (lane == 0) {
      double rec[PWC];
#pragma unroll
      for (int e = 0; e < PWC; ++e) rec[e] = w.L[e];
      dsh = R64<3>::mahalanobis(rec);
    }
    __syncthreads();
    const double d = dsh;
    if (lane == 0) {
      const bool cons = d < thr;
      unsigned long long* wj = adj + (size_t)j * W + (i >> 6);
      unsigned long long* wi = adj + (size_t)i * W + (j >> 6);
      const unsigned long long bj = 1ull << (i & 63), bi = 1ull << (j & 63);
      const bool was = ((*wj) & bj) != 0ull;
      if (cons != was) {
        if (cons) {
          atomicOr(wj, bj);
          atomicOr(wi, bi);
          atomicAdd(n_consistent, 1ull);
        } else {
          atomicAnd(wj, ~bj);
          atomicAnd(wi, ~bi);
          atomicAdd(n_consistent, ~0ull);  // -1 modulo 2^64
        }
        atomicAdd(flips, 1ull);
      }
      if (dist) {
        dist[(size_t)j * m + i] = d;
        dist[(size_t)i * m + j] = d;
      }
    }
    __syncthreads();  // the scratch is reused by the next pair
  }
}

// ---------------------------------------------------------------------------
// streaming append (one update() of the reference after another)
// ---------------------------------------------------------------------------
// accepted[m_old + t] = l0 + tmp[t] and the compacted records, t < m_new
template <int DIM>
__global__ void k_gather_append(int m_old, int m_new, int l0, const int* __restrict__ tmp,
                                const int* __restrict__ from, const int* __restrict__ to,
                                const double* __restrict__ lcm, int* __restrict__ accepted,
                                int* __restrict__ cf, int* __restrict__ ct,
                                double* __restrict__ clcm) {
  constexpr int PWC = R64<DIM>::PWC;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m_new) return;
  const int k = l0 + tmp[t];
  const int o = m_old + t;
  accepted[o] = k;
  cf[o] = from[k];
  ct[o] = to[k];
  for (int e = 0; e < PWC; ++e) clcm[(size_t)o * PWC + e] = lcm[(size_t)k * PWC + e];
}

// copy rows [0, m_old) from stride W_old to stride W_new, zero-filling the
// new words; one thread per destination word
__global__ void k_adj_relayout(int m_old, int W_old, int W_new,
                               const unsigned long long* __restrict__ src,
                               unsigned long long* __restrict__ dst) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)m_old * W_new) return;
  const int r = (int)(t / W_new), w = (int)(t % W_new);
  dst[t] = w < W_old ? src[(size_t)r * W_old + w] : 0ull;
}

// New rows j in [m_old, m) against every i < j: one wavefront per (row,
// word) with word <= j/64, lane = i - 64 word.  Both triangles are written
// with atomicOr (rows of the new block receive bits from each other); the
// new rows are zero on entry.  counter += consistent pairs.
template <int DIM, bool SIMPLE>
__global__ void __launch_bounds__(256) k_pairs_append(int m_old, int m, int W,
                                                      const double* __restrict__ traj,
                                                      const int* __restrict__ cf,
                                                      const int* __restrict__ ct,
                                                      const double* __restrict__ clcm,
                                                      double thr, double thr_r,
                                                      unsigned long long* __restrict__ adj,
                                                      unsigned long long* __restrict__ counter) {
  constexpr int PWC = R64<DIM>::PWC;
  const long long wave = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (wave >= (long long)(m - m_old) * W) return;
  const int j = m_old + (int)(wave / W), w = (int)(wave % W);
  if (w > (j >> 6)) return;
  const int lane = threadIdx.x & 63;
  const int i = w * 64 + lane;
  bool cons = false;
  if (i < j) {
    if (SIMPLE) {
      double tn, rn;
      R64<DIM>::simple_pair(traj, cf[i], ct[i], clcm + (size_t)i * PWC, cf[j], ct[j],
                            clcm + (size_t)j * PWC, &tn, &rn);
      cons = tn < thr && rn < thr_r;
    } else {
      const double d = R64<DIM>::pair_dist(traj, cf[i], ct[i], clcm + (size_t)i * PWC, cf[j],
                                           ct[j], clcm + (size_t)j * PWC);
      cons = d < thr;
    }
  }
  if (cons) atomicOr(adj + (size_t)i * W + (j >> 6), 1ull << (j & 63));
  const unsigned long long bits = __ballot(cons);
  if (lane == 0 && bits) {
    atomicOr(adj + (size_t)j * W + w, bits);
    atomicAdd(counter, (unsigned long long)__popcll(bits));
  }
}

// ---------------------------------------------------------------------------
// host drivers
// ---------------------------------------------------------------------------
static Chains chains_of(const PcmState& s) {
  return Chains{s.chain_off.as<int>(), s.n_chains, s.landmark ? 1 : 0};
}

int pcm_recheck64(rpgo_ctx* ctx, int n_queue, double lc_thr, bool want_dist) {
  PcmState& s = ctx->pcm;
  if (n_queue <= 0) return RPGO_OK;
  const int m = s.m, W = (m + 63) / 64;
  double* dist = want_dist ? s.dist.as<double>() : nullptr;
  unsigned long long* flips = s.counters.as<unsigned long long>() + 2;
  if (s.dim == 3)
    k_recheck64<3><<<(n_queue + 63) / 64, 64, 0, ctx->stream>>>(
        n_queue, s.queue.as<int2>(), s.traj.as<double>(), s.c_from.as<int>(), s.c_to.as<int>(),
        s.c_lcm.as<double>(), lc_thr, m, W, s.adj.as<unsigned long long>(), dist, flips, chains_of(s));
  else
    k_recheck64<2><<<(n_queue + 63) / 64, 64, 0, ctx->stream>>>(
        n_queue, s.queue.as<int2>(), s.traj.as<double>(), s.c_from.as<int>(), s.c_to.as<int>(),
        s.c_lcm.as<double>(), lc_thr, m, W, s.adj.as<unsigned long long>(), dist, flips, chains_of(s));
  RPGO_HIP_CHECK(hipGetLastError());
  return RPGO_OK;
}

int pcm_measure_range(rpgo_ctx* ctx, int o0, int o1, int l0, int l1) {
  PcmState& s = ctx->pcm;
  const int pwc = pwc_len(s.dim), D = tan_dim(s.dim), PL = pose_len(s.dim);
  const int nd = o1 - o0, nl = l1 - l0;
  if (nd > 0) {
    const double* pose = s.odom.as<double>() + (size_t)o0 * PL;
    const double* cov = s.odom_cov.as<double>() + (size_t)o0 * D * D;
    double* pwc_out = s.delta.as<double>() + (size_t)o0 * pwc;
    double* h_out = s.hinv.as<double>() + (size_t)o0 * D * D;
    if (s.dim == 3)
      k_measurements<3><<<(nd + 255) / 256, 256, 0, ctx->stream>>>(nd, pose, cov, pwc_out, h_out);
    else
      k_measurements<2><<<(nd + 255) / 256, 256, 0, ctx->stream>>>(nd, pose, cov, pwc_out, h_out);
    RPGO_HIP_CHECK(hipGetLastError());
  }
  if (nl > 0) {
    const double* pose = s.lc_pose.as<double>() + (size_t)l0 * PL;
    const double* cov = s.lc_cov.as<double>() + (size_t)l0 * D * D;
    double* pwc_out = s.lcm.as<double>() + (size_t)l0 * pwc;
    if (s.dim == 3)
      k_measurements<3><<<(nl + 255) / 256, 256, 0, ctx->stream>>>(nl, pose, cov, pwc_out, nullptr);
    else
      k_measurements<2><<<(nl + 255) / 256, 256, 0, ctx->stream>>>(nl, pose, cov, pwc_out, nullptr);
    RPGO_HIP_CHECK(hipGetLastError());
  }
  return RPGO_OK;
}

int pcm_recheck64_device(rpgo_ctx* ctx, double lc_thr, bool want_dist) {
  PcmState& s = ctx->pcm;
  const int m = s.m, W = (m + 63) / 64;
  double* dist = want_dist ? s.dist.as<double>() : nullptr;
  unsigned long long* cnt = s.counters.as<unsigned long long>();
  const unsigned int* qcount = s.counters.as<unsigned int>() + 16;
  // enough wavefronts for the longest queue seen in practice (0.2-0.3 % of
  // the pairs at 50k), grid-stride beyond
  const unsigned blocks = 2048;
  if (s.dim == 3) {
    // the queue length picks on the device: a wavefront or a lane per pair
    k_recheck64_wave<<<blocks, 64, 0, ctx->stream>>>(qcount, kRecheckQueueCap, s.queue.as<int2>(),
                                                     s.traj.as<double>(), s.c_from.as<int>(), s.c_to.as<int>(),
                                                     s.c_lcm.as<double>(), lc_thr, m, W,
                                                     s.adj.as<unsigned long long>(), dist, cnt + 2, cnt,
                                                     chains_of(s));
  }
  else
    k_recheck64_dev<2><<<blocks, 64, 0, ctx->stream>>>(qcount, kRecheckQueueCap, s.queue.as<int2>(),
                                                       s.traj.as<double>(), s.c_from.as<int>(), s.c_to.as<int>(),
                                                       s.c_lcm.as<double>(), lc_thr, m, W,
                                                       s.adj.as<unsigned long long>(), dist, cnt + 2, cnt, chains_of(s));
  RPGO_HIP_CHECK(hipGetLastError());
  return RPGO_OK;
}

int pcm_prepare(rpgo_ctx* ctx) {
  PcmState& s = ctx->pcm;
  const int nd = s.n_poses - (s.n_chains > 1 ? s.n_chains : 1);
  const int pwc = pwc_len(s.dim), D = tan_dim(s.dim);
  int rc;
  if ((rc = s.delta.reserve(sizeof(double) * (size_t)(nd > 0 ? nd : 1) * pwc))) return rc;
  if ((rc = s.hinv.reserve(sizeof(double) * (size_t)(nd > 0 ? nd : 1) * D * D))) return rc;
  if ((rc = s.lcm.reserve(sizeof(double) * (size_t)(s.n_lc > 0 ? s.n_lc : 1) * pwc))) return rc;
  return pcm_measure_range(ctx, 0, nd, 0, s.n_lc);
}

int pcm_fold(rpgo_ctx* ctx) {
  PcmState& s = ctx->pcm;
  int rc;
  if ((rc = s.traj.reserve(sizeof(double) * (size_t)s.n_poses * pwc_len(s.dim)))) return rc;
  if (s.n_chains > 1) {
    if (s.dim == 3)
      k_fold_chains<3><<<s.n_chains, 64, 0, ctx->stream>>>(chains_of(s), s.pose0.as<double>(), s.delta.as<double>(),
                                                           s.hinv.as<double>(), s.traj.as<double>());
    else
      k_fold_chains<2><<<s.n_chains, 64, 0, ctx->stream>>>(chains_of(s), s.pose0.as<double>(), s.delta.as<double>(),
                                                           s.hinv.as<double>(), s.traj.as<double>());
    RPGO_HIP_CHECK(hipGetLastError());
    return RPGO_OK;
  }
  if (s.dim == 3)
    k_fold<3><<<1, 64, 0, ctx->stream>>>(0, s.n_poses, s.pose0.as<double>(), s.delta.as<double>(),
                                         s.hinv.as<double>(), s.traj.as<double>());
  else
    k_fold<2><<<1, 64, 0, ctx->stream>>>(0, s.n_poses, s.pose0.as<double>(), s.delta.as<double>(),
                                         s.hinv.as<double>(), s.traj.as<double>());
  RPGO_HIP_CHECK(hipGetLastError());
  return RPGO_OK;
}

int pcm_fold_from(rpgo_ctx* ctx, int start) {
  PcmState& s = ctx->pcm;
  if (start + 1 >= s.n_poses) return RPGO_OK;
  if (s.dim == 3)
    k_fold<3><<<1, 64, 0, ctx->stream>>>(start, s.n_poses, nullptr, s.delta.as<double>(),
                                         s.hinv.as<double>(), s.traj.as<double>());
  else
    k_fold<2><<<1, 64, 0, ctx->stream>>>(start, s.n_poses, nullptr, s.delta.as<double>(),
                                         s.hinv.as<double>(), s.traj.as<double>());
  RPGO_HIP_CHECK(hipGetLastError());
  return RPGO_OK;
}

int pcm_odom_and_compact(rpgo_ctx* ctx, double odom_thr, int* m_out, const SimpleThr* simple) {
  PcmState& s = ctx->pcm;
  const int n = s.n_lc;
  int rc;
  const int pwc = pwc_len(s.dim);
  if ((rc = s.odom_dist.reserve(sizeof(double) * (size_t)(n + 1)))) return rc;
  if ((rc = s.flags.reserve(sizeof(int) * (size_t)(n + 1)))) return rc;
  if ((rc = s.accepted.reserve(sizeof(int) * (size_t)(n + 1)))) return rc;
  if ((rc = s.c_from.reserve(sizeof(int) * (size_t)(n + 1)))) return rc;
  if ((rc = s.c_to.reserve(sizeof(int) * (size_t)(n + 1)))) return rc;
  if ((rc = s.c_lcm.reserve(sizeof(double) * (size_t)(n + 1) * pwc))) return rc;
  if (n == 0) {
    *m_out = 0;
    return RPGO_OK;
  }
  if (simple) {
    const bool check = simple->trans >= 0.0 && simple->rot >= 0.0;
    if (s.dim == 3)
      k_odom_simple<3><<<(n + 255) / 256, 256, 0, ctx->stream>>>(
          n, s.traj.as<double>(), s.lc_from.as<int>(), s.lc_to.as<int>(), s.lcm.as<double>(),
          check, simple->trans, simple->rot, s.odom_dist.as<double>(), s.flags.as<int>(), chains_of(s));
    else
      k_odom_simple<2><<<(n + 255) / 256, 256, 0, ctx->stream>>>(
          n, s.traj.as<double>(), s.lc_from.as<int>(), s.lc_to.as<int>(), s.lcm.as<double>(),
          check, simple->trans, simple->rot, s.odom_dist.as<double>(), s.flags.as<int>(), chains_of(s));
  } else if (s.dim == 3) {
    k_odom_check<3><<<(n + 255) / 256, 256, 0, ctx->stream>>>(
        n, s.traj.as<double>(), s.lc_from.as<int>(), s.lc_to.as<int>(), s.lcm.as<double>(),
        odom_thr, s.odom_dist.as<double>(), s.flags.as<int>(), chains_of(s));
  } else {
    k_odom_check<2><<<(n + 255) / 256, 256, 0, ctx->stream>>>(
        n, s.traj.as<double>(), s.lc_from.as<int>(), s.lc_to.as<int>(), s.lcm.as<double>(),
        odom_thr, s.odom_dist.as<double>(), s.flags.as<int>(), chains_of(s));
  }
  RPGO_HIP_CHECK(hipGetLastError());
  k_compact<<<1, 1024, 0, ctx->stream>>>(n, s.flags.as<int>(), s.accepted.as<int>());
  RPGO_HIP_CHECK(hipGetLastError());
  int m = 0;
  RPGO_HIP_CHECK(hipMemcpyAsync(&m, s.accepted.as<int>() + n, sizeof(int), hipMemcpyDeviceToHost,
                                ctx->stream));
  RPGO_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  if (m > 0) {
    if (s.dim == 3)
      k_gather<3><<<(m + 255) / 256, 256, 0, ctx->stream>>>(
          m, s.accepted.as<int>(), s.lc_from.as<int>(), s.lc_to.as<int>(), s.lcm.as<double>(),
          s.c_from.as<int>(), s.c_to.as<int>(), s.c_lcm.as<double>());
    else
      k_gather<2><<<(m + 255) / 256, 256, 0, ctx->stream>>>(
          m, s.accepted.as<int>(), s.lc_from.as<int>(), s.lc_to.as<int>(), s.lcm.as<double>(),
          s.c_from.as<int>(), s.c_to.as<int>(), s.c_lcm.as<double>());
    RPGO_HIP_CHECK(hipGetLastError());
  }
  *m_out = m;
  return RPGO_OK;
}

int pcm_pairs_fp64(rpgo_ctx* ctx, double lc_thr, bool want_dist) {
  PcmState& s = ctx->pcm;
  const int m = s.m;
  if (m < 2) return RPGO_OK;
  const int W = (m + 63) / 64;
  const long long G = W;
  const long long n_tiles = 32ll * G * (G + 1);
  const long long blocks = (n_tiles + 3) / 4;
  if (s.dim == 3)
    k_pairs64<3><<<(unsigned)blocks, 256, 0, ctx->stream>>>(
        m, n_tiles, s.traj.as<double>(), s.c_from.as<int>(), s.c_to.as<int>(),
        s.c_lcm.as<double>(), lc_thr, W, s.adj.as<unsigned long long>(),
        want_dist ? s.dist.as<double>() : nullptr, s.counters.as<unsigned long long>(), chains_of(s));
  else
    k_pairs64<2><<<(unsigned)blocks, 256, 0, ctx->stream>>>(
        m, n_tiles, s.traj.as<double>(), s.c_from.as<int>(), s.c_to.as<int>(),
        s.c_lcm.as<double>(), lc_thr, W, s.adj.as<unsigned long long>(),
        want_dist ? s.dist.as<double>() : nullptr, s.counters.as<unsigned long long>(), chains_of(s));
  RPGO_HIP_CHECK(hipGetLastError());
  return RPGO_OK;
}

int pcm_pairs_simple(rpgo_ctx* ctx, const SimpleThr& thr, bool want_dist) {
  PcmState& s = ctx->pcm;
  const int m = s.m;
  if (m < 2) return RPGO_OK;
  const int W = (m + 63) / 64;
  const long long G = W;
  const long long n_tiles = 32ll * G * (G + 1);
  const long long blocks = (n_tiles + 3) / 4;
  if (s.dim == 3)
    k_pairs_simple<3><<<(unsigned)blocks, 256, 0, ctx->stream>>>(
        m, n_tiles, s.traj.as<double>(), s.c_from.as<int>(), s.c_to.as<int>(),
        s.c_lcm.as<double>(), thr.trans, thr.rot, W, s.adj.as<unsigned long long>(),
        want_dist ? s.dist.as<double>() : nullptr, s.counters.as<unsigned long long>(), chains_of(s));
  else
    k_pairs_simple<2><<<(unsigned)blocks, 256, 0, ctx->stream>>>(
        m, n_tiles, s.traj.as<double>(), s.c_from.as<int>(), s.c_to.as<int>(),
        s.c_lcm.as<double>(), thr.trans, thr.rot, W, s.adj.as<unsigned long long>(),
        want_dist ? s.dist.as<double>() : nullptr, s.counters.as<unsigned long long>(), chains_of(s));
  RPGO_HIP_CHECK(hipGetLastError());
  return RPGO_OK;
}

int pcm_symmetrize(rpgo_ctx* ctx) {
  PcmState& s = ctx->pcm;
  const int m = s.m;
  if (m < 2) return RPGO_OK;
  const int W = (m + 63) / 64;
  k_symmetrize<<<dim3(W, W), 64, 0, ctx->stream>>>(m, W, s.adj.as<unsigned long long>());
  RPGO_HIP_CHECK(hipGetLastError());
  return RPGO_OK;
}

int pcm_odom_append(rpgo_ctx* ctx, int l0, double odom_thr, const SimpleThr* simple, int* m_new) {
  PcmState& s = ctx->pcm;
  const int n = s.n_lc, nn = n - l0, m_old = s.m, pwc = pwc_len(s.dim);
  int rc;
  *m_new = 0;
  if (nn <= 0) return RPGO_OK;
  hipStream_t st = ctx->stream;
  if ((rc = s.odom_dist.grow_keep(sizeof(double) * (size_t)(n + 1), sizeof(double) * l0, st))) return rc;
  if ((rc = s.flags.grow_keep(sizeof(int) * (size_t)(n + 1), 0, st))) return rc;
  if ((rc = s.tmp.reserve(sizeof(int) * (size_t)(nn + 1)))) return rc;
  const size_t mcap = (size_t)m_old + nn + 1;
  if ((rc = s.accepted.grow_keep(sizeof(int) * (size_t)(n + 1 > (int)mcap ? n + 1 : mcap),
                                 sizeof(int) * m_old, st)))
    return rc;
  if ((rc = s.c_from.grow_keep(sizeof(int) * mcap, sizeof(int) * m_old, st))) return rc;
  if ((rc = s.c_to.grow_keep(sizeof(int) * mcap, sizeof(int) * m_old, st))) return rc;
  if ((rc = s.c_lcm.grow_keep(sizeof(double) * mcap * pwc, sizeof(double) * m_old * pwc, st))) return rc;
  const int* from = s.lc_from.as<int>() + l0;
  const int* to = s.lc_to.as<int>() + l0;
  const double* lcm = s.lcm.as<double>() + (size_t)l0 * pwc;
  double* od = s.odom_dist.as<double>() + l0;
  int* fl = s.flags.as<int>() + l0;
  const unsigned g = (unsigned)((nn + 255) / 256);
  if (simple) {
    const bool check = simple->trans >= 0.0 && simple->rot >= 0.0;
    if (s.dim == 3)
      k_odom_simple<3><<<g, 256, 0, st>>>(nn, s.traj.as<double>(), from, to, lcm, check, simple->trans,
                                          simple->rot, od, fl, Chains{nullptr, 1});
    else
      k_odom_simple<2><<<g, 256, 0, st>>>(nn, s.traj.as<double>(), from, to, lcm, check, simple->trans,
                                          simple->rot, od, fl, Chains{nullptr, 1});
  } else if (s.dim == 3) {
    k_odom_check<3><<<g, 256, 0, st>>>(nn, s.traj.as<double>(), from, to, lcm, odom_thr, od, fl, Chains{nullptr, 1});
  } else {
    k_odom_check<2><<<g, 256, 0, st>>>(nn, s.traj.as<double>(), from, to, lcm, odom_thr, od, fl, Chains{nullptr, 1});
  }
  RPGO_HIP_CHECK(hipGetLastError());
  k_compact<<<1, 1024, 0, st>>>(nn, fl, s.tmp.as<int>());
  RPGO_HIP_CHECK(hipGetLastError());
  int mn = 0;
  RPGO_HIP_CHECK(hipMemcpyAsync(&mn, s.tmp.as<int>() + nn, sizeof(int), hipMemcpyDeviceToHost, st));
  RPGO_HIP_CHECK(hipStreamSynchronize(st));
  if (mn > 0) {
    if (s.dim == 3)
      k_gather_append<3><<<(mn + 255) / 256, 256, 0, st>>>(
          m_old, mn, l0, s.tmp.as<int>(), s.lc_from.as<int>(), s.lc_to.as<int>(), s.lcm.as<double>(),
          s.accepted.as<int>(), s.c_from.as<int>(), s.c_to.as<int>(), s.c_lcm.as<double>());
    else
      k_gather_append<2><<<(mn + 255) / 256, 256, 0, st>>>(
          m_old, mn, l0, s.tmp.as<int>(), s.lc_from.as<int>(), s.lc_to.as<int>(), s.lcm.as<double>(),
          s.accepted.as<int>(), s.c_from.as<int>(), s.c_to.as<int>(), s.c_lcm.as<double>());
    RPGO_HIP_CHECK(hipGetLastError());
  }
  *m_new = mn;
  return RPGO_OK;
}

int pcm_adj_append(rpgo_ctx* ctx, int m_old, double lc_thr, const SimpleThr* simple) {
  PcmState& s = ctx->pcm;
  const int m = s.m;
  if (m <= m_old) return RPGO_OK;
  hipStream_t st = ctx->stream;
  const int W_old = (m_old + 63) / 64, W = (m + 63) / 64;
  int rc;
  const size_t need = sizeof(uint64_t) * (size_t)m * W;
  if (W != W_old && m_old > 0) {
    // new stride: re-lay the old rows out into a fresh buffer
    DevBuf fresh;
    const size_t cap = need * 2;
    if ((rc = fresh.reserve(cap))) return rc;
    const long long words = (long long)m_old * W;
    k_adj_relayout<<<(unsigned)((words + 255) / 256), 256, 0, st>>>(
        m_old, W_old, W, s.adj.as<unsigned long long>(), fresh.as<unsigned long long>());
    RPGO_HIP_CHECK(hipGetLastError());
    RPGO_HIP_CHECK(hipStreamSynchronize(st));
    s.adj.release();
    s.adj = fresh;
  } else {
    if ((rc = s.adj.grow_keep(need, sizeof(uint64_t) * (size_t)m_old * W, st))) return rc;
  }
  RPGO_HIP_CHECK(hipMemsetAsync(s.adj.as<unsigned long long>() + (size_t)m_old * W, 0,
                                sizeof(uint64_t) * (size_t)(m - m_old) * W, st));
  const bool check = simple ? (simple->trans >= 0.0 && simple->rot >= 0.0) : lc_thr >= 0.0;
  if (!check || m < 2) return RPGO_OK;
  const long long waves = (long long)(m - m_old) * W;
  const unsigned blocks = (unsigned)((waves + 3) / 4);
  unsigned long long* cnt = s.counters.as<unsigned long long>();
  const double t1 = simple ? simple->trans : lc_thr, t2 = simple ? simple->rot : 0.0;
  if (s.dim == 3) {
    if (simple)
      k_pairs_append<3, true><<<blocks, 256, 0, st>>>(m_old, m, W, s.traj.as<double>(), s.c_from.as<int>(),
                                                      s.c_to.as<int>(), s.c_lcm.as<double>(), t1, t2,
                                                      s.adj.as<unsigned long long>(), cnt);
    else
      k_pairs_append<3, false><<<blocks, 256, 0, st>>>(m_old, m, W, s.traj.as<double>(), s.c_from.as<int>(),
                                                       s.c_to.as<int>(), s.c_lcm.as<double>(), t1, t2,
                                                       s.adj.as<unsigned long long>(), cnt);
  } else {
    if (simple)
      k_pairs_append<2, true><<<blocks, 256, 0, st>>>(m_old, m, W, s.traj.as<double>(), s.c_from.as<int>(),
                                                      s.c_to.as<int>(), s.c_lcm.as<double>(), t1, t2,
                                                      s.adj.as<unsigned long long>(), cnt);
    else
      k_pairs_append<2, false><<<blocks, 256, 0, st>>>(m_old, m, W, s.traj.as<double>(), s.c_from.as<int>(),
                                                       s.c_to.as<int>(), s.c_lcm.as<double>(), t1, t2,
                                                       s.adj.as<unsigned long long>(), cnt);
  }
  RPGO_HIP_CHECK(hipGetLastError());
  return RPGO_OK;
}

}  // namespace rpgo
