// Batched PCM: many independent single-prefix pose graphs of uneven size in
// one launch per stage.  Graph g owns poses [pose_off[g], pose_off[g+1]),
// odometry rows [pose_off[g] - g, pose_off[g+1] - g - 1) and loop closures
// [lc_off[g], lc_off[g+1]).  Every stage is the single-graph stage of
// replay64.hip (the FP64 replay of Pcm.h's fold, odometry check and pairwise
// check, bit-exact with the oracle) with a graph lookup in front:
//   measurements   one thread per odometry row / loop closure (all graphs)
//   fold           one wavefront per graph
//   odometry check one thread per loop closure
//   compaction     one workgroup per graph (stable, input order)
//   pairs          one wavefront per (row, 64-column word) tile; the tiles of
//                  all graphs are numbered by a prefix sum over graphs (sized
//                  by each graph's loop-closure count, rows beyond the
//                  accepted count exit) and a wavefront finds its graph by
//                  binary search
//   symmetrize     one wavefront per 64 x 64 block, numbered the same way
// With RPGO_PCM_FP32_RECHECK an SE(3) batch takes the FP32 pair stage of
// pcm_fp32.hip instead (every graph's anchor tables and records, one pair
// launch over all graphs writing the lower triangle), and the pairs it queues
// near the threshold are replayed here in FP64 by the same pair_dist as the
// FP64 stage, so the adjacency equals the FP64 stage's bit for bit.
// The heuristic clique (FMC::maxCliqueHeu) then runs per graph on the
// device-resident adjacency.  Like replay64.hip this unit is built with
// -ffp-contract=off: a contracted multiply-add would round the fold and the
// between covariances (a difference of nearly equal matrices whose LLT test
// picks the formula) differently from the single-graph path.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "replay64.h"
#include "rpgo_internal.h"

namespace rpgo {

struct BatchState {
  int dim = 0, n_graphs = 0;
  long long n_poses = 0, n_lc = 0, n_odom = 0, adj_words = 0;
  std::vector<long long> pose_off, lc_off, adj_off, tile_off, sym_off;
  DevBuf d_pose_off, d_lc_off, d_adj_off, d_tile_off, d_sym_off, d_lc_graph;
  DevBuf pose0, odom, odom_cov, lc_from, lc_to, lc_pose, lc_cov;
  DevBuf delta, hinv, traj, lcm, odom_dist, flags, accepted, m, c_from, c_to, c_lcm, adj, counters;
  DevBuf deg, cidx, ck;  // clique stage: degrees, members, sizes
  DevBuf cb_state, cb_icc, cb_dm;  // clique stage: chip-wide rounds
  // clique rounds' graph placement: graphs grouped by XCD (G entries), then
  // the 9 group offsets; cb_xmax = graphs of the largest group
  DevBuf cb_xorder;
  int cb_xmax = 0;
  std::vector<int> m_host;
  // FP32 pair stage (SE(3), RPGO_PCM_FP32_RECHECK): per-graph anchor blocks,
  // anchor-table entries and pair workgroups by prefix, and the stage's tables
  bool fp32 = false;
  int fp32_S = 4;
  std::vector<long long> blk_off, sq_off, pblk_off;
  DevBuf d_blk_off, d_sq_off, d_pblk_off;
  DevBuf fp_prec, fp_lcr, fp_pl64, fp_pb64, fp_bend, fp_ta, fp_sanc, fp_terms, fp_queue;
  long long fp_qcap = 0;
  // launch shapes chosen by the autotuner, per tune_slot() (0: not tuned yet)
  int tuned_pair[8] = {}, tuned_clique[8] = {};
  // captured stages and the layout they were captured for
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<long long> graph_key;
};

// PoseWithCovariance records of the measurements (GeometryUtils.h:91-115) and
// the fold Jacobians Ad(delta^-1), one thread per row of every graph
template <int DIM>
__global__ void k_meas_batch(long long n, const double* __restrict__ pose, const double* __restrict__ cov,
                             double* __restrict__ pwc, double* __restrict__ hinv) {
  constexpr int PL = R64<DIM>::PL, D = R64<DIM>::D, PWC = R64<DIM>::PWC;
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
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

__device__ __forceinline__ int graph_of(const long long* __restrict__ off, int n, long long t) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <int DIM>
__global__ void __launch_bounds__(64) k_fold_batch(int n_graphs, const long long* __restrict__ pose_off,
                                                  const double* __restrict__ pose0, const double* __restrict__ delta,
                                                  const double* __restrict__ hinv, double* __restrict__ traj) {
  constexpr int PL = R64<DIM>::PL, D = R64<DIM>::D, PWC = R64<DIM>::PWC, DD = D * D;
  __shared__ double S[DD], T[DD], P[PL], Hs[DD], Dl[PWC], ob[kFoldAhead * PWC];
  __shared__ int ri;
  const int g = blockIdx.x;
  if (g >= n_graphs) return;
  const long long p0 = pose_off[g];
  const int n_poses = (int)(pose_off[g + 1] - p0);
  const long long o0 = p0 - g;  // odometry rows of graph g
  double* tr = traj + (size_t)p0 * PWC;
  const int l = threadIdx.x;
  if (l < DD) S[l] = 0.0;
  if (l < PL) P[l] = pose0[(size_t)g * PL + l];
  if (l == 0) ri = 1;
  __syncthreads();
  for (int e = l; e < PWC; e += 64) tr[e] = e < PL ? P[e] : (e < PL + DD ? 0.0 : 1.0);
  if (n_poses > 1)
    fold_run<DIM>(n_poses - 1, hinv + (size_t)o0 * DD, delta + (size_t)o0 * PWC, tr + PWC, S, T, P, Hs, Dl, &ri,
                  ob, l);
}

template <int DIM>
__global__ void __launch_bounds__(256) k_odom_batch(long long n_lc, const int* __restrict__ lc_graph, const long long* __restrict__ pose_off,
                             const double* __restrict__ traj, const int* __restrict__ from,
                             const int* __restrict__ to, const double* __restrict__ lcm, double thr,
                             double* __restrict__ dist, int* __restrict__ flag) {
  constexpr int PWC = R64<DIM>::PWC;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_lc) return;
  const double* tr = traj + (size_t)pose_off[lc_graph[t]] * PWC;
  const double d = R64<DIM>::odom_dist(tr + (size_t)from[t] * PWC, tr + (size_t)to[t] * PWC, lcm + (size_t)t * PWC);
  dist[t] = d;
  flag[t] = (thr < 0.0 || d < thr) ? 1 : 0;
}

// per-graph stable compaction + gather of the accepted loop closures (local indices)
template <int DIM>
__global__ void __launch_bounds__(1024) k_compact_batch(int n_graphs, const long long* __restrict__ lc_off,
                                                        const int* __restrict__ flag, const int* __restrict__ from,
                                                        const int* __restrict__ to, const double* __restrict__ lcm,
                                                        int* __restrict__ accepted, int* __restrict__ m_out,
                                                        int* __restrict__ cf, int* __restrict__ ct,
                                                        double* __restrict__ clcm) {
  constexpr int PWC = R64<DIM>::PWC;
  __shared__ int wsum[16];
  __shared__ int base;
  const int g = blockIdx.x;
  if (g >= n_graphs) return;
  const long long o = lc_off[g];
  const int n = (int)(lc_off[g + 1] - o);
  if (threadIdx.x == 0) base = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int start = 0; start < n; start += 1024) {
    const int k = start + threadIdx.x;
    const int v = k < n ? flag[o + k] : 0;
    int incl = v;
    for (int off = 1; off < 64; off <<= 1) {
      const int u = __shfl_up(incl, off);
      if (lane >= off) incl += u;
    }
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    int pre = base;
    for (int w = 0; w < wid; ++w) pre += wsum[w];
    if (k < n && v) {
      const int dst = pre + incl - 1;
      accepted[o + dst] = k;
      cf[o + dst] = from[o + k];
      ct[o + dst] = to[o + k];
      for (int e = 0; e < PWC; ++e) clcm[(size_t)(o + dst) * PWC + e] = lcm[(size_t)(o + k) * PWC + e];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int s = 0;
      for (int w = 0; w < 16; ++w) s += wsum[w];
      base += s;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) m_out[g] = base;
}

__device__ inline void tile_of_b(long long t, int* j, int* w) {
  long long q = (long long)((sqrt(1.0 + (double)t / 8.0) - 1.0) * 0.5);
  while (32ll * (q + 1) * (q + 2) <= t) ++q;
  while (q > 0 && 32ll * q * (q + 1) > t) --q;
  const long long rem = t - 32ll * q * (q + 1);
  *j = (int)(64 * q + rem / (q + 1));
  *w = (int)(rem % (q + 1));
}

// STRIDE: a grid smaller than the tiles strides over them (the launch gated
// behind the FP32 stage); without it one wavefront takes one tile and the
// kernel keeps its registers (a loop here spills 40 % more)
template <int DIM, int WPB, bool STRIDE>
__global__ void __launch_bounds__(64 * WPB) k_pairs_batch(int n_graphs, long long n_tiles, const long long* __restrict__ tile_off,
                                                     const long long* __restrict__ pose_off,
                                                     const long long* __restrict__ lc_off,
                                                     const long long* __restrict__ adj_off, const int* __restrict__ m_arr,
                                                     const double* __restrict__ traj, const int* __restrict__ cf,
                                                     const int* __restrict__ ct, const double* __restrict__ clcm,
                                                     double thr, unsigned long long* __restrict__ adj,
                                                     unsigned long long* __restrict__ n_consistent,
                                                     const int* __restrict__ gate) {
  constexpr int PWC = R64<DIM>::PWC;
  // gate: with the FP32 stage, this FP64 stage runs only when the FP32 one
  // declined (a covariance without rotation information)
  if (gate && *gate == 0) return;
  const int lane = threadIdx.x & 63;
  for (long long tile = (long long)blockIdx.x * WPB + (threadIdx.x >> 6); tile < n_tiles;
       tile += STRIDE ? (long long)gridDim.x * WPB : n_tiles) {
    const int g = graph_of(tile_off, n_graphs, tile);
    int j, w;
    tile_of_b(tile - tile_off[g], &j, &w);
    const int m = m_arr[g];
    if (j >= m) continue;
    const long long o = lc_off[g];
    const int W = (m + 63) / 64;  // rows of the accepted set, as the clique reads them
    const double* tr = traj + (size_t)pose_off[g] * PWC;
    const int i = w * 64 + lane;
    const bool active = i < j;
    double d = 0.0;
    if (active)
      d = R64<DIM>::pair_dist(tr, cf[o + i], ct[o + i], clcm + (size_t)(o + i) * PWC, cf[o + j], ct[o + j],
                              clcm + (size_t)(o + j) * PWC);
    const bool cons = active && d < thr;
    const unsigned long long bits = __ballot(cons);
    if (lane == 0) {
      adj[adj_off[g] + (size_t)j * W + w] = bits;
      if (bits) atomicAdd(n_consistent, (unsigned long long)__popcll(bits));
    }
  }
}

__global__ void __launch_bounds__(64) k_sym_batch(int n_graphs, long long n_blocks, const long long* __restrict__ sym_off,
                                                  const long long* __restrict__ lc_off,
                                                  const long long* __restrict__ adj_off, const int* __restrict__ m_arr,
                                                  unsigned long long* __restrict__ adj) {
  const long long b = blockIdx.x;
  if (b >= n_blocks) return;
  const int g = graph_of(sym_off, n_graphs, b);
  const long long local = b - sym_off[g];
  // lower-triangle block (J, Wd), Wd <= J, numbered row by row
  int J = (int)((sqrt(8.0 * (double)local + 1.0) - 1.0) * 0.5);
  while ((long long)(J + 1) * (J + 2) / 2 <= local) ++J;
  while ((long long)J * (J + 1) / 2 > local) --J;
  const int Wd = (int)(local - (long long)J * (J + 1) / 2);
  const int m = m_arr[g];
  if (64 * J >= m) return;  // block beyond the accepted rows
  const int W = (m + 63) / 64;
  unsigned long long* A = adj + adj_off[g];
  const int lane = threadIdx.x;
  const int row = 64 * J + lane;
  const unsigned long long word = row < m ? A[(size_t)row * W + Wd] : 0ull;
  unsigned long long mine = 0ull;
  for (int c = 0; c < 64; ++c) {
    const unsigned long long bb = __ballot((word >> c) & 1ull);
    if (lane == c) mine = bb;
  }
  const int orow = 64 * Wd + lane;
  if (orow < m) {
    if (Wd == J) A[(size_t)orow * W + J] |= mine;
    else A[(size_t)orow * W + J] = mine;
  }
}

// FP64 replay of the pairs the FP32 stage queued: the FP64 stage's pair_dist
// decides each bit (lower triangle), and the consistent count follows every
// overturned FP32 decision.  counters: u64[0] consistent pairs, u32[2] queue
// length, u64[2] overturned decisions.
template <int DIM>
__global__ void __launch_bounds__(256) k_recheck_batch(int n_graphs, const long long* __restrict__ pose_off,
                                const long long* __restrict__ lc_off, const long long* __restrict__ adj_off,
                                const int* __restrict__ m_arr, const double* __restrict__ traj,
                                const int* __restrict__ cf, const int* __restrict__ ct,
                                const double* __restrict__ clcm, double thr, const int2* __restrict__ queue,
                                long long qcap, unsigned long long* __restrict__ adj,
                                unsigned long long* __restrict__ counters) {
  constexpr int PWC = R64<DIM>::PWC;
  const unsigned int* qcount = reinterpret_cast<const unsigned int*>(counters + 1);
  const long long n = min((long long)*qcount, qcap);
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x) {
    const int2 e = queue[q];
    const int g = graph_of(lc_off, n_graphs, e.x);
    const long long o = lc_off[g];
    const int i = (int)(e.x - o), j = (int)(e.y - o);
    const int W = (m_arr[g] + 63) / 64;
    const double* tr = traj + (size_t)pose_off[g] * PWC;
    const double d = R64<DIM>::pair_dist(tr, cf[o + i], ct[o + i], clcm + (size_t)(o + i) * PWC, cf[o + j],
                                         ct[o + j], clcm + (size_t)(o + j) * PWC);
    const bool bit = d < thr;
    unsigned long long* word = adj + adj_off[g] + (size_t)j * W + (i >> 6);
    const unsigned long long mask = 1ull << (i & 63);
    // only this thread changes this bit: the word's other bits may move meanwhile
    const bool old = (__hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & mask) != 0ull;
    if (bit != old) {
      if (bit) atomicOr(word, mask);
      else atomicAnd(word, ~mask);
      atomicAdd(&counters[0], bit ? 1ull : ~0ull);
      atomicAdd(&counters[2], 1ull);
    }
  }
}

__global__ void k_zero_u64(long long n, unsigned long long* __restrict__ a) {
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
    a[k] = 0ull;
}

// zero fill as a kernel: the stages may be captured in a hipGraph, where
// memset nodes are avoided
static void zero_u64(unsigned long long* a, long long n, hipStream_t st) {
  if (n <= 0) return;
  k_zero_u64<<<(unsigned)std::min<long long>((n + 255) / 256, 4096), 256, 0, st>>>(n, a);
}

void batch_destroy(rpgo_ctx* ctx) {
  if (!ctx->batch) return;
  BatchState& s = *ctx->batch;
  for (DevBuf* b : {&s.d_pose_off, &s.d_lc_off, &s.d_adj_off, &s.d_tile_off, &s.d_sym_off, &s.d_lc_graph, &s.pose0,
                    &s.odom, &s.odom_cov, &s.lc_from, &s.lc_to, &s.lc_pose, &s.lc_cov, &s.delta, &s.hinv, &s.traj,
                    &s.lcm, &s.odom_dist, &s.flags, &s.accepted, &s.m, &s.c_from, &s.c_to, &s.c_lcm, &s.adj,
                    &s.counters, &s.deg, &s.cidx, &s.ck, &s.cb_state, &s.cb_icc, &s.cb_dm, &s.cb_xorder, &s.d_blk_off,
                    &s.d_sq_off, &s.d_pblk_off, &s.fp_prec, &s.fp_lcr, &s.fp_pl64, &s.fp_pb64, &s.fp_bend, &s.fp_ta,
                    &s.fp_sanc, &s.fp_terms, &s.fp_queue})
    b->release();
  if (s.exec) (void)hipGraphExecDestroy(s.exec);
  if (s.graph) (void)hipGraphDestroy(s.graph);
  delete ctx->batch;
  ctx->batch = nullptr;
}

static int up(DevBuf& b, const void* h, size_t bytes, hipStream_t st) {
  int rc;
  if ((rc = b.reserve(bytes ? bytes : 8))) return rc;
  if (bytes) RPGO_HIP_CHECK(hipMemcpyAsync(b.ptr, h, bytes, hipMemcpyHostToDevice, st));
  return RPGO_OK;
}

constexpr int kDefaultPairWaves = 4;
constexpr long long kBatchQueueMax = 1ll << 29;  // recheck queue entries (int2): 4 GiB
constexpr int kDefaultCliqueWaves = 16;

template <int DIM>
static void launch_pairs_dim(rpgo_ctx* ctx, BatchState& s, double thr, int pw, const int* gate) {
  hipStream_t st = ctx->stream;
  const long long n_tiles = s.tile_off[s.n_graphs];
  // gated behind the FP32 stage (it runs only when that stage declines): a
  // small grid that strides over the tiles, so the usual early exit is cheap
  const long long full = (n_tiles + pw - 1) / pw;
  const unsigned blocks = (unsigned)(gate ? std::min<long long>(full, 2048) : full);
#define RPGO_PAIRS(WPB)                                                                                       \
  if (gate)                                                                                                  \
    k_pairs_batch<DIM, WPB, true><<<blocks, 64 * WPB, 0, st>>>(                                               \
        s.n_graphs, n_tiles, s.d_tile_off.as<long long>(), s.d_pose_off.as<long long>(),                      \
        s.d_lc_off.as<long long>(), s.d_adj_off.as<long long>(), s.m.as<int>(), s.traj.as<double>(),          \
        s.c_from.as<int>(), s.c_to.as<int>(), s.c_lcm.as<double>(), thr, s.adj.as<unsigned long long>(),      \
        s.counters.as<unsigned long long>(), gate);                                                           \
  else                                                                                                       \
  k_pairs_batch<DIM, WPB, false><<<blocks, 64 * WPB, 0, st>>>(                                               \
      s.n_graphs, n_tiles, s.d_tile_off.as<long long>(), s.d_pose_off.as<long long>(),                        \
      s.d_lc_off.as<long long>(), s.d_adj_off.as<long long>(), s.m.as<int>(), s.traj.as<double>(),            \
      s.c_from.as<int>(), s.c_to.as<int>(), s.c_lcm.as<double>(), thr, s.adj.as<unsigned long long>(),        \
      s.counters.as<unsigned long long>(), gate)
  switch (pw) {
    case 1: RPGO_PAIRS(1); break;
    case 2: RPGO_PAIRS(2); break;
    case 8: RPGO_PAIRS(8); break;
    default: RPGO_PAIRS(4); break;
  }
#undef RPGO_PAIRS
}

// tuned shapes are kept per dim and pair stage (the FP64 stage's launch
// shape is only timed when it is the stage that runs)
static int tune_slot(const BatchState& s) { return s.dim + (s.fp32 ? 4 : 0); }

static BatchFp32Args fp32_args(BatchState& s) {
  BatchFp32Args a;
  const int G = s.n_graphs;
  a.G = G;
  a.n_poses = s.n_poses;
  a.n_lc = s.n_lc;
  a.n_odom = s.n_odom;
  a.n_blocks = s.blk_off[G];
  a.n_sq = s.sq_off[G];
  a.n_pair_blocks = s.pblk_off[G];
  a.S = s.fp32_S;
  a.pose_off = s.d_pose_off.as<long long>();
  a.lc_off = s.d_lc_off.as<long long>();
  a.adj_off = s.d_adj_off.as<long long>();
  a.blk_off = s.d_blk_off.as<long long>();
  a.sq_off = s.d_sq_off.as<long long>();
  a.pblk_off = s.d_pblk_off.as<long long>();
  a.m = s.m.as<int>();
  a.delta = s.delta.as<double>();
  a.hinv = s.hinv.as<double>();
  a.lcm = s.lcm.as<double>();
  a.pose0 = s.pose0.as<double>();
  a.c_from = s.c_from.as<int>();
  a.c_to = s.c_to.as<int>();
  a.c_lcm = s.c_lcm.as<double>();
  a.prec = s.fp_prec.ptr;
  a.lcr = s.fp_lcr.ptr;
  a.sanc = s.fp_sanc.ptr;
  a.pl64 = s.fp_pl64.as<double>();
  a.pb64 = s.fp_pb64.as<double>();
  a.bend = s.fp_bend.as<double>();
  a.ta = s.fp_ta.as<double>();
  a.terms = s.fp_terms.as<double>();
  a.adj = s.adj.as<unsigned long long>();
  a.queue = s.fp_queue.as<int2>();
  a.qcount = reinterpret_cast<unsigned int*>(s.counters.as<unsigned long long>() + 1);
  a.qcap = s.fp_qcap;
  a.n_consistent = s.counters.as<unsigned long long>();
  a.no_rotation = reinterpret_cast<int*>(s.counters.as<unsigned long long>() + 3);
  return a;
}

// pair checks + symmetrize over every graph's tiles (adjacency zeroed first)
static int launch_pairs(rpgo_ctx* ctx, BatchState& s, const rpgo_pcm_params* prm, int pw) {
  hipStream_t st = ctx->stream;
  const int G = s.n_graphs;
  const long long n_tiles = s.tile_off[G];
  zero_u64(s.adj.as<unsigned long long>(), std::max<long long>(s.adj_off[G], 1), st);
  zero_u64(s.counters.as<unsigned long long>(), 4, st);
  // lc_threshold < 0 turns the pairwise check off (Pcm.h:79-82): the reference
  // then never grows the matrix, and as in rpgo_pcm_run the adjacency stays zero
  if (n_tiles > 0 && prm->lc_threshold >= 0.0) {
    if (s.fp32) {
      BatchFp32Args a = fp32_args(s);
      int rc;
      if ((rc = pcm_fp32_batch_launch(st, a, prm->lc_threshold, prm->recheck_margin, prm->recheck_chain_coef))) return rc;
      launch_pairs_dim<3>(ctx, s, prm->lc_threshold, pw, a.no_rotation);
      k_recheck_batch<3><<<(unsigned)std::min<long long>((s.fp_qcap + 255) / 256, 4096), 256, 0, st>>>(
          G, s.d_pose_off.as<long long>(), s.d_lc_off.as<long long>(), s.d_adj_off.as<long long>(), s.m.as<int>(),
          s.traj.as<double>(), s.c_from.as<int>(), s.c_to.as<int>(), s.c_lcm.as<double>(), prm->lc_threshold,
          s.fp_queue.as<int2>(), s.fp_qcap, s.adj.as<unsigned long long>(), s.counters.as<unsigned long long>());
    } else if (s.dim == 3) {
      launch_pairs_dim<3>(ctx, s, prm->lc_threshold, pw, nullptr);
    } else {
      launch_pairs_dim<2>(ctx, s, prm->lc_threshold, pw, nullptr);
    }
    const long long n_blocks = s.sym_off[G];
    k_sym_batch<<<(unsigned)n_blocks, 64, 0, st>>>(G, n_blocks, s.d_sym_off.as<long long>(), s.d_lc_off.as<long long>(),
                                                   s.d_adj_off.as<long long>(), s.m.as<int>(),
                                                   s.adj.as<unsigned long long>());
  }
  RPGO_HIP_CHECK(hipGetLastError());
  return RPGO_OK;
}

// Every stage of the batch, in stream order, with no host synchronisation
// and no memset/memcpy node (capturable).  events: record ev[0..3], ev[5]
// between the stages for per-stage times.
static CliqueSpread spread_of(BatchState& s) {
  CliqueSpread sp;
  sp.n_slots = s.n_lc;
  sp.max_slots = 0;
  for (int g = 0; g < s.n_graphs; ++g) sp.max_slots = std::max(sp.max_slots, s.lc_off[g + 1] - s.lc_off[g]);
  sp.slot_graph = s.d_lc_graph.as<int>();
  sp.state = s.cb_state.ptr;
  sp.icc = s.cb_icc.as<int>();
  sp.dm = s.cb_dm.as<unsigned long long>();
  sp.xorder = s.cb_xorder.as<int>();
  sp.xmax = s.cb_xmax;
  return sp;
}

static int launch_stages(rpgo_ctx* ctx, BatchState& s, const rpgo_pcm_params* prm, int pw, int cw, bool events) {
  hipStream_t st = ctx->stream;
  hipEvent_t* ev = ctx->ev;
  const int G = s.n_graphs, dim = s.dim;
  const size_t nlc = (size_t)s.n_lc, nod = (size_t)s.n_odom;
  if (events) RPGO_HIP_CHECK(hipEventRecord(ev[0], st));
  // measurements of every odometry row and loop closure
  if (nod) {
    if (dim == 3)
      k_meas_batch<3><<<(unsigned)((nod + 255) / 256), 256, 0, st>>>((long long)nod, s.odom.as<double>(), s.odom_cov.as<double>(),
                                                                     s.delta.as<double>(), s.hinv.as<double>());
    else
      k_meas_batch<2><<<(unsigned)((nod + 255) / 256), 256, 0, st>>>((long long)nod, s.odom.as<double>(), s.odom_cov.as<double>(),
                                                                     s.delta.as<double>(), s.hinv.as<double>());
  }
  if (nlc) {
    if (dim == 3)
      k_meas_batch<3><<<(unsigned)((nlc + 255) / 256), 256, 0, st>>>((long long)nlc, s.lc_pose.as<double>(), s.lc_cov.as<double>(),
                                                                     s.lcm.as<double>(), nullptr);
    else
      k_meas_batch<2><<<(unsigned)((nlc + 255) / 256), 256, 0, st>>>((long long)nlc, s.lc_pose.as<double>(), s.lc_cov.as<double>(),
                                                                     s.lcm.as<double>(), nullptr);
  }
  // fold: one wavefront per graph
  if (dim == 3)
    k_fold_batch<3><<<G, 64, 0, st>>>(G, s.d_pose_off.as<long long>(), s.pose0.as<double>(), s.delta.as<double>(),
                                      s.hinv.as<double>(), s.traj.as<double>());
  else
    k_fold_batch<2><<<G, 64, 0, st>>>(G, s.d_pose_off.as<long long>(), s.pose0.as<double>(), s.delta.as<double>(),
                                      s.hinv.as<double>(), s.traj.as<double>());
  RPGO_HIP_CHECK(hipGetLastError());
  if (events) RPGO_HIP_CHECK(hipEventRecord(ev[1], st));
  // odometry check (one thread per loop closure) and per-graph compaction;
  // the compaction also writes m[g] = 0 for graphs without loop closures
  if (nlc) {
    if (dim == 3)
      k_odom_batch<3><<<(unsigned)((nlc + 255) / 256), 256, 0, st>>>(
          (long long)nlc, s.d_lc_graph.as<int>(), s.d_pose_off.as<long long>(), s.traj.as<double>(), s.lc_from.as<int>(),
          s.lc_to.as<int>(), s.lcm.as<double>(), prm->odom_threshold, s.odom_dist.as<double>(), s.flags.as<int>());
    else
      k_odom_batch<2><<<(unsigned)((nlc + 255) / 256), 256, 0, st>>>(
          (long long)nlc, s.d_lc_graph.as<int>(), s.d_pose_off.as<long long>(), s.traj.as<double>(), s.lc_from.as<int>(),
          s.lc_to.as<int>(), s.lcm.as<double>(), prm->odom_threshold, s.odom_dist.as<double>(), s.flags.as<int>());
  }
  if (dim == 3)
    k_compact_batch<3><<<G, 1024, 0, st>>>(G, s.d_lc_off.as<long long>(), s.flags.as<int>(), s.lc_from.as<int>(),
                                           s.lc_to.as<int>(), s.lcm.as<double>(), s.accepted.as<int>(), s.m.as<int>(),
                                           s.c_from.as<int>(), s.c_to.as<int>(), s.c_lcm.as<double>());
  else
    k_compact_batch<2><<<G, 1024, 0, st>>>(G, s.d_lc_off.as<long long>(), s.flags.as<int>(), s.lc_from.as<int>(),
                                           s.lc_to.as<int>(), s.lcm.as<double>(), s.accepted.as<int>(), s.m.as<int>(),
                                           s.c_from.as<int>(), s.c_to.as<int>(), s.c_lcm.as<double>());
  RPGO_HIP_CHECK(hipGetLastError());
  if (events) RPGO_HIP_CHECK(hipEventRecord(ev[2], st));
  int rc;
  if ((rc = launch_pairs(ctx, s, prm, pw))) return rc;
  if (events) RPGO_HIP_CHECK(hipEventRecord(ev[3], st));
  if (cw) {
    const CliqueSpread sp = spread_of(s);
    if ((rc = clique_batch_launch(ctx, G, cw, s.d_adj_off.as<long long>(), s.d_lc_off.as<long long>(),
                                  s.m.as<int>(), s.adj.as<unsigned long long>(), s.deg.as<int>(), s.cidx.as<int>(),
                                  s.ck.as<int>(), &sp)))
      return rc;
  }
  if (events) RPGO_HIP_CHECK(hipEventRecord(ev[5], st));
  return RPGO_OK;
}

// Launch-shape autotuner: runs the stages once, then times on this batch the
// pair stage for 1, 2, 4 and 8 wavefronts per workgroup (best of two launches)
// and the clique stage for 4, 8 and 16 wavefronts per graph (one launch each:
// a clique launch of a large batch takes long enough to time once), and keeps
// the fastest per dim.  Every candidate computes the same bits.
static int autotune(rpgo_ctx* ctx, BatchState& s, const rpgo_pcm_params* prm, bool clique, float* ms_tune) {
  hipStream_t st = ctx->stream;
  hipEvent_t* ev = ctx->ev;
  int rc;
  RPGO_HIP_CHECK(hipEventRecord(ev[6], st));
  if ((rc = launch_stages(ctx, s, prm, kDefaultPairWaves, clique ? kDefaultCliqueWaves : 0, false))) return rc;
  auto timed = [&](auto&& launch, float* t_min, int reps = 2) -> int {
    *t_min = 1e30f;
    for (int rep = 0; rep < reps; ++rep) {
      RPGO_HIP_CHECK(hipEventRecord(ev[7], st));
      const int r = launch();
      if (r) return r;
      RPGO_HIP_CHECK(hipEventRecord(ev[8], st));
      RPGO_HIP_CHECK(hipEventSynchronize(ev[8]));
      float t = 0.f;
      (void)hipEventElapsedTime(&t, ev[7], ev[8]);
      *t_min = std::min(*t_min, t);
    }
    return RPGO_OK;
  };
  float best_t = 1e30f;
  int best_pw = kDefaultPairWaves;
  for (int pw : {1, 2, 4, 8}) {
    if (s.fp32) break;  // the FP64 stage's shape does not matter there
    float t;
    if ((rc = timed([&] { return launch_pairs(ctx, s, prm, pw); }, &t))) return rc;
    if (t < best_t) {
      best_t = t;
      best_pw = pw;
    }
  }
  s.tuned_pair[tune_slot(s)] = best_pw;
  if (clique) {
    best_t = 1e30f;
    int best_cw = kDefaultCliqueWaves;
    auto clique_time = [&](int cw, float* t) {
      return timed(
          [&] {
            const CliqueSpread sp = spread_of(s);
            return clique_batch_launch(ctx, s.n_graphs, cw, s.d_adj_off.as<long long>(),
                                       s.d_lc_off.as<long long>(), s.m.as<int>(), s.adj.as<unsigned long long>(),
                                       s.deg.as<int>(), s.cidx.as<int>(), s.ck.as<int>(), &sp);
          },
          t, 1);
    };
    for (int cw : {4, 8, 16}) {
      float t;
      if ((rc = clique_time(cw, &t))) return rc;
      if (t < best_t) {
        best_t = t;
        best_cw = cw;
      }
    }
    s.tuned_clique[tune_slot(s)] = best_cw;
  }
  RPGO_HIP_CHECK(hipEventRecord(ev[9], st));
  RPGO_HIP_CHECK(hipEventSynchronize(ev[9]));
  (void)hipEventElapsedTime(ms_tune, ev[6], ev[9]);
  return RPGO_OK;
}

}  // namespace rpgo

using namespace rpgo;

extern "C" {

int rpgo_pcm_batch_run(rpgo_ctx* ctx, const rpgo_pcm_batch* B, const rpgo_pcm_params* prm, int32_t* n_accepted,
                       int32_t* accepted, int32_t* clique_size, int32_t* clique_idx, rpgo_pcm_batch_stats* stats) {
  if (!ctx || !B || !prm || B->n_graphs < 1 || !B->pose_off || !B->lc_off || !B->pose0) {
    set_error("rpgo_pcm_batch_run: invalid argument");
    return RPGO_ERR_ARG;
  }
  if (B->dim != 3 && B->dim != 2) {
    set_error("rpgo_pcm_batch_run: dim must be 2 or 3");
    return RPGO_ERR_ARG;
  }
  const int G = B->n_graphs, dim = B->dim;
  const int PL = dim == 3 ? 12 : 4, D = dim == 3 ? 6 : 3, PWC = dim == 3 ? R64<3>::PWC : R64<2>::PWC;
  // validate the layout on the host
  if (B->pose_off[0] != 0 || B->lc_off[0] != 0) {
    set_error("rpgo_pcm_batch_run: offsets must start at 0");
    return RPGO_ERR_ARG;
  }
  for (int g = 0; g < G; ++g) {
    const long long np = B->pose_off[g + 1] - B->pose_off[g], nl = B->lc_off[g + 1] - B->lc_off[g];
    if (np < 1 || nl < 0) {
      set_error("rpgo_pcm_batch_run: graph %d has %lld poses, %lld loop closures", g, np, nl);
      return RPGO_ERR_ARG;
    }
    for (long long t = B->lc_off[g]; t < B->lc_off[g + 1]; ++t)
      if (B->lc_from[t] < 0 || B->lc_from[t] >= np || B->lc_to[t] < 0 || B->lc_to[t] >= np) {
        set_error("rpgo_pcm_batch_run: loop closure %lld of graph %d outside [0, %lld)", t - B->lc_off[g], g, np);
        return RPGO_ERR_ARG;
      }
  }
  RPGO_HIP_CHECK(hipSetDevice(ctx->device));
  if (!ctx->batch) ctx->batch = new BatchState();
  BatchState& s = *ctx->batch;
  hipStream_t st = ctx->stream;
  s.dim = dim;
  s.n_graphs = G;
  s.pose_off.assign(B->pose_off, B->pose_off + G + 1);
  s.lc_off.assign(B->lc_off, B->lc_off + G + 1);
  s.n_poses = s.pose_off[G];
  s.n_lc = s.lc_off[G];
  s.n_odom = s.n_poses - G;
  // adjacency, tile and symmetrize-block offsets sized by each graph's loop-closure count
  s.adj_off.assign(G + 1, 0);
  s.tile_off.assign(G + 1, 0);
  s.sym_off.assign(G + 1, 0);
  for (int g = 0; g < G; ++g) {
    const long long nl = s.lc_off[g + 1] - s.lc_off[g], W = (nl + 63) / 64, q = W;  // 64-row groups
    s.adj_off[g + 1] = s.adj_off[g] + nl * W;
    s.tile_off[g + 1] = s.tile_off[g] + 32ll * q * (q + 1);
    s.sym_off[g + 1] = s.sym_off[g] + q * (q + 1) / 2;
  }
  // FP32 pair stage: SE(3) only, and a recheck queue that holds every pair of
  // the batch (so no pair near the threshold can be dropped)
  long long pair_bound = 0;
  for (int g = 0; g < G; ++g) {
    const long long nl = s.lc_off[g + 1] - s.lc_off[g];
    pair_bound += nl * (nl - 1) / 2;
  }
  s.fp32 = prm->precision == RPGO_PCM_FP32_RECHECK && dim == 3 && pair_bound <= kBatchQueueMax;
  if (s.fp32) {
    s.blk_off.assign(G + 1, 0);
    s.sq_off.assign(G + 1, 0);
    s.pblk_off.assign(G + 1, 0);
    long long words = 0;
    for (int g = 0; g < G; ++g) {
      const long long nl = s.lc_off[g + 1] - s.lc_off[g], JT = (nl + 63) / 64;
      words += JT * (JT + 1) / 2;
    }
    // runs per word as in the single-graph stage: about 16 rounds of resident
    // wavefronts, at most 16 runs
    int S = 4;
    while (S < 16 && words * S < 128ll * ctx->n_cu) S *= 2;
    s.fp32_S = S;
    for (int g = 0; g < G; ++g) {
      const long long np = s.pose_off[g + 1] - s.pose_off[g], nb = (np + 63) / 64;
      const long long nl = s.lc_off[g + 1] - s.lc_off[g], JT = (nl + 63) / 64;
      s.blk_off[g + 1] = s.blk_off[g] + nb;
      s.sq_off[g + 1] = s.sq_off[g] + nb * nb;
      s.pblk_off[g + 1] = s.pblk_off[g] + (long long)(S / 4) * JT * (JT + 1) / 2;
    }
    s.fp_qcap = std::max<long long>(pair_bound, 1);
  }
  std::vector<int> lc_graph((size_t)std::max<long long>(s.n_lc, 1));
  for (int g = 0; g < G; ++g)
    for (long long t = s.lc_off[g]; t < s.lc_off[g + 1]; ++t) lc_graph[t] = g;
  int rc;
  const size_t nlc = (size_t)s.n_lc, nod = (size_t)s.n_odom;
  if ((rc = up(s.d_pose_off, s.pose_off.data(), sizeof(long long) * (G + 1), st)) ||
      (rc = up(s.d_lc_off, s.lc_off.data(), sizeof(long long) * (G + 1), st)) ||
      (rc = up(s.d_adj_off, s.adj_off.data(), sizeof(long long) * (G + 1), st)) ||
      (rc = up(s.d_tile_off, s.tile_off.data(), sizeof(long long) * (G + 1), st)) ||
      (rc = up(s.d_sym_off, s.sym_off.data(), sizeof(long long) * (G + 1), st)) ||
      (rc = up(s.d_lc_graph, lc_graph.data(), sizeof(int) * nlc, st)) ||
      (rc = up(s.pose0, B->pose0, sizeof(double) * PL * G, st)) ||
      (rc = up(s.odom, B->odom, sizeof(double) * PL * nod, st)) ||
      (rc = up(s.odom_cov, B->odom_cov, sizeof(double) * D * D * nod, st)) ||
      (rc = up(s.lc_from, B->lc_from, sizeof(int) * nlc, st)) || (rc = up(s.lc_to, B->lc_to, sizeof(int) * nlc, st)) ||
      (rc = up(s.lc_pose, B->lc_pose, sizeof(double) * PL * nlc, st)) ||
      (rc = up(s.lc_cov, B->lc_cov, sizeof(double) * D * D * nlc, st)))
    return rc;
  const size_t nl1 = nlc ? nlc : 1, no1 = nod ? nod : 1;
  if ((rc = s.delta.reserve(sizeof(double) * PWC * no1)) || (rc = s.hinv.reserve(sizeof(double) * D * D * no1)) ||
      (rc = s.traj.reserve(sizeof(double) * PWC * (size_t)s.n_poses)) || (rc = s.lcm.reserve(sizeof(double) * PWC * nl1)) ||
      (rc = s.odom_dist.reserve(sizeof(double) * nl1)) || (rc = s.flags.reserve(sizeof(int) * nl1)) ||
      (rc = s.accepted.reserve(sizeof(int) * nl1)) || (rc = s.m.reserve(sizeof(int) * G)) ||
      (rc = s.c_from.reserve(sizeof(int) * nl1)) || (rc = s.c_to.reserve(sizeof(int) * nl1)) ||
      (rc = s.c_lcm.reserve(sizeof(double) * PWC * nl1)) ||
      (rc = s.adj.reserve(sizeof(unsigned long long) * (size_t)std::max<long long>(s.adj_off[G], 1))) ||
      (rc = s.counters.reserve(sizeof(unsigned long long) * 4)))
    return rc;
  if (s.fp32) {
    const size_t nb = (size_t)std::max<long long>(s.blk_off[G], 1), nsq = (size_t)std::max<long long>(s.sq_off[G], 1);
    if ((rc = up(s.d_blk_off, s.blk_off.data(), sizeof(long long) * (G + 1), st)) ||
        (rc = up(s.d_sq_off, s.sq_off.data(), sizeof(long long) * (G + 1), st)) ||
        (rc = up(s.d_pblk_off, s.pblk_off.data(), sizeof(long long) * (G + 1), st)) ||
        (rc = s.fp_prec.reserve(pcm_fp32_pose_rec_bytes() * (size_t)s.n_poses)) ||
        (rc = s.fp_lcr.reserve(pcm_fp32_lc_rec_bytes() * nl1)) ||
        (rc = s.fp_pl64.reserve(sizeof(double) * 12 * (size_t)s.n_poses)) ||
        (rc = s.fp_pb64.reserve(sizeof(double) * 12 * nb)) || (rc = s.fp_ta.reserve(sizeof(double) * 12 * nb)) ||
        (rc = s.fp_bend.reserve(sizeof(double) * 36 * nb)) ||
        (rc = s.fp_sanc.reserve(pcm_fp32_anchor_entry_bytes() * nsq)) ||
        (rc = s.fp_terms.reserve(sizeof(double) * 36 * nsq)) ||
        (rc = s.fp_queue.reserve(sizeof(int2) * (size_t)s.fp_qcap)))
      return rc;
  }
  const bool want_clique = clique_size != nullptr;
  if (want_clique) {
    for (int g = 0; g < G; ++g)
      if (s.lc_off[g + 1] - s.lc_off[g] > clique_max_vertices()) {
        set_error("rpgo_pcm_batch_run: graph %d has %lld loop closures; the clique stage takes at most %d", g,
                  s.lc_off[g + 1] - s.lc_off[g], clique_max_vertices());
        return RPGO_ERR_ARG;
      }
    if ((rc = s.deg.reserve(sizeof(int) * nl1)) || (rc = s.cidx.reserve(sizeof(int) * nl1)) ||
        (rc = s.ck.reserve(sizeof(int) * (size_t)G)) || (rc = s.cb_state.reserve(clique_spread_state_bytes(G))) ||
        (rc = s.cb_icc.reserve(clique_spread_icc_bytes(G))) ||
        (rc = s.cb_dm.reserve(sizeof(unsigned long long) * nl1)))
      return rc;
    // The rounds walk each graph's adjacency rows many times over.  Each graph
    // goes to one XCD, so its rows stay in that XCD's L2, and the graphs are
    // spread over the 8 XCDs largest first onto the least loaded (work ~ n^2).
    std::vector<int> xorder((size_t)G + 9);
    {
      std::vector<int> byn(G);
      for (int g = 0; g < G; ++g) byn[g] = g;
      std::stable_sort(byn.begin(), byn.end(), [&](int a, int b) {
        return s.lc_off[a + 1] - s.lc_off[a] > s.lc_off[b + 1] - s.lc_off[b];
      });
      std::vector<std::vector<int>> grp(8);
      double load[8] = {};
      for (int g : byn) {
        int x = 0;
        for (int k = 1; k < 8; ++k)
          if (load[k] < load[x]) x = k;
        const double n = (double)(s.lc_off[g + 1] - s.lc_off[g]);
        load[x] += n * n;
        grp[x].push_back(g);
      }
      int o = 0;
      s.cb_xmax = 0;
      for (int x = 0; x < 8; ++x) {
        xorder[(size_t)G + x] = o;
        for (int g : grp[x]) xorder[o++] = g;
        s.cb_xmax = std::max<int>(s.cb_xmax, (int)grp[x].size());
      }
      xorder[(size_t)G + 8] = o;
    }
    if ((rc = up(s.cb_xorder, xorder.data(), sizeof(int) * xorder.size(), st))) return rc;
  }
  float ms_tune = 0.f;
  const int ts = tune_slot(s);
  if ((B->flags & RPGO_BATCH_AUTOTUNE) && (!s.tuned_pair[ts] || (want_clique && !s.tuned_clique[ts]))) {
    if ((rc = autotune(ctx, s, prm, want_clique, &ms_tune))) return rc;
  }
  const int pw = s.tuned_pair[ts] ? s.tuned_pair[ts] : kDefaultPairWaves;
  const int cw = want_clique ? (s.tuned_clique[ts] ? s.tuned_clique[ts] : kDefaultCliqueWaves) : 0;
  hipEvent_t* ev = ctx->ev;
  int replayed = 0;
  if (B->flags & RPGO_BATCH_GRAPH) {
    // thresholds are baked into the captured launches: the key holds their bits
    long long bo, bl;
    std::memcpy(&bo, &prm->odom_threshold, 8);
    std::memcpy(&bl, &prm->lc_threshold, 8);
    // every host-side size a captured launch bakes in: the adjacency words the
    // zeroing kernel clears (adj_off[G]) and the clique rounds (largest graph)
    long long bm, bc;
    std::memcpy(&bm, &prm->recheck_margin, 8);
    std::memcpy(&bc, &prm->recheck_chain_coef, 8);
    std::vector<long long> key = {dim, G, (long long)nlc, (long long)nod, s.tile_off[G], s.sym_off[G],
                                  s.adj_off[G], spread_of(s).max_slots, pw, cw, bo, bl, s.fp32 ? 1 : 0,
                                  bm, bc, s.cb_xmax};
    if (s.fp32) {
      key.insert(key.end(), {s.fp32_S, s.blk_off[G], s.sq_off[G], s.pblk_off[G], s.fp_qcap});
      for (DevBuf* b : {&s.d_blk_off, &s.d_sq_off, &s.d_pblk_off, &s.fp_prec, &s.fp_lcr, &s.fp_pl64, &s.fp_pb64,
                        &s.fp_bend, &s.fp_ta, &s.fp_sanc, &s.fp_terms, &s.fp_queue})
        key.push_back((long long)(uintptr_t)b->ptr);
    }
    for (DevBuf* b : {&s.d_pose_off, &s.d_lc_off, &s.d_adj_off, &s.d_tile_off, &s.d_sym_off, &s.d_lc_graph,
                      &s.pose0, &s.odom, &s.odom_cov, &s.lc_from, &s.lc_to, &s.lc_pose, &s.lc_cov, &s.delta,
                      &s.hinv, &s.traj, &s.lcm, &s.odom_dist, &s.flags, &s.accepted, &s.m, &s.c_from, &s.c_to,
                      &s.c_lcm, &s.adj, &s.counters, &s.deg, &s.cidx, &s.ck, &s.cb_state, &s.cb_icc, &s.cb_dm,
                      &s.cb_xorder})
      key.push_back((long long)(uintptr_t)b->ptr);
    if (s.exec && key == s.graph_key) {
      replayed = 1;
    } else {
      if (s.exec) (void)hipGraphExecDestroy(s.exec);
      if (s.graph) (void)hipGraphDestroy(s.graph);
      s.exec = nullptr;
      s.graph = nullptr;
      s.graph_key.clear();
      RPGO_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
      const int rc_cap = launch_stages(ctx, s, prm, pw, cw, false);
      hipGraph_t gr = nullptr;
      const hipError_t e_end = hipStreamEndCapture(st, &gr);
      if (rc_cap) {
        if (gr) (void)hipGraphDestroy(gr);
        return rc_cap;
      }
      RPGO_HIP_CHECK(e_end);
      s.graph = gr;
      RPGO_HIP_CHECK(hipGraphInstantiate(&s.exec, s.graph, nullptr, nullptr, 0));
      s.graph_key = key;
    }
    RPGO_HIP_CHECK(hipEventRecord(ev[0], st));
    RPGO_HIP_CHECK(hipGraphLaunch(s.exec, st));
    RPGO_HIP_CHECK(hipEventRecord(ev[5], st));
  } else {
    if ((rc = launch_stages(ctx, s, prm, pw, cw, true))) return rc;
  }
  s.m_host.assign(G, 0);
  RPGO_HIP_CHECK(hipMemcpyAsync(s.m_host.data(), s.m.ptr, sizeof(int) * G, hipMemcpyDeviceToHost, st));
  unsigned long long cnt[4];
  RPGO_HIP_CHECK(hipMemcpyAsync(cnt, s.counters.ptr, sizeof(cnt), hipMemcpyDeviceToHost, st));
  if (accepted && nlc)
    RPGO_HIP_CHECK(hipMemcpyAsync(accepted, s.accepted.ptr, sizeof(int) * nlc, hipMemcpyDeviceToHost, st));
  std::vector<int> k_host(want_clique ? G : 0);
  if (want_clique) {
    RPGO_HIP_CHECK(hipMemcpyAsync(k_host.data(), s.ck.ptr, sizeof(int) * G, hipMemcpyDeviceToHost, st));
    if (clique_idx && nlc)
      RPGO_HIP_CHECK(hipMemcpyAsync(clique_idx, s.cidx.ptr, sizeof(int) * nlc, hipMemcpyDeviceToHost, st));
  }
  RPGO_HIP_CHECK(hipStreamSynchronize(st));
  long long pairs = 0;
  for (int g = 0; g < G; ++g) pairs += (long long)s.m_host[g] * (s.m_host[g] - 1) / 2;
  if (n_accepted) std::memcpy(n_accepted, s.m_host.data(), sizeof(int) * G);
  if (want_clique) {
    for (int g = 0; g < G; ++g) {
      if (k_host[g] == INT_MIN) {
        set_error("rpgo_pcm_batch_run: clique replay of graph %d disagrees with its scan", g);
        return RPGO_ERR_STATE;
      }
      clique_size[g] = k_host[g];
    }
  }
  if (stats) {
    *stats = rpgo_pcm_batch_stats{};
    stats->n_pairs = pairs;
    stats->n_consistent = (long long)cnt[0];
    stats->fp32 = s.fp32 ? 1 : 0;
    if (s.fp32) {
      const unsigned int q = (unsigned int)(cnt[1] & 0xffffffffull);
      stats->n_rechecked = std::min<long long>(q, s.fp_qcap);
      stats->n_flipped = (long long)cnt[2];
      if ((int)(cnt[3] & 0xffffffffull)) stats->fp32 = 0;  // a covariance without rotation: FP64 pairs ran
    }
    if (B->flags & RPGO_BATCH_GRAPH) {
      float t = 0.f;
      (void)hipEventElapsedTime(&t, ev[0], ev[5]);
      stats->ms_total = t;
    } else {
      float a = 0.f, b = 0.f, c = 0.f, d = 0.f;
      (void)hipEventElapsedTime(&a, ev[0], ev[1]);
      (void)hipEventElapsedTime(&b, ev[1], ev[2]);
      (void)hipEventElapsedTime(&c, ev[2], ev[3]);
      (void)hipEventElapsedTime(&d, ev[3], ev[5]);
      stats->ms_fold = a;
      stats->ms_odom = b;
      stats->ms_pairs = c;
      stats->ms_clique = want_clique ? d : 0.0;
      stats->ms_total = a + b + c + d;
    }
    stats->ms_tune = ms_tune;
    stats->pair_waves = pw;
    stats->clique_waves = cw;
    stats->clique_prefetch = want_clique ? clique_walk_chunk() : 0;
    stats->graph_replayed = replayed;
  }
  return RPGO_OK;
}

int rpgo_pcm_batch_download_adj(rpgo_ctx* ctx, int32_t g, uint64_t* adj_bits) {
  if (!ctx || !ctx->batch || g < 0 || g >= ctx->batch->n_graphs || !adj_bits) {
    set_error("rpgo_pcm_batch_download_adj: invalid argument or no batch run");
    return RPGO_ERR_ARG;
  }
  BatchState& s = *ctx->batch;
  // graph g's region is sized for all its loop closures; the adjacency in it
  // has the accepted count's rows and stride
  const long long m = s.m_host[g], words = m * ((m + 63) / 64);
  if (words > 0)
    RPGO_HIP_CHECK(hipMemcpy(adj_bits, s.adj.as<unsigned long long>() + s.adj_off[g], sizeof(uint64_t) * words,
                             hipMemcpyDeviceToHost));
  return RPGO_OK;
}

}  // extern "C"
