// Internal declarations shared by the HIP translation units of librpgo_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "rpgo_hip.h"

namespace rpgo {

void set_error(const char* fmt, ...);

#define RPGO_HIP_CHECK(expr)                                                   \
  do {                                                                         \
    hipError_t _e = (expr);                                                    \
    if (_e != hipSuccess) {                                                    \
      ::rpgo::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                        __FILE__, __LINE__);                                   \
      return RPGO_ERR_HIP;                                                     \
    }                                                                          \
  } while (0)

// Grow-only device buffer.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  int reserve(size_t n) {
    if (n <= bytes) return RPGO_OK;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
    size_t want = n < 256 ? 256 : n;
    hipError_t e = hipMalloc(&ptr, want);
    if (e != hipSuccess) {
      set_error("hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
      ptr = nullptr;
      return RPGO_ERR_NOMEM;
    }
    bytes = want;
    return RPGO_OK;
  }
  // grow to at least n bytes (doubling), keeping the first `keep` bytes
  int grow_keep(size_t n, size_t keep, hipStream_t st) {
    if (n <= bytes) return RPGO_OK;
    size_t want = bytes * 2 > n ? bytes * 2 : n;
    if (want < 256) want = 256;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
      set_error("hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
      return RPGO_ERR_NOMEM;
    }
    if (ptr && keep) {
      e = hipMemcpyAsync(p, ptr, keep < bytes ? keep : bytes, hipMemcpyDeviceToDevice, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      if (e != hipSuccess) {
        set_error("grow_keep copy failed: %s", hipGetErrorString(e));
        (void)hipFree(p);
        return RPGO_ERR_HIP;
      }
    }
    if (ptr) (void)hipFree(ptr);
    ptr = p;
    bytes = want;
    return RPGO_OK;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const { return static_cast<T*>(ptr); }
};

inline int pose_len(int dim) { return dim == 3 ? 12 : 4; }
inline int tan_dim(int dim) { return dim == 3 ? 6 : 3; }
// PoseWithCovariance record: pose, covariance, rotation_info flag (as double)
inline int pwc_len(int dim) { return pose_len(dim) + tan_dim(dim) * tan_dim(dim) + 1; }

struct PcmState {
  int dim = 0;
  int n_poses = 0;
  int n_lc = 0;
  int n_chains = 1;  // robot trajectories of the uploaded problem
  bool landmark = false;  // loop closures are observations of one landmark
  DevBuf chain_off;  // n_chains + 1 ints (several chains only)
  bool uploaded = false;
  bool ran = false;
  int m = 0;  // accepted loop closures of the last run
  bool have_dist = false;
  int precision = 0;     // precision of the last run
  bool simple = false;   // last run used the PcmSimple (PoseWithNode) checks
  bool fp32_ready = false;  // FP32 tables built for the current accepted set
  bool fp32_fallback = false;  // FP32 path declined (SE(2) or no rotation info)
  int fp32_nb = 0;       // anchor blocks of the FP32 tables
  // inputs
  DevBuf pose0, odom, odom_cov, lc_from, lc_to, lc_pose, lc_cov;
  // derived
  DevBuf delta;      // n_poses-1 pwc records of the odometry measurements
  DevBuf hinv;       // n_poses-1 Ad(delta^-1) matrices
  DevBuf traj;       // n_poses pwc records (FP64 fold)
  DevBuf lcm;        // n_lc pwc records of the loop-closure measurements
  DevBuf odom_dist;  // n_lc doubles
  DevBuf flags;      // n_lc ints
  DevBuf accepted;   // n_lc ints (+1 for the count)
  DevBuf c_from, c_to, c_lcm;  // compacted accepted loop closures
  DevBuf adj;        // m*W uint64
  DevBuf dist;       // m*m double
  DevBuf counters;   // device counters (recheck, flips, consistent)
  DevBuf fp32_buf;   // FP32-mode per-pose / per-LC tables
  DevBuf queue;      // FP32-mode recheck queue
  DevBuf tmp;        // append: compaction of the new loop closures
};

struct CliqueState {
  DevBuf adj;      // caller-provided adjacency (rpgo_max_clique_bits)
  DevBuf deg, dmask, icc, out, scratch, best;
  DevBuf cidx, ccol, cdmask, cadj, cn;  // the heuristic rounds' compacted graph
  DevBuf ex_adj, ex_ctl, ex_stack, ex_path;  // exact mode (clique_exact.hip)
};

struct SolverRun;  // solver.hip / rpgo_solver.h
struct SolveLanes; // solver_capi.hip: streams and solver states of rpgo_solve_batched
struct MlpState;   // learned.hip
struct BatchState; // batch.hip

}  // namespace rpgo

struct rpgo_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[16] = {};
  rpgo::PcmState pcm;
  rpgo::CliqueState clq;
  rpgo::SolverRun* solver = nullptr;  // created on the first rpgo_solve
  rpgo::SolveLanes* solve_lanes = nullptr;  // created by rpgo_solve_batched
  rpgo::MlpState* mlp = nullptr;      // created by rpgo_mlp_set_params
  rpgo::BatchState* batch = nullptr;  // created by rpgo_pcm_batch_run
  void* host_pinned = nullptr;        // 256 B pinned host buffer for counter readbacks
  int n_cu = 256;                     // compute units of the device (launch shapes)
  bool stage_timing = true;           // HIP events around the PCM pass stages (ms_* of the stats)
};

namespace rpgo {
// PcmSimple thresholds of one check (translation, rotation per node);
// a negative value disables the check
struct SimpleThr {
  double trans, rot;
};
// replay64.hip
int pcm_prepare(rpgo_ctx* ctx);                  // delta/hinv/lcm precompute
int pcm_fold(rpgo_ctx* ctx);                     // sequential FP64 fold
// odometry check (PCM Mahalanobis, or PcmSimple when simple != nullptr) and
// stable compaction of the accepted loop closures
int pcm_odom_and_compact(rpgo_ctx* ctx, double odom_thr, int* m_out,
                         const SimpleThr* simple = nullptr);
int pcm_pairs_fp64(rpgo_ctx* ctx, double lc_thr, bool want_dist);
int pcm_pairs_simple(rpgo_ctx* ctx, const SimpleThr& thr, bool want_dist);
// measurement records of odometry [o0, o1) and loop closures [l0, l1)
int pcm_measure_range(rpgo_ctx* ctx, int o0, int o1, int l0, int l1);
// continue the fold from chain record `start` up to record n_poses-1
int pcm_fold_from(rpgo_ctx* ctx, int start);
// odometry check of loop closures [l0, s.n_lc); accepted ones are appended
// to the compacted set (s.accepted, c_from, c_to, c_lcm); *m_new receives
// their count.  s.m is not changed.
int pcm_odom_append(rpgo_ctx* ctx, int l0, double odom_thr, const SimpleThr* simple, int* m_new);
// adjacency rows [m_old, s.m) against every earlier accepted loop closure,
// both triangles; the adjacency is re-laid out when W = ceil(m/64) grows
int pcm_adj_append(rpgo_ctx* ctx, int m_old, double lc_thr, const SimpleThr* simple);
int pcm_symmetrize(rpgo_ctx* ctx);
// FP64 replay of queued (i, j) pairs; patches the adjacency bits, counts flips
int pcm_recheck64(rpgo_ctx* ctx, int n_queue, double lc_thr, bool want_dist);
// the same for the queue the FP32 kernel left on the device: the length is
// read on the device; a queue longer than kRecheckQueueCap is left untouched
// (the host sees the length after the stage and replays every pair)
int pcm_recheck64_device(rpgo_ctx* ctx, double lc_thr, bool want_dist);
// pcm_fp32.hip
int pcm_fp32_prepare(rpgo_ctx* ctx, bool* use_fp64);
int pcm_pairs_fp32_launch(rpgo_ctx* ctx, double lc_thr, double margin, double chain_coef,
                          bool want_dist);
int pcm_count_consistent(rpgo_ctx* ctx);
constexpr int kRecheckQueueCap = 1 << 22;

// Batched FP32 pair stage (batch.hip drives it, pcm_fp32.hip runs it): every
// graph of an SE(3) batch gets the single-graph FP32 tables (anchor blocks of
// 64 poses, records, anchor table) at the offsets below, and one pair launch
// covers all graphs.  Device pointers; offsets are per-graph prefix sums.
struct BatchFp32Args {
  int G = 0;
  long long n_poses = 0, n_lc = 0, n_odom = 0;
  long long n_blocks = 0;       // sum of anchor blocks ceil(N_g / 64)
  long long n_sq = 0;           // sum of nb_g^2 (anchor table entries)
  long long n_pair_blocks = 0;  // workgroups of the pair launch
  int S = 4;                    // runs per 64-wide word of older loop closures
  const long long *pose_off = nullptr, *lc_off = nullptr, *adj_off = nullptr, *blk_off = nullptr,
                  *sq_off = nullptr, *pblk_off = nullptr;
  const int* m = nullptr;       // accepted loop closures per graph
  const double *delta = nullptr, *hinv = nullptr, *lcm = nullptr, *pose0 = nullptr;
  const int *c_from = nullptr, *c_to = nullptr;
  const double* c_lcm = nullptr;
  void *prec = nullptr, *lcr = nullptr, *sanc = nullptr;  // PoseRec, LcRec, AncRec (pcm_fp32.hip)
  double *pl64 = nullptr, *pb64 = nullptr, *bend = nullptr, *ta = nullptr, *terms = nullptr;
  unsigned long long* adj = nullptr;   // lower triangle written (row j, bits i < j)
  int2* queue = nullptr;               // (lc_off[g] + i, lc_off[g] + j) of queued pairs
  unsigned int* qcount = nullptr;
  long long qcap = 0;
  unsigned long long* n_consistent = nullptr;
  int* no_rotation = nullptr;          // set when a covariance lacks rotation information
};
// bytes of the per-pose, per-LC, per-block and per-entry records
size_t pcm_fp32_pose_rec_bytes();
size_t pcm_fp32_lc_rec_bytes();
size_t pcm_fp32_anchor_entry_bytes();
// rotation-flag check, tables and pair launch; the pair launch does nothing
// when *no_rotation is set (the caller then runs the FP64 pairs)
int pcm_fp32_batch_launch(hipStream_t st, const BatchFp32Args& a, double lc_thr, double margin,
                          double chain_coef);

// counters buffer layout (256 bytes, zeroed per pairwise pass)
//   u64[0] consistent pairs i < j (pair kernels, moved by the FP64 recheck)
//   u64[1] popcount of the adjacency
//   u64[2] FP64 recheck flips               u32[16] recheck queue length
//   i32[32] "some covariance has no rotation information" flag
// solver_capi.hip
void solver_destroy(rpgo_ctx* ctx);
void mlp_destroy(rpgo_ctx* ctx);  // learned.hip
void batch_destroy(rpgo_ctx* ctx);  // batch.hip
int clique_incremental(rpgo_ctx* ctx, const uint64_t* d_adj_u64, int n, int num_new, int prev_size,
                       int32_t* idx_out, int32_t* k_out);
// clique_exact.hip: exact maximum clique given a clique size known to exist
// batched FMC::maxCliqueHeu, one workgroup of `waves` wavefronts (4, 8, 16)
// per graph; out_k[g] = -1 for an empty graph, INT_MIN on an internal error
// Optional chip-wide rounds before the per-graph scan (spread != NULL):
// vertex slots are indexed like the loop closures (n_slots of them, graph of
// each in slot_graph); state, icc and dm are scratch of
// clique_spread_state_bytes(n_graphs), clique_spread_icc_bytes(n_graphs) and
// 8 * n_slots bytes.
struct CliqueSpread {
  long long n_slots;
  long long max_slots;  // slots of the largest graph
  const int* slot_graph;
  void* state;
  int* icc;
  unsigned long long* dm;
  const int* xorder;  // graphs grouped by XCD (n_graphs entries), then 9 group offsets
  int xmax;           // graphs of the largest group
};
size_t clique_spread_state_bytes(int n_graphs);
size_t clique_spread_icc_bytes(int n_graphs);
int clique_batch_launch(rpgo_ctx* ctx, int n_graphs, int waves, const long long* d_adj_off,
                        const long long* d_lc_off, const int* d_m, const unsigned long long* adj, int* deg,
                        int* out_idx, int* out_k, const CliqueSpread* spread = nullptr);
int clique_max_vertices();
int clique_walk_chunk();  // rows the greedy walk loads together
int clique_exact_run(rpgo_ctx* ctx, const unsigned long long* adj, int n, int lower_bound, int32_t* idx_out,
                     int32_t* k_out, long long node_budget);
// clique.hip
int clique_run(rpgo_ctx* ctx, const uint64_t* d_adj, int n, int mode, int32_t* idx_out,
               int32_t* k_out);
// allocation-free heuristic scan on the per-device scratch array
// (clique_noalloc_init once per device, with that device current)
int clique_noalloc_init(int device);
size_t clique_noalloc_bytes();
int clique_run_noalloc(rpgo_ctx* ctx, const uint64_t* d_adj, int n, int32_t* idx_out, int32_t* k_out);
}  // namespace rpgo
