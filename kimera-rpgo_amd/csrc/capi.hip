// C ABI entry points of librpgo_hip (declared in include/rpgo_hip.h).
#include <algorithm>
#include <cstdarg>
#include <cstring>
#include <vector>

#include "rpgo_internal.h"

namespace rpgo {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

static int upload(DevBuf& b, const void* src, size_t bytes, hipStream_t st) {
  int rc = b.reserve(bytes > 0 ? bytes : 8);
  if (rc) return rc;
  if (bytes) RPGO_HIP_CHECK(hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, st));
  return RPGO_OK;
}

static float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, a, b);
  return ms;
}

}  // namespace rpgo

using namespace rpgo;

extern "C" {

const char* rpgo_last_error(void) { return rpgo::g_err; }

const char* rpgo_version(void) { return "rpgo-mi355x 0.1.0"; }

int rpgo_ctx_create(int device, uint32_t flags, rpgo_ctx** out) {
  (void)flags;
  if (!out) {
    set_error("rpgo_ctx_create: out is NULL");
    return RPGO_ERR_ARG;
  }
  int count = 0;
  RPGO_HIP_CHECK(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) {
    set_error("rpgo_ctx_create: device %d out of range (%d devices)", device, count);
    return RPGO_ERR_ARG;
  }
  RPGO_HIP_CHECK(hipSetDevice(device));
  rpgo_ctx* c = new rpgo_ctx();
  c->device = device;
  if (hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
      c->n_cu <= 0)
    c->n_cu = 256;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    set_error("hipStreamCreate failed: %s", hipGetErrorString(e));
    delete c;
    return RPGO_ERR_HIP;
  }
  for (auto& ev : c->ev) {
    e = hipEventCreate(&ev);
    if (e != hipSuccess) {
      set_error("hipEventCreate failed: %s", hipGetErrorString(e));
      return RPGO_ERR_HIP;
    }
  }
  if (c->pcm.counters.reserve(256)) {
    delete c;
    return RPGO_ERR_NOMEM;
  }
  if (int rc = clique_noalloc_init(device)) {
    c->pcm.counters.release();
    delete c;
    return rc;
  }
  e = hipHostMalloc(&c->host_pinned, 256, hipHostMallocDefault);
  if (e != hipSuccess) {
    set_error("hipHostMalloc failed: %s", hipGetErrorString(e));
    c->host_pinned = nullptr;
    c->pcm.counters.release();
    delete c;
    return RPGO_ERR_NOMEM;
  }
  *out = c;
  return RPGO_OK;
}

int rpgo_ctx_destroy(rpgo_ctx* ctx) {
  if (!ctx) return RPGO_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  PcmState& s = ctx->pcm;
  for (DevBuf* b : {&s.pose0, &s.odom, &s.odom_cov, &s.lc_from, &s.lc_to, &s.lc_pose, &s.lc_cov,
                    &s.delta, &s.hinv, &s.traj, &s.lcm, &s.odom_dist, &s.flags, &s.accepted,
                    &s.c_from, &s.c_to, &s.c_lcm, &s.adj, &s.dist, &s.counters, &s.fp32_buf,
                    &s.queue, &s.tmp, &s.chain_off})
    b->release();
  CliqueState& q = ctx->clq;
  for (DevBuf* b : {&q.adj, &q.deg, &q.dmask, &q.icc, &q.out, &q.scratch, &q.best, &q.ex_adj,
                    &q.ex_ctl, &q.ex_stack, &q.ex_path})
    b->release();
  solver_destroy(ctx);
  mlp_destroy(ctx);
  batch_destroy(ctx);
  if (ctx->host_pinned) (void)hipHostFree(ctx->host_pinned);
  for (auto& ev : ctx->ev)
    if (ev) (void)hipEventDestroy(ev);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return RPGO_OK;
}

int rpgo_ctx_sync(rpgo_ctx* ctx) {
  if (!ctx) return RPGO_ERR_ARG;
  RPGO_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return RPGO_OK;
}

int rpgo_ctx_set_stage_timing(rpgo_ctx* ctx, int enabled) {
  if (!ctx) {
    set_error("rpgo_ctx_set_stage_timing: NULL context");
    return RPGO_ERR_ARG;
  }
  ctx->stage_timing = enabled != 0;
  return RPGO_OK;
}

int rpgo_pcm_upload(rpgo_ctx* ctx, const rpgo_pcm_problem* p) {
  if (!ctx || !p) {
    set_error("rpgo_pcm_upload: NULL argument");
    return RPGO_ERR_ARG;
  }
  if (p->dim != 2 && p->dim != 3) {
    set_error("rpgo_pcm_upload: dim must be 2 or 3 (got %d)", p->dim);
    return RPGO_ERR_ARG;
  }
  const int nc = p->n_chains > 1 ? p->n_chains : 1;
  const bool landmark = p->landmark != 0;
  if (p->n_poses < nc || p->n_lc < 0 || !p->pose0 ||
      (p->n_poses > nc && (!p->odom || !p->odom_cov)) ||
      (p->n_lc > 0 && (!p->lc_from || (!landmark && !p->lc_to) || !p->lc_pose || !p->lc_cov)) ||
      (nc > 1 && !p->chain_off)) {
    set_error("rpgo_pcm_upload: inconsistent sizes or NULL arrays (n_poses=%d n_lc=%d n_chains=%d)",
              p->n_poses, p->n_lc, p->n_chains);
    return RPGO_ERR_ARG;
  }
  if (landmark && nc > 1) {
    set_error("rpgo_pcm_upload: landmark observations come from one trajectory (n_chains <= 1)");
    return RPGO_ERR_ARG;
  }
  std::vector<int> chain_of;
  if (nc > 1) {
    if (p->chain_off[0] != 0 || p->chain_off[nc] != p->n_poses) {
      set_error("rpgo_pcm_upload: chain_off must run from 0 to n_poses");
      return RPGO_ERR_ARG;
    }
    chain_of.resize(p->n_poses);
    for (int c = 0; c < nc; ++c) {
      if (p->chain_off[c + 1] <= p->chain_off[c]) {
        set_error("rpgo_pcm_upload: chain %d is empty or chain_off decreases", c);
        return RPGO_ERR_ARG;
      }
      for (int k = p->chain_off[c]; k < p->chain_off[c + 1]; ++k) chain_of[k] = c;
    }
  }
  for (int k = 0; k < p->n_lc; ++k) {
    if (p->lc_from[k] < 0 || p->lc_from[k] >= p->n_poses ||
        (!landmark && (p->lc_to[k] < 0 || p->lc_to[k] >= p->n_poses))) {
      set_error("rpgo_pcm_upload: loop closure %d references pose (%d,%d) outside [0,%d)", k,
                p->lc_from[k], landmark ? p->lc_from[k] : p->lc_to[k], p->n_poses);
      return RPGO_ERR_ARG;
    }
    if (nc > 1) {
      // one group: every loop closure joins the same (unordered) pair of chains
      const int a0 = chain_of[p->lc_from[0]], b0 = chain_of[p->lc_to[0]];
      const int a = chain_of[p->lc_from[k]], b = chain_of[p->lc_to[k]];
      if (!((a == a0 && b == b0) || (a == b0 && b == a0))) {
        set_error("rpgo_pcm_upload: loop closure %d joins chains (%d,%d), the group joins (%d,%d)", k, a, b,
                  a0, b0);
        return RPGO_ERR_ARG;
      }
    }
  }
  RPGO_HIP_CHECK(hipSetDevice(ctx->device));
  PcmState& s = ctx->pcm;
  const int PL = pose_len(p->dim), D = tan_dim(p->dim);
  const size_t nd = (size_t)(p->n_poses - nc), nl = (size_t)p->n_lc;
  int rc;
  hipStream_t st = ctx->stream;
  if ((rc = upload(s.pose0, p->pose0, sizeof(double) * PL * nc, st))) return rc;
  if (nc > 1 && (rc = upload(s.chain_off, p->chain_off, sizeof(int) * (nc + 1), st))) return rc;
  if ((rc = upload(s.odom, p->odom, sizeof(double) * PL * nd, st))) return rc;
  if ((rc = upload(s.odom_cov, p->odom_cov, sizeof(double) * D * D * nd, st))) return rc;
  if ((rc = upload(s.lc_from, p->lc_from, sizeof(int) * nl, st))) return rc;
  // landmark observations: lc_to is not read; the pose index stands in for it
  if ((rc = upload(s.lc_to, landmark ? p->lc_from : p->lc_to, sizeof(int) * nl, st))) return rc;
  if ((rc = upload(s.lc_pose, p->lc_pose, sizeof(double) * PL * nl, st))) return rc;
  if ((rc = upload(s.lc_cov, p->lc_cov, sizeof(double) * D * D * nl, st))) return rc;
  RPGO_HIP_CHECK(hipStreamSynchronize(st));
  s.dim = p->dim;
  s.n_poses = p->n_poses;
  s.n_lc = p->n_lc;
  s.n_chains = nc;
  s.landmark = landmark;
  s.uploaded = true;
  s.ran = false;
  return RPGO_OK;
}

// Zero a pass's counters (16 words: the consistent-pair, flip and queue
// counts) and its adjacency (n words) in one launch on the context stream;
// the two hipMemsetAsync calls it replaces took about 16 us per pass at 1,000
// accepted loop closures.
__global__ void k_clear_pass(unsigned long long* __restrict__ counters, unsigned long long* __restrict__ adj,
                             long long n) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 16) counters[t] = 0ull;
  for (long long w = t; w < n; w += (long long)gridDim.x * blockDim.x) adj[w] = 0ull;
}

// The pairwise stage (A3 + A4 of the reference: areLoopsConsistent over every
// pair, adjacency fill) on the accepted set of the last run.  Events ev[8..11]
// bracket the pair kernel and the recheck on the context stream.
static int pairs_pass(rpgo_ctx* ctx, const rpgo_pcm_params* prm, rpgo_pcm_stats* stats,
                      const SimpleThr* simple = nullptr) {
  PcmState& s = ctx->pcm;
  hipStream_t st = ctx->stream;
  hipEvent_t* ev = ctx->ev;
  const bool timing = ctx->stage_timing;
  auto rec = [&](int k) { return timing ? hipEventRecord(ev[k], st) : hipSuccess; };
  const int m = s.m, W = (m + 63) / 64;
  const bool want_dist = prm->want_dist != 0;
  const bool check = simple ? (simple->trans >= 0.0 && simple->rot >= 0.0) : prm->lc_threshold >= 0.0;
  int rc;
  if ((rc = s.adj.reserve(sizeof(uint64_t) * (size_t)(m > 0 ? m : 1) * (W > 0 ? W : 1)))) return rc;
  if (want_dist && (rc = s.dist.reserve(sizeof(double) * (size_t)(m > 0 ? m : 1) * (m > 0 ? m : 1))))
    return rc;
  {
    const long long nw = m > 0 ? (long long)m * W : 0;
    const unsigned blocks = (unsigned)std::max(1ll, std::min((nw + 255) / 256, 1024ll));
    k_clear_pass<<<blocks, 256, 0, st>>>(static_cast<unsigned long long*>(s.counters.ptr),
                                         static_cast<unsigned long long*>(s.adj.ptr), nw);
    RPGO_HIP_CHECK(hipGetLastError());
  }
  if (want_dist && m > 0)
    RPGO_HIP_CHECK(hipMemsetAsync(s.dist.ptr, 0, sizeof(double) * (size_t)m * m, st));
  const bool fp32 = !simple && prm->precision == RPGO_PCM_FP32_RECHECK && !s.fp32_fallback;
  int64_t n_recheck = 0;
  RPGO_HIP_CHECK(rec(8));
  if (!(check && m > 1)) {  // no pairs: empty stage
    RPGO_HIP_CHECK(rec(9));
    RPGO_HIP_CHECK(rec(10));
  } else {
    if (simple) {
      if ((rc = pcm_pairs_simple(ctx, *simple, want_dist))) return rc;
      if ((rc = pcm_symmetrize(ctx))) return rc;
      RPGO_HIP_CHECK(rec(9));
      RPGO_HIP_CHECK(rec(10));
    } else if (!fp32) {
      if ((rc = pcm_pairs_fp64(ctx, prm->lc_threshold, want_dist))) return rc;
      if ((rc = pcm_symmetrize(ctx))) return rc;
      RPGO_HIP_CHECK(rec(9));
      RPGO_HIP_CHECK(rec(10));
    } else {
      // FP32 pairs, then the FP64 replay of the queued near-threshold pairs;
      // the queue length stays on the device (no host round trip)
      if ((rc = pcm_pairs_fp32_launch(ctx, prm->lc_threshold, prm->recheck_margin,
                                      prm->recheck_chain_coef, want_dist)))
        return rc;
      RPGO_HIP_CHECK(rec(9));
      if ((rc = pcm_recheck64_device(ctx, prm->lc_threshold, want_dist))) return rc;
      RPGO_HIP_CHECK(rec(10));
    }
  }
  RPGO_HIP_CHECK(rec(11));
  // counters (u64[0..3]) and the queue length (u32[16]) in one copy to pinned
  // memory, one wait
  RPGO_HIP_CHECK(hipMemcpyAsync(ctx->host_pinned, s.counters.ptr, 80, hipMemcpyDeviceToHost, st));
  RPGO_HIP_CHECK(hipStreamSynchronize(st));
  unsigned long long cnt[4];
  unsigned int nq = 0;
  std::memcpy(cnt, ctx->host_pinned, sizeof(cnt));
  std::memcpy(&nq, static_cast<const char*>(ctx->host_pinned) + 64, sizeof(nq));
  if (fp32 && check && m > 1) {
    n_recheck = nq;
    if (nq > (unsigned)kRecheckQueueCap) {
      // band too crowded for the queue (the device replay skipped it): replay
      // every pair in FP64
      RPGO_HIP_CHECK(hipMemsetAsync(s.adj.ptr, 0, sizeof(uint64_t) * (size_t)m * W, st));
      RPGO_HIP_CHECK(hipMemsetAsync(s.counters.ptr, 0, sizeof(unsigned long long) * 3, st));
      if ((rc = pcm_pairs_fp64(ctx, prm->lc_threshold, want_dist))) return rc;
      if ((rc = pcm_symmetrize(ctx))) return rc;
      RPGO_HIP_CHECK(rec(10));
      RPGO_HIP_CHECK(rec(11));
      RPGO_HIP_CHECK(hipStreamSynchronize(st));
      RPGO_HIP_CHECK(hipMemcpy(cnt, s.counters.ptr, sizeof(cnt), hipMemcpyDeviceToHost));
      n_recheck = (int64_t)m * (m - 1) / 2;
    }
  }
  s.ran = true;
  s.have_dist = want_dist;
  s.precision = simple ? RPGO_PCM_FP64_REPLAY : prm->precision;
  s.simple = simple != nullptr;
  if (stats) {
    stats->n_accepted = m;
    stats->reserved = 0;
    stats->n_pairs = check ? (int64_t)m * (m - 1) / 2 : 0;
    stats->n_rechecked = n_recheck;
    stats->n_flipped = (int64_t)cnt[2];
    stats->n_consistent = (int64_t)cnt[0];
    stats->ms_pairs = timing ? elapsed(ev[8], ev[9]) : 0.f;
    stats->ms_recheck = timing ? elapsed(ev[9], ev[10]) : 0.f;
  }
  return RPGO_OK;
}

int rpgo_pcm_run(rpgo_ctx* ctx, const rpgo_pcm_params* prm, rpgo_pcm_stats* stats) {
  if (!ctx || !prm) {
    set_error("rpgo_pcm_run: NULL argument");
    return RPGO_ERR_ARG;
  }
  PcmState& s = ctx->pcm;
  if (!s.uploaded) {
    set_error("rpgo_pcm_run: no problem uploaded");
    return RPGO_ERR_STATE;
  }
  if (prm->precision != RPGO_PCM_FP64_REPLAY && prm->precision != RPGO_PCM_FP32_RECHECK) {
    set_error("rpgo_pcm_run: unknown precision mode %d", prm->precision);
    return RPGO_ERR_ARG;
  }
  RPGO_HIP_CHECK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  hipEvent_t* ev = ctx->ev;
  int rc;
  RPGO_HIP_CHECK(hipMemsetAsync(s.counters.ptr, 0, 256, st));
  RPGO_HIP_CHECK(hipEventRecord(ev[0], st));
  if ((rc = pcm_prepare(ctx))) return rc;
  if ((rc = pcm_fold(ctx))) return rc;
  RPGO_HIP_CHECK(hipEventRecord(ev[1], st));
  int m = 0;
  if ((rc = pcm_odom_and_compact(ctx, prm->odom_threshold, &m))) return rc;
  RPGO_HIP_CHECK(hipEventRecord(ev[2], st));
  s.m = m;
  s.fp32_ready = false;
  s.fp32_fallback = false;
  const bool check = prm->lc_threshold >= 0.0;
  if (prm->precision == RPGO_PCM_FP32_RECHECK && check && m > 1) {
    bool use_fp64 = false;
    if ((rc = pcm_fp32_prepare(ctx, &use_fp64))) return rc;
    s.fp32_ready = !use_fp64;
    s.fp32_fallback = use_fp64;
  }
  RPGO_HIP_CHECK(hipEventRecord(ev[3], st));
  if ((rc = pairs_pass(ctx, prm, stats))) return rc;
  if (stats) {
    stats->ms_fold = elapsed(ev[0], ev[1]);
    stats->ms_odom = elapsed(ev[1], ev[2]);
    stats->ms_total = ctx->stage_timing ? elapsed(ev[0], ev[11]) : 0.f;
  }
  return RPGO_OK;
}

int rpgo_pcm_pairs(rpgo_ctx* ctx, const rpgo_pcm_params* prm, rpgo_pcm_stats* stats) {
  if (!ctx || !prm) {
    set_error("rpgo_pcm_pairs: NULL argument");
    return RPGO_ERR_ARG;
  }
  PcmState& s = ctx->pcm;
  if (!s.ran) {
    set_error("rpgo_pcm_pairs: call rpgo_pcm_run first (fold and odometry check)");
    return RPGO_ERR_STATE;
  }
  if (prm->precision != RPGO_PCM_FP64_REPLAY && prm->precision != RPGO_PCM_FP32_RECHECK) {
    set_error("rpgo_pcm_pairs: unknown precision mode %d", prm->precision);
    return RPGO_ERR_ARG;
  }
  RPGO_HIP_CHECK(hipSetDevice(ctx->device));
  int rc;
  if (prm->precision == RPGO_PCM_FP32_RECHECK && !s.fp32_ready && !s.fp32_fallback && s.m > 1) {
    bool use_fp64 = false;
    if ((rc = pcm_fp32_prepare(ctx, &use_fp64))) return rc;
    s.fp32_ready = !use_fp64;
    s.fp32_fallback = use_fp64;
  }
  if (ctx->stage_timing) RPGO_HIP_CHECK(hipEventRecord(ctx->ev[12], ctx->stream));
  if ((rc = pairs_pass(ctx, prm, stats))) return rc;
  if (stats) {
    stats->ms_fold = 0.0;
    stats->ms_odom = 0.0;
    stats->ms_total = ctx->stage_timing ? elapsed(ctx->ev[12], ctx->ev[11]) : 0.f;
  }
  return RPGO_OK;
}

int rpgo_pcm_simple_run(rpgo_ctx* ctx, const rpgo_pcm_simple_params* prm, rpgo_pcm_stats* stats) {
  if (!ctx || !prm) {
    set_error("rpgo_pcm_simple_run: NULL argument");
    return RPGO_ERR_ARG;
  }
  PcmState& s = ctx->pcm;
  if (!s.uploaded) {
    set_error("rpgo_pcm_simple_run: no problem uploaded");
    return RPGO_ERR_STATE;
  }
  RPGO_HIP_CHECK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  hipEvent_t* ev = ctx->ev;
  int rc;
  const SimpleThr odom{prm->odom_trans_threshold, prm->odom_rot_threshold};
  const SimpleThr loop{prm->dist_trans_threshold, prm->dist_rot_threshold};
  RPGO_HIP_CHECK(hipMemsetAsync(s.counters.ptr, 0, 256, st));
  RPGO_HIP_CHECK(hipEventRecord(ev[0], st));
  if ((rc = pcm_prepare(ctx))) return rc;
  if ((rc = pcm_fold(ctx))) return rc;
  RPGO_HIP_CHECK(hipEventRecord(ev[1], st));
  int m = 0;
  if ((rc = pcm_odom_and_compact(ctx, -1.0, &m, &odom))) return rc;
  RPGO_HIP_CHECK(hipEventRecord(ev[2], st));
  s.m = m;
  s.fp32_ready = false;
  s.fp32_fallback = false;
  rpgo_pcm_params pp{};
  pp.lc_threshold = -1.0;
  pp.precision = RPGO_PCM_FP64_REPLAY;
  pp.want_dist = prm->want_dist;
  if ((rc = pairs_pass(ctx, &pp, stats, &loop))) return rc;
  if (stats) {
    stats->ms_fold = elapsed(ev[0], ev[1]);
    stats->ms_odom = elapsed(ev[1], ev[2]);
    stats->ms_total = ctx->stage_timing ? elapsed(ev[0], ev[11]) : 0.f;
  }
  return RPGO_OK;
}

int rpgo_pcm_simple_adjacency(rpgo_ctx* ctx, const rpgo_pcm_problem* prob,
                              const rpgo_pcm_simple_params* params, int32_t* accepted,
                              double* odom_dist, uint64_t* adj_bits, double* dist,
                              rpgo_pcm_stats* stats) {
  int rc = rpgo_pcm_upload(ctx, prob);
  if (rc) return rc;
  if ((rc = rpgo_pcm_simple_run(ctx, params, stats))) return rc;
  return rpgo_pcm_download(ctx, accepted, odom_dist, adj_bits, dist);
}

}  // extern "C"

namespace {

// One reference update() after another (Pcm.h:148-281 with
// parseAndIncrementAdjMatrix :414-502): fold the new odometry onto the chain
// end, odometry-check the new loop closures against the extended chain and
// grow the adjacency by their rows (incrementAdjMatrix :725-768).
int pcm_append_impl(rpgo_ctx* ctx, const rpgo_pcm_increment* inc, double odom_thr, double lc_thr,
                    const SimpleThr* s_odom, const SimpleThr* s_loop, int32_t want_dist,
                    rpgo_pcm_stats* stats, const char* name) {
  if (!ctx || !inc) {
    set_error("%s: NULL argument", name);
    return RPGO_ERR_ARG;
  }
  PcmState& s = ctx->pcm;
  if (!s.ran) {
    set_error("%s: start the stream with a full run (rpgo_pcm_run / rpgo_pcm_simple_run)", name);
    return RPGO_ERR_STATE;
  }
  if (s.simple != (s_odom != nullptr)) {
    set_error("%s: the last run used the %s checks", name, s.simple ? "PcmSimple" : "PCM");
    return RPGO_ERR_STATE;
  }
  if (s.n_chains > 1 || s.landmark) {
    set_error("%s: streaming takes a problem with one chain and no landmark", name);
    return RPGO_ERR_STATE;
  }
  if (want_dist) {
    set_error("%s: the distance matrix is not kept while streaming (want_dist must be 0)", name);
    return RPGO_ERR_ARG;
  }
  const int n0 = s.n_poses, l0 = s.n_lc;
  if (inc->n_odom < 0 || inc->n_lc < 0 || (inc->n_odom > 0 && (!inc->odom || !inc->odom_cov)) ||
      (inc->n_lc > 0 && (!inc->lc_from || !inc->lc_to || !inc->lc_pose || !inc->lc_cov))) {
    set_error("%s: inconsistent sizes or NULL arrays (n_odom=%d n_lc=%d)", name, inc->n_odom, inc->n_lc);
    return RPGO_ERR_ARG;
  }
  const int n1 = n0 + inc->n_odom, l1 = l0 + inc->n_lc;
  for (int k = 0; k < inc->n_lc; ++k) {
    if (inc->lc_from[k] < 0 || inc->lc_from[k] >= n1 || inc->lc_to[k] < 0 || inc->lc_to[k] >= n1) {
      set_error("%s: loop closure %d references pose (%d,%d) outside [0,%d)", name, k, inc->lc_from[k],
                inc->lc_to[k], n1);
      return RPGO_ERR_ARG;
    }
  }
  RPGO_HIP_CHECK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  hipEvent_t* ev = ctx->ev;
  const size_t PL = pose_len(s.dim), D = tan_dim(s.dim), DD = D * D, PWC = pwc_len(s.dim);
  const size_t d0 = (size_t)(n0 - 1), d1 = (size_t)(n1 - 1);
  int rc;
  if ((rc = s.odom.grow_keep(8 * PL * (d1 > 0 ? d1 : 1), 8 * PL * d0, st))) return rc;
  if ((rc = s.odom_cov.grow_keep(8 * DD * (d1 > 0 ? d1 : 1), 8 * DD * d0, st))) return rc;
  if ((rc = s.delta.grow_keep(8 * PWC * (d1 > 0 ? d1 : 1), 8 * PWC * d0, st))) return rc;
  if ((rc = s.hinv.grow_keep(8 * DD * (d1 > 0 ? d1 : 1), 8 * DD * d0, st))) return rc;
  if ((rc = s.traj.grow_keep(8 * PWC * (size_t)n1, 8 * PWC * (size_t)n0, st))) return rc;
  const size_t L0 = (size_t)l0, L1 = (size_t)(l1 > 0 ? l1 : 1);
  if ((rc = s.lc_from.grow_keep(4 * L1, 4 * L0, st))) return rc;
  if ((rc = s.lc_to.grow_keep(4 * L1, 4 * L0, st))) return rc;
  if ((rc = s.lc_pose.grow_keep(8 * PL * L1, 8 * PL * L0, st))) return rc;
  if ((rc = s.lc_cov.grow_keep(8 * DD * L1, 8 * DD * L0, st))) return rc;
  if ((rc = s.lcm.grow_keep(8 * PWC * L1, 8 * PWC * L0, st))) return rc;
  if (inc->n_odom > 0) {
    RPGO_HIP_CHECK(hipMemcpyAsync(s.odom.as<double>() + PL * d0, inc->odom, 8 * PL * inc->n_odom,
                                  hipMemcpyHostToDevice, st));
    RPGO_HIP_CHECK(hipMemcpyAsync(s.odom_cov.as<double>() + DD * d0, inc->odom_cov, 8 * DD * inc->n_odom,
                                  hipMemcpyHostToDevice, st));
  }
  if (inc->n_lc > 0) {
    RPGO_HIP_CHECK(hipMemcpyAsync(s.lc_from.as<int>() + L0, inc->lc_from, 4 * (size_t)inc->n_lc,
                                  hipMemcpyHostToDevice, st));
    RPGO_HIP_CHECK(hipMemcpyAsync(s.lc_to.as<int>() + L0, inc->lc_to, 4 * (size_t)inc->n_lc,
                                  hipMemcpyHostToDevice, st));
    RPGO_HIP_CHECK(hipMemcpyAsync(s.lc_pose.as<double>() + PL * L0, inc->lc_pose, 8 * PL * inc->n_lc,
                                  hipMemcpyHostToDevice, st));
    RPGO_HIP_CHECK(hipMemcpyAsync(s.lc_cov.as<double>() + DD * L0, inc->lc_cov, 8 * DD * inc->n_lc,
                                  hipMemcpyHostToDevice, st));
  }
  s.n_poses = n1;
  s.n_lc = l1;
  s.fp32_ready = false;
  s.fp32_fallback = false;
  s.have_dist = false;
  RPGO_HIP_CHECK(hipMemsetAsync(s.counters.ptr, 0, 256, st));
  RPGO_HIP_CHECK(hipEventRecord(ev[0], st));
  if ((rc = pcm_measure_range(ctx, n0 - 1, n1 - 1, l0, l1))) return rc;
  if ((rc = pcm_fold_from(ctx, n0 - 1))) return rc;
  RPGO_HIP_CHECK(hipEventRecord(ev[1], st));
  int mn = 0;
  if ((rc = pcm_odom_append(ctx, l0, odom_thr, s_odom, &mn))) return rc;
  RPGO_HIP_CHECK(hipEventRecord(ev[2], st));
  const int m_old = s.m;
  s.m = m_old + mn;
  if ((rc = pcm_adj_append(ctx, m_old, lc_thr, s_loop))) return rc;
  RPGO_HIP_CHECK(hipEventRecord(ev[3], st));
  RPGO_HIP_CHECK(hipEventSynchronize(ev[3]));
  if (stats) {
    const bool check = s_loop ? (s_loop->trans >= 0.0 && s_loop->rot >= 0.0) : lc_thr >= 0.0;
    unsigned long long cnt = 0;
    RPGO_HIP_CHECK(hipMemcpy(&cnt, s.counters.ptr, sizeof(cnt), hipMemcpyDeviceToHost));
    const int m = s.m;
    *stats = rpgo_pcm_stats{};
    stats->n_accepted = m;
    stats->n_pairs = check ? ((int64_t)m * (m - 1) - (int64_t)m_old * (m_old - 1)) / 2 : 0;
    stats->n_consistent = (int64_t)cnt;
    stats->ms_fold = elapsed(ev[0], ev[1]);
    stats->ms_odom = elapsed(ev[1], ev[2]);
    stats->ms_pairs = elapsed(ev[2], ev[3]);
    stats->ms_total = elapsed(ev[0], ev[3]);
  }
  return RPGO_OK;
}

}  // namespace

extern "C" {

int rpgo_pcm_append(rpgo_ctx* ctx, const rpgo_pcm_increment* inc, const rpgo_pcm_params* prm,
                    rpgo_pcm_stats* stats) {
  if (!prm) {
    set_error("rpgo_pcm_append: NULL params");
    return RPGO_ERR_ARG;
  }
  return pcm_append_impl(ctx, inc, prm->odom_threshold, prm->lc_threshold, nullptr, nullptr,
                         prm->want_dist, stats, "rpgo_pcm_append");
}

int rpgo_pcm_simple_append(rpgo_ctx* ctx, const rpgo_pcm_increment* inc,
                           const rpgo_pcm_simple_params* prm, rpgo_pcm_stats* stats) {
  if (!prm) {
    set_error("rpgo_pcm_simple_append: NULL params");
    return RPGO_ERR_ARG;
  }
  const SimpleThr odom{prm->odom_trans_threshold, prm->odom_rot_threshold};
  const SimpleThr loop{prm->dist_trans_threshold, prm->dist_rot_threshold};
  return pcm_append_impl(ctx, inc, -1.0, -1.0, &odom, &loop, prm->want_dist, stats,
                         "rpgo_pcm_simple_append");
}

int rpgo_pcm_download(rpgo_ctx* ctx, int32_t* accepted, double* odom_dist, uint64_t* adj_bits,
                      double* dist) {
  if (!ctx) return RPGO_ERR_ARG;
  PcmState& s = ctx->pcm;
  if (!s.ran) {
    set_error("rpgo_pcm_download: nothing computed yet");
    return RPGO_ERR_STATE;
  }
  RPGO_HIP_CHECK(hipSetDevice(ctx->device));
  const int m = s.m, W = (m + 63) / 64;
  if (accepted && m > 0)
    RPGO_HIP_CHECK(hipMemcpy(accepted, s.accepted.ptr, sizeof(int) * m, hipMemcpyDeviceToHost));
  if (odom_dist && s.n_lc > 0)
    RPGO_HIP_CHECK(hipMemcpy(odom_dist, s.odom_dist.ptr, sizeof(double) * s.n_lc,
                             hipMemcpyDeviceToHost));
  if (adj_bits && m > 0)
    RPGO_HIP_CHECK(hipMemcpy(adj_bits, s.adj.ptr, sizeof(uint64_t) * (size_t)m * W,
                             hipMemcpyDeviceToHost));
  if (dist && m > 0) {
    if (!s.have_dist) {
      set_error("rpgo_pcm_download: run without want_dist");
      return RPGO_ERR_STATE;
    }
    RPGO_HIP_CHECK(hipMemcpy(dist, s.dist.ptr, sizeof(double) * (size_t)m * m,
                             hipMemcpyDeviceToHost));
  }
  return RPGO_OK;
}

int rpgo_pcm_download_traj(rpgo_ctx* ctx, double* traj) {
  if (!ctx || !traj) return RPGO_ERR_ARG;
  PcmState& s = ctx->pcm;
  if (!s.ran) {
    set_error("rpgo_pcm_download_traj: nothing computed yet");
    return RPGO_ERR_STATE;
  }
  RPGO_HIP_CHECK(hipSetDevice(ctx->device));
  RPGO_HIP_CHECK(hipMemcpy(traj, s.traj.ptr, sizeof(double) * (size_t)s.n_poses * pwc_len(s.dim),
                           hipMemcpyDeviceToHost));
  return RPGO_OK;
}

int rpgo_pcm_adjacency(rpgo_ctx* ctx, const rpgo_pcm_problem* prob, const rpgo_pcm_params* params,
                       int32_t* accepted, double* odom_dist, uint64_t* adj_bits, double* dist,
                       rpgo_pcm_stats* stats) {
  int rc = rpgo_pcm_upload(ctx, prob);
  if (rc) return rc;
  if ((rc = rpgo_pcm_run(ctx, params, stats))) return rc;
  return rpgo_pcm_download(ctx, accepted, odom_dist, adj_bits, dist);
}

int rpgo_max_clique(rpgo_ctx* ctx, int32_t mode, int32_t* idx_out, int32_t* k_out,
                    double* ms_out) {
  if (!ctx || !k_out) {
    set_error("rpgo_max_clique: NULL argument");
    return RPGO_ERR_ARG;
  }
  PcmState& s = ctx->pcm;
  if (!s.ran) {
    set_error("rpgo_max_clique: no adjacency computed yet");
    return RPGO_ERR_STATE;
  }
  RPGO_HIP_CHECK(hipSetDevice(ctx->device));
  RPGO_HIP_CHECK(hipEventRecord(ctx->ev[5], ctx->stream));
  int rc = clique_run(ctx, s.adj.as<uint64_t>(), s.m, mode, idx_out, k_out);
  if (rc) return rc;
  RPGO_HIP_CHECK(hipEventRecord(ctx->ev[6], ctx->stream));
  RPGO_HIP_CHECK(hipEventSynchronize(ctx->ev[6]));
  if (ms_out) *ms_out = elapsed(ctx->ev[5], ctx->ev[6]);
  return RPGO_OK;
}

int rpgo_max_clique_bits(rpgo_ctx* ctx, int32_t n, const uint64_t* adj_bits, int32_t mode,
                         int32_t* idx_out, int32_t* k_out) {
  if (!ctx || !k_out || n < 0 || (n > 0 && !adj_bits)) {
    set_error("rpgo_max_clique_bits: invalid argument");
    return RPGO_ERR_ARG;
  }
  if (n > clique_max_vertices()) {  // checked before the adjacency is copied
    set_error("rpgo_max_clique_bits: %d vertices exceed the %d supported", n, clique_max_vertices());
    return RPGO_ERR_ARG;
  }
  RPGO_HIP_CHECK(hipSetDevice(ctx->device));
  const int W = (n + 63) / 64;
  int rc = upload(ctx->clq.adj, adj_bits, sizeof(uint64_t) * (size_t)n * W, ctx->stream);
  if (rc) return rc;
  return clique_run(ctx, ctx->clq.adj.as<uint64_t>(), n, mode, idx_out, k_out);
}

int rpgo_max_clique_bits_noalloc(rpgo_ctx* ctx, int32_t n, const uint64_t* d_adj_bits, int32_t* idx_out,
                                 int32_t* k_out) {
  if (!ctx || !k_out || n < 0 || (n > 0 && !d_adj_bits)) {
    set_error("rpgo_max_clique_bits_noalloc: invalid argument");
    return RPGO_ERR_ARG;
  }
  RPGO_HIP_CHECK(hipSetDevice(ctx->device));
  return clique_run_noalloc(ctx, d_adj_bits, n, idx_out, k_out);
}

int rpgo_max_clique_incremental(rpgo_ctx* ctx, int32_t num_new, int32_t prev_size, int32_t* idx_out,
                                int32_t* k_out) {
  if (!ctx || !k_out) {
    set_error("rpgo_max_clique_incremental: NULL argument");
    return RPGO_ERR_ARG;
  }
  PcmState& s = ctx->pcm;
  if (!s.ran) {
    set_error("rpgo_max_clique_incremental: no adjacency computed yet");
    return RPGO_ERR_STATE;
  }
  RPGO_HIP_CHECK(hipSetDevice(ctx->device));
  return clique_incremental(ctx, s.adj.as<uint64_t>(), s.m, num_new, prev_size, idx_out, k_out);
}

int rpgo_max_clique_incremental_bits(rpgo_ctx* ctx, int32_t n, const uint64_t* adj_bits, int32_t num_new,
                                     int32_t prev_size, int32_t* idx_out, int32_t* k_out) {
  if (!ctx || !k_out || n < 0 || (n > 0 && !adj_bits)) {
    set_error("rpgo_max_clique_incremental_bits: invalid argument");
    return RPGO_ERR_ARG;
  }
  RPGO_HIP_CHECK(hipSetDevice(ctx->device));
  const int W = (n + 63) / 64;
  int rc = upload(ctx->clq.adj, adj_bits, sizeof(uint64_t) * (size_t)n * W, ctx->stream);
  if (rc) return rc;
  return clique_incremental(ctx, ctx->clq.adj.as<uint64_t>(), n, num_new, prev_size, idx_out, k_out);
}

}  // extern "C"
