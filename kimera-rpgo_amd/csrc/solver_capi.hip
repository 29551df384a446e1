// C ABI of the robust solve (rpgo_solve, rpgo_solve_batched) and its
// host-side set-up.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "rpgo_solver.h"

namespace rpgo {

// Regularized lower incomplete gamma P(a, x): power series for x < a + 1,
// Lentz continued fraction for the complement otherwise.
static double gamma_p(double a, double x) {
  if (x <= 0.0) return 0.0;
  const double lg = std::lgamma(a);
  if (x < a + 1.0) {
    double ap = a, sum = 1.0 / a, del = sum;
    for (int n = 0; n < 1000; ++n) {
      ap += 1.0;
      del *= x / ap;
      sum += del;
      if (std::fabs(del) < std::fabs(sum) * 1e-16) break;
    }
    return sum * std::exp(-x + a * std::log(x) - lg);
  }
  double b = x + 1.0 - a, c = 1.0 / 1e-300, d = 1.0 / b, h = d;
  for (int i = 1; i < 1000; ++i) {
    const double an = -i * (i - a);
    b += 2.0;
    d = an * d + b;
    if (std::fabs(d) < 1e-300) d = 1e-300;
    c = b + an / c;
    if (std::fabs(c) < 1e-300) c = 1e-300;
    d = 1.0 / d;
    const double del = d * c;
    h *= del;
    if (std::fabs(del - 1.0) < 1e-16) break;
  }
  return 1.0 - std::exp(-x + a * std::log(x) - lg) * h;
}

double chi2_quantile(double p, double dof) {
  if (p <= 0.0) return 0.0;
  if (p >= 1.0) return INFINITY;
  double lo = 0.0, hi = std::max(1.0, dof);
  while (gamma_p(0.5 * dof, 0.5 * hi) < p) hi *= 2.0;
  for (int k = 0; k < 200; ++k) {
    const double mid = 0.5 * (lo + hi);
    if (gamma_p(0.5 * dof, 0.5 * mid) < p) lo = mid;
    else hi = mid;
    if (hi - lo <= 1e-15 * hi) break;
  }
  return 0.5 * (lo + hi);
}

// One lane of rpgo_solve_batched: a stream, two events and a solver state.
struct SolveLane {
  hipStream_t st = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  SolverRun S;
};

struct SolveLanes {
  std::vector<SolveLane*> lanes;
};

void solver_destroy(rpgo_ctx* ctx) {
  if (ctx->solver) {
    ctx->solver->release();
    delete ctx->solver;
    ctx->solver = nullptr;
  }
  if (ctx->solve_lanes) {
    for (SolveLane* l : ctx->solve_lanes->lanes) {
      l->S.release();
      if (l->ev0) (void)hipEventDestroy(l->ev0);
      if (l->ev1) (void)hipEventDestroy(l->ev1);
      if (l->st) (void)hipStreamDestroy(l->st);
      delete l;
    }
    delete ctx->solve_lanes;
    ctx->solve_lanes = nullptr;
  }
}

// The checks of rpgo_solve on one problem (everything but the context).
static int check_problem(const char* who, const rpgo_solve_problem* pr, const rpgo_solve_params* prm) {
  const int N = pr->n_poses, F = pr->n_factors;
  if (N < 1 || F < 0 || !pr->poses || (F > 0 && (!pr->kind || !pr->key_i || !pr->key_j ||
                                                  !pr->measured || !pr->information))) {
    set_error("%s: inconsistent problem (n_poses=%d n_factors=%d)", who, N, F);
    return RPGO_ERR_ARG;
  }
  if (prm->solver != RPGO_SOLVER_LM && prm->solver != RPGO_SOLVER_GN) {
    set_error("%s: unknown solver %d", who, prm->solver);
    return RPGO_ERR_ARG;
  }
  if (!(prm->lambda_factor > 1.0) || !(prm->lambda_upper_bound > 0.0) || prm->lambda_initial < 0.0 ||
      prm->lm_max_iterations < 0 || prm->gnc_max_iterations < 0 || !(prm->gnc_mu_step > 0.0)) {
    set_error("%s: invalid LM/GNC parameters (lambda_factor must exceed 1)", who);
    return RPGO_ERR_ARG;
  }
  if (prm->pcg_coarse > kCoarseSegMax) {
    set_error("%s: pcg_coarse %d above the maximum of %d segments", who, prm->pcg_coarse,
              kCoarseSegMax);
    return RPGO_ERR_ARG;
  }
  if (prm->known_inliers < 0 || prm->known_inliers > F) {
    set_error("%s: known_inliers %d outside [0, %d]", who, prm->known_inliers, F);
    return RPGO_ERR_ARG;
  }
  for (int f = 0; f < F; ++f) {
    const int k = pr->kind[f];
    if ((k != RPGO_FACTOR_BETWEEN && k != RPGO_FACTOR_PRIOR) || pr->key_i[f] < 0 ||
        pr->key_i[f] >= N ||
        (k == RPGO_FACTOR_BETWEEN && (pr->key_j[f] < 0 || pr->key_j[f] >= N || pr->key_j[f] == pr->key_i[f]))) {
      set_error("%s: factor %d has kind %d keys (%d, %d) outside [0, %d)", who, f, k,
                pr->key_i[f], pr->key_j[f], N);
      return RPGO_ERR_ARG;
    }
  }
  return RPGO_OK;
}

// Solve one checked problem on stream st with solver state S; ev0/ev1
// bracket the device time.  Blocks until the results are on the host.
static int solve_one(hipStream_t st, hipEvent_t ev0, hipEvent_t ev1, SolverRun& S,
                     const rpgo_solve_problem* pr, const rpgo_solve_params* prm, double* poses_out,
                     double* weights_out, rpgo_solve_report* rep) {
  const int N = pr->n_poses, F = pr->n_factors;
  int rc;
  if ((rc = S.alloc(N, F > 0 ? F : 1))) return rc;
  S.F = F;
  S.prm = *prm;
  S.n_known = prm->known_inliers;
  // host-side set-up of the per-factor constants
  std::vector<int> kj(F);
  std::vector<double> w0(F), barc(F);
  std::vector<unsigned char> known(F);
  const double thr = prm->gnc_threshold_mode == RPGO_GNC_PROBABILITY
                         ? 0.5 * chi2_quantile(prm->gnc_inlier_threshold,
                                               prm->gnc_factor_dim > 0 ? prm->gnc_factor_dim : 6.0)
                         : prm->gnc_inlier_threshold;
  for (int f = 0; f < F; ++f) {
    kj[f] = pr->kind[f] == RPGO_FACTOR_BETWEEN ? pr->key_j[f] : pr->key_i[f];
    known[f] = f < prm->known_inliers ? 1 : 0;
    barc[f] = thr;
    if (prm->init_weights) w0[f] = prm->init_weights[f];
    else if (prm->use_gnc && prm->gnc_bias_odom) w0[f] = known[f] ? 1.0 : 0.0;
    else w0[f] = 1.0;
  }
  RPGO_HIP_CHECK(hipMemcpyAsync(S.x0.ptr, pr->poses, sizeof(double) * 12 * N, hipMemcpyHostToDevice, st));
  if (F > 0) {
    RPGO_HIP_CHECK(hipMemcpyAsync(S.kind.ptr, pr->kind, sizeof(int) * F, hipMemcpyHostToDevice, st));
    RPGO_HIP_CHECK(hipMemcpyAsync(S.fi.ptr, pr->key_i, sizeof(int) * F, hipMemcpyHostToDevice, st));
    RPGO_HIP_CHECK(hipMemcpyAsync(S.fj.ptr, kj.data(), sizeof(int) * F, hipMemcpyHostToDevice, st));
    RPGO_HIP_CHECK(hipMemcpyAsync(S.z.ptr, pr->measured, sizeof(double) * 12 * F, hipMemcpyHostToDevice, st));
    RPGO_HIP_CHECK(hipMemcpyAsync(S.info.ptr, pr->information, sizeof(double) * 36 * F,
                                  hipMemcpyHostToDevice, st));
    RPGO_HIP_CHECK(hipMemcpyAsync(S.w.ptr, w0.data(), sizeof(double) * F, hipMemcpyHostToDevice, st));
    RPGO_HIP_CHECK(hipMemcpyAsync(S.barc.ptr, barc.data(), sizeof(double) * F, hipMemcpyHostToDevice, st));
    RPGO_HIP_CHECK(hipMemcpyAsync(S.known.ptr, known.data(), F, hipMemcpyHostToDevice, st));
  }
  RPGO_HIP_CHECK(hipStreamSynchronize(st));
  RPGO_HIP_CHECK(hipEventRecord(ev0, st));
  if (F > 0) {
    if ((rc = S.run(st, rep))) return rc;
  } else {
    RPGO_HIP_CHECK(hipMemcpyAsync(S.x.ptr, S.x0.ptr, sizeof(double) * 12 * N, hipMemcpyDeviceToDevice, st));
  }
  RPGO_HIP_CHECK(hipEventRecord(ev1, st));
  std::vector<double> wout(F > 0 ? F : 1);
  RPGO_HIP_CHECK(hipMemcpyAsync(poses_out, S.x.ptr, sizeof(double) * 12 * N, hipMemcpyDeviceToHost, st));
  if (F > 0)
    RPGO_HIP_CHECK(hipMemcpyAsync(wout.data(), S.w.ptr, sizeof(double) * F, hipMemcpyDeviceToHost, st));
  RPGO_HIP_CHECK(hipStreamSynchronize(st));
  if (weights_out && F > 0) std::memcpy(weights_out, wout.data(), sizeof(double) * F);
  if (rep) {
    if (F == 0) std::memset(rep, 0, sizeof(*rep));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev0, ev1);
    rep->ms_total = ms;
    double sum = 0.0;
    for (int f = 0; f < F; ++f) sum += wout[f];
    rep->n_inliers = (int32_t)((size_t)sum - (size_t)prm->known_inliers);
  }
  return RPGO_OK;
}

}  // namespace rpgo

using namespace rpgo;

extern "C" {

void rpgo_solve_params_default(rpgo_solve_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->solver = RPGO_SOLVER_LM;
  p->use_gnc = 0;
  p->known_inliers = 0;
  p->gnc_threshold_mode = RPGO_GNC_PROBABILITY;
  p->gnc_inlier_threshold = 0.9;
  p->gnc_max_iterations = 100;
  p->gnc_mu_step = 1.4;
  p->gnc_relative_cost_tol = 1e-5;
  p->gnc_weights_tol = 1e-4;
  p->gnc_bias_odom = 0;
  p->lm_diagonal_damping = 1;
  p->init_weights = nullptr;
  p->init_mu = 0.0;
  p->lambda_initial = 1e-5;
  p->lambda_factor = 10.0;
  p->lambda_upper_bound = 1e5;
  p->min_model_fidelity = 1e-3;
  p->lm_max_iterations = 100;
  p->lm_relative_error_tol = 1e-5;
  p->lm_absolute_error_tol = 1e-5;
  p->pcg_max_iterations = 0;
  // Steps to 1e-6 relative residual keep LM on the direct solve's path: on
  // the 5k and 50k helices (GNC over LM) the GNC and LM iteration counts, the
  // weights and the poses against the FP64 oracle are the same as at 1e-8,
  // with 25 % / 21 % fewer PCG iterations
  p->pcg_rel_tol = 1e-6;
  p->polish_iterations = 0;
  p->pcg_segment = 0;
  p->gnc_warm_start = 0;
  p->pcg_coarse = 0;
  p->pcg_graphs = 1;
  p->gnc_factor_dim = 6;
  p->pose2_chart = 0;
}

double rpgo_gnc_threshold(double alpha, int32_t dof) {
  return 0.5 * chi2_quantile(alpha, dof > 0 ? dof : 6);
}

int rpgo_solve_errors(rpgo_ctx* ctx, const double* poses, double* err_out) {
  if (!ctx || !err_out) {
    set_error("rpgo_solve_errors: NULL argument");
    return RPGO_ERR_ARG;
  }
  if (!ctx->solver || ctx->solver->N <= 0) {
    set_error("rpgo_solve_errors: no rpgo_solve problem in this context");
    return RPGO_ERR_STATE;
  }
  RPGO_HIP_CHECK(hipSetDevice(ctx->device));
  SolverRun& S = *ctx->solver;
  hipStream_t st = ctx->stream;
  if (S.F <= 0) return RPGO_OK;
  const double* xp = S.x.as<double>();
  if (poses) {
    RPGO_HIP_CHECK(hipMemcpyAsync(S.xn.ptr, poses, sizeof(double) * 12 * S.N, hipMemcpyHostToDevice, st));
    xp = S.xn.as<double>();
  }
  int rc;
  if ((rc = S.compute_errors(st, xp))) return rc;
  RPGO_HIP_CHECK(hipMemcpyAsync(err_out, S.err.ptr, sizeof(double) * S.F, hipMemcpyDeviceToHost, st));
  RPGO_HIP_CHECK(hipStreamSynchronize(st));
  return RPGO_OK;
}

int rpgo_solve(rpgo_ctx* ctx, const rpgo_solve_problem* pr, const rpgo_solve_params* prm,
               double* poses_out, double* weights_out, rpgo_solve_report* rep) {
  if (!ctx || !pr || !prm || !poses_out) {
    set_error("rpgo_solve: NULL argument");
    return RPGO_ERR_ARG;
  }
  int rc;
  if ((rc = check_problem("rpgo_solve", pr, prm))) return rc;
  RPGO_HIP_CHECK(hipSetDevice(ctx->device));
  if (!ctx->solver) ctx->solver = new SolverRun();
  return solve_one(ctx->stream, ctx->ev[13], ctx->ev[14], *ctx->solver, pr, prm, poses_out, weights_out,
                   rep);
}

int rpgo_solve_batched(rpgo_ctx* ctx, const rpgo_solve_batch* b, const rpgo_solve_params* prm,
                       int32_t n_lanes, double* poses_out, double* weights_out,
                       rpgo_solve_report* reps, rpgo_solve_batch_report* brep) {
  const auto t0 = std::chrono::steady_clock::now();
  // the batch is checked before the context is touched
  if (!b || !prm || !poses_out) {
    set_error("rpgo_solve_batched: NULL argument");
    return RPGO_ERR_ARG;
  }
  const int G = b->n_problems;
  if (G < 0 || (G > 0 && (!b->pose_off || !b->factor_off || !b->poses))) {
    set_error("rpgo_solve_batched: n_problems %d with missing offsets or poses", G);
    return RPGO_ERR_ARG;
  }
  if (n_lanes == 0) n_lanes = 4;
  if (n_lanes < 1 || n_lanes > RPGO_SOLVE_MAX_LANES) {
    set_error("rpgo_solve_batched: n_lanes %d outside [1, %d]", n_lanes, RPGO_SOLVE_MAX_LANES);
    return RPGO_ERR_ARG;
  }
  // per-problem views, checked up front (nothing runs on a malformed batch)
  std::vector<rpgo_solve_problem> probs(G);
  std::vector<rpgo_solve_params> prms(G, *prm);
  if (G > 0 && (b->pose_off[0] != 0 || b->factor_off[0] != 0)) {
    set_error("rpgo_solve_batched: pose_off[0] and factor_off[0] must be 0");
    return RPGO_ERR_ARG;
  }
  for (int g = 0; g < G; ++g) {
    const int64_t p0 = b->pose_off[g], p1 = b->pose_off[g + 1];
    const int64_t f0 = b->factor_off[g], f1 = b->factor_off[g + 1];
    if (p1 - p0 < 1 || p1 - p0 > INT32_MAX || f1 < f0 || f1 - f0 > INT32_MAX) {
      set_error("rpgo_solve_batched: problem %d has pose range [%lld, %lld) and factor range [%lld, %lld)",
                g, (long long)p0, (long long)p1, (long long)f0, (long long)f1);
      return RPGO_ERR_ARG;
    }
    rpgo_solve_problem& q = probs[g];
    q.n_poses = (int32_t)(p1 - p0);
    q.poses = b->poses + p0 * 12;
    q.n_factors = (int32_t)(f1 - f0);
    q.kind = b->kind ? b->kind + f0 : nullptr;
    q.key_i = b->key_i ? b->key_i + f0 : nullptr;
    q.key_j = b->key_j ? b->key_j + f0 : nullptr;
    q.measured = b->measured ? b->measured + f0 * 12 : nullptr;
    q.information = b->information ? b->information + f0 * 36 : nullptr;
    if (b->known_inliers) prms[g].known_inliers = b->known_inliers[g];
    if (prm->init_weights) prms[g].init_weights = prm->init_weights + f0;
    char who[64];
    snprintf(who, sizeof(who), "rpgo_solve_batched: problem %d", g);
    int rc;
    if ((rc = check_problem(who, &q, &prms[g]))) return rc;
  }
  if (!ctx) {
    set_error("rpgo_solve_batched: NULL context");
    return RPGO_ERR_ARG;
  }
  RPGO_HIP_CHECK(hipSetDevice(ctx->device));
  const int L = std::min(n_lanes, std::max(G, 1));
  if (!ctx->solve_lanes) ctx->solve_lanes = new SolveLanes();
  auto& lanes = ctx->solve_lanes->lanes;
  while ((int)lanes.size() < L) {
    SolveLane* l = new SolveLane();
    hipError_t e = hipStreamCreateWithFlags(&l->st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&l->ev0);
    if (e == hipSuccess) e = hipEventCreate(&l->ev1);
    if (e != hipSuccess) {
      if (l->ev0) (void)hipEventDestroy(l->ev0);
      if (l->st) (void)hipStreamDestroy(l->st);
      delete l;
      set_error("rpgo_solve_batched: stream or event creation failed: %s", hipGetErrorString(e));
      return RPGO_ERR_HIP;
    }
    lanes.push_back(l);
  }
  // largest problems first, so the last ones to start are short
  std::vector<int> order(G);
  for (int g = 0; g < G; ++g) order[g] = g;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    return (int64_t)probs[x].n_poses + probs[x].n_factors > (int64_t)probs[y].n_poses + probs[y].n_factors;
  });
  std::atomic<int> next{0};
  std::atomic<bool> stop{false};
  std::mutex mu;
  int first_rc = RPGO_OK, failed = -1;
  std::string failed_msg;
  std::vector<rpgo_solve_report> rep_local(G);
  auto worker = [&](int li) {
    SolveLane& l = *lanes[li];
    if (hipSetDevice(ctx->device) != hipSuccess) {
      std::lock_guard<std::mutex> lock(mu);
      if (first_rc == RPGO_OK) first_rc = RPGO_ERR_HIP, failed_msg = "hipSetDevice failed in a lane";
      stop = true;
      return;
    }
    for (;;) {
      if (stop.load()) return;
      const int k = next.fetch_add(1);
      if (k >= G) return;
      const int g = order[k];
      std::memset(&rep_local[g], 0, sizeof(rpgo_solve_report));
      const int rc = solve_one(l.st, l.ev0, l.ev1, l.S, &probs[g], &prms[g],
                               poses_out + b->pose_off[g] * 12,
                               weights_out ? weights_out + b->factor_off[g] : nullptr, &rep_local[g]);
      if (rc != RPGO_OK) {
        std::lock_guard<std::mutex> lock(mu);
        if (first_rc == RPGO_OK) {
          first_rc = rc;
          failed = g;
          failed_msg = rpgo_last_error();
        }
        stop = true;
        return;
      }
    }
  };
  if (L == 1 || G <= 1) {
    worker(0);
  } else {
    std::vector<std::thread> th;
    th.reserve(L);
    for (int li = 0; li < L; ++li) th.emplace_back(worker, li);
    for (auto& t : th) t.join();
  }
  if (reps)
    for (int g = 0; g < G; ++g) reps[g] = rep_local[g];
  if (brep) {
    brep->ms_wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    double sum = 0.0;
    for (int g = 0; g < G; ++g) sum += rep_local[g].ms_total;
    brep->ms_device_sum = sum;
    brep->lanes = L;
    brep->failed_problem = failed;
  }
  if (first_rc != RPGO_OK) {
    set_error("rpgo_solve_batched: problem %d: %s", failed, failed_msg.c_str());
    return first_rc;
  }
  return RPGO_OK;
}

}  // extern "C"
