// C ABI of the robust solve (rpgo_solve) and its host-side set-up.
#include <chrono>
#include <cmath>
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

void solver_destroy(rpgo_ctx* ctx) {
  if (ctx->solver) {
    ctx->solver->release();
    delete ctx->solver;
    ctx->solver = nullptr;
  }
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
  const int N = pr->n_poses, F = pr->n_factors;
  if (N < 1 || F < 0 || !pr->poses || (F > 0 && (!pr->kind || !pr->key_i || !pr->key_j ||
                                                  !pr->measured || !pr->information))) {
    set_error("rpgo_solve: inconsistent problem (n_poses=%d n_factors=%d)", N, F);
    return RPGO_ERR_ARG;
  }
  if (prm->solver != RPGO_SOLVER_LM && prm->solver != RPGO_SOLVER_GN) {
    set_error("rpgo_solve: unknown solver %d", prm->solver);
    return RPGO_ERR_ARG;
  }
  if (!(prm->lambda_factor > 1.0) || !(prm->lambda_upper_bound > 0.0) || prm->lambda_initial < 0.0 ||
      prm->lm_max_iterations < 0 || prm->gnc_max_iterations < 0 || !(prm->gnc_mu_step > 0.0)) {
    set_error("rpgo_solve: invalid LM/GNC parameters (lambda_factor must exceed 1)");
    return RPGO_ERR_ARG;
  }
  if (prm->pcg_coarse > kCoarseSegMax) {
    set_error("rpgo_solve: pcg_coarse %d above the maximum of %d segments", prm->pcg_coarse,
              kCoarseSegMax);
    return RPGO_ERR_ARG;
  }
  if (prm->known_inliers < 0 || prm->known_inliers > F) {
    set_error("rpgo_solve: known_inliers %d outside [0, %d]", prm->known_inliers, F);
    return RPGO_ERR_ARG;
  }
  for (int f = 0; f < F; ++f) {
    const int k = pr->kind[f];
    if ((k != RPGO_FACTOR_BETWEEN && k != RPGO_FACTOR_PRIOR) || pr->key_i[f] < 0 ||
        pr->key_i[f] >= N ||
        (k == RPGO_FACTOR_BETWEEN && (pr->key_j[f] < 0 || pr->key_j[f] >= N || pr->key_j[f] == pr->key_i[f]))) {
      set_error("rpgo_solve: factor %d has kind %d keys (%d, %d) outside [0, %d)", f, k,
                pr->key_i[f], pr->key_j[f], N);
      return RPGO_ERR_ARG;
    }
  }
  RPGO_HIP_CHECK(hipSetDevice(ctx->device));
  if (!ctx->solver) ctx->solver = new SolverRun();
  SolverRun& S = *ctx->solver;
  int rc;
  if ((rc = S.alloc(N, F > 0 ? F : 1))) return rc;
  S.F = F;
  S.prm = *prm;
  S.n_known = prm->known_inliers;
  hipStream_t st = ctx->stream;
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
  RPGO_HIP_CHECK(hipEventRecord(ctx->ev[13], st));
  if (F > 0) {
    if ((rc = S.run(st, rep))) return rc;
  } else {
    RPGO_HIP_CHECK(hipMemcpyAsync(S.x.ptr, S.x0.ptr, sizeof(double) * 12 * N, hipMemcpyDeviceToDevice, st));
  }
  RPGO_HIP_CHECK(hipEventRecord(ctx->ev[14], st));
  RPGO_HIP_CHECK(hipEventSynchronize(ctx->ev[14]));
  RPGO_HIP_CHECK(hipMemcpy(poses_out, S.x.ptr, sizeof(double) * 12 * N, hipMemcpyDeviceToHost));
  std::vector<double> wout(F > 0 ? F : 1);
  if (F > 0) RPGO_HIP_CHECK(hipMemcpy(wout.data(), S.w.ptr, sizeof(double) * F, hipMemcpyDeviceToHost));
  if (weights_out && F > 0) std::memcpy(weights_out, wout.data(), sizeof(double) * F);
  if (rep) {
    if (F == 0) std::memset(rep, 0, sizeof(*rep));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev[13], ctx->ev[14]);
    rep->ms_total = ms;
    double sum = 0.0;
    for (int f = 0; f < F; ++f) sum += wout[f];
    rep->n_inliers = (int32_t)((size_t)sum - (size_t)prm->known_inliers);
  }
  return RPGO_OK;
}

}  // extern "C"
