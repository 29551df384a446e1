// Internal state of the robust pose-graph solve (solver.hip).
#pragma once

#include <cstring>

#include "rpgo_internal.h"

namespace rpgo {

constexpr int kPcgChunk = 8;       // PCG iterations per captured graph launch
constexpr int kCoarseSegMax = 256;  // rigid segments of the coarse preconditioner level (maximum)
constexpr int kCoarseSegSmall = 128;  // default below kCoarseLargeN poses
constexpr int kCoarseLargeN = 10000;  // poses from which the default is kCoarseSegMax
constexpr int kCoarseMax = 6 * kCoarseSegMax;

struct GraphKey {
  int N = 0, S = 0, Sc = 0, nseg_c = 0;
  const void *hist = nullptr, *p = nullptr, *Minv = nullptr, *Hij = nullptr, *inc = nullptr,
             *inc_q = nullptr, *Ainv = nullptr, *Pb = nullptr, *Ds = nullptr, *ext = nullptr;
  bool operator==(const GraphKey& o) const {
    return N == o.N && S == o.S && Sc == o.Sc && nseg_c == o.nseg_c && hist == o.hist && p == o.p &&
           Minv == o.Minv && Hij == o.Hij && inc == o.inc && inc_q == o.inc_q && Ainv == o.Ainv &&
           Pb == o.Pb && Ds == o.Ds && ext == o.ext;
  }
};

struct SolverRun {
  struct Stats {
    int lm_iterations = 0;
    int lm_rejections = 0;
    int pcg_iterations = 0;
    int linearizations = 0;
    double mu_init = -1.0;
  };
  int N = 0, F = 0, n_known = 0;
  rpgo_solve_params prm{};
  Stats stats;
  // problem (device)
  DevBuf x0, x, xn;             // FP64 poses
  DevBuf kind, fi, fj, z, info;  // factors
  DevBuf w, barc, known, err, cost;
  // Gauss-Newton system
  DevBuf Hii, Hjj, Hij, gi, gj;  // per-factor contributions
  DevBuf inc_off, inc_cnt, inc;  // pose -> incident factors
  DevBuf inc_q;                  // per incidence: the factor's other pose (-1: not a between factor)
  DevBuf D, Minv, hdiag, b64;
  DevBuf chain, Useg;             // odometry chain links, preconditioner off-diagonals
  DevBuf Ds, Cs;                  // entry-major diagonal and chain-link blocks (chain SpMV)
  DevBuf ext_off, ext, ext_q;     // per pose: between incidences that are not chain links
  DevBuf dx, r, zv, p, q, hx;     // PCG vectors (FP64)
  DevBuf sc;                      // FP64 scalars
  DevBuf hist;                    // PCG residual history (per iteration)
  DevBuf Pb, Ac, Lci, Ainv, cres, ccor;  // coarse level: prolongation blocks, coarse matrix,
                                     // its inverse factor, FP32 inverse, coarse vectors
  DevBuf LcT;                        // the coarse Cholesky factor transposed (column access)
  bool polish = false;               // refinement pass: a step that raises the cost is rejected
  hipGraphExec_t graph_exec = nullptr;  // captured PCG chunk
  GraphKey graph_key;
  bool use_graphs = true;
  // pinned host buffers of this run's readbacks (scalars; one PCG chunk's
  // residual history), allocated on first use
  double* h_scalars = nullptr;
  double* h_history = nullptr;

  int alloc(int N_, int F_);
  void spmv_dev(hipStream_t st, const double* lam, const double* x, double* y);
  int build_incidence(hipStream_t st);
  int linearize(hipStream_t st);
  void spmv(hipStream_t st, double lambda, const double* x, double* y);
  int pcg(hipStream_t st, double lambda, int* iters);
  int coarse_setup(hipStream_t st, double lambda, int* Sc, int* nseg);
  void pcg_step(hipStream_t st, int S, int Sc, int nseg_c, int first, int a);
  int model_decrease(hipStream_t st, double* out);
  int total_cost(hipStream_t st, const double* xp, double* out);
  int compute_errors(hipStream_t st, const double* xp);
  int nonlinear_solve(hipStream_t st, bool from_x0, double* cost, int* iters);
  int read_scalars(hipStream_t st, int count, double* out);
  double* pinned_scalars();
  double* pinned_history();
  int run(hipStream_t st, rpgo_solve_report* rep);
  double* cost_buf() { return cost.as<double>(); }
  void release() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    graph_exec = nullptr;
    if (h_scalars) (void)hipHostFree(h_scalars);
    if (h_history) (void)hipHostFree(h_history);
    h_scalars = h_history = nullptr;
    for (DevBuf* v : {&hist,&x0, &x, &xn, &kind, &fi, &fj, &z, &info, &w, &barc, &known, &err, &cost, &Hii,
                      &Hjj, &Hij, &gi, &gj, &inc_off, &inc_cnt, &inc, &inc_q, &D, &Minv, &hdiag, &b64,
                      &chain, &Useg, &Ds, &Cs, &ext_off, &ext, &ext_q, &Pb, &Ac, &Lci, &LcT, &Ainv, &cres,
                      &ccor,
                      &dx, &r, &zv, &p, &q, &hx, &sc})
      v->release();
  }
};

// chi^2 quantile (inverse regularized lower incomplete gamma), host side
double chi2_quantile(double p, double dof);

}  // namespace rpgo
