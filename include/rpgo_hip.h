/*
 * rpgo_hip.h -- C ABI of the MI355X-native Kimera-RPGO robust back end.
 *
 * Plain C: pointers, sizes and status codes; no C++ or GPU types cross it.
 * Every entry point returns RPGO_OK (0) or a negative status; the message of
 * the last failure on the calling thread is available from rpgo_last_error().
 * A context owns its device memory and one HIP stream; use one context per
 * host thread.  Host buffers are caller owned.
 *
 * Data layouts (row-major doubles):
 *   SE(3) pose  12 doubles: R (9, row-major) then t (3)     [gtsam::Pose3]
 *   SE(2) pose   4 doubles: cos, sin, x, y                   [gtsam::Pose2]
 *   SE(3) cov   36 doubles, tangent order (omega, v)         [Pose3 tangent]
 *   SE(2) cov    9 doubles, tangent order (x, y, theta)      [Pose2 tangent]
 *   adjacency   bit-packed rows: W = ceil(m/64) uint64 words per row,
 *               bit (i & 63) of word [j*W + i/64] is adj(j, i)
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   rpgo_pcm_*        Pcm<poseT,PoseWithCovariance>::parseAndIncrementAdjMatrix
 *                     include/KimeraRPGO/outlier/Pcm.h:414-502 with updateOdom
 *                     :516-557, isOdomConsistent :604-629, areLoopsConsistent
 *                     :670-718 and incrementAdjMatrix :725-768
 *   rpgo_pcm_simple_* the same stages of Pcm<poseT,PoseWithNode> (PcmSimple2D/3D,
 *                     Pcm.h:1169-1170): checkOdomConsistent :581-596,
 *                     checkLoopConsistent :654-662, PoseWithNode
 *                     include/KimeraRPGO/utils/GeometryUtils.h:193-289
 *   rpgo_max_clique*  findMaxCliqueHeu src/utils/GraphUtils.cpp:19-27
 *                     (FMC::maxCliqueHeu findCliqueHeu.cpp:32-117) and the
 *                     exact FMC::maxClique findClique.cpp:80-145
 */
#ifndef RPGO_HIP_H
#define RPGO_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RPGO_OK 0
#define RPGO_ERR_ARG (-1)     /* invalid argument or size                    */
#define RPGO_ERR_HIP (-2)     /* a HIP runtime call failed                   */
#define RPGO_ERR_NOMEM (-3)   /* device allocation failed                    */
#define RPGO_ERR_STATE (-4)   /* call out of order (e.g. run before upload)  */

/* PCM precision modes */
#define RPGO_PCM_FP64_REPLAY 0  /* every pair in FP64, reference operation order */
#define RPGO_PCM_FP32_RECHECK 1 /* FP32 pairs, FP64 replay of pairs near threshold */

/* max-clique modes */
#define RPGO_CLIQUE_HEU_REF 0   /* bit-exact FMC::maxCliqueHeu output (stale buffer) */
#define RPGO_CLIQUE_EXACT 1     /* exact maximum clique */

typedef struct rpgo_ctx rpgo_ctx;

/* Context on HIP device `device` (flags reserved, pass 0). */
int rpgo_ctx_create(int device, uint32_t flags, rpgo_ctx** out);
int rpgo_ctx_destroy(rpgo_ctx* ctx);
int rpgo_ctx_sync(rpgo_ctx* ctx);
/* Stage timing of the PCM passes (the ms_* fields of rpgo_pcm_stats), on by
 * default.  No counterpart in the reference (its PCM keeps no stage timers).
 * Off: the passes record no HIP events and ms_pairs, ms_recheck and ms_total
 * read 0 (ms_fold and ms_odom of rpgo_pcm_run are still measured); the five
 * event records cost a pairwise pass about 20 us at 1,000 accepted loop
 * closures.  Results are the same either way. */
int rpgo_ctx_set_stage_timing(rpgo_ctx* ctx, int enabled);
const char* rpgo_last_error(void);
/* Library version string ("rpgo-mi355x <semver>"). */
const char* rpgo_version(void);

/* One PCM loop-closure group: odometry chain 0..n_poses-1 plus loop closures.
 * Mirrors what Pcm::removeOutliers extracts from one update() call: the first
 * pose comes from the values (zero covariance, Pcm.h:534-542), odometry k is
 * the BetweenFactor k -> k+1, loop closure t is BetweenFactor(from, to, pose).
 *
 * Several robots (n_chains > 1): the poses are the robots' trajectories one
 * after the other, chain c covering [chain_off[c], chain_off[c+1]) from its
 * own start pose pose0[c]; odometry rows of chain c are [chain_off[c] - c,
 * chain_off[c+1] - c - 1) (n_poses - n_chains rows in all).  The loop
 * closures are one group of the reference (one ObservationId: the same two
 * robots, or one robot twice).  An inter-robot loop closure passes the
 * odometry check unchecked (Pcm.h:460-465; its odom_dist is 0), and for a pair
 * whose first endpoints lie on different chains the newer loop closure's
 * endpoints are swapped without inverting its measurement (Pcm.h:691-698).
 * Such problems run the FP64 replay; rpgo_pcm_append takes one chain only.
 *
 * Landmark observations (landmark != 0): the loop closures are the
 * observations of one landmark l, BetweenFactor(pose lc_from[k], l, lc_pose[k])
 * from one robot's trajectory (n_chains <= 1); lc_to is not read.  No odometry
 * check (every observation is accepted, odom_dist 0) and the pair check is
 * incrementLandmarkAdjMatrix's (Pcm.h:775-844): loop = (between(i, j) o
 * Z_jl)^-1 o Z_il, with the PCM or PcmSimple distance of the run.  FP64 replay;
 * not streamed. */
typedef struct {
  int32_t dim;            /* 3 = SE(3) (Pcm3D), 2 = SE(2) (Pcm2D)        */
  int32_t n_poses;        /* >= 1                                          */
  const double* pose0;    /* pose_len doubles                             */
  const double* odom;     /* (n_poses-1) * pose_len                        */
  const double* odom_cov; /* (n_poses-1) * d*d ; NaN rotation trace drops
                             the rotation block (GeometryUtils.h:102-113) */
  int32_t n_lc;           /* >= 0                                          */
  const int32_t* lc_from; /* n_lc, in [0, n_poses)                         */
  const int32_t* lc_to;   /* n_lc, in [0, n_poses)                         */
  const double* lc_pose;  /* n_lc * pose_len                               */
  const double* lc_cov;   /* n_lc * d*d                                    */
  int32_t n_chains;       /* 0 or 1: one chain; > 1: several robots         */
  const int32_t* chain_off; /* n_chains + 1 (n_chains > 1 only)            */
  int32_t landmark;       /* nonzero: observations of one landmark         */
} rpgo_pcm_problem;

typedef struct {
  double odom_threshold;  /* PcmParams::odom_threshold; < 0 disables (Pcm.h:74-77)  */
  double lc_threshold;    /* PcmParams::lc_threshold;   < 0 disables (Pcm.h:79-82)  */
  int32_t precision;      /* RPGO_PCM_FP64_REPLAY or RPGO_PCM_FP32_RECHECK           */
  int32_t want_dist;      /* keep the m*m distance matrix for rpgo_pcm_download      */
  double recheck_margin;  /* FP32 mode: relative band around the squared threshold that
                             is replayed in FP64 (<= 0 selects the default 2e-3)   */
  double recheck_chain_coef; /* FP32 mode: band term for the reference's own FP64
                             cancellation, coef * tr(chain covariances) / min pivot
                             (<= 0 selects the default 16 * 2^-52)                 */
} rpgo_pcm_params;

typedef struct {
  int32_t n_accepted;     /* loop closures that passed the odometry check         */
  int32_t reserved;
  int64_t n_pairs;        /* unique (i<j) pairs evaluated                          */
  int64_t n_rechecked;    /* FP32 mode: pairs replayed in FP64                      */
  int64_t n_flipped;      /* FP32 mode: FP32 decisions the FP64 replay overturned   */
  int64_t n_consistent;   /* pairs with d < lc_threshold                            */
  double ms_fold;         /* device time of each stage (HIP events, ctx stream)     */
  double ms_odom;
  double ms_pairs;
  double ms_recheck;
  double ms_total;
} rpgo_pcm_stats;

/* Copy a problem to device memory (kept in the context until the next upload). */
int rpgo_pcm_upload(rpgo_ctx* ctx, const rpgo_pcm_problem* prob);
/* Run the adjacency stage on the uploaded problem, device-resident; blocks
 * until done and fills *stats (may be NULL). */
int rpgo_pcm_run(rpgo_ctx* ctx, const rpgo_pcm_params* params, rpgo_pcm_stats* stats);
/* Recompute only the pairwise stage -- every pair check of the accepted loop
 * closures and the adjacency fill (Pcm.h:725-768 / areLoopsConsistent
 * :670-718) -- on the device-resident trajectory of the last rpgo_pcm_run,
 * e.g. with another threshold or precision.  Blocks; fills *stats. */
int rpgo_pcm_pairs(rpgo_ctx* ctx, const rpgo_pcm_params* params, rpgo_pcm_stats* stats);
/* Fetch results of the last run.  accepted: n_lc ints (indices of accepted loop
 * closures, in input order; n_accepted valid).  odom_dist: n_lc doubles or NULL.
 * adj_bits: m*W words or NULL.  dist: m*m doubles or NULL (needs want_dist). */
int rpgo_pcm_download(rpgo_ctx* ctx, int32_t* accepted, double* odom_dist,
                      uint64_t* adj_bits, double* dist);
/* upload + run + download in one call. */
int rpgo_pcm_adjacency(rpgo_ctx* ctx, const rpgo_pcm_problem* prob,
                       const rpgo_pcm_params* params, int32_t* accepted,
                       double* odom_dist, uint64_t* adj_bits, double* dist,
                       rpgo_pcm_stats* stats);
/* The FP64 trajectory fold of the last run: n_poses * (pose_len + d*d + 1)
 * doubles per pose (pose, covariance, rotation_info flag). */
int rpgo_pcm_download_traj(rpgo_ctx* ctx, double* traj);

/* PcmSimple: the same stages with the PoseWithNode distances
 * (Pcm<poseT, PoseWithNode>, GeometryUtils.h:193-289, Pcm.h:581-596,654-662).
 * The odometry check passes a loop closure (i, j) when the per-node
 * translation and rotation norms of getBetween(i,j) * LC^-1 (node |j-i|+1)
 * are below odom_*; a pair passes when those of the loop
 * (a->c * LC_cd)^-1 * LC_ab * (b->d) (node |c-a|+|d-b|+2) are below dist_*.
 * The norms are |tail(translation size)| and |head(rotation dim)| of Logmap
 * divided by the node count (for Pose2: |(v_y, theta)| and |v_x|); the
 * rotation norm is 0 without rotation information (NaN rotation variance).
 * A negative threshold in a pair disables that check (Pcm.h:74-82: every
 * loop closure passes / no adjacency; the caller keeps every accepted loop
 * closure as an inlier).  Covariances only feed the rotation_info flags.
 * Results are read with rpgo_pcm_download (odom_dist and dist hold the
 * translation norms) and rpgo_max_clique. */
typedef struct {
  double odom_trans_threshold; /* PcmParams::odom_trans_threshold */
  double odom_rot_threshold;   /* PcmParams::odom_rot_threshold   */
  double dist_trans_threshold; /* PcmParams::dist_trans_threshold */
  double dist_rot_threshold;   /* PcmParams::dist_rot_threshold   */
  int32_t want_dist;           /* keep the m*m translation-norm matrix */
  int32_t reserved;
} rpgo_pcm_simple_params;

int rpgo_pcm_simple_run(rpgo_ctx* ctx, const rpgo_pcm_simple_params* params,
                        rpgo_pcm_stats* stats);
/* upload + rpgo_pcm_simple_run + download in one call. */
int rpgo_pcm_simple_adjacency(rpgo_ctx* ctx, const rpgo_pcm_problem* prob,
                              const rpgo_pcm_simple_params* params, int32_t* accepted,
                              double* odom_dist, uint64_t* adj_bits, double* dist,
                              rpgo_pcm_stats* stats);

/* Streaming: one reference update() after another (Pcm::removeOutliers
 * Pcm.h:148-281 called once per increment, parseAndIncrementAdjMatrix
 * :414-502, incrementAdjMatrix :725-768).  Starts from the state of a full
 * rpgo_pcm_run / rpgo_pcm_simple_run (which may hold a single pose and no
 * loop closure).  Each call appends n_odom odometry factors at the chain end
 * (poses n_poses .. n_poses+n_odom-1, folded from the stored last record),
 * odometry-checks the new loop closures against the extended chain and adds
 * the accepted ones, in input order, as new adjacency rows checked against
 * every earlier accepted loop closure: O(m) pairs per new loop closure
 * instead of a full O(m^2) pass.  The result equals a full run over the
 * concatenated problem bit for bit.  Pairs are evaluated in FP64
 * (params->precision is ignored); want_dist must be 0.  Loop-closure indices
 * in rpgo_pcm_download's accepted list count every loop closure of the
 * stream in arrival order.  Stats: n_accepted = total accepted; n_pairs and
 * n_consistent count this call's pairs.  Follow with rpgo_max_clique, or
 * rpgo_max_clique_incremental (num_new = accepted loop closures of this call)
 * for PcmParams::incremental. */
typedef struct {
  int32_t n_odom;          /* odometry factors appended at the chain end          */
  const double* odom;      /* n_odom * pose_len                                   */
  const double* odom_cov;  /* n_odom * d*d                                        */
  int32_t n_lc;            /* new loop closures, endpoints in the extended chain  */
  const int32_t* lc_from;
  const int32_t* lc_to;
  const double* lc_pose;   /* n_lc * pose_len */
  const double* lc_cov;    /* n_lc * d*d      */
} rpgo_pcm_increment;

int rpgo_pcm_append(rpgo_ctx* ctx, const rpgo_pcm_increment* inc, const rpgo_pcm_params* params,
                    rpgo_pcm_stats* stats);
int rpgo_pcm_simple_append(rpgo_ctx* ctx, const rpgo_pcm_increment* inc,
                           const rpgo_pcm_simple_params* params, rpgo_pcm_stats* stats);

/* Max clique over the adjacency of the last rpgo_pcm_run (device resident).
 * idx_out: room for m ints; *k_out: clique size (FMC returns -1 for m == 0). */
int rpgo_max_clique(rpgo_ctx* ctx, int32_t mode, int32_t* idx_out, int32_t* k_out,
                    double* ms_out);
/* Max clique of a caller-provided bit-packed symmetric adjacency (n vertices). */
int rpgo_max_clique_bits(rpgo_ctx* ctx, int32_t n, const uint64_t* adj_bits,
                         int32_t mode, int32_t* idx_out, int32_t* k_out);
/* Allocation-free variant of rpgo_max_clique_bits(ctx, n, adj, RPGO_CLIQUE_HEU_REF,
 * idx_out, k_out).  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (serving loops, harnesses that forbid
 * hipMalloc, code that must not grow the device heap).  The adjacency is read
 * from caller-owned DEVICE memory (d_adj_bits: n rows of ceil(n/64) words,
 * same layout as above); idx_out (host, room for n ints, may be NULL) and
 * *k_out receive what rpgo_max_clique_bits returns.  The scan's working
 * memory (degrees, degree filter, candidate sizes, the replayed buffer) is a
 * fixed library-owned device array of about 400 KB per GPU, held until the
 * process ends; no device memory is allocated or freed by the call.  n above
 * RPGO_NOALLOC_MAX_VERTICES is rejected with RPGO_ERR_ARG.  Every call on one
 * GPU shares that GPU's array: calls that overlap on one GPU (several
 * contexts or threads) race and must be serialised by the caller. */
#define RPGO_NOALLOC_MAX_VERTICES 32768
int rpgo_max_clique_bits_noalloc(rpgo_ctx* ctx, int32_t n, const uint64_t* d_adj_bits,
                                 int32_t* idx_out, int32_t* k_out);
/* Incremental heuristic over the last rpgo_pcm_run adjacency, replacing
 * findMaxCliqueHeuIncremental (src/utils/GraphUtils.cpp:30-44, FMC::
 * maxCliqueHeuIncremental findCliqueHeu.cpp:119-209): only the last num_new
 * vertices are candidates and the bound starts at prev_size; *k_out is the new
 * size when it strictly exceeds prev_size, else 0 ("keep the previous set"). */
int rpgo_max_clique_incremental(rpgo_ctx* ctx, int32_t num_new, int32_t prev_size,
                                int32_t* idx_out, int32_t* k_out);
int rpgo_max_clique_incremental_bits(rpgo_ctx* ctx, int32_t n, const uint64_t* adj_bits,
                                     int32_t num_new, int32_t prev_size, int32_t* idx_out,
                                     int32_t* k_out);

/* Batched PCM: n_graphs independent single-prefix graphs of uneven size,
 * packed by prefix offsets, one device launch per stage (the C4 batch).
 * Graph g: poses [pose_off[g], pose_off[g+1]) (pose0 row g), odometry rows
 * [pose_off[g] - g, pose_off[g+1] - g - 1), loop closures [lc_off[g],
 * lc_off[g+1]) with pose indices local to the graph.  The fold, odometry
 * check and compaction are the FP64 replay of the single-graph path; the pairs
 * are too with params->precision RPGO_PCM_FP64_REPLAY, and with
 * RPGO_PCM_FP32_RECHECK an SE(3) batch runs the single-graph FP32 stage per
 * graph with the FP64 replay of the pairs near the threshold.  Either way the
 * adjacency of each graph equals rpgo_pcm_run's on that graph alone, bit for
 * bit (all zero when lc_threshold < 0). */
typedef struct {
  int32_t dim;              /* 3 or 2                                         */
  int32_t n_graphs;
  const int64_t* pose_off;  /* n_graphs + 1, pose_off[0] = 0                  */
  const double* pose0;      /* n_graphs * pose_len                            */
  const double* odom;       /* (pose_off[n] - n) * pose_len                   */
  const double* odom_cov;   /* (pose_off[n] - n) * d * d                      */
  const int64_t* lc_off;    /* n_graphs + 1, lc_off[0] = 0                    */
  const int32_t* lc_from;   /* lc_off[n], local indices                       */
  const int32_t* lc_to;
  const double* lc_pose;
  const double* lc_cov;
  int32_t flags;            /* RPGO_BATCH_* bits (0: stages launched one by one,
                               per-stage times in the stats)                  */
} rpgo_pcm_batch;

/* rpgo_pcm_batch.flags */
#define RPGO_BATCH_GRAPH 1    /* capture every stage (measurements .. clique) in one
                                 hipGraph, replayed while the batch layout (sizes,
                                 launch shapes) is unchanged; stats give ms_total only */
#define RPGO_BATCH_AUTOTUNE 2 /* before the first run of a dim on this context, time
                                 the candidate launch shapes of the pair and clique
                                 stages on this batch and keep the fastest          */

typedef struct {
  int64_t n_pairs;          /* sum over graphs of m (m - 1) / 2               */
  int64_t n_consistent;     /* consistent pairs (lower triangle)              */
  double ms_fold;           /* measurements + fold (all graphs)               */
  double ms_odom;           /* odometry check + compaction                    */
  double ms_pairs;          /* pair checks + symmetrize                       */
  double ms_clique;         /* heuristic clique of every graph                */
  double ms_total;          /* device time of the whole batch                 */
  double ms_tune;           /* time spent by RPGO_BATCH_AUTOTUNE in this call  */
  int32_t pair_waves;       /* wavefronts per workgroup of the pair stage      */
  int32_t clique_waves;     /* wavefronts per graph of the clique stage        */
  int32_t graph_replayed;   /* 1 when an existing hipGraph was replayed        */
  int32_t clique_prefetch;  /* rows loaded together in the clique walk         */
  int32_t fp32;             /* 1 when the FP32 pair stage ran (precision FP32_RECHECK,
                               dim 3, no covariance without rotation information,
                               at most 2^29 pairs in the batch)                 */
  int64_t n_rechecked;      /* FP32 stage: pairs replayed in FP64              */
  int64_t n_flipped;        /* FP32 stage: FP32 decisions the replay overturned */
} rpgo_pcm_batch_stats;

/* Run the batch.  n_accepted: n_graphs counts; accepted: lc_off[n] entries,
 * graph g's accepted local loop-closure indices at [lc_off[g], +n_accepted[g]);
 * clique_size / clique_idx (may be NULL): FMC::maxCliqueHeu per graph, sizes
 * per graph and members (indices into that graph's accepted set) at lc_off[g].
 * The clique stage runs on the device for every graph in one launch (one
 * workgroup per graph); graphs with more than 32768 accepted loop closures are
 * rejected.  Blocks until the results are on the host. */
int rpgo_pcm_batch_run(rpgo_ctx* ctx, const rpgo_pcm_batch* batch, const rpgo_pcm_params* params,
                       int32_t* n_accepted, int32_t* accepted, int32_t* clique_size,
                       int32_t* clique_idx, rpgo_pcm_batch_stats* stats);
/* Adjacency of graph g of the last batch: n_accepted[g] rows of
 * ceil(n_accepted[g] / 64) words. */
int rpgo_pcm_batch_download_adj(rpgo_ctx* ctx, int32_t g, uint64_t* adj_bits);

/* ------------------------------------------------------------------------
 * Robust solve: GNC-TLS over Levenberg-Marquardt / Gauss-Newton with a PCG
 * linear step (block-tridiagonal segment preconditioner along the odometry
 * chain plus a rigid-segment coarse level).  Replaces RobustSolver::optimize
 * (src/RobustSolver.cpp:115-306): gtsam::GncOptimizer with known inliers
 * (:105-113, :130-180, :225-275) or plain LevenbergMarquardtOptimizer /
 * GaussNewtonOptimizer (:181-218, :276-296).  SE(3) poses only.
 * ---------------------------------------------------------------------- */
#define RPGO_SOLVER_LM 0         /* Solver::LM (SolverParams.h:13) */
#define RPGO_SOLVER_GN 1         /* Solver::GN */
#define RPGO_FACTOR_BETWEEN 0    /* BetweenFactor<Pose3>(i, j, z) */
#define RPGO_FACTOR_PRIOR 1      /* PriorFactor<Pose3>(i, prior)  */
#define RPGO_GNC_COST 0          /* GncParams::GncThresholdMode::COST */
#define RPGO_GNC_PROBABILITY 1   /* GncParams::GncThresholdMode::PROBABILITY */

typedef struct {
  int32_t n_poses;
  const double* poses;        /* n_poses * 12 initial values                  */
  int32_t n_factors;
  const int32_t* kind;        /* RPGO_FACTOR_BETWEEN | RPGO_FACTOR_PRIOR       */
  const int32_t* key_i;       /* pose index of the (first) key                 */
  const int32_t* key_j;       /* second pose of a between factor (else ignored) */
  const double* measured;     /* n_factors * 12: z of a between, prior pose     */
  const double* information;  /* n_factors * 36, (omega, v) tangent order       */
} rpgo_solve_problem;

typedef struct {
  int32_t solver;               /* RPGO_SOLVER_LM | RPGO_SOLVER_GN                 */
  int32_t use_gnc;              /* RobustSolverParams::use_gnc_                    */
  int32_t known_inliers;        /* the first k factors are known inliers (odometry
                                   then special factors, RobustSolver.cpp:105-113) */
  int32_t gnc_threshold_mode;   /* RPGO_GNC_COST | RPGO_GNC_PROBABILITY            */
  double gnc_inlier_threshold;  /* cost, or probability alpha                      */
  int32_t gnc_max_iterations;   /* GncParams::max_iterations_ (100)               */
  double gnc_mu_step;           /* 1.4                                            */
  double gnc_relative_cost_tol; /* 1e-5                                           */
  double gnc_weights_tol;       /* 1e-4                                           */
  int32_t gnc_bias_odom;        /* initial weights 1 on known inliers, 0 elsewhere */
  int32_t lm_diagonal_damping;  /* RobustSolverParams::lm_diagonal_damping (1)     */
  const double* init_weights;   /* NULL, or n_factors initial GNC weights          */
  double init_mu;               /* <= 0: GNC's own mu initialisation               */
  double lambda_initial;        /* LM defaults: 1e-5, 10, 1e5, 1e-3                */
  double lambda_factor;
  double lambda_upper_bound;
  double min_model_fidelity;
  int32_t lm_max_iterations;    /* 100 */
  double lm_relative_error_tol; /* 1e-5 */
  double lm_absolute_error_tol; /* 1e-5 */
  int32_t pcg_max_iterations;   /* <= 0: 6 * n_poses */
  double pcg_rel_tol;           /* relative residual of each linear solve */
  int32_t polish_iterations;    /* final Gauss-Newton refinement steps (0: none) */
  int32_t pcg_segment;          /* poses per block-tridiagonal preconditioner segment
                                   (default 0: automatic, 16 below 10,000 poses and
                                   32 from there; 1: plain block-Jacobi)          */
  int32_t gnc_warm_start;       /* 0: every inner solve restarts from the initial
                                   values (GncOptimizer); 1: from the last result  */
  int32_t pcg_coarse;           /* rigid-segment coarse level of the preconditioner:
                                   < 0 off; 0 auto (up to 128 segments below 10,000
                                   poses, 256 from there); > 0 at most that many
                                   segments (<= 256)                              */
  int32_t pcg_graphs;           /* 1: replay PCG iterations as captured HIP graphs;
                                   0: direct launches of the same kernels          */
  int32_t gnc_factor_dim;       /* degrees of freedom of the chi^2 threshold in
                                   PROBABILITY mode: 6 (Pose3), 3 for Pose2 graphs
                                   embedded in the plane (<= 0 selects 6)          */
  int32_t pose2_chart;          /* 0: Pose3 factors (Logmap errors, Expmap
                                   retraction: the reference's EXPMAP build).
                                   1: Pose2 graphs embedded in the plane (theta ->
                                   omega_z, x -> v_x, y -> v_y, R = Rz(theta)) with
                                   GTSAM's default Pose2 chart: errors are the
                                   (x, y, theta) of z^-1 T_i^-1 T_j (of x^-1 prior,
                                   negated, for a prior) and values move by
                                   x * Pose2(dx) (Pose2::ChartAtOrigin)          */
} rpgo_solve_params;

typedef struct {
  int32_t gnc_iterations;
  int32_t lm_iterations;
  int32_t pcg_iterations;
  int32_t linearizations;
  double mu_init;
  double mu_final;
  double final_cost;            /* weighted cost of the returned poses */
  double ms_total;              /* device-stream time of the solve     */
  int32_t n_inliers;            /* truncated weight sum minus known inliers
                                   (RobustSolver.cpp:170-171)           */
  int32_t n_tls_disagree;       /* GNC: factors whose final weight differs from
                                   the TLS decision at the final poses (0 at a
                                   GNC fixed point; the check a learned start
                                   must pass)                           */
} rpgo_solve_report;

/* 1/2 e^T Lambda e of every factor of the last rpgo_solve problem at poses
 * (n_poses * 12 doubles), or at the returned solution when poses is NULL. */
int rpgo_solve_errors(rpgo_ctx* ctx, const double* poses, double* err_out);
/* GNC's TLS threshold in PROBABILITY mode: 0.5 * chi^2 quantile(alpha, dof)
 * (GncOptimizer::setInlierCostThresholdsAtProbability). */
double rpgo_gnc_threshold(double alpha, int32_t dof);
/* Fill *p with the reference defaults (SolverParams.h:60-94, GTSAM LM). */
void rpgo_solve_params_default(rpgo_solve_params* p);
/* Solve; poses_out: n_poses * 12, weights_out: n_factors (may be NULL). */
int rpgo_solve(rpgo_ctx* ctx, const rpgo_solve_problem* prob, const rpgo_solve_params* params,
               double* poses_out, double* weights_out, rpgo_solve_report* report);

/* Batched robust solve: n_problems independent problems of uneven size (the
 * C4 batch after PCM), packed by prefix offsets.  Problem g: poses
 * [pose_off[g], pose_off[g+1]) (rows of 12), factors [factor_off[g],
 * factor_off[g+1]) with pose keys local to the problem.  Each problem is
 * solved exactly as rpgo_solve solves it alone (RobustSolver::optimize per
 * graph, src/RobustSolver.cpp:115-306); up to n_lanes problems run at once,
 * each on its own stream and solver state of this context, largest first. */
typedef struct {
  int32_t n_problems;
  const int64_t* pose_off;      /* n_problems + 1, pose_off[0] = 0              */
  const int64_t* factor_off;    /* n_problems + 1, factor_off[0] = 0            */
  const double* poses;          /* pose_off[n] * 12 initial values              */
  const int32_t* kind;          /* factor_off[n] entries, as rpgo_solve_problem */
  const int32_t* key_i;         /* local to the problem                         */
  const int32_t* key_j;
  const double* measured;       /* factor_off[n] * 12                           */
  const double* information;    /* factor_off[n] * 36                           */
  const int32_t* known_inliers; /* n_problems counts, or NULL: params' value    */
} rpgo_solve_batch;

typedef struct {
  double ms_wall;               /* host time of the whole call                  */
  double ms_device_sum;         /* sum of the problems' ms_total                */
  int32_t lanes;                /* problems in flight at once                   */
  int32_t failed_problem;       /* -1, or the problem whose solve failed        */
} rpgo_solve_batch_report;

#define RPGO_SOLVE_MAX_LANES 16

/* params apply to every problem (params->init_weights, when not NULL, holds
 * factor_off[n] packed weights).  poses_out: pose_off[n] * 12; weights_out:
 * factor_off[n] (may be NULL); reports: n_problems (may be NULL).  n_lanes in
 * [1, RPGO_SOLVE_MAX_LANES], 0 selects 4.  The first failing problem stops
 * the batch; its status is returned and rpgo_last_error() names it. */
int rpgo_solve_batched(rpgo_ctx* ctx, const rpgo_solve_batch* batch, const rpgo_solve_params* params,
                       int32_t n_lanes, double* poses_out, double* weights_out,
                       rpgo_solve_report* reports, rpgo_solve_batch_report* batch_report);

/* ------------------------------------------------------------------------
 * Learned GNC initialisation.  No counterpart in the reference: an addition
 * of this back end that predicts, per loop closure, the GNC weight the plain
 * method converges to (as a logit) and log(mu) of the final GNC iteration,
 * from features of the PCM state.  Plain GNC stays the arbiter: callers start
 * rpgo_solve from the prediction (init_weights, init_mu) and fall back to the
 * plain solve when the result fails their check.
 * MLP 16 -> 128 -> 128 -> 2 with ReLU; BF16 matrix-core products with FP32
 * accumulation; Adam on FP32 master weights.  Parameter layout (FP32, this
 * order): W1[16][128], b1[128], W2[128][128], b2[128], W3[128][2], b3[2].
 * ---------------------------------------------------------------------- */
#define RPGO_LC_FEATURES 16
#define RPGO_MLP_HIDDEN 128
#define RPGO_MLP_OUT 2
#define RPGO_MLP_NPARAMS (16 * 128 + 128 + 128 * 128 + 128 + 128 * 2 + 2)

/* Features of the accepted loop closures of the last rpgo_pcm_run (SE(3)),
 * in the compacted (accepted) order: feat m * RPGO_LC_FEATURES floats.
 * clique_idx: the k clique members (indices into the accepted set);
 * lc_resid: NULL, or m residuals 1/2 e^T Lambda e of the loop closures after
 * GNC's first solve (all weights 1), with barc the TLS threshold. */
int rpgo_lc_features(rpgo_ctx* ctx, const int32_t* clique_idx, int32_t k, const double* lc_resid,
                     double barc, float* feat);
/* Load parameters and the feature normalisation (mean[16] then 1/std[16]);
 * resets the Adam state. */
int rpgo_mlp_set_params(rpgo_ctx* ctx, const float* params, const float* norm);
int rpgo_mlp_get_params(rpgo_ctx* ctx, float* params);
/* y[n][2] = MLP(x[n][16]). */
int rpgo_mlp_forward(rpgo_ctx* ctx, int32_t n, const float* x, float* y);
/* Gradient of the mean loss over the n rows (BCE on the weight logit against
 * target[:,0], mu_weight/2 (y1 - target[:,1])^2 on log mu, each scaled by
 * sample_w, which may be NULL) with respect to the parameters. */
int rpgo_mlp_gradients(rpgo_ctx* ctx, int32_t n, const float* x, const float* target,
                       const float* sample_w, float mu_weight, float* grad, double* loss);
/* Adam training on the device-resident set: `epochs` passes over shuffled
 * full batches of `batch` rows; loss_hist[e] = mean batch loss of epoch e. */
int rpgo_mlp_train(rpgo_ctx* ctx, int32_t n, const float* x, const float* target,
                   const float* sample_w, float mu_weight, int32_t batch, int32_t epochs,
                   float lr, uint64_t seed, double* loss_hist);

#ifdef __cplusplus
}
#endif
#endif /* RPGO_HIP_H */
