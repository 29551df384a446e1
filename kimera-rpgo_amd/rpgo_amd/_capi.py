"""ctypes binding of librpgo_hip.so (the C ABI in include/rpgo_hip.h).

The HIP library is the only compute path: loading fails loudly when the
built library is missing, and nothing here computes on the CPU.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "lib", "librpgo_hip.so")

RPGO_OK = 0
RPGO_ERR_ARG, RPGO_ERR_HIP, RPGO_ERR_NOMEM, RPGO_ERR_STATE = -1, -2, -3, -4
PCM_FP64_REPLAY = 0
PCM_FP32_RECHECK = 1
CLIQUE_HEU_REF = 0
CLIQUE_EXACT = 1
NOALLOC_MAX_VERTICES = 32768  # RPGO_NOALLOC_MAX_VERTICES

_d = ctypes.POINTER(ctypes.c_double)
_i32 = ctypes.POINTER(ctypes.c_int32)
_u64 = ctypes.POINTER(ctypes.c_uint64)
_f = ctypes.POINTER(ctypes.c_float)


class PcmProblem(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("n_poses", ctypes.c_int32), ("pose0", _d),
                ("odom", _d), ("odom_cov", _d), ("n_lc", ctypes.c_int32), ("lc_from", _i32),
                ("lc_to", _i32), ("lc_pose", _d), ("lc_cov", _d), ("n_chains", ctypes.c_int32),
                ("chain_off", _i32), ("landmark", ctypes.c_int32)]


class PcmParams(ctypes.Structure):
    _fields_ = [("odom_threshold", ctypes.c_double), ("lc_threshold", ctypes.c_double),
                ("precision", ctypes.c_int32), ("want_dist", ctypes.c_int32),
                ("recheck_margin", ctypes.c_double), ("recheck_chain_coef", ctypes.c_double)]


class PcmSimpleParams(ctypes.Structure):
    _fields_ = [("odom_trans_threshold", ctypes.c_double), ("odom_rot_threshold", ctypes.c_double),
                ("dist_trans_threshold", ctypes.c_double), ("dist_rot_threshold", ctypes.c_double),
                ("want_dist", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class PcmIncrement(ctypes.Structure):
    _fields_ = [("n_odom", ctypes.c_int32), ("odom", _d), ("odom_cov", _d), ("n_lc", ctypes.c_int32),
                ("lc_from", _i32), ("lc_to", _i32), ("lc_pose", _d), ("lc_cov", _d)]


class PcmStats(ctypes.Structure):
    _fields_ = [("n_accepted", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("n_pairs", ctypes.c_int64), ("n_rechecked", ctypes.c_int64),
                ("n_flipped", ctypes.c_int64), ("n_consistent", ctypes.c_int64),
                ("ms_fold", ctypes.c_double), ("ms_odom", ctypes.c_double),
                ("ms_pairs", ctypes.c_double), ("ms_recheck", ctypes.c_double),
                ("ms_total", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


SOLVER_LM = 0
SOLVER_GN = 1
FACTOR_BETWEEN = 0
FACTOR_PRIOR = 1
GNC_COST = 0
GNC_PROBABILITY = 1


class PcmBatch(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("n_graphs", ctypes.c_int32), ("pose_off", ctypes.POINTER(ctypes.c_int64)),
                ("pose0", _d), ("odom", _d), ("odom_cov", _d), ("lc_off", ctypes.POINTER(ctypes.c_int64)),
                ("lc_from", _i32), ("lc_to", _i32), ("lc_pose", _d), ("lc_cov", _d), ("flags", ctypes.c_int32)]


BATCH_GRAPH = 1
BATCH_AUTOTUNE = 2


class PcmBatchStats(ctypes.Structure):
    _fields_ = [("n_pairs", ctypes.c_int64), ("n_consistent", ctypes.c_int64), ("ms_fold", ctypes.c_double),
                ("ms_odom", ctypes.c_double), ("ms_pairs", ctypes.c_double), ("ms_clique", ctypes.c_double),
                ("ms_total", ctypes.c_double), ("ms_tune", ctypes.c_double), ("pair_waves", ctypes.c_int32),
                ("clique_waves", ctypes.c_int32), ("graph_replayed", ctypes.c_int32),
                ("clique_prefetch", ctypes.c_int32), ("fp32", ctypes.c_int32), ("n_rechecked", ctypes.c_int64),
                ("n_flipped", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SolveProblem(ctypes.Structure):
    _fields_ = [("n_poses", ctypes.c_int32), ("poses", _d), ("n_factors", ctypes.c_int32),
                ("kind", _i32), ("key_i", _i32), ("key_j", _i32), ("measured", _d),
                ("information", _d)]


class SolveParams(ctypes.Structure):
    _fields_ = [("solver", ctypes.c_int32), ("use_gnc", ctypes.c_int32),
                ("known_inliers", ctypes.c_int32), ("gnc_threshold_mode", ctypes.c_int32),
                ("gnc_inlier_threshold", ctypes.c_double), ("gnc_max_iterations", ctypes.c_int32),
                ("gnc_mu_step", ctypes.c_double), ("gnc_relative_cost_tol", ctypes.c_double),
                ("gnc_weights_tol", ctypes.c_double), ("gnc_bias_odom", ctypes.c_int32),
                ("lm_diagonal_damping", ctypes.c_int32), ("init_weights", _d),
                ("init_mu", ctypes.c_double), ("lambda_initial", ctypes.c_double),
                ("lambda_factor", ctypes.c_double), ("lambda_upper_bound", ctypes.c_double),
                ("min_model_fidelity", ctypes.c_double), ("lm_max_iterations", ctypes.c_int32),
                ("lm_relative_error_tol", ctypes.c_double), ("lm_absolute_error_tol", ctypes.c_double),
                ("pcg_max_iterations", ctypes.c_int32), ("pcg_rel_tol", ctypes.c_double),
                ("polish_iterations", ctypes.c_int32), ("pcg_segment", ctypes.c_int32),
                ("gnc_warm_start", ctypes.c_int32), ("pcg_coarse", ctypes.c_int32),
                ("pcg_graphs", ctypes.c_int32), ("gnc_factor_dim", ctypes.c_int32),
                ("pose2_chart", ctypes.c_int32)]


class SolveReport(ctypes.Structure):
    _fields_ = [("gnc_iterations", ctypes.c_int32), ("lm_iterations", ctypes.c_int32),
                ("pcg_iterations", ctypes.c_int32), ("linearizations", ctypes.c_int32),
                ("mu_init", ctypes.c_double), ("mu_final", ctypes.c_double),
                ("final_cost", ctypes.c_double), ("ms_total", ctypes.c_double),
                ("n_inliers", ctypes.c_int32), ("n_tls_disagree", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_i64 = ctypes.POINTER(ctypes.c_int64)


class SolveBatch(ctypes.Structure):
    _fields_ = [("n_problems", ctypes.c_int32), ("pose_off", _i64), ("factor_off", _i64), ("poses", _d),
                ("kind", _i32), ("key_i", _i32), ("key_j", _i32), ("measured", _d), ("information", _d),
                ("known_inliers", _i32)]


class SolveBatchReport(ctypes.Structure):
    _fields_ = [("ms_wall", ctypes.c_double), ("ms_device_sum", ctypes.c_double), ("lanes", ctypes.c_int32),
                ("failed_problem", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


SOLVE_MAX_LANES = 16

# (name, argtypes, restype) of every exported entry point
SIGNATURES = [
    ("rpgo_solve_params_default", [ctypes.POINTER(SolveParams)], None),
    ("rpgo_solve", [ctypes.c_void_p, ctypes.POINTER(SolveProblem), ctypes.POINTER(SolveParams), _d, _d,
                    ctypes.POINTER(SolveReport)], ctypes.c_int),
    ("rpgo_solve_batched", [ctypes.c_void_p, ctypes.POINTER(SolveBatch), ctypes.POINTER(SolveParams),
                            ctypes.c_int32, _d, _d, ctypes.POINTER(SolveReport),
                            ctypes.POINTER(SolveBatchReport)], ctypes.c_int),
    ("rpgo_ctx_create", [ctypes.c_int, ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
    ("rpgo_ctx_destroy", [ctypes.c_void_p], ctypes.c_int),
    ("rpgo_ctx_sync", [ctypes.c_void_p], ctypes.c_int),
    ("rpgo_ctx_set_stage_timing", [ctypes.c_void_p, ctypes.c_int], ctypes.c_int),
    ("rpgo_last_error", [], ctypes.c_char_p),
    ("rpgo_version", [], ctypes.c_char_p),
    ("rpgo_pcm_upload", [ctypes.c_void_p, ctypes.POINTER(PcmProblem)], ctypes.c_int),
    ("rpgo_pcm_run", [ctypes.c_void_p, ctypes.POINTER(PcmParams), ctypes.POINTER(PcmStats)], ctypes.c_int),
    ("rpgo_pcm_pairs", [ctypes.c_void_p, ctypes.POINTER(PcmParams), ctypes.POINTER(PcmStats)], ctypes.c_int),
    ("rpgo_pcm_download", [ctypes.c_void_p, _i32, _d, _u64, _d], ctypes.c_int),
    ("rpgo_pcm_adjacency", [ctypes.c_void_p, ctypes.POINTER(PcmProblem), ctypes.POINTER(PcmParams),
                            _i32, _d, _u64, _d, ctypes.POINTER(PcmStats)], ctypes.c_int),
    ("rpgo_pcm_download_traj", [ctypes.c_void_p, _d], ctypes.c_int),
    ("rpgo_pcm_append", [ctypes.c_void_p, ctypes.POINTER(PcmIncrement), ctypes.POINTER(PcmParams),
                         ctypes.POINTER(PcmStats)], ctypes.c_int),
    ("rpgo_pcm_simple_append", [ctypes.c_void_p, ctypes.POINTER(PcmIncrement),
                                ctypes.POINTER(PcmSimpleParams), ctypes.POINTER(PcmStats)], ctypes.c_int),
    ("rpgo_pcm_simple_run", [ctypes.c_void_p, ctypes.POINTER(PcmSimpleParams), ctypes.POINTER(PcmStats)],
     ctypes.c_int),
    ("rpgo_pcm_simple_adjacency", [ctypes.c_void_p, ctypes.POINTER(PcmProblem),
                                   ctypes.POINTER(PcmSimpleParams), _i32, _d, _u64, _d,
                                   ctypes.POINTER(PcmStats)], ctypes.c_int),
    ("rpgo_max_clique", [ctypes.c_void_p, ctypes.c_int32, _i32, _i32, _d], ctypes.c_int),
    ("rpgo_max_clique_bits", [ctypes.c_void_p, ctypes.c_int32, _u64, ctypes.c_int32, _i32, _i32],
     ctypes.c_int),
    ("rpgo_max_clique_bits_noalloc", [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, _i32, _i32],
     ctypes.c_int),
    ("rpgo_max_clique_incremental", [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, _i32, _i32],
     ctypes.c_int),
    ("rpgo_max_clique_incremental_bits", [ctypes.c_void_p, ctypes.c_int32, _u64, ctypes.c_int32,
                                          ctypes.c_int32, _i32, _i32], ctypes.c_int),
    ("rpgo_pcm_batch_run", [ctypes.c_void_p, ctypes.POINTER(PcmBatch), ctypes.POINTER(PcmParams), _i32, _i32,
                            _i32, _i32, ctypes.POINTER(PcmBatchStats)], ctypes.c_int),
    ("rpgo_pcm_batch_download_adj", [ctypes.c_void_p, ctypes.c_int32, _u64], ctypes.c_int),
    ("rpgo_lc_features", [ctypes.c_void_p, _i32, ctypes.c_int32, _d, ctypes.c_double, _f], ctypes.c_int),
    ("rpgo_solve_errors", [ctypes.c_void_p, _d, _d], ctypes.c_int),
    ("rpgo_gnc_threshold", [ctypes.c_double, ctypes.c_int32], ctypes.c_double),
    ("rpgo_mlp_set_params", [ctypes.c_void_p, _f, _f], ctypes.c_int),
    ("rpgo_mlp_get_params", [ctypes.c_void_p, _f], ctypes.c_int),
    ("rpgo_mlp_forward", [ctypes.c_void_p, ctypes.c_int32, _f, _f], ctypes.c_int),
    ("rpgo_mlp_gradients", [ctypes.c_void_p, ctypes.c_int32, _f, _f, _f, ctypes.c_float, _f, _d],
     ctypes.c_int),
    ("rpgo_mlp_train", [ctypes.c_void_p, ctypes.c_int32, _f, _f, _f, ctypes.c_float, ctypes.c_int32,
                        ctypes.c_int32, ctypes.c_float, ctypes.c_uint64, _d], ctypes.c_int),
]

LC_FEATURES = 16
MLP_HIDDEN = 128
MLP_OUT = 2
MLP_NPARAMS = 16 * 128 + 128 + 128 * 128 + 128 + 128 * 2 + 2

_lib = None


def lib():
    """Load librpgo_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with kimera-rpgo_amd/build.py "
                               "(the HIP library is the only compute path)")
        L = ctypes.CDLL(LIB_PATH)
        for name, argt, rest in SIGNATURES:
            f = getattr(L, name)
            f.argtypes = argt
            f.restype = rest
        _lib = L
    return _lib


class RpgoError(RuntimeError):
    pass


def check(rc):
    if rc != RPGO_OK:
        msg = lib().rpgo_last_error()
        raise RpgoError(f"rpgo status {rc}: {msg.decode() if msg else ''}")


def _ptr(a, t=_d):
    return None if a is None else a.ctypes.data_as(t)


class Context:
    """Owns an rpgo_ctx (device memory + one HIP stream)."""

    def __init__(self, device: int = 0):
        self._h = ctypes.c_void_p()
        check(lib().rpgo_ctx_create(device, 0, ctypes.byref(self._h)))
        self._keep = []

    def close(self):
        if self._h:
            lib().rpgo_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def sync(self):
        check(lib().rpgo_ctx_sync(self._h))

    def set_stage_timing(self, enabled: bool):
        """HIP events around the PCM pass stages (the ms_* stats; on by default)."""
        check(lib().rpgo_ctx_set_stage_timing(self._h, 1 if enabled else 0))

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- PCM
    def pcm_upload(self, dim, pose0, odom, odom_cov, lc_from, lc_to, lc_pose, lc_cov, chain_off=None,
                   landmark=False):
        """chain_off: None for one chain, else the n_chains + 1 chain offsets (several
        robots; pose0 then holds one start pose per chain).  landmark: the loop
        closures are observations of one landmark from poses lc_from (lc_to is
        not read and may be None)."""
        if landmark and lc_to is None:
            lc_to = lc_from
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        arrs = dict(pose0=f64(pose0), odom=f64(odom), odom_cov=f64(odom_cov), lc_from=i32(lc_from),
                    lc_to=i32(lc_to), lc_pose=f64(lc_pose), lc_cov=f64(lc_cov))
        nc = 1 if chain_off is None else len(chain_off) - 1
        pl, dd = (12, 36) if dim == 3 else (4, 9)
        n_odom = arrs["odom"].size // pl
        n_lc = arrs["lc_from"].shape[0]
        _expect_size("pose0", arrs["pose0"], nc * pl)
        _expect_size("odom", arrs["odom"], n_odom * pl)
        _expect_size("odom_cov", arrs["odom_cov"], n_odom * dd)
        _expect_size("lc_to", arrs["lc_to"], n_lc)
        _expect_size("lc_pose", arrs["lc_pose"], n_lc * pl)
        _expect_size("lc_cov", arrs["lc_cov"], n_lc * dd)
        n_poses = n_odom + nc
        if chain_off is not None:
            arrs["chain_off"] = i32(chain_off)
            if int(arrs["chain_off"][-1]) != n_poses:
                raise ValueError("chain_off does not match the odometry rows")
        p = PcmProblem(dim=dim, n_poses=n_poses, pose0=_ptr(arrs["pose0"]), odom=_ptr(arrs["odom"]),
                       odom_cov=_ptr(arrs["odom_cov"]), n_lc=arrs["lc_from"].shape[0],
                       lc_from=_ptr(arrs["lc_from"], _i32), lc_to=_ptr(arrs["lc_to"], _i32),
                       lc_pose=_ptr(arrs["lc_pose"]), lc_cov=_ptr(arrs["lc_cov"]), n_chains=nc,
                       chain_off=_ptr(arrs["chain_off"], _i32) if chain_off is not None else None,
                       landmark=1 if landmark else 0)
        check(lib().rpgo_pcm_upload(self._h, ctypes.byref(p)))
        self._n_lc = arrs["lc_from"].shape[0]
        self._n_poses = n_poses
        self._dim = dim

    def pcm_run(self, odom_thr, lc_thr, precision=PCM_FP64_REPLAY, want_dist=False, margin=0.0,
                chain_coef=0.0):
        prm = PcmParams(odom_threshold=odom_thr, lc_threshold=lc_thr, precision=precision,
                        want_dist=1 if want_dist else 0, recheck_margin=margin,
                        recheck_chain_coef=chain_coef)
        st = PcmStats()
        check(lib().rpgo_pcm_run(self._h, ctypes.byref(prm), ctypes.byref(st)))
        self._m = st.n_accepted
        return st

    def pcm_simple_run(self, odom_trans, odom_rot, dist_trans, dist_rot, want_dist=False):
        """PcmSimple stages (PoseWithNode distances) on the uploaded problem."""
        prm = PcmSimpleParams(odom_trans_threshold=odom_trans, odom_rot_threshold=odom_rot,
                              dist_trans_threshold=dist_trans, dist_rot_threshold=dist_rot,
                              want_dist=1 if want_dist else 0)
        st = PcmStats()
        check(lib().rpgo_pcm_simple_run(self._h, ctypes.byref(prm), ctypes.byref(st)))
        self._m = st.n_accepted
        return st

    def _increment(self, odom=None, odom_cov=None, lc_from=(), lc_to=(), lc_pose=None, lc_cov=None):
        pl, dd = (12, 36) if self._dim == 3 else (4, 9)
        f64 = lambda a, w: np.ascontiguousarray(np.zeros((0, w)) if a is None else a, dtype=np.float64).reshape(-1, w)
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32).reshape(-1)
        arrs = dict(odom=f64(odom, pl), odom_cov=f64(odom_cov, dd), lc_from=i32(lc_from), lc_to=i32(lc_to),
                    lc_pose=f64(lc_pose, pl), lc_cov=f64(lc_cov, dd))
        inc = PcmIncrement(n_odom=arrs["odom"].shape[0], odom=_ptr(arrs["odom"]), odom_cov=_ptr(arrs["odom_cov"]),
                           n_lc=arrs["lc_from"].shape[0], lc_from=_ptr(arrs["lc_from"], _i32),
                           lc_to=_ptr(arrs["lc_to"], _i32), lc_pose=_ptr(arrs["lc_pose"]),
                           lc_cov=_ptr(arrs["lc_cov"]))
        return inc, arrs

    def pcm_append(self, odom_thr, lc_thr, **increment):
        """Stream one increment (new odometry at the chain end, new loop closures) into
        the device-resident PCM state; returns the stats of this call."""
        inc, arrs = self._increment(**increment)
        prm = PcmParams(odom_threshold=odom_thr, lc_threshold=lc_thr, precision=PCM_FP64_REPLAY, want_dist=0)
        st = PcmStats()
        check(lib().rpgo_pcm_append(self._h, ctypes.byref(inc), ctypes.byref(prm), ctypes.byref(st)))
        self._after_append(arrs, st)
        return st

    def pcm_simple_append(self, odom_trans, odom_rot, dist_trans, dist_rot, **increment):
        inc, arrs = self._increment(**increment)
        prm = PcmSimpleParams(odom_trans_threshold=odom_trans, odom_rot_threshold=odom_rot,
                              dist_trans_threshold=dist_trans, dist_rot_threshold=dist_rot, want_dist=0)
        st = PcmStats()
        check(lib().rpgo_pcm_simple_append(self._h, ctypes.byref(inc), ctypes.byref(prm), ctypes.byref(st)))
        self._after_append(arrs, st)
        return st

    def _after_append(self, arrs, st):
        self._n_poses += arrs["odom"].shape[0]
        self._n_lc += arrs["lc_from"].shape[0]
        self._m = st.n_accepted

    def pcm_pairs(self, lc_thr, precision=PCM_FP64_REPLAY, want_dist=False, margin=0.0,
                  chain_coef=0.0):
        """Re-run only the pairwise stage on the last run's device-resident data."""
        prm = PcmParams(odom_threshold=-1.0, lc_threshold=lc_thr, precision=precision,
                        want_dist=1 if want_dist else 0, recheck_margin=margin,
                        recheck_chain_coef=chain_coef)
        st = PcmStats()
        check(lib().rpgo_pcm_pairs(self._h, ctypes.byref(prm), ctypes.byref(st)))
        return st

    def pcm_download(self, want_dist=False):
        m, n = self._m, self._n_lc
        W = (m + 63) // 64
        acc = np.zeros(max(n, 1), np.int32)
        od = np.zeros(max(n, 1))
        adj = np.zeros(max(m * W, 1), np.uint64)
        dist = np.zeros(max(m * m, 1)) if want_dist else None
        check(lib().rpgo_pcm_download(self._h, _ptr(acc, _i32), _ptr(od), _ptr(adj, _u64), _ptr(dist)))
        out = dict(accepted=acc[:m].copy(), odom_dist=od[:n].copy(),
                   adj_bits=adj[: m * W].reshape(m, W).copy())
        if want_dist:
            out["dist"] = dist[: m * m].reshape(m, m).copy()
        return out

    def pcm_traj(self):
        pl, d = (12, 6) if self._dim == 3 else (4, 3)
        t = np.zeros((self._n_poses, pl + d * d + 1))
        check(lib().rpgo_pcm_download_traj(self._h, _ptr(t)))
        return t

    def max_clique(self, mode=CLIQUE_HEU_REF):
        idx = np.zeros(max(self._m, 1), np.int32)
        k = ctypes.c_int32()
        ms = ctypes.c_double()
        check(lib().rpgo_max_clique(self._h, mode, _ptr(idx, _i32), ctypes.byref(k), ctypes.byref(ms)))
        return int(k.value), idx[: max(k.value, 0)].copy(), ms.value

    # ---------------------------------------------------------------- solve
    def solve(self, poses, kind, key_i, key_j, measured, information, **params):
        """rpgo_solve on an SE(3) factor list; params override the reference
        defaults (rpgo_solve_params_default).  Returns (poses, weights, report)."""
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        poses, measured, information = f64(poses).reshape(-1, 12), f64(measured), f64(information)
        kind, key_i, key_j = i32(kind), i32(key_i), i32(key_j)
        N, F = poses.shape[0], kind.shape[0]
        _expect_size("key_i", key_i, F)
        _expect_size("key_j", key_j, F)
        _expect_size("measured", measured, 12 * F)
        _expect_size("information", information, 36 * F)
        prob = SolveProblem(n_poses=N, poses=_ptr(poses), n_factors=F, kind=_ptr(kind, _i32),
                            key_i=_ptr(key_i, _i32), key_j=_ptr(key_j, _i32), measured=_ptr(measured),
                            information=_ptr(information))
        prm = SolveParams()
        lib().rpgo_solve_params_default(ctypes.byref(prm))
        keep = None
        for k, v in params.items():
            if k == "init_weights":
                if v is not None:
                    keep = f64(v)
                    _expect_size("init_weights", keep, F)
                    prm.init_weights = _ptr(keep)
                continue
            if not hasattr(prm, k):
                raise KeyError(k)
            setattr(prm, k, v)
        out = np.zeros((N, 12))
        w = np.zeros(max(F, 1))
        rep = SolveReport()
        check(lib().rpgo_solve(self._h, ctypes.byref(prob), ctypes.byref(prm), _ptr(out), _ptr(w),
                               ctypes.byref(rep)))
        return out, w[:F].copy(), rep

    def solve_batched(self, problems, known_inliers=None, lanes=0, **params):
        """rpgo_solve_batched over a list of (poses, kind, key_i, key_j, measured,
        information) tuples (the arguments of solve); known_inliers: one count per
        problem, or None for the params' value.  Returns (list of poses, list of
        weights, list of SolveReport, SolveBatchReport)."""
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        G = len(problems)
        Ns = [np.asarray(p[0]).reshape(-1, 12).shape[0] for p in problems]
        Fs = [np.asarray(p[1]).shape[0] for p in problems]
        pose_off = np.zeros(G + 1, np.int64)
        factor_off = np.zeros(G + 1, np.int64)
        pose_off[1:] = np.cumsum(Ns)
        factor_off[1:] = np.cumsum(Fs)
        cat = lambda k, f, shape: (f(np.concatenate([np.asarray(p[k]).reshape(shape) for p in problems]))
                                   if G else f(np.zeros((0,) + shape[1:])))
        poses = cat(0, f64, (-1, 12))
        kind, key_i, key_j = cat(1, i32, (-1,)), cat(2, i32, (-1,)), cat(3, i32, (-1,))
        measured, information = cat(4, f64, (-1, 12)), cat(5, f64, (-1, 36))
        for g, p in enumerate(problems):
            for name, k, per in (("key_i", 2, 1), ("key_j", 3, 1), ("measured", 4, 12), ("information", 5, 36)):
                _expect_size(f"problem {g} {name}", np.asarray(p[k]).reshape(-1), per * Fs[g])
        F = int(factor_off[-1])
        ki = None
        if known_inliers is not None:
            ki = i32(known_inliers)
            _expect_size("known_inliers", ki, G)
        b = SolveBatch(n_problems=G, pose_off=pose_off.ctypes.data_as(_i64),
                       factor_off=factor_off.ctypes.data_as(_i64), poses=_ptr(poses), kind=_ptr(kind, _i32),
                       key_i=_ptr(key_i, _i32), key_j=_ptr(key_j, _i32), measured=_ptr(measured),
                       information=_ptr(information), known_inliers=_ptr(ki, _i32) if ki is not None else None)
        prm = SolveParams()
        lib().rpgo_solve_params_default(ctypes.byref(prm))
        keep = None
        for k, v in params.items():
            if k == "init_weights":
                if v is not None:
                    keep = f64(v)
                    _expect_size("init_weights", keep, F)
                    prm.init_weights = _ptr(keep)
                continue
            if not hasattr(prm, k):
                raise KeyError(k)
            setattr(prm, k, v)
        out = np.zeros((max(int(pose_off[-1]), 1), 12))
        w = np.zeros(max(F, 1))
        reps = (SolveReport * max(G, 1))()
        brep = SolveBatchReport()
        check(lib().rpgo_solve_batched(self._h, ctypes.byref(b), ctypes.byref(prm), int(lanes), _ptr(out), _ptr(w),
                                       reps, ctypes.byref(brep)))
        poses_l = [out[pose_off[g]:pose_off[g + 1]].copy() for g in range(G)]
        w_l = [w[factor_off[g]:factor_off[g + 1]].copy() for g in range(G)]
        return poses_l, w_l, [reps[g] for g in range(G)], brep

    # ------------------------------------------------------ batched PCM
    def pcm_batch_run(self, graphs, odom_thr, lc_thr, clique=True, flags=0, precision=None):
        """graphs: sequence of synth.PoseGraph (same dim).  Returns dict with per-graph
        accepted lists, clique sizes / members and the batch stats.  precision:
        PCM_FP64_REPLAY (the default) or PCM_FP32_RECHECK (SE(3) batches)."""
        dim = graphs[0].dim
        G = len(graphs)
        npz = np.array([g.n_poses for g in graphs], np.int64)
        nlc = np.array([g.n_lc for g in graphs], np.int64)
        pose_off = np.concatenate([[0], np.cumsum(npz)]).astype(np.int64)
        lc_off = np.concatenate([[0], np.cumsum(nlc)]).astype(np.int64)
        cat = lambda xs, dt: np.ascontiguousarray(np.concatenate(xs) if len(xs) else np.zeros(0), dtype=dt)
        arrs = dict(pose0=cat([np.asarray(g.pose0).reshape(1, -1) for g in graphs], np.float64),
                    odom=cat([np.asarray(g.odom).reshape(g.n_poses - 1, -1) for g in graphs], np.float64),
                    odom_cov=cat([np.asarray(g.odom_cov).reshape(g.n_poses - 1, -1) for g in graphs], np.float64),
                    lc_from=cat([g.lc_from for g in graphs], np.int32), lc_to=cat([g.lc_to for g in graphs], np.int32),
                    lc_pose=cat([np.asarray(g.lc_pose).reshape(g.n_lc, -1) for g in graphs], np.float64),
                    lc_cov=cat([np.asarray(g.lc_cov).reshape(g.n_lc, -1) for g in graphs], np.float64))
        b = PcmBatch(dim=dim, n_graphs=G, pose_off=pose_off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                     pose0=_ptr(arrs["pose0"]), odom=_ptr(arrs["odom"]), odom_cov=_ptr(arrs["odom_cov"]),
                     lc_off=lc_off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                     lc_from=_ptr(arrs["lc_from"], _i32), lc_to=_ptr(arrs["lc_to"], _i32),
                     lc_pose=_ptr(arrs["lc_pose"]), lc_cov=_ptr(arrs["lc_cov"]), flags=int(flags))
        prm = PcmParams(odom_threshold=odom_thr, lc_threshold=lc_thr,
                        precision=PCM_FP64_REPLAY if precision is None else int(precision), want_dist=0)
        n_acc = np.zeros(G, np.int32)
        acc = np.zeros(max(int(lc_off[-1]), 1), np.int32)
        ksz = np.zeros(G, np.int32)
        kid = np.zeros(max(int(lc_off[-1]), 1), np.int32)
        st = PcmBatchStats()
        check(lib().rpgo_pcm_batch_run(self._h, ctypes.byref(b), ctypes.byref(prm), _ptr(n_acc, _i32),
                                       _ptr(acc, _i32), _ptr(ksz, _i32) if clique else None,
                                       _ptr(kid, _i32) if clique else None, ctypes.byref(st)))
        out = {"n_accepted": n_acc, "stats": st,
               "accepted": [acc[lc_off[g]:lc_off[g] + n_acc[g]].copy() for g in range(G)]}
        if clique:
            out["clique_size"] = ksz
            out["clique_idx"] = [kid[lc_off[g]:lc_off[g] + max(ksz[g], 0)].copy() for g in range(G)]
        self._batch_m = n_acc
        return out

    def pcm_batch_adj(self, g):
        m = int(self._batch_m[g])
        W = (m + 63) // 64
        a = np.zeros(max(m * W, 1), np.uint64)
        check(lib().rpgo_pcm_batch_download_adj(self._h, int(g), _ptr(a, _u64)))
        return a[: m * W].reshape(m, W) if m else np.zeros((0, 0), np.uint64)

    # ------------------------------------------------------ learned initialisation
    def lc_features(self, clique_idx, lc_resid=None, barc=1.0):
        """(m, 16) float32 features of the accepted loop closures of the last PCM run;
        lc_resid: residuals of the m accepted loop closures after GNC's first solve."""
        ci = np.ascontiguousarray(clique_idx, dtype=np.int32)
        m = self.pcm_download()["accepted"].shape[0]
        r = None if lc_resid is None else np.ascontiguousarray(lc_resid, dtype=np.float64)
        if r is not None and r.shape[0] != m:
            raise ValueError(f"lc_resid has {r.shape[0]} entries, expected {m}")
        out = np.zeros((max(m, 1), LC_FEATURES), np.float32)
        check(lib().rpgo_lc_features(self._h, _ptr(ci, _i32), ci.shape[0], _ptr(r), float(barc), _ptr(out, _f)))
        return out[:m]

    def solve_errors(self, n_factors, poses=None):
        p = None if poses is None else np.ascontiguousarray(poses, dtype=np.float64)
        out = np.zeros(max(n_factors, 1))
        check(lib().rpgo_solve_errors(self._h, _ptr(p), _ptr(out)))
        return out[:n_factors]

    def mlp_set_params(self, params, norm):
        p = np.ascontiguousarray(params, dtype=np.float32).reshape(MLP_NPARAMS)
        n = np.ascontiguousarray(norm, dtype=np.float32).reshape(2 * LC_FEATURES)
        check(lib().rpgo_mlp_set_params(self._h, _ptr(p, _f), _ptr(n, _f)))

    def mlp_get_params(self):
        out = np.zeros(MLP_NPARAMS, np.float32)
        check(lib().rpgo_mlp_get_params(self._h, _ptr(out, _f)))
        return out

    def mlp_forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, LC_FEATURES)
        y = np.zeros((x.shape[0], MLP_OUT), np.float32)
        check(lib().rpgo_mlp_forward(self._h, x.shape[0], _ptr(x, _f), _ptr(y, _f)))
        return y

    def mlp_gradients(self, x, target, sample_w=None, mu_weight=1.0):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, LC_FEATURES)
        t = np.ascontiguousarray(target, dtype=np.float32).reshape(-1, MLP_OUT)
        sw = None if sample_w is None else np.ascontiguousarray(sample_w, dtype=np.float32)
        g = np.zeros(MLP_NPARAMS, np.float32)
        loss = ctypes.c_double()
        check(lib().rpgo_mlp_gradients(self._h, x.shape[0], _ptr(x, _f), _ptr(t, _f), _ptr(sw, _f),
                                       float(mu_weight), _ptr(g, _f), ctypes.byref(loss)))
        return g, loss.value

    def mlp_train(self, x, target, sample_w=None, mu_weight=1.0, batch=1024, epochs=10, lr=1e-3, seed=0):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, LC_FEATURES)
        t = np.ascontiguousarray(target, dtype=np.float32).reshape(-1, MLP_OUT)
        sw = None if sample_w is None else np.ascontiguousarray(sample_w, dtype=np.float32)
        hist = np.zeros(max(epochs, 1))
        check(lib().rpgo_mlp_train(self._h, x.shape[0], _ptr(x, _f), _ptr(t, _f), _ptr(sw, _f),
                                   float(mu_weight), int(batch), int(epochs), float(lr), int(seed),
                                   _ptr(hist)))
        return hist[:epochs]

    def max_clique_incremental(self, num_new, prev_size):
        idx = np.zeros(max(self._m, 1), np.int32)
        k = ctypes.c_int32()
        check(lib().rpgo_max_clique_incremental(self._h, int(num_new), int(prev_size), _ptr(idx, _i32),
                                                ctypes.byref(k)))
        return int(k.value), idx[: max(k.value, 0)].copy()

    def max_clique_incremental_bits(self, adj_bits, n, num_new, prev_size):
        return _incremental_bits(self, adj_bits, n, num_new, prev_size)

    def max_clique_bits(self, adj_bits, n, mode=CLIQUE_HEU_REF):
        a = np.ascontiguousarray(adj_bits, dtype=np.uint64).reshape(-1)
        _expect_size("adj_bits", a, n * ((n + 63) // 64), at_least=True)
        idx = np.zeros(max(n, 1), np.int32)
        k = ctypes.c_int32()
        check(lib().rpgo_max_clique_bits(self._h, n, _ptr(a, _u64), mode, _ptr(idx, _i32), ctypes.byref(k)))
        return int(k.value), idx[: max(k.value, 0)].copy()

    def max_clique_bits_noalloc(self, d_adj_bits, n):
        """rpgo_max_clique_bits_noalloc: heuristic clique of the adjacency at
        device address d_adj_bits (n rows of ceil(n/64) words, caller owned)."""
        idx = np.zeros(max(n, 1), np.int32)
        k = ctypes.c_int32()
        check(lib().rpgo_max_clique_bits_noalloc(self._h, n, ctypes.c_void_p(int(d_adj_bits)),
                                                 _ptr(idx, _i32), ctypes.byref(k)))
        return int(k.value), idx[: max(k.value, 0)].copy()


def _expect_size(name, arr, n, at_least=False):
    """The C ABI reads arrays by count: a size mismatch (or, with at_least, too
    few entries) is an error here."""
    size = np.asarray(arr).size
    if size < n or (size != n and not at_least):
        raise ValueError(f"{name} has {size} entries, expected {'at least ' if at_least else ''}{n}")


def _incremental_bits(ctx, adj_bits, n, num_new, prev_size):
    a = np.ascontiguousarray(adj_bits, dtype=np.uint64).reshape(-1)
    idx = np.zeros(max(n, 1), np.int32)
    k = ctypes.c_int32()
    check(lib().rpgo_max_clique_incremental_bits(ctx._h, n, _ptr(a, _u64), int(num_new), int(prev_size),
                                                 _ptr(idx, _i32), ctypes.byref(k)))
    return int(k.value), idx[: max(k.value, 0)].copy()


def unpack_bits(adj_bits, m):
    """(m, W) uint64 rows -> (m, m) uint8 matrix."""
    a = np.ascontiguousarray(adj_bits, dtype=np.uint64).reshape(m, -1)
    b = np.unpackbits(a.view(np.uint8), axis=1, bitorder="little")
    return b[:, :m].copy()


def pack_bits(adj):
    """(m, m) 0/1 matrix -> (m, W) uint64 rows (bit i&63 of word i>>6)."""
    adj = np.asarray(adj, dtype=np.uint8)
    m = adj.shape[0]
    W = (m + 63) // 64
    pad = np.zeros((m, W * 64), np.uint8)
    pad[:, :m] = adj
    return np.packbits(pad, axis=1, bitorder="little").view(np.uint64).reshape(m, W)
