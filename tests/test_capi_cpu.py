"""CPU checks of the C ABI library: it loads and exports every symbol that
include/rpgo_hip.h declares (no compute calls: there is no GPU here)."""
import ctypes
import os
import re

import numpy as np

from rpgo_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(rpgo_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    L = _capi.lib()
    declared = _declared()
    assert len(declared) >= 12
    for name in sorted(declared):
        assert hasattr(L, name), name
    bound = {n for n, _, _ in _capi.SIGNATURES}
    assert declared <= bound, declared - bound


def test_version_and_error_string_without_gpu():
    L = _capi.lib()
    assert L.rpgo_version().decode().startswith("rpgo-mi355x")
    assert L.rpgo_last_error() is not None


def test_stage_timing_rejects_null_context():
    L = _capi.lib()
    assert L.rpgo_ctx_set_stage_timing(None, 0) == _capi.RPGO_ERR_ARG
    assert b"NULL context" in L.rpgo_last_error()


def test_bit_packing_roundtrip():
    rng = np.random.default_rng(0)
    for m in (1, 5, 63, 64, 65, 200):
        A = (rng.uniform(size=(m, m)) < 0.3).astype(np.uint8)
        bits = _capi.pack_bits(A)
        assert bits.shape == (m, (m + 63) // 64)
        assert np.array_equal(_capi.unpack_bits(bits, m), A)
        # bit (i & 63) of word i >> 6
        j, i = (int(v) for v in rng.integers(0, m, 2))
        assert ((int(bits[j, i >> 6]) >> (i & 63)) & 1) == A[j, i]


def test_noalloc_clique_rejects_null_context():
    L = _capi.lib()
    k = ctypes.c_int32()
    assert L.rpgo_max_clique_bits_noalloc(None, 10, None, None, ctypes.byref(k)) == _capi.RPGO_ERR_ARG
    assert b"invalid argument" in L.rpgo_last_error()


def test_binding_rejects_mis_sized_arrays_before_the_c_call():
    """The C ABI reads arrays by count; the binding checks sizes first (no GPU
    call is reached: the checks run before the context handle is used)."""
    import pytest

    class NoCtx:
        _h = None

    eye = np.concatenate([np.eye(3).reshape(9), np.zeros(3)])
    poses = np.tile(eye, (3, 1))
    with pytest.raises(ValueError, match="information"):
        _capi.Context.solve(NoCtx(), poses, [0, 0], [0, 1], [1, 2], np.zeros((2, 12)), np.zeros((1, 36)))
    with pytest.raises(ValueError, match="init_weights"):
        _capi.Context.solve(NoCtx(), poses, [0], [0], [1], np.zeros((1, 12)), np.zeros((1, 36)),
                            init_weights=np.ones(3))
    with pytest.raises(ValueError, match="lc_cov"):
        _capi.Context.pcm_upload(NoCtx(), 3, eye, np.tile(eye, (2, 1)), np.zeros((2, 36)), [0], [2],
                                 eye[None], np.zeros((1, 35)))
    with pytest.raises(ValueError, match="odom_cov"):
        _capi.Context.pcm_upload(NoCtx(), 3, eye, np.tile(eye, (2, 1)), np.zeros((1, 36)), [0], [2],
                                 eye[None], np.zeros((1, 36)))


def _batch_arrays(sizes):
    """Packed chain problems (odometry between factors + a prior on pose 0)."""
    eye = np.concatenate([np.eye(3).reshape(9), np.zeros(3)])
    pose_off = np.zeros(len(sizes) + 1, np.int64)
    factor_off = np.zeros(len(sizes) + 1, np.int64)
    pose_off[1:] = np.cumsum(sizes)
    factor_off[1:] = np.cumsum(sizes)  # n - 1 odometry + 1 prior
    kind = np.concatenate([np.r_[np.zeros(n - 1), 1] for n in sizes]).astype(np.int32)
    ki = np.concatenate([np.r_[np.arange(n - 1), 0] for n in sizes]).astype(np.int32)
    kj = np.concatenate([np.r_[np.arange(1, n), 0] for n in sizes]).astype(np.int32)
    F = int(factor_off[-1])
    return dict(pose_off=pose_off, factor_off=factor_off, poses=np.tile(eye, (int(pose_off[-1]), 1)),
                kind=kind, key_i=ki, key_j=kj, measured=np.tile(eye, (F, 1)),
                information=np.tile(np.eye(6).reshape(36), (F, 1)))


def _call_batched(a, n_lanes=0, ctx=None):
    L = _capi.lib()
    i64 = ctypes.POINTER(ctypes.c_int64)
    b = _capi.SolveBatch(n_problems=len(a["pose_off"]) - 1, pose_off=a["pose_off"].ctypes.data_as(i64),
                         factor_off=a["factor_off"].ctypes.data_as(i64), poses=_capi._ptr(a["poses"]),
                         kind=_capi._ptr(a["kind"], _capi._i32), key_i=_capi._ptr(a["key_i"], _capi._i32),
                         key_j=_capi._ptr(a["key_j"], _capi._i32), measured=_capi._ptr(a["measured"]),
                         information=_capi._ptr(a["information"]), known_inliers=None)
    prm = _capi.SolveParams()
    L.rpgo_solve_params_default(ctypes.byref(prm))
    out = np.zeros((max(int(a["pose_off"][-1]), 1), 12))
    rc = L.rpgo_solve_batched(ctx, ctypes.byref(b), ctypes.byref(prm), n_lanes, _capi._ptr(out), None, None,
                              None)
    return rc, L.rpgo_last_error().decode()


def test_solve_batched_checks_the_batch_before_the_context():
    """rpgo_solve_batched rejects a malformed batch, naming the problem, before
    it touches the context (so these run without a GPU); a well-formed batch
    with a NULL context is rejected last."""
    a = _batch_arrays([5, 3, 7])
    rc, msg = _call_batched(a)
    assert rc == _capi.RPGO_ERR_ARG and "NULL context" in msg
    bad = dict(a, pose_off=a["pose_off"].copy())
    bad["pose_off"][2] = bad["pose_off"][1]  # problem 1 without poses
    rc, msg = _call_batched(bad)
    assert rc == _capi.RPGO_ERR_ARG and "problem 1" in msg
    bad = dict(a, key_j=a["key_j"].copy())
    bad["key_j"][5 + 1] = 3  # problem 1, factor 1: key 3 of a 3-pose problem
    rc, msg = _call_batched(bad)
    assert rc == _capi.RPGO_ERR_ARG and "problem 1" in msg and "factor 1" in msg
    rc, msg = _call_batched(a, n_lanes=_capi.SOLVE_MAX_LANES + 1)
    assert rc == _capi.RPGO_ERR_ARG and "n_lanes" in msg
    bad = dict(a, factor_off=a["factor_off"].copy())
    bad["factor_off"][0] = 1
    rc, msg = _call_batched(bad)
    assert rc == _capi.RPGO_ERR_ARG and "factor_off[0]" in msg


def test_solve_batched_binding_checks_sizes():
    import pytest

    class NoCtx:
        _h = None

    eye = np.concatenate([np.eye(3).reshape(9), np.zeros(3)])
    ok = (np.tile(eye, (3, 1)), [0, 1], [0, 0], [1, 0], np.zeros((2, 12)), np.zeros((2, 36)))
    bad = (np.tile(eye, (3, 1)), [0, 1], [0, 0], [1, 0], np.zeros((2, 12)), np.zeros((1, 36)))
    with pytest.raises(ValueError, match="problem 1 information"):
        _capi.Context.solve_batched(NoCtx(), [ok, bad])
    with pytest.raises(ValueError, match="known_inliers"):
        _capi.Context.solve_batched(NoCtx(), [ok, ok], known_inliers=[1])
