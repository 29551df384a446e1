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
