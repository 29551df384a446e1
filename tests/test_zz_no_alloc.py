"""rpgo_max_clique_bits_noalloc against rpgo_max_clique_bits (HEU_REF) through
the C ABI: the same size and members, on two calls in a row (the second with
a smaller graph, so stale scratch contents from the first call must not leak
into it), with the adjacency in caller-owned device memory; a size beyond
RPGO_NOALLOC_MAX_VERTICES is rejected before any device work."""
import ctypes

import numpy as np
import pytest

import oracle as O
from rpgo_amd import _capi

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime librpgo_hip.so is linked against (already loaded)."""
    _capi.lib()
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


def _pcm_like(m, seed):
    """Dense-but-incomplete inlier block plus sparse outliers (PCM-shaped)."""
    rng = np.random.default_rng(seed)
    inl = rng.uniform(size=m) < 0.7
    p = np.where(inl[:, None] & inl[None, :], 0.9, 0.03)
    A = np.triu(rng.uniform(size=(m, m)) < p, 1)
    A = (A | A.T).astype(np.uint8)
    return A


class DevAdj:
    """Caller-owned device copy of a bit-packed adjacency (test-side hipMalloc)."""

    def __init__(self, hip, bits):
        self.hip = hip
        self.ptr = ctypes.c_void_p()
        nbytes = max(bits.nbytes, 8)
        assert hip.hipMalloc(ctypes.byref(self.ptr), nbytes) == 0
        assert hip.hipMemcpy(self.ptr, bits.ctypes.data, bits.nbytes, 1) == 0  # host to device

    def free(self):
        self.hip.hipDeviceSynchronize()
        self.hip.hipFree(self.ptr)


def test_noalloc_matches_clique_bits_on_two_calls():
    hip = _hip()
    with _capi.Context(0) as ctx:
        for m, seed in ((3000, 11), (700, 12)):
            A = _pcm_like(m, seed)
            bits = np.ascontiguousarray(_capi.pack_bits(A))
            k_ref, idx_ref = ctx.max_clique_bits(bits, m, _capi.CLIQUE_HEU_REF)
            d = DevAdj(hip, bits)
            try:
                print(f"[noalloc] m={m} seed={seed}", flush=True)
                k, idx = ctx.max_clique_bits_noalloc(d.ptr.value, m)
            finally:
                d.free()
            assert k == k_ref and np.array_equal(idx, idx_ref), (m, k, k_ref)
            ko, idxo = O.max_clique_heu_dense(A)
            assert k == ko and np.array_equal(idx, idxo)
        # an empty graph: FMC's maxCliqueHeu returns -1
        k, idx = ctx.max_clique_bits_noalloc(0, 0)
        assert k == -1 and len(idx) == 0


def test_noalloc_rejects_size_above_maximum():
    with _capi.Context(0) as ctx:
        n = _capi.NOALLOC_MAX_VERTICES + 1
        with pytest.raises(_capi.RpgoError, match="exceed the maximum"):
            # rejected on the host before any device access: the address is never read
            ctx.max_clique_bits_noalloc(8, n)
