"""rpgo_max_clique_bits_noalloc against rpgo_max_clique_bits (HEU_REF) through
the C ABI: the same size and members, on two calls in a row (the second with
a smaller graph, so stale scratch contents from the first call must not leak
into it), with the adjacency in caller-owned device memory; a size beyond
RPGO_NOALLOC_MAX_VERTICES is rejected before any device work.  Above 64
words per row the two scans walk different row layouts (compacted and full),
which the comparison covers too."""
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


def _pcm_like_bits(m, seed, rows=512):
    """Packed symmetric PCM-shaped adjacency for large m, built in row chunks
    (no m x m float arrays): pair (i, j) is an edge when a splitmix64 hash of
    (min, max, seed) falls below 0.99 inside the inlier block, 0.03 elsewhere."""
    rng = np.random.default_rng(seed)
    inl = rng.uniform(size=m) < 0.7
    W = (m + 63) // 64
    out = np.zeros((m, W), np.uint64)
    j = np.arange(m, dtype=np.uint64)
    s = np.uint64(seed * 0x9E3779B97F4A7C15 % (1 << 64))
    for r0 in range(0, m, rows):
        i = np.arange(r0, min(r0 + rows, m), dtype=np.uint64)[:, None]
        lo, hi = np.minimum(i, j), np.maximum(i, j)
        z = lo * np.uint64(m) + hi + s
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
        u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
        p = np.where(inl[r0:r0 + rows, None] & inl[None, :], 0.99, 0.03)
        pad = np.zeros((len(i), W * 64), np.uint8)
        pad[:, :m] = (u < p) & (lo != hi)
        out[r0:r0 + rows] = np.packbits(pad, axis=1, bitorder="little").view(np.uint64).reshape(-1, W)
    return out


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


def test_compacted_rounds_match_full_rows():
    """Above 64 words per row, rpgo_max_clique_bits walks each round and the
    winner's replay against the subgraph of vertices with deg >= best,
    renumbered in order (the buffer mapped back after the replay); the
    allocation-free scan walks the full rows.  Both give the same clique and
    buffer (the full-row scan equals the FP64 oracle's in the test above)."""
    hip = _hip()
    with _capi.Context(0) as ctx:
        for m, seed in ((5000, 21), (9000, 22)):  # 79 and 141 words per row
            bits = _pcm_like_bits(m, seed)
            print(f"[compacted] m={m} seed={seed}", flush=True)
            k_ref, idx_ref = ctx.max_clique_bits(bits, m, _capi.CLIQUE_HEU_REF)
            d = DevAdj(hip, bits)
            try:
                k, idx = ctx.max_clique_bits_noalloc(d.ptr.value, m)
            finally:
                d.free()
            print(f"[compacted] m={m} clique {k_ref} (full rows {k})", flush=True)
            assert k == k_ref, (m, k, k_ref)
            assert np.array_equal(idx, idx_ref), m
            assert k > 200, (m, k)  # a dense block spanning many words


def test_noalloc_rejects_size_above_maximum():
    with _capi.Context(0) as ctx:
        n = _capi.NOALLOC_MAX_VERTICES + 1
        with pytest.raises(_capi.RpgoError, match="exceed the maximum"):
            # rejected on the host before any device access: the address is never read
            ctx.max_clique_bits_noalloc(8, n)
