"""Python front end of the MI355X-native Kimera-RPGO robust back end.

The compute path is librpgo_hip.so (HIP kernels for gfx950 behind the C ABI
of include/rpgo_hip.h); this package only binds it (ctypes), generates
synthetic workloads and reads g2o files.
"""
from ._capi import (CLIQUE_EXACT, CLIQUE_HEU_REF, PCM_FP32_RECHECK, PCM_FP64_REPLAY, Context,
                    RpgoError, lib, pack_bits, unpack_bits)

__all__ = ["Context", "RpgoError", "lib", "pack_bits", "unpack_bits", "PCM_FP64_REPLAY",
           "PCM_FP32_RECHECK", "CLIQUE_HEU_REF", "CLIQUE_EXACT"]
