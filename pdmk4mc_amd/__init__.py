"""MI355X-native DMK Monte Carlo path (drop-in for hpdmk's init -> energy ->
trial dE -> accept/update path).  The compute runs in libhpdmk_amd.so (HIP
kernels for gfx950 behind the C ABI of include/hpdmk.h); this package is the
Python mirror of the reference's HPDMKPtTree interface."""
from .hpdmk import (DIRECT, PROXY, HPDMK_FP64, HPDMK_MIXED, Ewald, HPDMKError, HPDMKParams, HPDMKPtTree, HPDMKReplicas, copy_probe,
                    lib_path, load_library, plan_tree, pswf_constants)

__all__ = ["DIRECT", "PROXY", "HPDMK_FP64", "HPDMK_MIXED", "Ewald", "HPDMKError", "HPDMKParams", "HPDMKPtTree",
           "HPDMKReplicas", "copy_probe", "lib_path", "load_library", "plan_tree", "pswf_constants"]
