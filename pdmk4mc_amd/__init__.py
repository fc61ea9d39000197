"""MI355X-native DMK Monte Carlo path (drop-in for hpdmk's init -> energy ->
trial dE -> accept/update path).  The compute runs in libhpdmk_amd.so (HIP
kernels for gfx950 behind the C ABI of include/hpdmk.h); this package is the
Python mirror of the reference's HPDMKPtTree interface."""
from .hpdmk import (DIRECT, HPDMK_FP64, HPDMK_MIXED, MLP_FEATURES, MLP_PARAMS, PROXY, Ewald, HPDMKError,
                    HPDMKParams, HPDMKPtTree, HPDMKReplicas, copy_probe, lib_path, load_library, mlp_eval,
                    mlp_init, mlp_train, plan_tree, plan_tree_device, pswf_constants, read_particle_info)

__all__ = ["DIRECT", "PROXY", "HPDMK_FP64", "HPDMK_MIXED", "MLP_FEATURES", "MLP_PARAMS", "Ewald", "HPDMKError",
           "HPDMKParams", "HPDMKPtTree", "HPDMKReplicas", "copy_probe", "lib_path", "load_library", "mlp_eval",
           "mlp_init", "mlp_train", "plan_tree", "plan_tree_device", "pswf_constants", "read_particle_info"]
