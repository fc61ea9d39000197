"""The library's hipcc flags keep packed FP32 instructions (v_pk_fma_f32,
v_pk_mul_f32, v_pk_add_f32) out of the device code.

The feature flag in pdmk4mc_amd/build.py is read by hipcc's device pass; its
host pass prints "not a recognized feature for this target" and ignores it.
This checks on the AMD assembly of code built here: a float2 multiply-add
compiles to a packed FMA without the flag and to scalar FMAs with it, and a
library source built with the library's flags has no packed FP32 op."""
import re
import shutil
import subprocess

import pytest

from pdmk4mc_amd import build as B

PACKED = re.compile(r"\bv_pk_(fma|mul|add)_f32\b")
PROBE = r"""
#include <hip/hip_runtime.h>
__global__ void k(float2* a, const float2* b, const float2* c, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) { float2 x = a[i], y = b[i], z = c[i]; x.x = x.x * y.x + z.x; x.y = x.y * y.y + z.y; a[i] = x; }
}
"""

pytestmark = pytest.mark.skipif(shutil.which(B.HIPCC) is None and not B.pathlib.Path(B.HIPCC).exists(),
                                reason="hipcc not available")


def device_asm(src, tmp_path, flags):
    out = tmp_path / (src.stem + ".s")
    cmd = [B.HIPCC, *flags, "--cuda-device-only", "-S", "-o", str(out), str(src)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-2000:]
    return out.read_text()


def device_flags(with_feature):
    fl = [f for f in B.FLAGS if f not in ("-shared", "-fPIC")]
    if not with_feature:
        i = fl.index("-packed-fp32-ops")
        del fl[i - 3:i + 1]
    return fl


def test_feature_flag_removes_packed_fp32(tmp_path):
    src = tmp_path / "probe.hip"
    src.write_text(PROBE)
    assert PACKED.search(device_asm(src, tmp_path, device_flags(False)))
    assert not PACKED.search(device_asm(src, tmp_path, device_flags(True)))


def test_library_source_has_no_packed_fp32(tmp_path):
    asm = device_asm(B.CSRC / "mlp_train.hip", tmp_path, device_flags(True))
    assert "v_fma" in asm or "v_mul" in asm
    assert not PACKED.search(asm)
