"""Compiles libhpdmk_amd.so for gfx950 with hipcc (no GPU needed)."""
import os
import pathlib
import shutil
import subprocess

PKG = pathlib.Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libhpdmk_amd.so"
SOURCES = ["dmk_mc.hip", "dmk_build.hip", "dmk_proxy.hip", "dmk_nudft_mfma.hip", "ewald.hip", "mlp_train.hip", "hpdmk_capi.cpp", "hpdmk_replicas.cpp",
           "hpdmk_ewald.cpp", "hpdmk_mlp.cpp", "dmk_host.cpp", "hpdmk_noalloc.hip", "dmk_octree.hip"]
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
FLAGS = [
    "-std=c++17", "-O3", "--offload-arch=gfx950", "-fPIC", "-shared",
    "-munsafe-fp-atomics",
    # no packed FP32 instructions (v_pk_fma_f32 and kin) in the device code.
    # The device pass honours the feature; the host pass warns that it does
    # not know it ("not a recognized feature for this target"), which is
    # harmless.  tests/test_build_flags.py checks the device assembly.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


ROOT = PKG.parent
CPP_CHECK_SRC = ROOT / "tests" / "cpp" / "tree_check.cpp"
CPP_CHECK = ROOT / "tests" / "cpp" / "_build" / "tree_check"


def build_cpp_check(verbose=False):
    """Links the C++ caller of include/hpdmk_tree.hpp against the library
    (host compiler only; the binary runs wherever the library does)."""
    CPP_CHECK.parent.mkdir(parents=True, exist_ok=True)
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", str(ROOT / "include"), str(CPP_CHECK_SRC),
           "-o", str(CPP_CHECK), "-L", str(PKG), "-lhpdmk_amd", "-Wl,-rpath,$ORIGIN/../../../pdmk4mc_amd",
           "-Wl,-rpath-link,/opt/rocm/lib"]
    out = subprocess.run(cmd, capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("g++ failed:\n" + " ".join(cmd) + "\n" + out.stderr[-4000:])
    if verbose:
        print("built", CPP_CHECK)
    return CPP_CHECK


OBJ = PKG / "_build"


def _run(cmd):
    out = subprocess.run(cmd, capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + out.stderr[-6000:])


def build(verbose=False):
    """Compiles every source to an object in parallel (each file launches its
    own kernels, so no relocatable device code is needed), then links."""
    from concurrent.futures import ThreadPoolExecutor

    OBJ.mkdir(exist_ok=True)
    cflags = [f for f in FLAGS if f != "-shared"]
    objs = [OBJ / (s + ".o") for s in SOURCES]
    jobs = min(len(SOURCES), max(1, min(16, os.cpu_count() or 1)))
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(_run, [[HIPCC, *cflags, "-c", "-o", str(o), str(CSRC / s)] for s, o in zip(SOURCES, objs)]))
    _run([HIPCC, *FLAGS, "-o", str(LIB) + ".tmp", *[str(o) for o in objs]])
    os.replace(str(LIB) + ".tmp", LIB)
    if verbose:
        print("built", LIB)
    build_cpp_check(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
