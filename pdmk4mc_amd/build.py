"""Compiles libhpdmk_amd.so for gfx950 with hipcc (no GPU needed)."""
import os
import pathlib
import shutil
import subprocess

PKG = pathlib.Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libhpdmk_amd.so"
SOURCES = ["dmk_mc.hip", "dmk_build.hip", "dmk_proxy.hip", "dmk_nudft_mfma.hip", "ewald.hip", "mlp_train.hip", "hpdmk_capi.cpp", "hpdmk_replicas.cpp",
           "hpdmk_ewald.cpp", "hpdmk_mlp.cpp", "dmk_host.cpp"]
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
FLAGS = [
    "-std=c++17", "-O3", "--offload-arch=gfx950", "-fPIC", "-shared",
    "-munsafe-fp-atomics",
    # no packed FP32 instructions (they are avoided on this machine pool)
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


def build(verbose=False):
    cmd = [HIPCC, *FLAGS, "-o", str(LIB) + ".tmp", *[str(CSRC / s) for s in SOURCES]]
    out = subprocess.run(cmd, capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + out.stderr[-6000:])
    os.replace(str(LIB) + ".tmp", LIB)
    if verbose:
        print("built", LIB)
    build_cpp_check(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
