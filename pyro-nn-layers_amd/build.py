"""Build libpyronn_mi.so (HIP, gfx950) and the torch op library, in-tree.

    python pyro-nn-layers_amd/build.py            # both libraries
    python pyro-nn-layers_amd/build.py --hip-only

The HIP library has no torch dependency; the op library (pyronn_layers/_ops*.so)
links against it and libtorch.
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
import time

PKG_ROOT = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG_ROOT)
CSRC = os.path.join(PKG_ROOT, "csrc")
INCLUDE = os.path.join(REPO, "include")
OUT_DIR = os.path.join(PKG_ROOT, "pyronn_layers", "lib")
HIP_LIB = os.path.join(OUT_DIR, "libpyronn_mi.so")
OPS_LIB = os.path.join(OUT_DIR, "libpyronn_torch_ops.so")
TESTKIT_SRC = os.path.join(REPO, "tests", "native", "rcp_check.hip")
TESTKIT_LIB = os.path.join(REPO, "tests", "native", "libpyronn_testkit.so")
RATES_SRC = os.path.join(REPO, "tests", "native", "unit_rates.hip")
RATES_BIN = os.path.join(REPO, "tests", "native", "unit_rates")

HIP_SOURCES = ["abi_common.cpp", "cone.hip", "planar.hip", "fdk.hip"]
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _run(cmd, cwd=None):
    t0 = time.time()
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd[:3]) + " ...")
    return time.time() - t0, r.stderr


def hipcc() -> str:
    return shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def build_hip(verbose=False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    objs = []
    obj_dir = os.path.join(PKG_ROOT, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    for src in HIP_SOURCES:
        obj = os.path.join(obj_dir, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc()] + HIP_FLAGS + ["-I", INCLUDE, "-I", CSRC, "-c", os.path.join(CSRC, src), "-o", obj]
        dt, err = _run(cmd)
        if verbose:
            print(f"  {src}: {dt:.1f}s")
            if err.strip():
                print(err)
        objs.append(obj)
    tmp = HIP_LIB + ".tmp"
    _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs +
         ["-L/opt/rocm/lib", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, HIP_LIB)
    return HIP_LIB


def build_testkit(verbose=False) -> str:
    """Device-side checks used only by the test suite (tests/native), built with the
    library's flags so that they exercise the same code generation."""
    tmp = TESTKIT_LIB + ".tmp"
    dt, _ = _run([hipcc()] + HIP_FLAGS + ["-shared", "-I", INCLUDE, "-I", CSRC, TESTKIT_SRC, "-o", tmp])
    if verbose:
        print(f"  rcp_check.hip: {dt:.1f}s")
    os.replace(tmp, TESTKIT_LIB)
    return TESTKIT_LIB


def build_rates(verbose=False) -> str:
    """The unit-rate measurement program (load and VALU issue rates per CU; tests/native)."""
    tmp = RATES_BIN + ".tmp"
    dt, _ = _run([hipcc(), "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                  "-packed-fp32-ops", RATES_SRC, "-o", tmp])
    if verbose:
        print(f"  unit_rates.hip: {dt:.1f}s")
    os.replace(tmp, RATES_BIN)
    return RATES_BIN


def build_ops(verbose=False) -> str:
    """g++ build of the TORCH_LIBRARY op library against libtorch and libpyronn_mi."""
    import torch
    from torch.utils import cpp_extension as ce

    incs = ce.include_paths() + [INCLUDE, CSRC, "/opt/rocm/include"]
    libdirs = ce.library_paths() + [OUT_DIR]
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    tmp = OPS_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", f"-D_GLIBCXX_USE_CXX11_ABI={abi}",
           "-D__HIP_PLATFORM_AMD__", "-DUSE_ROCM",
           os.path.join(CSRC, "torch_ops.cpp"), "-o", tmp]
    cmd += [f"-I{p}" for p in incs] + [f"-L{p}" for p in libdirs]
    cmd += ["-lpyronn_mi", "-lc10", "-ltorch", "-ltorch_cpu", "-lc10_hip", "-ltorch_hip",
            "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + ce.library_paths()[0],
            "-L/opt/rocm/lib", "-lamdhip64"]
    dt, err = _run(cmd)
    if verbose:
        print(f"  torch_ops.cpp: {dt:.1f}s")
    os.replace(tmp, OPS_LIB)
    return OPS_LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--hip-only", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args(argv)
    print("building", build_hip(a.verbose))
    if os.path.exists(TESTKIT_SRC):
        print("building", build_testkit(a.verbose))
    if os.path.exists(RATES_SRC):
        print("building", build_rates(a.verbose))
    if not a.hip_only and os.path.exists(os.path.join(CSRC, "torch_ops.cpp")):
        print("building", build_ops(a.verbose))


if __name__ == "__main__":
    main()
