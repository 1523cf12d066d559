"""Builds libumcp_core.so in-tree with hipcc for gfx950.

    python -m umcp_amd.build        (or umcp_amd.build.build() from Python)

All HIP device code is compiled with packed FP32 instructions disabled
(``-Xclang -target-feature -Xclang -packed-fp32-ops``); the host pass warns
that it does not know the feature, which is harmless.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libumcp_core.so")
SOURCES = ["umcp_kernels.hip", "umcp_faces64.hip", "umcp_faces32.hip", "umcp_faces_mx.hip", "umcp_noalloc.hip", "umcp_sort.hip", "surface_host.cpp", "effm_host.cpp", "umcp_lipids.cpp", "umcp_api.cpp", "umcp_ensemble.cpp",
           "umcp_server.cpp", "umcp_io.cpp", "umcp_points.cpp"]
ARCH = os.environ.get("UMCP_OFFLOAD_ARCH", "gfx950")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found; the HIP toolchain is required to build libumcp_core.so")


def build(verbose: bool = False) -> str:
    objs = []
    cmds = []
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    common = ["-O3", "-fPIC", "-std=c++17", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
              "-Wall", "-Wno-unused-function"]
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(build_dir, src.rsplit(".", 1)[0] + ".o")
        if src.endswith(".hip"):
            cmd = [hipcc(), "-x", "hip", "-c", path, "-o", obj] + common + [
                "--offload-arch=" + ARCH, "-ffp-contract=fast",
                "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
        else:  # host-only C++: no device pass
            cmd = [os.environ.get("CXX", "g++"), "-c", path, "-o", obj] + common + [
                "-pthread", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include")]
        if verbose:
            print(" ".join(cmd), flush=True)
        cmds.append(cmd)
        objs.append(obj)
    # the translation units are independent: compile them side by side
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "8"))))
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        for f in [ex.submit(subprocess.run, c, check=True) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    # rocSOLVER (potrf + potri of the effective-mass set-up) and rocBLAS
    libs = ["-L" + os.path.join(ROCM, "lib"), "-lrocsolver", "-lrocblas",
            "-Wl,-rpath," + os.path.join(ROCM, "lib")]
    cmd = [hipcc(), "-shared", "-fPIC", "-pthread", "--offload-arch=" + ARCH, "-o", tmp] + objs + libs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
