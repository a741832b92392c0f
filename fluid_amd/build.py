"""Build recipe of the HIP library (gfx950) -- run by __graft_entry__.build().

Produces fluid_amd/_lib/libfluid_projection.so in-tree with hipcc.  Every
HIP source is compiled with packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "libfluid_projection.so")
SOURCES = ["fp_uniform.hip", "fp_serve.hip", "fp_tall.hip", "fp_tall_remesh.hip", "fp_cm11a.hip", "fp_extrap.hip", "fp_advect.hip", "fp_api.hip"]
ARCH = os.environ.get("FP_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc"), shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wall", "-Wno-unused-result", "-Wno-unused-variable", "-pthread",
           "-I", os.path.join(ROOT, "include"), *srcs, "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    if verbose:
        print(res.stderr)
    os.replace(tmp, LIB)
    build_node_addon()
    return LIB


NODE_ADDON = os.path.join(LIBDIR, "fluid_projection.node")


def build_node_addon() -> str | None:
    """The N-API addon (plain g++ against the node headers) over the C ABI."""
    inc = "/usr/include/node"
    if not os.path.exists(os.path.join(inc, "node_api.h")) or not shutil.which("g++"):
        return None
    tmp = NODE_ADDON + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-DNODE_GYP_MODULE_NAME=fluid_projection",
           "-I", inc, os.path.join(HERE, "node", "fp_addon.cc"), "-L", LIBDIR, "-lfluid_projection",
           "-Wl,-rpath,$ORIGIN", "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("N-API addon build failed:\n" + res.stdout + res.stderr)
    os.replace(tmp, NODE_ADDON)
    return NODE_ADDON


if __name__ == "__main__":
    print(build(verbose=True))
