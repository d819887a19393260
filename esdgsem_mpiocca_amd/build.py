"""Build the native libraries in-tree.

  libesdg_setup.so  host setup (basis, meshes, geometry, states); g++ only
  libesdg.so        gfx950 HIP kernels + C ABI + deviceclass mirror; hipcc, links libesdg_setup
  oracle/liboracle.so and oracle/_ref/ (test oracle; built by oracle/Makefile)

Run as ``python -m esdgsem_mpiocca_amd.build`` or through ``__graft_entry__.build()``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
ORACLE = os.path.join(ROOT, "oracle")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("ESDG_OFFLOAD_ARCH", "gfx950")

# gfx950 code only, packed FP32 instructions disabled.  Multiply-adds are not contracted into FMA: every kernel then performs the reference's IEEE
# operations in the reference's order, which makes the device results reproduce the x86-64
# reference arithmetic (the kernels are bandwidth-bound, so the extra multiplies are hidden).
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    "-munsafe-fp-atomics",
]

HIP_SOURCES = [
    "esdg_runtime.hip",
    "esdg_kernels.hip",
    "esdg_analysis.hip",
    "esdg_volume.hip",
    "esdg_fused.hip",
    "esdg_stage.hip",
    "esdg_ensemble.hip",
    "esdg_mlp.hip",
    "esdg_geometry.hip",
    "esdg_deviceclass.hip",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_setup() -> str:
    out = os.path.join(PKG, "libesdg_setup.so")
    src = os.path.join(CSRC, "host", "esdg_setup.cpp")
    _run(["g++", "-O2", "-ffp-contract=off", "-std=c++17", "-fPIC", "-shared", "-Wall",
          f"-I{INCLUDE}", src, "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_hip() -> str:
    """Each source is compiled to an object of its own, in parallel (no device symbol crosses
    translation units), then the objects are linked into libesdg.so."""
    from concurrent.futures import ThreadPoolExecutor
    out = os.path.join(PKG, "libesdg.so")
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    if not srcs:
        raise RuntimeError("no HIP sources found")
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    compile_flags = [f for f in HIP_FLAGS if f != "-shared"]
    objs = [os.path.join(objdir, os.path.basename(s) + ".o") for s in srcs]
    jobs = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(jobs) as ex:
        for f in [ex.submit(_run, [HIPCC] + compile_flags + [f"-I{INCLUDE}", f"-I{CSRC}", "-c", s, "-o", o])
                  for s, o in zip(srcs, objs)]:
            f.result()
    _run([HIPCC] + HIP_FLAGS + objs + ["-o", out + ".tmp", f"-L{PKG}", "-lesdg_setup", "-Wl,-rpath,$ORIGIN"])
    os.replace(out + ".tmp", out)
    return out


def build_oracle() -> None:
    _run(["make", "-s", "-C", ORACLE, "all"])


def build_all() -> None:
    build_setup()
    build_hip()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "setup": build_setup, "hip": build_hip, "oracle": build_oracle}[w]()
