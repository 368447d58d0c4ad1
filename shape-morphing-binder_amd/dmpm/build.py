"""Builds libdmpm.so (the HIP library behind include/dmpm.h) for gfx950.

    python shape-morphing-binder_amd/dmpm/build.py

hipcc cross-compiles without a GPU.  The library lands next to this file
(dmpm/libdmpm.so), in-tree, so it travels with the repository snapshot.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libdmpm.so")

SOURCES = ["dmpm_kernels.hip", "dmpm_runtime.cpp", "dmpm_io.cpp", "dmpm_mesh.cpp", "dmpm_noalloc.cpp"]
# dmpm_kernels.hip is compiled once per part (DMPM_PART = 0..6, see the file's
# header), in parallel: each part instantiates one group of kernels.
KERNEL_PARTS = 7

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # hardware float atomics (global_atomic_add_f32/f64) instead of CAS loops
    "-munsafe-fp-atomics",
    # keep the compiler from emitting packed FP32 (v_pk_*_f32) instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = True, extra: list | None = None) -> str:
    objdir = os.path.join(ROOT, "_build", "dmpm")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    t0 = time.time()
    units = []
    for src in SOURCES:
        if src == "dmpm_kernels.hip":
            units += [(src, f"dmpm_kernels_p{k}.o", [f"-DDMPM_PART={k}"]) for k in range(KERNEL_PARTS)]
        else:
            units.append((src, os.path.splitext(src)[0] + ".o", []))
    for src, oname, defs in units:
        obj = os.path.join(objdir, oname)
        cmd = [hipcc(), *HIPCC_FLAGS, *(extra or []), *defs, "-I", INCLUDE, "-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd.insert(1, "-xhip")
        procs.append((oname, cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    for src, cmd, p in procs:
        out, _ = p.communicate()
        txt = out.decode(errors="replace")
        # the host pass does not know the packed-fp32 feature; that warning is expected
        lines = [ln for ln in txt.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
        if p.returncode != 0:
            sys.stderr.write("\n".join(lines) + "\n")
            raise RuntimeError(f"hipcc failed on {src}")
        if verbose and lines:
            sys.stderr.write("\n".join(lines[:40]) + "\n")
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} in {time.time() - t0:.1f}s")
    return LIB


if __name__ == "__main__":
    build()
