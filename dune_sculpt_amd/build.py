"""Build recipe for libsculpt_gpu.so (hipcc, gfx950) -- run by __graft_entry__.build().

Every source is compiled with -ffp-contract=off so FP32 expressions round exactly as
written (the brush spec fixes the order of every operation), and with packed FP32
instructions disabled.  Objects go to dune_sculpt_amd/_build/, the library next to
this file so it travels with the tree.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "_build"
LIB = PKG / "libsculpt_gpu.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-result",
    "-I", str(PKG.parent / "include"),
]

SOURCES = ["kernels.hip", "learned.hip", "batch.hip", "batch_mesh.hip", "ray.hip", "draw.hip", "undo.hip", "proxy.hip",
           "grids.hip", "dyntopo.hip", "api.cpp", "pbvh_build.cpp", "meshgen.cpp", "subdiv_ccg.cpp", "grids_api.cpp",
           "dyntopo_api.cpp"]


def _run(cmds):
    procs = [(c, subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)) for c in cmds]
    failed = []
    for c, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append((c, text))
        elif text.strip() and os.environ.get("SG_BUILD_VERBOSE"):
            sys.stderr.write(text)
    if failed:
        msg = "\n".join(" ".join(c) + "\n" + t for c, t in failed)
        raise RuntimeError("hipcc failed:\n" + msg)


def build(verbose: bool = False) -> Path:
    OUT.mkdir(exist_ok=True)
    srcs = [s for s in SOURCES if (CSRC / s).exists()]
    objs = []
    cmds = []
    for s in srcs:
        o = OUT / (s + ".o")
        objs.append(o)
        cmds.append([HIPCC, "-x", "hip", *COMMON, "-c", str(CSRC / s), "-o", str(o)])
    _run(cmds)
    tmp = LIB.with_suffix(".so.tmp")
    _run([[HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
           "-lpthread"]])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
