"""Build libqtt.so (the HIP engine) in-tree for gfx950 with hipcc.

Usage: python build.py   (or call build() from __graft_entry__.build()).
       python build.py --phase-timing   builds libqtt_phase.so beside it: the
       same engine with per-move timestamps in the truncation sweep
       (-DQTT_PHASE_TIMING, read by phases.py); never loaded by the package.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
SOURCES = ["capi.hip", "eval.hip", "sweep.hip", "factor.hip", "grad.hip", "algebra.hip", "io.hip", "bp.hip"]
LIB = os.path.join(HERE, "libqtt.so")
PHASE_LIB = os.path.join(HERE, "libqtt_phase.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # no packed FP32 instructions in any kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    return "hipcc"


def build(verbose: bool = False, defines=(), lib: str = LIB, build_dir: str = "_build") -> str:
    objs = []
    odir = os.path.join(HERE, build_dir)
    os.makedirs(odir, exist_ok=True)
    procs = []
    for src in SOURCES:
        obj = os.path.join(odir, src.replace(".hip", ".o"))
        cmd = [hipcc(), *HIPCC_FLAGS, *[f"-D{d}" for d in defines], "-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((src, cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    failed = []
    for src, cmd, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # the host pass does not know the packed-fp32 feature; that warning is expected
        lines = [ln for ln in text.splitlines() if "packed-fp32-ops" not in ln or "error" in ln]
        if verbose or p.returncode:
            sys.stderr.write("\n".join(lines) + ("\n" if lines else ""))
        if p.returncode:
            failed.append(src)
    if failed:
        raise RuntimeError(f"hipcc failed for {failed}")
    tmp = lib + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    subprocess.run(cmd, check=True)
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    if "--phase-timing" in sys.argv[1:]:
        print(build(verbose=True, defines=("QTT_PHASE_TIMING",), lib=PHASE_LIB, build_dir="_build_phase"))
    else:
        print(build(verbose=True))
