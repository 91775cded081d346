"""Per-move timing of the truncation sweep of operate: the time of each move
of the sweep split into its phases, read from the timestamps that the engine
built with -DQTT_PHASE_TIMING writes (wall clock, 100 MHz), with the Jacobi
sweeps of each SVD move.  Workloads: laplacian_2d (2 x 20 interleaved digits
on a 40-site chain, Laplacian, cutoff 1e-10, maxdim 128; the default) and
comb_3d (comb 3 x 16, bond 128, d/dx, maxdim 128) and fredholm_member (the
costliest network of the 512-function Fredholm batch: 23 sites, bond 95,
the bond-3 kernel operator, cutoff 1e-10).

    python itensornumericalanalysis.jl_amd/build.py --phase-timing
    python itensornumericalanalysis.jl_amd/phases.py [laplacian_2d|comb_3d|fredholm_member]

Phases of a move (csrc/factor.hip, run_move):
  gather   the move's matrix copied from the network into the workspace
  factor   the Householder QR (QR moves; the QR of a tall SVD matrix)
  x        Q formed (QR moves) / Jacobi SVD, cutoff and U (SVD moves)
  u        R (or S V^T) multiplied into the neighbouring core
  end      bond bookkeeping
A diagnostic: it loads libqtt_phase.so in place of libqtt.so.
"""
from __future__ import annotations

import ctypes
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
TICK_US = 0.01  # wall_clock64 runs at 100 MHz


def main():
    sys.path.insert(0, ROOT)
    import __graft_entry__ as ge
    q = ge._load_package()
    from importlib import import_module
    capi = import_module(q.__name__ + "._capi")
    workloads = import_module(q.__name__ + ".workloads")
    capi.LIBPATH = os.path.join(HERE, "libqtt_phase.so")
    if not os.path.exists(capi.LIBPATH):
        raise SystemExit(f"{capi.LIBPATH} is not built (python build.py --phase-timing)")
    import torch

    which = sys.argv[1] if len(sys.argv) > 1 else "laplacian_2d"
    if which == "laplacian_2d":
        w = workloads.laplacian_2d(npts=16)
        fc = q.truncate(w.f, cutoff=1e-16, maxdim=w.extra["compress_maxdim"])
        reps = 5
    elif which == "comb_3d":
        w = workloads.comb_3d()
        fc = w.f
        reps = 2
    elif which == "fredholm_member":
        batch = workloads.fredholm_batch()
        w = batch[97]
        fc = w.f
        reps = 3
    else:
        raise SystemExit(f"unknown workload {which}")
    for _ in range(2):
        q.operate(w.op, fc, cutoff=w.cutoff, maxdim=w.maxdim or None)
    times = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g = q.operate(w.op, fc, cutoff=w.cutoff, maxdim=w.maxdim or None)
        torch.cuda.synchronize()
        times.append((time.perf_counter() - t0) * 1e3)
    L = capi.lib()
    n = 1024  # records read back (a sweep over a tree may revisit edges)
    L.qtt_debug_times.argtypes = [ctypes.POINTER(ctypes.c_ulonglong), ctypes.c_int]
    L.qtt_debug_sweeps.argtypes = [ctypes.POINTER(ctypes.c_ulonglong), ctypes.c_int]
    tb = (ctypes.c_ulonglong * (n * 8))()
    sb = (ctypes.c_ulonglong * n)()
    L.qtt_debug_times(tb, n)
    L.qtt_debug_sweeps(sb, n)
    t = np.array(tb, dtype=np.int64).reshape(n, 8)
    sw = np.array(sb, dtype=np.int64)
    print(f"{which}: operate+truncate wall ms ({reps} calls): {' '.join(f'{x:.2f}' for x in times)}")
    print("move  rows   b  kind     total_us  gather  factor       x       u   end  jacobi_sweeps")
    sums = {"qr": np.zeros(5), "svd": np.zeros(5)}
    nmov = {"qr": 0, "svd": 0}
    sweeps = 0
    for s in range(n):
        r = t[s]
        if r[0] == 0 or r[5] == 0:
            continue
        ph = np.array([(r[k] - r[k - 1]) * TICK_US for k in range(1, 6)])
        kind = "svd" if (r[7] >> 20) else "qr"
        sums[kind] += ph
        nmov[kind] += 1
        sweeps += int(sw[s] & 0xFFFF)
        print(f"{s:4d} {int(r[6]):5d} {int(r[7] & 0xFFFFF):3d}  {kind:4s} {ph.sum():12.1f} "
              + " ".join(f"{x:7.1f}" for x in ph[:4]) + f" {ph[4]:5.1f} {int(sw[s] & 0xFFFF):6d}")
    for kind in ("qr", "svd"):
        p = sums[kind]
        print(f"{kind} sweep: {nmov[kind]} moves, {p.sum():.0f} us  (gather {p[0]:.0f}, factor {p[1]:.0f}, "
              f"x {p[2]:.0f}, u {p[3]:.0f}, end {p[4]:.0f})")
    print(f"jacobi sweeps in all SVD moves: {sweeps}")


if __name__ == "__main__":
    main()
