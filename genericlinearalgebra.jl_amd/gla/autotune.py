"""Autotuner for the batched uneven Cholesky+solve launch shape.

Times the batched call over a synthetic batch (default: 10,000 SPD
matrices, n uniform in [8, 256], A = BᵀB + n·I) for every combination of
  wave_max  — largest n handled by the wave kernel (one wave per matrix,
              16×16 MFMA tiles); larger ones go to the workgroup kernel, one
              matrix per workgroup; 0 sends every matrix there, 256 none,
  grid_mult — persistent workgroups per CU (four waves each),
per element type, and writes the fastest to gla/batched_tuning.json, which
gla.BatchPlan reads.

Run on the GPU:  python -m gla.autotune  (from genericlinearalgebra.jl_amd/)
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
if os.path.dirname(HERE) not in sys.path:
    sys.path.insert(0, os.path.dirname(HERE))

import gla  # noqa: E402

WAVE_MAX = (0, 128, 256)
GRID_MULT = (1, 2, 3, 4, 6, 8, 12, 16)


def make_batch(count: int, dtype, seed: int = 0):
    rng = np.random.default_rng(seed)
    sizes = rng.integers(8, 257, size=count)
    pb = gla.PackedBatch(sizes, dtype)
    # A = BᵀB + n·I built on the device in groups of equal size
    for n in np.unique(sizes):
        idx = np.nonzero(sizes == n)[0]
        Bm = torch.randn(len(idx), n, n, dtype=dtype, device="cuda")
        A = torch.bmm(Bm.transpose(1, 2).contiguous(), Bm) + n * torch.eye(n, dtype=dtype, device="cuda")
        for k, i in enumerate(idx):
            pb.matrix(int(i)).copy_(A[k])
        bv = torch.randn(len(idx), n, dtype=dtype, device="cuda")
        for k, i in enumerate(idx):
            pb.vector(int(i)).copy_(bv[k])
    return sizes, pb


def time_config(sizes, pb, A0, b0, wave_max, grid_mult, reps=5):
    plan = gla.BatchPlan(sizes, pb.dtype, wave_max=wave_max, grid_mult=grid_mult)
    ts = []
    for _ in range(reps + 1):
        pb.A.copy_(A0)
        pb.b.copy_(b0)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        gla.chol_solve_batched_(plan, pb.A_ptrs, pb.lda, pb.b_ptrs, pb.info)
        e1.record()
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1) / 1e3)
    return min(ts[1:])


def tune(count: int = 10000, out: str | None = None) -> dict:
    table = {}
    for dtype, key in ((torch.float64, "float64"), (torch.float32, "float32")):
        sizes, pb = make_batch(count, dtype)
        A0, b0 = pb.A.clone(), pb.b.clone()
        best = None
        results = []
        for wm in WAVE_MAX:
            for gm in GRID_MULT:
                t = time_config(sizes, pb, A0, b0, wm, gm)
                results.append({"wave_max": wm, "grid_mult": gm, "seconds": t})
                print(f"{key} wave_max={wm:2d} grid_mult={gm}: {t * 1e3:.3f} ms", flush=True)
                if best is None or t < best[0]:
                    best = (t, wm, gm)
        table[key] = {"wave_max": best[1], "grid_mult": best[2], "seconds": best[0],
                      "batch": count, "sizes": "uniform integer [8, 256]", "sweep": results}
    path = out or os.path.join(HERE, "batched_tuning.json")
    with open(path, "w") as f:
        json.dump(table, f, indent=1)
    return table


if __name__ == "__main__":
    out = sys.argv[1] if len(sys.argv) > 1 else None
    tune(out=out)
