"""Benchmark: FP64 blocked Cholesky and Householder QR on one MI355X.

Workload (one "step"):
  * cholRecursive!(copy(A), Val{:L}) with A = BᵀB + n·I, n = 16384,
    B iid N(0,1) — algorithmic work n³/3 flop;
  * qrBlocked!(copy(M), 128) with M 8192×4096 iid N(0,1) — 2mn² − 2n³/3 flop.
Inputs are generated on the device and are resident in HBM before timing;
copy(A) is inside the step, as in the reference's own timing
(`@time cholRecursive!(copy(A), Val{:L})`, README.md:36-39).

value = (chol flop + qr flop) / step time, TFLOP/s.  Extra fields give each
factorisation alone (HIP events on the stream they run on), size-independent
parity checks at full size, the roofline of the dominant kernel (the MFMA
SYRK/GEMM engine) and the CPU baseline (the oracle's restatement of the
reference on the host cores, on a bounded sample).
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "genericlinearalgebra.jl_amd"), os.path.join(ROOT, "oracle")]

FP64_PEAK_TFLOPS = 78.6
N_CHOL = 16384
M_QR, N_QR, NB_QR = 8192, 4096, 128


def chol_flops(n):
    return n ** 3 / 3.0


def qr_flops(m, n):
    return 2.0 * m * n * n - 2.0 * n ** 3 / 3.0


def cpu_info():
    cores = os.cpu_count() or 1
    model = ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                model = line.split(":", 1)[1].strip()
    except OSError:
        pass
    return cores, model


def cpu_baseline():
    """The oracle (plain-C restatement of src/cholesky.jl and src/qr.jl, one
    thread, generic loops as in the reference) on a bounded sample."""
    import numpy as np

    import oracle as O

    rng = np.random.default_rng(1)
    n = 1536
    Bm = rng.standard_normal((n, n))
    A = Bm.T @ Bm + n * np.eye(n)
    O.lib()
    t0 = time.perf_counter()
    O.chol_recursive(A, 1)
    tc = time.perf_counter() - t0
    m, k = 1536, 768
    M = rng.standard_normal((m, k))
    t0 = time.perf_counter()
    O.qr_blocked(M, 128)
    tq = time.perf_counter() - t0
    fl = chol_flops(n) + qr_flops(m, k)
    return {
        "value": fl / (tc + tq) / 1e12,
        "unit": "TFLOP/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"oracle cholRecursive!(copy(A), Val{{:L}}) n={n} ({tc:.2f}s) + "
                   f"qrBlocked!(copy(M), 128) {m}x{k} ({tq:.2f}s), FP64, single thread"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the factorisations run on one MI355X); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    import gla

    dev = torch.device("cuda:0")
    torch.manual_seed(0)
    n = N_CHOL
    Bm = torch.randn(n, n, dtype=torch.float64, device=dev)
    A0 = torch.addmm(torch.eye(n, dtype=torch.float64, device=dev), Bm.t(), Bm, beta=n, alpha=1.0)
    del Bm
    A0 = A0.t().contiguous().t()                    # column-major (symmetric anyway)
    Aw = torch.empty_like(A0)
    M0 = torch.randn(N_QR, M_QR, dtype=torch.float64, device=dev).t()
    Mw = torch.empty(N_QR, M_QR, dtype=torch.float64, device=dev).t()
    tau = torch.zeros(N_QR, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    def chol():
        Aw.copy_(A0)
        gla.cholRecursive_(Aw, gla.Val.L)

    def qr():
        Mw.copy_(M0)
        gla.qrBlocked_(Mw, NB_QR, tau)

    def step():
        chol()
        qr()

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    flops = chol_flops(n) + qr_flops(M_QR, N_QR)

    # each factorisation alone, HIP events on the current stream
    def ev_time(fn, reps=2):
        ts = []
        for _ in range(reps):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1) / 1e3)
        return min(ts)

    t_copy_a = ev_time(lambda: Aw.copy_(A0))
    t_copy_m = ev_time(lambda: Mw.copy_(M0))
    t_chol = ev_time(chol)       # leaves the factors in Aw, Mw for the checks
    t_qr = ev_time(qr)

    # size-independent parity checks at full size (Aw, Mw hold the last factors)
    idx = torch.randperm(n, device=dev)[:64]
    L = torch.tril(Aw)
    chol_bwd = ((L[idx] @ L.t() - A0[idx]).norm() / A0[idx].norm()).item()
    del L
    x = torch.randn(N_QR, 1, dtype=torch.float64, device=dev)
    Rx = torch.triu(Mw[:N_QR, :N_QR]) @ x
    QRx = torch.zeros(M_QR, 1, dtype=torch.float64, device=dev)
    QRx[:N_QR] = Rx
    # Q·(R x) through F[Tuple{:QBlocked}] (the compact-WY block), against M x
    Hq = gla.QR2(Mw, tau)[gla.QBlocked]
    QRx = Hq @ gla.colmajor(QRx)
    qr_bwd = ((QRx - M0 @ x).norm() / (M0 @ x).norm()).item()

    # roofline of the dominant kernel: the MFMA SYRK of the top recursion
    # level of cholRecursive! at n = 16384 (C 8192×8192 lower -= A21 A21ᵀ,
    # K = 8192), one launch timed with HIP events on the stream it runs on
    h = n // 2
    Cs = torch.empty_like(A0)[:h, :h]
    Xs = A0[h:, :h]
    t_syrk = ev_time(lambda: gla.rankUpdate_(gla.Hermitian(Cs, "L"), Xs, -1.0), reps=3)
    syrk_flops = float(h) * float(h) * float(h)       # n(n+1)k ≈ n²k for the lower triangle
    achieved = syrk_flops / t_syrk / 1e12

    out = {
        "metric": "FP64 Cholesky/QR TFLOP/s & %peak; mixed+IR vs FP64 solve time at 1e-12 bwd err",
        "value": flops / dt / 1e12,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (device RNG): A = BᵀB + n·I, B iid N(0,1); M iid N(0,1)",
        "config": {
            "workload": (f"cholRecursive!(copy(A), Val{{:L}}) n={n} + qrBlocked!(copy(M), {NB_QR}) "
                         f"{M_QR}x{N_QR}, Float64, one MI355X"),
            "chol_n": n, "qr_m": M_QR, "qr_n": N_QR, "qr_blocksize": NB_QR,
        },
        "pct_fp64_peak": 100.0 * flops / dt / 1e12 / FP64_PEAK_TFLOPS,
        "chol": {"ms": t_chol * 1e3, "tflops": chol_flops(n) / (t_chol - t_copy_a) / 1e12,
                 "tflops_incl_copy": chol_flops(n) / t_chol / 1e12,
                 "pct_peak": 100 * chol_flops(n) / (t_chol - t_copy_a) / 1e12 / FP64_PEAK_TFLOPS,
                 "backward_error_rows": chol_bwd},
        "qr": {"ms": t_qr * 1e3, "tflops": qr_flops(M_QR, N_QR) / (t_qr - t_copy_m) / 1e12,
               "pct_peak": 100 * qr_flops(M_QR, N_QR) / (t_qr - t_copy_m) / 1e12 / FP64_PEAK_TFLOPS,
               "residual_QRx_vs_Mx": qr_bwd},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "kernel": "gemm_kernel<double,128,128,N,T,lower> (SYRK 8192x8192, K=8192)",
                     "ms_per_launch": t_syrk * 1e3},
    }
    if not args.no_cpu:
        cores, model = cpu_info()
        cb = cpu_baseline()
        cb["host_cpu"] = model
        cb["host_cores_available"] = cores
        out["cpu_baseline"] = cb
    print(json.dumps(out))


if __name__ == "__main__":
    main()
