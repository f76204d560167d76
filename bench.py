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

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "genericlinearalgebra.jl_amd"), os.path.join(ROOT, "oracle")]

FP64_PEAK_TFLOPS = 78.6
N_CHOL = 16384
M_QR, N_QR, NB_QR = 8192, 4096, 128
# the look-ahead step width cholRecursive! takes at n = 16384 (handle knob
# chol_block −1: 1024 from n = 14336 up); the first step's trailing SYRK has K = SYRK_K
SYRK_K = 1024


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


def host_threads():
    """Host threads this run may use: the box's CPU share (OMP_NUM_THREADS)
    or every core."""
    cores = os.cpu_count() or 1
    try:
        return max(1, min(cores, int(os.environ.get("OMP_NUM_THREADS", cores))))
    except ValueError:
        return cores


def cpu_baseline():
    """The reference's CPU path (oracle/reference_cpu.c: the package's generic
    loops single-threaded as Julia runs them, the steps Julia hands to BLAS —
    rdiv!, mul!, lmul! with a triangle, dot, inv! — on numpy's OpenBLAS with
    all host threads) on a bounded sample of the step: cholRecursive!(copy(A))
    n = 4096 and qrBlocked!(copy(M), 128) 4096×2048.  Also config C1:
    cholRecursive!(copy(A), Val{:L}) at n = 1000, median of 20, beside the
    reference's published 0.012098 s (README.md:36-39)."""
    import numpy as np

    import reference_cpu as R

    threads = R.set_threads(host_threads())
    rng = np.random.default_rng(1)
    n = 4096
    Bm = rng.standard_normal((n, n))
    A = np.asfortranarray(Bm.T @ Bm + n * np.eye(n))
    del Bm
    W = np.empty_like(A, order="F")
    t0 = time.perf_counter()
    W[...] = A
    info = R.chol_recursive_inplace(W, 1)
    tc = time.perf_counter() - t0
    m, k = 4096, 2048
    M = rng.standard_normal((m, k))
    t0 = time.perf_counter()
    R.qr_blocked(M, 128)
    tq = time.perf_counter() - t0
    fl = chol_flops(n) + qr_flops(m, k)
    # C1
    n1 = 1000
    B1 = rng.standard_normal((n1, n1))
    A1 = np.asfortranarray(B1.T @ B1 + n1 * np.eye(n1))
    ts = []
    for _ in range(21):
        t0 = time.perf_counter()
        W1 = A1.copy(order="F")
        R.chol_recursive_inplace(W1, 1)
        ts.append(time.perf_counter() - t0)
    t1 = float(np.median(ts[1:]))
    L1 = np.tril(W1)
    c1_res = float(np.linalg.norm(A1 - L1 @ L1.T) / np.linalg.norm(A1))
    return {
        "value": fl / (tc + tq) / 1e12,
        "unit": "TFLOP/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"reference CPU path (generic loops 1 thread, BLAS steps on OpenBLAS with "
                   f"{threads} threads): cholRecursive!(copy(A), Val{{:L}}) n={n} ({tc:.2f}s, "
                   f"info {info}) + qrBlocked!(copy(M), 128) {m}x{k} ({tq:.2f}s), FP64"),
        "c1": {"workload": "cholRecursive!(copy(A), Val{:L}) n=1000, A = BᵀB + n·I",
               "s": t1, "gflops": chol_flops(n1) / t1 / 1e9,
               "residual": c1_res, "published_s": 0.012098,
               "published_note": "README.md:36-39, author's machine (unstated)"},
        "lapack": lapack_line(),
    }


def lapack_line():
    """Host LAPACK (scipy's dpotrf + dgeqrf on its bundled OpenBLAS, all BLAS
    threads) on the full bench step: Cholesky n = 16384 and QR 8192×4096,
    each timed once (about 15 s of host time in all).  A second CPU line
    beside the port."""
    import numpy as np
    from scipy.linalg import lapack

    rng = np.random.default_rng(2)
    n = N_CHOL
    # SPD by diagonal dominance ((R + Rᵀ)/2 + n·I, eigenvalues ≥ n − 2√n):
    # O(n²) to build, and LAPACK's time does not depend on the values
    A = rng.standard_normal((n, n))
    A += A.T
    A *= 0.5
    A[np.diag_indices(n)] += n
    A = np.asfortranarray(A)
    M = np.asfortranarray(rng.standard_normal((M_QR, N_QR)))
    threads = os.cpu_count()
    try:
        import threadpoolctl

        threads = max(i.get("num_threads", 0) for i in threadpoolctl.threadpool_info()) or threads
    except Exception:  # noqa: BLE001  (thread count is informational)
        pass

    def med(fn):
        t0 = time.perf_counter()
        fn()
        return time.perf_counter() - t0

    tc = med(lambda: lapack.dpotrf(A, lower=1, overwrite_a=0))
    tq = med(lambda: lapack.dgeqrf(M, overwrite_a=0))
    fl = chol_flops(n) + qr_flops(M_QR, N_QR)
    return {"value": fl / (tc + tq) / 1e12, "unit": "TFLOP/s", "cores": threads,
            "kind": "lapack",
            "sample": (f"scipy dpotrf n={n} ({tc:.2f}s, incl. the copy of A) + dgeqrf "
                       f"{M_QR}x{N_QR} ({tq:.2f}s), FP64, one run each")}


def profiled_traffic(name):
    """HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE and
    WRITE_SIZE (FETCH_SIZE scaled by the plain-copy calibration), as recorded
    in profiles/<name>; None if the summary is absent."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def progress(msg):
    """One line on stderr as each part of the run ends (the JSON result is
    the only line on stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def ev_time(torch, fn, reps=2, mean=False):
    """Best (or, with mean=True, the mean) of `reps` HIP-event timings of fn
    on the current stream (s)."""
    ts = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1) / 1e3)
    return sum(ts) / len(ts) if mean else min(ts)


def mixed_section(torch, gla, n=N_CHOL, nrhs=16):
    """Mixed precision + FP64 refinement at n=16384: A = Q diag(σ) Qᵀ, Q Haar
    (QR of a Gaussian with the sign fix), σ log-spaced 1 → 1/κ; 16 N(0,1)
    right-hand sides.  Time to η ≤ 1e-12 against the plain FP64 path
    (FP64 factor + solve + residual check, same entry point)."""
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(3)
    G = torch.randn(n, n, dtype=torch.float64, device=dev, generator=g)
    Q, Rq = torch.linalg.qr(G)
    Q *= torch.sign(torch.diagonal(Rq))
    del G, Rq
    B = torch.randn(nrhs, n, dtype=torch.float64, device=dev, generator=g).t()
    out = {"n": n, "nrhs": nrhs, "tol": 1e-12, "cases": []}
    for kappa in (1e2, 1e4, 1e6):
        sig = torch.logspace(0, -float(np.log10(kappa)), n, dtype=torch.float64, device=dev)
        A = (Q * sig) @ Q.t()
        A = A.t().contiguous().t()
        row = {"kappa": kappa}
        for prec in ("fp64", "auto", "fp32", "bf16", "fp16"):
            try:
                res = {}

                def run():
                    X, r = gla.chol_solve_mixed(A, B, precision=prec, tol=1e-12, max_iter=30)
                    res["r"] = r

                t = ev_time(torch, run, reps=2)
                r = res["r"]
                row[prec] = {"ms": t * 1e3, "precision": r.precision, "iters": r.iters,
                             "berr": r.berr, "converged": r.converged}
            except gla.DomainError as e:
                row[prec] = {"error": f"DomainError (pivot {e})"}
        for prec in ("auto", "fp32", "bf16", "fp16"):
            if "ms" in row["fp64"] and "ms" in row[prec] and row[prec]["converged"]:
                row[prec]["speedup_vs_fp64"] = row["fp64"]["ms"] / row[prec]["ms"]
        if "speedup_vs_fp64" in row["auto"]:
            row["speedup_auto_vs_fp64"] = row["auto"]["speedup_vs_fp64"]
        out["cases"].append(row)
        del A
        progress(f"mixed kappa {kappa:g} done")
    return out


def batched_section(torch, gla, count=10000):
    """10,000 SPD matrices, n uniform in [8, 256], A = BᵀB + n·I, one RHS,
    factor + solve in one persistent launch, captured in a graph and
    replayed; FP64 and FP32."""
    sys.path.insert(0, os.path.join(ROOT, "genericlinearalgebra.jl_amd"))
    from gla import autotune
    out = {"count": count}
    for dt, key in ((torch.float64, "fp64"), (torch.float32, "fp32")):
        sizes, pb = autotune.make_batch(count, dt, seed=0)
        A0, b0 = pb.A.clone(), pb.b.clone()
        plan = gla.BatchPlan(sizes, dt)
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            gla.chol_solve_batched_(plan, pb.A_ptrs, pb.lda, pb.b_ptrs, pb.info)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            gla.chol_solve_batched_(plan, pb.A_ptrs, pb.lda, pb.b_ptrs, pb.info)
        ts = []
        for _ in range(5):
            pb.A.copy_(A0)
            pb.b.copy_(b0)
            torch.cuda.synchronize()
            ts.append(ev_time(torch, graph.replay, reps=1))
        t = sorted(ts)[len(ts) // 2]
        nn = sizes.astype(np.float64)
        esz = 8 if key == "fp64" else 4
        flops = float(np.sum(nn ** 3 / 3 + 2 * nn ** 2))
        nbytes = float(np.sum(esz * (nn * (nn + 1) + 2 * nn)))
        bad = int((pb.info != 0).sum().item())
        out[key] = {"ms": t * 1e3, "tflops": flops / t / 1e12, "gbs": nbytes / t / 1e9,
                    "hbm_frac": nbytes / t / 8e12, "info_nonzero": bad,
                    "wave_max": plan.wave_max, "grid_mult": plan.grid_mult}
        if key == "fp64":
            # measured HBM bytes of the same call (rocprofv3 counters)
            tb = profiled_traffic("batched_chol_10000_fp64.md")
            if tb:
                out[key]["traffic_bytes_profiled"] = tb
                out[key]["traffic_tbs_at_this_time"] = tb / t / 1e12
    return out


def adjoint_section(torch, gla):
    """Cholesky adjoint for 512 matrices n=128 (FP64, factors recomputed)
    and one n=8192 (FP32); QR adjoint for 4096×256 (FP64)."""
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(5)
    out = {}
    n, cnt = 128, 512
    Bs = torch.randn(cnt, n, n, dtype=torch.float64, device=dev, generator=g)
    As = torch.bmm(Bs.transpose(1, 2).contiguous(), Bs) + n * torch.eye(n, dtype=torch.float64,
                                                                         device=dev)
    Ws = torch.randn(cnt, n, n, dtype=torch.float64, device=dev, generator=g)
    Acm, Wcm = As.transpose(1, 2), Ws.transpose(1, 2)     # column-major matrices
    outb = torch.empty_like(As).transpose(1, 2)
    t = ev_time(torch, lambda: gla.chol_adjoint_batched(Acm, Wcm, "S", out=outb), reps=3)
    out["chol_512x128_fp64"] = {"ms": t * 1e3,
                                "tflops": cnt * 7 * n ** 3 / 3 / t / 1e12,
                                "note": "one batched call; factors recomputed inside (the "
                                        "n^3/3 per matrix of the recompute is not counted)"}
    del As, Ws, Bs, outb
    n = 8192
    Bm = torch.randn(n, n, dtype=torch.float32, device=dev, generator=g)
    A = (Bm.t() @ Bm + n * torch.eye(n, device=dev)).t().contiguous().t()
    del Bm
    L = A.clone()
    gla.cholRecursive_(L, gla.Val.L)
    W = gla.colmajor(torch.tril(torch.randn(n, n, device=dev, generator=g)))
    Ab = gla.empty_colmajor(n, n, torch.float32)
    t = ev_time(torch, lambda: gla.chol_adjoint(L, W, "S", out=Ab), reps=2)
    out["chol_8192_fp32"] = {"ms": t * 1e3, "tflops": 7 * n ** 3 / 3 / t / 1e12}
    del A, L, W, Ab
    m, k = 4096, 256
    M = gla.colmajor(torch.randn(m, k, dtype=torch.float64, device=dev, generator=g))
    q = gla.qrBlocked_(M, 128)
    Rb = gla.colmajor(torch.triu(torch.randn(k, k, dtype=torch.float64, device=dev, generator=g)))
    Abq = gla.empty_colmajor(m, k, torch.float64)
    t = ev_time(torch, lambda: gla.qr_adjoint(q.factors, q.tau, Rb, out=Abq), reps=3)
    fl = k ** 3 / 3 + k ** 3 + (4 * m * k * k - 2 * k ** 3)
    out["qr_4096x256_fp64"] = {"ms": t * 1e3, "tflops": fl / t / 1e12}
    return out


def next_rows_section(torch, gla):
    """The rows beyond Cholesky/QR, each timed with HIP events and checked by
    a size-independent property: blocked LDLᵀ (FP64, n = 8192, indefinite),
    symtri! (FP64, n = 4096), bidiagonalize! (FP64, 4096²), complex
    Cholesky (ComplexF64, n = 8192) and complex qrBlocked! (ComplexF64,
    4096×2048)."""
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(17)
    out = {}

    def sampled_rows(n, k=64):
        return torch.from_numpy(np.sort(np.random.default_rng(0).choice(n, k, replace=False))).to(dev)

    # LDLᵀ: A = G + Gᵀ + 3√n·diag(±1), about half the pivots negative
    n = 8192
    G = torch.randn(n, n, dtype=torch.float64, device=dev, generator=g)
    A = G + G.t()
    A.diagonal().add_(3 * n ** 0.5 * torch.sign(torch.randn(n, dtype=torch.float64, device=dev,
                                                            generator=g)))
    del G
    W = gla.empty_colmajor(n, n, torch.float64)

    def ldlt():
        W.copy_(A)
        gla.ldlt_(gla.Hermitian(W, "L"))

    t_copy = ev_time(torch, lambda: W.copy_(A), reps=2)
    t = ev_time(torch, ldlt, reps=2) - t_copy
    F = gla.LDLt(W, "L")
    rows = sampled_rows(n)
    L = F.L
    res = ((L[rows] * F.d) @ L.t() - A[rows]).norm() / A[rows].norm()
    out["ldlt_8192_fp64"] = {"ms": t * 1e3, "tflops": n ** 3 / 3 / t / 1e12,
                             "pct_fp64_peak": 100 * n ** 3 / 3 / t / 1e12 / FP64_PEAK_TFLOPS,
                             "residual_rows": res.item()}
    del A, W, F, L

    # symtri!: FP64 n = 4096; the dominant kernel is the symmetric product,
    # which reads the trailing lower triangle once per column: Σ (n−c)²/2·8 B
    n = 4096
    G = torch.randn(n, n, dtype=torch.float64, device=dev, generator=g)
    A = G + G.t()
    del G
    W = gla.empty_colmajor(n, n, torch.float64)
    tau = torch.zeros(n - 1, dtype=torch.float64, device=dev)

    def tri():
        W.copy_(A)
        gla.symtri_(gla.Hermitian(W, "L"), tau)

    t_copy = ev_time(torch, lambda: W.copy_(A), reps=2)
    t = ev_time(torch, tri, reps=2) - t_copy
    d = torch.diagonal(W).cpu().numpy()
    e = torch.diagonal(W, -1).cpu().numpy()
    import scipy.linalg as _sl
    lam_t = _sl.eigvalsh_tridiagonal(d, e)
    lam_a = torch.linalg.eigvalsh(A).cpu().numpy()
    symv_bytes = sum((n - c - 1) ** 2 / 2 * 8 for c in range(n - 2))
    out["symtri_4096_fp64"] = {"ms": t * 1e3, "tflops": 4 * n ** 3 / 3 / t / 1e12,
                               "symv_bytes": symv_bytes, "symv_gbs_if_alone": symv_bytes / t / 1e9,
                               "eigenvalue_max_abs_diff_rel": float(np.max(np.abs(lam_t - lam_a))
                                                                    / np.max(np.abs(lam_a)))}
    del A, W

    # bidiagonalize!: FP64 4096×4096; two matrix-vector products per column
    m = n = 4096
    A = torch.randn(n, m, dtype=torch.float64, device=dev, generator=g).t()
    W = torch.empty(n, m, dtype=torch.float64, device=dev).t()

    def bd():
        W.copy_(A)
        gla.bidiagonalize_(W)

    t_copy = ev_time(torch, lambda: W.copy_(A), reps=2)
    t = ev_time(torch, bd, reps=2) - t_copy
    F = gla.BidiagonalFactorization(W, None, None)
    dv, ev = F.bidiagonal
    B = torch.diag(dv) + torch.diag(ev, 1)
    s_b = torch.linalg.svdvals(B).cpu().numpy()
    s_a = torch.linalg.svdvals(A).cpu().numpy()
    fl = 4 * m * n * n - 4 * n ** 3 / 3
    out["bidiag_4096x4096_fp64"] = {"ms": t * 1e3, "tflops": fl / t / 1e12,
                                    "svdvals_max_abs_diff_rel": float(np.max(np.abs(s_b - s_a))
                                                                      / s_a[0])}
    del A, W, B

    # complex Cholesky: ComplexF64 n = 8192, A = BᴴB + n·I; 4n³/3 real flops
    n = 8192
    Bc = torch.complex(torch.randn(n, n, dtype=torch.float64, device=dev, generator=g),
                       torch.randn(n, n, dtype=torch.float64, device=dev, generator=g))
    A = Bc.conj().t() @ Bc
    del Bc
    A.diagonal().add_(n)
    A = (0.5 * (A + A.conj().t())).t().contiguous().t()
    W = gla.empty_colmajor(n, n, torch.complex128)

    def zchol():
        W.copy_(A)
        gla.cholRecursive_(W, gla.Val.L)

    t_copy = ev_time(torch, lambda: W.copy_(A), reps=2)
    t = ev_time(torch, zchol, reps=2) - t_copy
    rows = sampled_rows(n)
    L = torch.tril(W)
    res = (L[rows] @ L.conj().t() - A[rows]).norm() / A[rows].norm()
    out["zchol_8192_c128"] = {"ms": t * 1e3, "tflops": 4 * n ** 3 / 3 / t / 1e12,
                              "pct_fp64_peak": 100 * 4 * n ** 3 / 3 / t / 1e12 / FP64_PEAK_TFLOPS,
                              "residual_rows": res.item()}
    del A, W, L

    # complex qrBlocked!: ComplexF64 4096×2048; 4 × (2mn² − 2n³/3) real flops
    m, n = 4096, 2048
    A = torch.complex(torch.randn(n, m, dtype=torch.float64, device=dev, generator=g),
                      torch.randn(n, m, dtype=torch.float64, device=dev, generator=g)).t()
    W = torch.empty(n, m, dtype=torch.complex128, device=dev).t()
    tauc = torch.zeros(n, dtype=torch.complex128, device=dev)

    def zqr():
        W.copy_(A)
        gla.qrBlocked_(W, 32, tauc)

    t_copy = ev_time(torch, lambda: W.copy_(A), reps=2)
    t = ev_time(torch, zqr, reps=2) - t_copy
    Rr = torch.linalg.qr(A, mode="r")[1]
    R = torch.triu(W[:n, :n])
    ph = torch.diagonal(R) / torch.diagonal(Rr)
    diffr = ((R - ph[:, None] * Rr).norm() / Rr.norm()).item()
    fl = 4 * (2 * m * n * n - 2 * n ** 3 / 3)
    out["zqr_4096x2048_c128"] = {"ms": t * 1e3, "tflops": fl / t / 1e12,
                                 "R_rel_diff_vs_rocsolver_phase_normalised": diffr}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the mixed-precision, batched and adjoint sections")
    ap.add_argument("--profile-panel", action="store_true",
                    help="only run qrBlocked!'s panel (one 8192x128 block column, the "
                         "reconstructed-Householder launch sequence) and a plain copy of the "
                         "same panel five times each, for hardware-counter runs")
    ap.add_argument("--profile-column-panel", action="store_true",
                    help="the same with qrUnblocked! of the panel (the column-by-column "
                         "panel kernel, one launch)")
    ap.add_argument("--profile-syrk", action="store_true",
                    help="only run the trailing SYRK of the Cholesky's first look-ahead step "
                         "(C 14336x14336 lower -= X X^T, K=1024) and a plain 64 MiB copy five "
                         "times each, for hardware-counter runs")
    ap.add_argument("--profile-batched", action="store_true",
                    help="only run the batched Cholesky+solve of the benchmark's FP64 batch "
                         "(10,000 matrices, n uniform in [8, 256]) five times, each after a "
                         "plain device copy restoring the matrices, for hardware-counter runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the factorisations run on one MI355X); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    import gla

    dev = torch.device("cuda:0")
    torch.manual_seed(0)
    if args.profile_syrk:
        kk, mm = SYRK_K, N_CHOL - 2 * SYRK_K
        X = torch.randn(kk, mm, dtype=torch.float64, device=dev).t()
        Cs = torch.randn(mm, mm, dtype=torch.float64, device=dev).t()
        Cc = torch.empty(8 * 1024 * 1024, dtype=torch.float64, device=dev)
        Cd = torch.empty_like(Cc)
        for _ in range(5):
            Cd.copy_(Cc)
            gla.rankUpdate_(gla.Hermitian(Cs, "L"), X, -1.0)
        torch.cuda.synchronize()
        print(json.dumps({"profile_syrk": f"{mm}x{mm} K={kk}", "launches": 5}))
        return
    if args.profile_batched:
        sys.path.insert(0, os.path.join(ROOT, "genericlinearalgebra.jl_amd"))
        from gla import autotune
        sizes, pb = autotune.make_batch(10000, torch.float64, seed=0)
        A0, b0 = pb.A.clone(), pb.b.clone()
        plan = gla.BatchPlan(sizes, torch.float64)
        for _ in range(5):
            pb.A.copy_(A0)
            pb.b.copy_(b0)
            gla.chol_solve_batched_(plan, pb.A_ptrs, pb.lda, pb.b_ptrs, pb.info)
        torch.cuda.synchronize()
        print(json.dumps({"profile_batched": "10000 FP64, n in [8, 256]", "launches": 5,
                          "grid_mult": plan.grid_mult, "matrix_bytes": int(pb.A.numel() * 8)}))
        return
    if args.profile_panel or args.profile_column_panel:
        P0 = torch.randn(NB_QR, M_QR, dtype=torch.float64, device=dev).t()
        Pw = torch.empty_like(P0)
        ptau = torch.zeros(NB_QR, dtype=torch.float64, device=dev)
        for _ in range(5):
            Pw.copy_(P0)
            if args.profile_panel:
                gla.qrBlocked_(Pw, NB_QR)
            else:
                gla.set_tuning("qr_hr", 0)
                gla.qrUnblocked_(Pw, ptau)
                gla.set_tuning("qr_hr", -1)
        torch.cuda.synchronize()
        print(json.dumps({"profile_panel": f"{M_QR}x{NB_QR}", "launches": 5}))
        return
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
    progress(f"timed {args.steps} steps: {dt * 1e3:.2f} ms/step")

    t_copy_a = ev_time(torch, lambda: Aw.copy_(A0))
    t_copy_m = ev_time(torch, lambda: Mw.copy_(M0))
    t_chol = ev_time(torch, chol)       # leaves the factors in Aw, Mw for the checks
    t_qr = ev_time(torch, qr)

    def qr64():
        Mw.copy_(M0)
        gla.qrBlocked_(Mw, 64, tau)

    t_qr64 = ev_time(torch, qr64)                 # the other block size of the config

    # config C1 on the device: cholRecursive!(copy(A), Val{:L}) at n = 1000,
    # copy included, median of 21 HIP-event timings
    g1 = torch.Generator(device=dev).manual_seed(11)
    B1 = torch.randn(1000, 1000, dtype=torch.float64, device=dev, generator=g1)
    A1 = (B1.t() @ B1 + 1000 * torch.eye(1000, dtype=torch.float64, device=dev))
    A1 = A1.t().contiguous().t()
    W1 = torch.empty_like(A1)

    def c1():
        W1.copy_(A1)
        gla.cholRecursive_(W1, gla.Val.L)

    c1()
    t_c1 = sorted(ev_time(torch, c1, reps=1) for _ in range(21))[10]
    L1 = torch.tril(W1)
    c1_res = ((A1 - L1 @ L1.t()).norm() / A1.norm()).item()
    del B1, A1, W1, L1
    progress("single-factorisation timings and C1 done")

    # parity at full size: the factors against rocSOLVER's (an independent
    # library), to 10·n·eps, plus size-independent residuals
    eps = float(np.finfo(np.float64).eps)
    L = torch.tril(Aw)
    Lref = torch.linalg.cholesky(A0)
    chol_diff = ((L - Lref).norm() / Lref.norm()).item()
    del Lref
    idx = torch.randperm(n, device=dev)[:64]
    chol_bwd = ((L[idx] @ L.t() - A0[idx]).norm() / A0[idx].norm()).item()
    del L
    Rref = torch.linalg.qr(M0, mode="r")[1]
    R = torch.triu(Mw[:N_QR, :N_QR])
    sgn = torch.sign(torch.diagonal(R)) * torch.sign(torch.diagonal(Rref))
    qr_diff = ((R - sgn[:, None] * Rref).norm() / Rref.norm()).item()
    del Rref
    x = torch.randn(N_QR, 1, dtype=torch.float64, device=dev)
    QRx = torch.zeros(M_QR, 1, dtype=torch.float64, device=dev)
    QRx[:N_QR] = R @ x
    Hq = gla.QR2(Mw, tau)[gla.QBlocked]           # Q·(R x) through the compact-WY block
    QRx = Hq @ gla.colmajor(QRx)
    qr_bwd = ((QRx - M0 @ x).norm() / (M0 @ x).norm()).item()
    del R, QRx

    # the QR panel: qrBlocked! of one fresh 8192×128 block column is exactly
    # the panel's launch sequence (the reconstructed-Householder path, no
    # trailing update); qrUnblocked! of the same panel with that path off is
    # one launch of the column-by-column kernel.  Algorithmic bytes 16·m·nb (read + write of
    # the panel).
    P0 = M0[:, :NB_QR].contiguous().t().contiguous().t()
    Pw = torch.empty_like(P0)
    ptau = torch.zeros(NB_QR, dtype=torch.float64, device=dev)
    t_pc = ev_time(torch, lambda: Pw.copy_(P0), reps=3)
    t_panel = ev_time(torch, lambda: (Pw.copy_(P0), gla.qrBlocked_(Pw, NB_QR)), reps=5,
                      mean=True) - t_pc
    gla.set_tuning("qr_hr", 0)          # the column-by-column panel kernel
    try:
        t_colpanel = ev_time(torch, lambda: (Pw.copy_(P0), gla.qrUnblocked_(Pw, ptau)),
                             reps=3) - t_pc
    finally:
        gla.set_tuning("qr_hr", -1)
    panel_bytes = 16.0 * M_QR * NB_QR
    achieved = panel_bytes / t_panel / 1e9
    # the dominant kernel (largest share of the step's kernel time): the
    # FP64 direct-to-LDS MFMA SYRK gemm_dl_kernel<8,3,lower> of the
    # look-ahead Cholesky, timed on its first step's shape (C 14336×14336
    # lower -= X Xᵀ, K = 1024, the step width cholRecursive! takes at
    # n = 16384; rankUpdate! of that shape is exactly one launch);
    # algorithmic work mm²·K flop
    kk, mm = SYRK_K, n - 2 * SYRK_K
    Cs = torch.empty_like(A0)[:mm, :mm]
    Xs = A0[2 * kk:, :kk]
    # one warm launch, then the mean of five, as the rocprof summary averages
    # its five dispatches
    gla.rankUpdate_(gla.Hermitian(Cs, "L"), Xs, -1.0)
    t_syrk = ev_time(torch, lambda: gla.rankUpdate_(gla.Hermitian(Cs, "L"), Xs, -1.0), reps=5,
                     mean=True)
    syrk_tf = float(mm) * float(mm) * kk / t_syrk / 1e12
    del Cs

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
        "data": "synthetic (device RNG, seed 0): A = BᵀB + n·I, B iid N(0,1); M iid N(0,1)",
        "config": {
            "workload": (f"cholRecursive!(copy(A), Val{{:L}}) n={n} + qrBlocked!(copy(M), {NB_QR}) "
                         f"{M_QR}x{N_QR}, Float64, one MI355X"),
            "chol_n": n, "qr_m": M_QR, "qr_n": N_QR, "qr_blocksize": NB_QR,
        },
        "pct_fp64_peak": 100.0 * flops / dt / 1e12 / FP64_PEAK_TFLOPS,
        "chol": {"ms": t_chol * 1e3, "tflops": chol_flops(n) / (t_chol - t_copy_a) / 1e12,
                 "tflops_incl_copy": chol_flops(n) / t_chol / 1e12,
                 "pct_peak": 100 * chol_flops(n) / (t_chol - t_copy_a) / 1e12 / FP64_PEAK_TFLOPS,
                 "factor_rel_diff_vs_rocsolver": chol_diff, "bar_10_n_eps": 10 * n * eps,
                 "backward_error_rows": chol_bwd},
        "qr": {"ms": t_qr * 1e3, "tflops": qr_flops(M_QR, N_QR) / (t_qr - t_copy_m) / 1e12,
               "pct_peak": 100 * qr_flops(M_QR, N_QR) / (t_qr - t_copy_m) / 1e12 / FP64_PEAK_TFLOPS,
               "blocksize64_ms": t_qr64 * 1e3,
               "blocksize64_tflops": qr_flops(M_QR, N_QR) / (t_qr64 - t_copy_m) / 1e12,
               "R_rel_diff_vs_rocsolver_sign_normalised": qr_diff,
               "bar_10_n_eps": 10 * M_QR * eps, "residual_QRx_vs_Mx": qr_bwd},
        "c1": {"workload": "cholRecursive!(copy(A), Val{:L}) n=1000, A = BᵀB + n·I, copy "
                           "included, median of 21",
               "ms": t_c1 * 1e3, "gflops": chol_flops(1000) / t_c1 / 1e9,
               "residual": c1_res, "reference_published_ms": 12.098},
        "roofline": {"bound": "mfma", "achieved": syrk_tf, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": syrk_tf / FP64_PEAK_TFLOPS,
                     "traffic": profiled_traffic("syrk_14336_k1024.md"),
                     "kernel": ("gemm_dl_kernel<8,3,lower> (FP64 direct-to-LDS MFMA SYRK "
                                "14336x14336, K=1024)"),
                     "flops_per_launch": float(mm) * float(mm) * kk,
                     "us_per_launch": t_syrk * 1e6},
        "roofline_panel": {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                           "frac": achieved / 8000.0,
                           "traffic": profiled_traffic("qr_hr_panel_8192x128.md"),
                           "kernel": "reconstructed-Householder panel of qrBlocked! (block "
                                     "column 8192x128; its 13-launch sequence timed as one)",
                           "bytes_per_launch": panel_bytes, "us_per_launch": t_panel * 1e6,
                           "column_kernel_us": t_colpanel * 1e6},
    }
    del A0, Aw, M0, Mw
    torch.cuda.empty_cache()
    if not args.no_extras:
        for key, fn in (("mixed", mixed_section), ("batched", batched_section),
                        ("adjoint", adjoint_section), ("next_rows", next_rows_section)):
            try:
                out[key] = fn(torch, gla)
            except Exception as e:  # noqa: BLE001  (a section failure is reported, not fatal)
                out[key] = {"error": f"{type(e).__name__}: {e}"}
            torch.cuda.empty_cache()
            progress(f"{key} section done")
    if not args.no_cpu:
        cores, model = cpu_info()
        cb = cpu_baseline()
        cb["c1"]["device_ms"] = out["c1"]["ms"]
        cb["host_cpu"] = model
        cb["host_cores_available"] = cores
        out["cpu_baseline"] = cb
        progress("cpu baseline done")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
