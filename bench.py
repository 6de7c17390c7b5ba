"""Benchmark: FP64-accurate dgesv at n=16384 -- plain FP64 LU vs FP32 LU +
FP64 iterative refinement -- plus DGEMM %peak, on one MI355X.

Workload (the dense-solve configuration): A = U*diag(sigma)*V^T, n=16384,
U and V Haar-random orthogonal (QR of a seeded Gaussian, signs fixed by
diag(R)), sigma geometric from 1 down to 1/kappa (kappa=1e6), one
right-hand side b ~ U[-1,1).  A step is one full solve of A x = b with the
inputs already resident in HBM:
  * plain FP64: copy A, clapack-style getrf + getrs on the GPU;
  * mixed: atlmx_dsgesv (FP32 LU on MFMA, FP64 residual/update/stop test).
GFLOP/s counts the FP64 dgesv flops (2/3 n^3 + n^2 (nrhs + 1/2),
bin/slvtst.c:644) for both.  `value` is the mixed solve's rate when its
normwise backward error is <= n*eps64, else the FP64 rate.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 16384]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FP64-accurate dgesv GFLOP/s n=16384: FP64 LU vs low-prec LU+IR; DGEMM %peak"
PEAK_FP64_MFMA = 78.6e12
PEAK_FP32_MFMA = 157.3e12
PEAK_HBM = 8.0e12
EPS64 = 2.0 ** -52


def dgesv_flops(n, nrhs):
    return 2.0 / 3.0 * n ** 3 + n * n * (nrhs + 0.5)


def make_system(n, kappa, nrhs, seed, torch, X):
    """A = U diag(sigma) V^T with Haar U, V (QR of Gaussian) and b ~ U[-1,1)."""
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(seed)

    def haar():
        G = torch.randn(n, n, dtype=torch.float64, device=dev, generator=g)
        Q, R = torch.linalg.qr(G)
        Q *= torch.sign(torch.diagonal(R)).unsqueeze(0)
        return Q

    U = haar()
    V = haar()
    sigma = kappa ** (-torch.arange(n, dtype=torch.float64, device=dev) / (n - 1))
    U *= sigma.unsqueeze(0)                       # U diag(sigma)
    # A = (U diag(sigma)) V^T on the library's DGEMM.  torch tensors are
    # row-major: the column-major view of a row-major M is M^T, so compute
    # A^T = V (U diag(sigma))^T  <=>  col-major C = opA * opB with
    # opA = V (col-major view of V^T storage transposed) ...
    A = torch.empty(n, n, dtype=torch.float64, device=dev)
    X.use_torch_stream()
    # column-major views: Vc = V^T, Uc = (U diag s)^T, Ac = A^T.
    # A^T = V diag(s) U^T = (Vc)^T * Uc  -> gemm(Trans, NoTrans, Vc, Uc)
    X.lib().atlmx_dgemm(X.Trans, X.NoTrans, n, n, n, 1.0, X.dptr(V), n, X.dptr(U), n, 0.0,
                        X.dptr(A), n)
    torch.cuda.synchronize()
    del U, V
    b = torch.rand(nrhs, n, dtype=torch.float64, device=dev, generator=g) * 2 - 1
    # the column-major matrix the library sees is A.T (storage of A); its
    # condition number is the same.
    return A, b


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` as last measured with
    two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE, each scaled by the
    same counter's reading for a 1 GiB device copy in the same run), kept as
    `kernel=... workload=... traffic_bytes_per_launch=...` lines in
    profiles/traffic.txt.  None when that file holds no such line."""
    path = os.path.join(ROOT, "profiles", "traffic.txt")
    try:
        lines = open(path).read().splitlines()
    except OSError:
        return None
    for ln in lines:
        kv = dict(f.split("=", 1) for f in ln.split() if "=" in f)
        if kv.get("kernel") == kernel and kv.get("workload") == workload:
            return float(kv["traffic_bytes_per_launch"])
    return None


def time_steps(fn, steps, warmup, torch):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def cpu_baseline(n_cpu, seed):
    """The oracle's restatement of ATLAS's CPU dgesv (recursive getrfC with a
    threaded blocked GEMM, then getrs), timed on this host's cores on a
    bounded sample of the workload (one n_cpu x n_cpu random system)."""
    import numpy as np
    import oracle as O
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    O.set_threads(cores)
    A = O.gegen(n_cpu, n_cpu, seed=seed)
    B = O.gegen(n_cpu, 1, seed=seed + 1)
    ipiv = np.zeros(n_cpu, dtype=np.int32)
    t0 = time.perf_counter()
    info = O.lib().orc_dgesv(O.COL, n_cpu, 1, A.ctypes.data, n_cpu, ipiv.ctypes.data,
                             B.ctypes.data, n_cpu)
    t = time.perf_counter() - t0
    O.set_threads(1)
    return {"value": round(dgesv_flops(n_cpu, 1) / t / 1e9, 3), "unit": "GFLOP/s",
            "cores": cores, "kind": "port",
            "sample": "oracle dgesv (recursive ATL_getrfC restatement + threaded GEMM) on one "
                      "%dx%d system, 1 RHS, %.1f s, info=%d" % (n_cpu, n_cpu, t, info)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--nrhs", type=int, default=1)
    ap.add_argument("--kappa", type=float, default=1e6)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--gemm-n", type=int, default=8192)
    ap.add_argument("--cpu-n", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-run", action="store_true",
                    help="one warm mixed solve plus a 1 GiB device copy (the byte-counter "
                         "calibration) only: the run rocprofv3 --pmc passes wrap")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import atlas_amd as X
    L = X.lib()
    torch.cuda.set_device(0)
    X.use_torch_stream()
    n, nrhs = args.n, args.nrhs
    log = lambda *a: print(*a, file=sys.stderr, flush=True)   # noqa: E731

    if args.profile_run:
        # uniform[-1,1) A: the same launch sequence as the Haar system (the
        # byte counters depend on shapes, not values) without the 2 QRs'
        # ~10^5 small dispatches that a counter pass serialises
        g = torch.Generator(device="cuda")
        g.manual_seed(args.seed)
        A = torch.rand(n, n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
        b = torch.rand(nrhs, n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
        xmp = torch.empty_like(b)
        X.dsgesv_dev(A, b, xmp)
        src = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)                  # reads 1 GiB, writes 1 GiB
        X.dsgesv_dev(A, b, xmp)
        torch.cuda.synchronize()
        log("profile run done")
        return

    # ---- DGEMM %peak (n=8192, one launch, HIP events on the library stream)
    gn = args.gemm_n
    Ga = torch.rand(gn, gn, dtype=torch.float64, device="cuda") - 0.5
    Gb = torch.rand(gn, gn, dtype=torch.float64, device="cuda") - 0.5
    Gc = torch.rand(gn, gn, dtype=torch.float64, device="cuda") - 0.5
    dg = lambda: L.atlmx_dgemm(X.NoTrans, X.NoTrans, gn, gn, gn, 1.0, X.dptr(Ga), gn,  # noqa
                               X.dptr(Gb), gn, 0.5, X.dptr(Gc), gn)
    dg()
    torch.cuda.synchronize()
    X.prof_enable(1 << X.PROF_GEMM)
    for _ in range(3):
        dg()
    ms, cnt, work = X.prof_query(X.PROF_GEMM)
    X.prof_enable(0)
    dgemm_tflops = work / (ms * 1e-3) / 1e12
    log("dgemm n=%d: %.2f TFLOP/s (%.1f%% of FP64 MFMA peak)" %
        (gn, dgemm_tflops, 100 * dgemm_tflops * 1e12 / PEAK_FP64_MFMA))
    del Ga, Gb, Gc

    # ---- the dense-solve workload
    t0 = time.perf_counter()
    A, b = make_system(n, args.kappa, nrhs, args.seed, torch, X)
    log("generated A (n=%d, kappa=%g) in %.1f s" % (n, args.kappa, time.perf_counter() - t0))
    Aw = torch.empty_like(A)
    x64 = torch.empty_like(b)
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    info = [0]

    def fp64_step():
        Aw.copy_(A)
        x64.copy_(b)
        info[0] = L.atlmx_dgetrf(n, n, X.dptr(Aw), n, X.dptr(ipiv))
        L.atlmx_dgetrs(X.NoTrans, n, nrhs, X.dptr(Aw), n, X.dptr(ipiv), X.dptr(x64), n)

    t64 = time_steps(fp64_step, args.steps, args.warmup, torch)
    be64 = L.atlmx_dbackward_error(n, nrhs, X.dptr(A), n, X.dptr(b), n, X.dptr(x64), n)
    g64 = dgesv_flops(n, nrhs) / t64 / 1e9
    log("fp64 dgesv: %.1f ms, %.1f GFLOP/s, info %d, backward error %.3g" %
        (t64 * 1e3, g64, info[0], be64))

    xmp = torch.empty_like(b)
    st = [None]

    def mp_step():
        info[0], st[0] = X.dsgesv_dev(A, b, xmp)

    tmp = time_steps(mp_step, args.steps, args.warmup, torch)
    # the same steps again with a HIP event pair around every library launch
    # (on the library's stream): per-kernel-class times and the roofline
    X.prof_enable(-1)
    tmp_prof = time_steps(mp_step, args.steps, 0, torch)
    classes = {}
    for cls, name in enumerate(["gemm", "panel", "trsm_inv", "laswp", "potf2", "misc", "trsv",
                                "update"]):
        ms, cnt, work, nbytes = X.prof_query(cls, with_bytes=True)
        if cnt:
            classes[name] = {"ms_per_step": round(ms / args.steps, 3),
                             "launches_per_step": cnt // args.steps}
            if cls == X.PROF_UPDATE:
                gemm_ms, gemm_cnt, gemm_work, gemm_bytes = ms, cnt, work, nbytes
    X.prof_enable(0)
    bemp = L.atlmx_dbackward_error(n, nrhs, X.dptr(A), n, X.dptr(b), n, X.dptr(xmp), n)
    gmp = dgesv_flops(n, nrhs) / tmp / 1e9
    log("mixed dsgesv: %.1f ms, %.1f GFLOP/s, iters %d, converged %d, fallback %d, "
        "backward error %.3g" % (tmp * 1e3, gmp, st[0].iterations, st[0].converged,
                                 st[0].fallback, bemp))
    ok_mp = (bemp <= n * EPS64) and not st[0].fallback
    value = gmp if ok_mp else g64

    # roofline of the dominant compute kernel of the headline solve: the FP32
    # MFMA trailing update of the LU (2*M*N*K flops per launch); the panels'
    # pivot choice runs on a second stream, overlapped with these updates,
    # so only the first panel's appears under "panel"
    gemm_avg_ms = gemm_ms / gemm_cnt
    gemm_flops_per_launch = gemm_work / gemm_cnt
    achieved = gemm_flops_per_launch / (gemm_avg_ms * 1e-3)
    roofline = {"bound": "mfma", "kernel": "update_kernel<float,32> (FP32 LU trailing updates)",
                "achieved": round(achieved / 1e12, 3), "peak": PEAK_FP32_MFMA / 1e12,
                "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_MFMA, 4),
                "traffic": measured_traffic("update_kernel<float,32>", "tslu_n%d" % n),
                "launches": gemm_cnt,
                "avg_launch_ms": round(gemm_avg_ms, 4),
                "algorithmic_bytes_per_launch": round(gemm_bytes / gemm_cnt)}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_n, args.seed)
        log("cpu baseline: %s" % cpu)

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "GFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(tmp * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "dgesv n=%d kappa=%g nrhs=%d, A=U diag(s) V^T Haar" %
                   (n, args.kappa, nrhs), "n": n, "nrhs": nrhs, "kappa": args.kappa,
                   "factor_precision": "fp32" if ok_mp else "fp64"},
        "fp64_dgesv_gflops": round(g64, 2), "fp64_ms_per_step": round(t64 * 1e3, 3),
        "fp64_backward_error": be64,
        "mixed_dgesv_gflops": round(gmp, 2), "mixed_backward_error": bemp,
        "mixed_iterations": st[0].iterations, "mixed_fallback": st[0].fallback,
        "mixed_factorization": {1: "fp32 LU, tournament pivoting",
                                2: "fp32 partial-pivoting LU",
                                3: "fp64 LU (fallback)"}.get(st[0].method, "?"),
        "mixed_vs_fp64": round(gmp / g64, 3),
        "dgemm_n": gn, "dgemm_tflops": round(dgemm_tflops, 3),
        "dgemm_pct_peak": round(100 * dgemm_tflops * 1e12 / PEAK_FP64_MFMA, 2),
        "kernel_time_per_step": classes, "profiled_ms_per_step": round(tmp_prof * 1e3, 3),
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
