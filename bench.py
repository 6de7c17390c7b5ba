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
    from atlas_amd import synth
    return synth.haar_system(n, kappa, nrhs, seed, torch, X)


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


def _cpu_has(*flags):
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    have = set(line.split(":", 1)[1].split())
                    return all(fl in have for fl in flags)
    except OSError:
        pass
    return False


def cpu_baseline(n_cpu, seed, torch, X):
    """The oracle's restatement of ATLAS's CPU dgesv -- recursive getrfC
    (LU2 leaves, recursive TRSM, blocked GEMM with a 2-vector x 4-column
    register kernel) then getrs -- timed on this host's cores on a bounded
    sample of the workload: the same Haar construction at order n_cpu,
    kappa 1e6, one right-hand side.  The AVX2/FMA build
    (oracle/liboracle_fast.so) runs when the CPU has those instructions,
    else the portable one; threads = the CPUs this process may use, at most
    16 (the host's share for one GPU)."""
    import ctypes as C
    import numpy as np
    from atlas_amd import synth
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    fast = _cpu_has("avx2", "fma")
    lib = C.CDLL(os.path.join(ROOT, "oracle", "liboracle_fast.so" if fast else "liboracle.so"))
    lib.orc_set_threads.argtypes = [C.c_int]
    lib.orc_dgesv.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                              C.c_void_p, C.c_int]
    lib.orc_dgesv.restype = C.c_int
    lib.orc_set_threads(cores)
    Ad, bd = synth.haar_system(n_cpu, 1e6, 1, seed + 7, torch, X)
    A = np.asfortranarray(Ad.cpu().numpy().T)       # the column-major matrix itself
    B = np.asfortranarray(bd.cpu().numpy().T)
    del Ad, bd
    ipiv = np.zeros(n_cpu, dtype=np.int32)
    t0 = time.perf_counter()
    info = lib.orc_dgesv(102, n_cpu, 1, A.ctypes.data, n_cpu, ipiv.ctypes.data, B.ctypes.data,
                         n_cpu)
    t = time.perf_counter() - t0
    lib.orc_set_threads(1)
    return {"value": round(dgesv_flops(n_cpu, 1) / t / 1e9, 3), "unit": "GFLOP/s",
            "cores": cores, "kind": "port",
            "sample": "oracle dgesv (recursive ATL_getrfC restatement, recursive TRSM, "
                      "blocked GEMM; %s build) on the Haar system at n=%d, kappa 1e6, "
                      "1 RHS: %.1f s, info=%d" % ("AVX2/FMA" if fast else "portable", n_cpu, t,
                                                  info)}


def _views(big, offs, shapes):
    return [big[int(offs[i]):int(offs[i + 1])].view(*shapes[i]) for i in range(len(shapes))]


def bench_batched(count, seed, torch, X, log, reps=3):
    """The batched configurations, each as one graph-replayed persistent
    launch: `count` SPD systems of order U{8..512} (Cholesky factor + solve,
    one right-hand side), `count` general systems of order U{8..256} (LU with
    partial pivoting + solve), `count` GEMMs with M, N, K = 16 U{1..24}; and
    the differentiated configuration: 512 SPD and 512 general systems of
    order 256 with 8 right-hand sides, forward solve + backward (adjoints)
    through atlas_amd.autograd.  Inputs are synthetic and diagonally
    dominant (SPD through the lower triangle the Cholesky reads).  Every
    timed replay starts from the pristine inputs (copied back outside the
    timed region); plan creation (the autotuner) is not timed."""
    import numpy as np
    rng = np.random.default_rng(seed)
    res = {}

    def timed(plan, restore):
        ts = []
        for _ in range(reps):
            restore()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            plan.run()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        return min(ts)

    def matrices(sizes, dt, dominant):
        """column-major matrices in one buffer: U[-1,1), with n+1 on the
        diagonal when dominant"""
        sz = np.asarray(sizes, np.int64)
        offs = np.concatenate([[0], np.cumsum(sz * sz)])
        big = torch.rand(int(offs[-1]), dtype=dt, device="cuda") * 2 - 1
        if dominant:
            diag = np.concatenate([o + np.arange(n) * (n + 1) for o, n in zip(offs[:-1], sz)])
            dval = np.repeat(sz + 1.0, sz)
            big[torch.tensor(diag, device="cuda")] = torch.tensor(dval, dtype=dt, device="cuda")
        return big, offs

    for dt, tag in ((torch.float64, ""), (torch.float32, "_fp32")):
        for kind, lo, hi in (("posv", 8, 512), ("gesv", 8, 256)):
            sizes = rng.integers(lo, hi + 1, count)
            big, offs = matrices(sizes, dt, kind == "posv")
            pristine = big.clone()
            As = _views(big, offs, [(n, n) for n in sizes])
            boffs = np.concatenate([[0], np.cumsum(sizes)])
            bbig = torch.rand(int(boffs[-1]), dtype=dt, device="cuda") - 0.5
            bpr = bbig.clone()
            Bs = _views(bbig, boffs, [(1, n) for n in sizes])
            info = torch.zeros(count, dtype=torch.int32, device="cuda")
            if kind == "posv":
                plan = X.posv_batch(As, Bs, info)
                flops = float(np.sum(sizes ** 3 / 3.0 + 2.0 * sizes ** 2))
            else:
                piv = _views(torch.zeros(int(boffs[-1]), dtype=torch.int32, device="cuda"), boffs,
                             [(n,) for n in sizes])
                plan = X.gesv_batch(As, piv, Bs, info)
                flops = float(np.sum(2.0 * sizes ** 3 / 3.0 + 2.0 * sizes ** 2))

            def restore():
                big.copy_(pristine)
                bbig.copy_(bpr)
            with plan:
                ms = timed(plan, restore)
                nb, grid = plan.config()
            bad = int((info != 0).sum().item())
            key = kind + tag
            res[key] = {"count": count, "n_range": [lo, hi], "nrhs": 1, "ms": round(ms, 3),
                        "gflops": round(flops / (ms * 1e-3) / 1e9, 1), "panel_width": nb,
                        "grid": grid, "failed_instances": bad,
                        "inputs": "U[-1,1) + (n+1)I, lower triangle read (SPD)" if kind == "posv"
                        else "U[-1,1)"}
            log("batched %s x%d: %.2f ms, %.1f GFLOP/s" % (key, count, ms, res[key]["gflops"]))
            del big, pristine, bbig, bpr, As, Bs

        M, N, K = (16 * rng.integers(1, 25, count) for _ in range(3))
        offa = np.concatenate([[0], np.cumsum(M * K)])
        offb = np.concatenate([[0], np.cumsum(K * N)])
        offc = np.concatenate([[0], np.cumsum(M * N)])
        ga = torch.rand(int(offa[-1]), dtype=dt, device="cuda") - 0.5
        gb = torch.rand(int(offb[-1]), dtype=dt, device="cuda") - 0.5
        gc = torch.rand(int(offc[-1]), dtype=dt, device="cuda") - 0.5
        gcp = gc.clone()
        As = _views(ga, offa, [(k, m) for m, k in zip(M, K)])
        Bs = _views(gb, offb, [(n, k) for k, n in zip(K, N)])
        Cs = _views(gc, offc, [(n, m) for m, n in zip(M, N)])
        with X.gemm_batch(X.NoTrans, X.NoTrans, 1.0, As, Bs, 0.5, Cs) as plan:
            ms = timed(plan, lambda: gc.copy_(gcp))
        flops = float(np.sum(2.0 * M * N * K))
        res["gemm" + tag] = {"count": count, "mnk": "16*U{1..24}", "ms": round(ms, 3),
                             "gflops": round(flops / (ms * 1e-3) / 1e9, 1)}
        log("batched gemm%s x%d: %.2f ms, %.1f GFLOP/s" % (tag, count, ms,
                                                          res["gemm" + tag]["gflops"]))
        del ga, gb, gc, gcp, As, Bs, Cs

    from atlas_amd.autograd import CholeskySolve, LUSolve
    b, n, r = 512, 256, 8
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    Bm = torch.randn(b, n, n, dtype=torch.float64, device="cuda", generator=g)
    Aspd = Bm @ Bm.transpose(1, 2).contiguous() / n + torch.eye(n, dtype=torch.float64, device="cuda")
    Agen = torch.rand(b, n, n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1 + \
        4 * torch.eye(n, dtype=torch.float64, device="cuda")
    Bv = torch.randn(b, n, r, dtype=torch.float64, device="cuda", generator=g)
    W = torch.randn(b, n, r, dtype=torch.float64, device="cuda", generator=g)
    for name, f, A0 in (("adjoint_posv", CholeskySolve.apply, Aspd), ("adjoint_gesv", LUSolve.apply, Agen)):
        ts = []
        for _ in range(reps + 1):
            A1 = A0.clone().requires_grad_(True)
            B1 = Bv.clone().requires_grad_(True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            (f(A1, B1) * W).sum().backward()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        res[name] = {"batch": b, "n": n, "nrhs": r, "ms_forward_backward": round(min(ts[1:]) * 1e3, 3)}
        log("%s %dx%d: %.2f ms forward+backward" % (name, b, n, res[name]["ms_forward_backward"]))
    return res


def bench_lapack(n, seed, torch, X, log, reps=3):
    """The plain FP64 factorizations at order n: getrf (partial pivoting,
    A ~ U[-1,1)) with getrs on 1 and 64 right-hand sides, potrf (Lower) with
    potrs on 1 and 64, A = B B^T / n + I with B standard normal.  Device
    operands; best of reps, each on a fresh copy of A."""
    L = X.lib()
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    A0 = torch.rand(n, n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    Bm = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
    S0 = torch.empty(n, n, dtype=torch.float64, device="cuda")
    L.atlmx_dgemm(X.NoTrans, X.Trans, n, n, n, 1.0 / n, X.dptr(Bm), n, X.dptr(Bm), n, 0.0,
                  X.dptr(S0), n)
    S0 += torch.eye(n, dtype=torch.float64, device="cuda")
    del Bm
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    R = {1: torch.rand(1, n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1,
         64: torch.rand(64, n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1}
    Aw = torch.empty_like(A0)

    def best(fn, prep):
        ts = []
        for _ in range(reps):
            prep()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return min(ts)

    out = {"n": n}
    t = best(lambda: L.atlmx_dgetrf(n, n, X.dptr(Aw), n, X.dptr(ipiv)), lambda: Aw.copy_(A0))
    fl = 2.0 * n ** 3 / 3 - n ** 2 / 2.0 - n / 6.0
    out["getrf"] = {"ms": round(t * 1e3, 2), "gflops": round(fl / t / 1e9, 1)}
    # one more factorization with HIP events around the panel launches
    # (getf2_panel_lds_kernel: bytes = the panel loaded and stored once) and
    # the GEMM launches (trailing updates and the panel recursion's, flops)
    Aw.copy_(A0)
    torch.cuda.synchronize()
    X.prof_enable((1 << X.PROF_PANEL) | (1 << X.PROF_GEMM))
    L.atlmx_dgetrf(n, n, X.dptr(Aw), n, X.dptr(ipiv))
    pms, pcnt, pwork = X.prof_query(X.PROF_PANEL)
    gms, gcnt, gwork = X.prof_query(X.PROF_GEMM)
    X.prof_enable(0)
    if pcnt:
        bw = pwork / pcnt / (pms / pcnt * 1e-3)
        out["roofline_panel"] = {
            "bound": "hbm", "kernel": "getf2_panel_lds_kernel (FP64 LU panel)",
            "achieved": round(bw / 1e9, 2), "peak": PEAK_HBM / 1e9, "unit": "GB/s",
            "frac": round(bw / PEAK_HBM, 5), "launches": pcnt,
            "avg_launch_ms": round(pms / pcnt, 4), "ms_total": round(pms, 2),
            "algorithmic_bytes_per_launch": round(pwork / pcnt)}
    if gcnt:
        fr = gwork / (gms * 1e-3)
        out["roofline_gemm"] = {
            "bound": "mfma", "kernel": "gemm_kernel<double> (FP64 LU updates)",
            "achieved": round(fr / 1e12, 3), "peak": PEAK_FP64_MFMA / 1e12, "unit": "TFLOP/s",
            "frac": round(fr / PEAK_FP64_MFMA, 4), "launches": gcnt,
            "ms_total": round(gms, 2)}
    for nr, rhs in R.items():
        Xw = torch.empty_like(rhs)
        t = best(lambda: L.atlmx_dgetrs(X.NoTrans, n, nr, X.dptr(Aw), n, X.dptr(ipiv), X.dptr(Xw), n),
                 lambda: Xw.copy_(rhs))
        out["getrs_%d" % nr] = {"ms": round(t * 1e3, 3), "gflops": round(2.0 * n * n * nr / t / 1e9, 1)}
    t = best(lambda: L.atlmx_dpotrf(X.Lower, n, X.dptr(Aw), n), lambda: Aw.copy_(S0))
    fl = n ** 3 / 3.0 + n ** 2 / 2.0 + n / 6.0
    out["potrf"] = {"ms": round(t * 1e3, 2), "gflops": round(fl / t / 1e9, 1)}
    for nr, rhs in R.items():
        Xw = torch.empty_like(rhs)
        t = best(lambda: L.atlmx_dpotrs(X.Lower, n, nr, X.dptr(Aw), n, X.dptr(Xw), n),
                 lambda: Xw.copy_(rhs))
        out["potrs_%d" % nr] = {"ms": round(t * 1e3, 3), "gflops": round(2.0 * n * n * nr / t / 1e9, 1)}
    log("lapack n=%d: %s" % (n, json.dumps(out)))
    return out


def bench_complex(n, seed, torch, X, log, reps=3):
    """The complex routines at order n on device operands: zgemm and cgemm
    (NoTrans x ConjTrans, alpha = 1 - 0.5i, beta = 0.25i; 8 n^3 flops),
    zgetrf (A ~ U[-1,1) + i U[-1,1); 8/3 n^3 flops) with zgetrs on 64
    right-hand sides, and zpotrf (Lower, A = B B^H / n + I; 4/3 n^3 flops)
    with zpotrs on 64, then zherk, zher2k, zhemm and ztrmm.  Best of reps."""
    import numpy as np
    L = X.lib()
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    out = {"n": n}

    def best(fn, prep):
        ts = []
        for _ in range(reps):
            prep()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return min(ts)

    for p, dt, peak in (("z", torch.complex128, PEAK_FP64_MFMA),
                        ("c", torch.complex64, PEAK_FP32_MFMA)):
        rt = np.float64 if p == "z" else np.float32
        A = torch.randn(n, n, dtype=dt, device="cuda", generator=g)
        B = torch.randn(n, n, dtype=dt, device="cuda", generator=g)
        Cm = torch.randn(n, n, dtype=dt, device="cuda", generator=g)
        al = np.array([1.0, -0.5], dtype=rt)
        be = np.array([0.0, 0.25], dtype=rt)
        f = getattr(L, "cblas_%sgemm" % p)
        t = best(lambda: f(X.ColMajor, X.NoTrans, X.ConjTrans, n, n, n, al.ctypes.data,
                           A.data_ptr(), n, B.data_ptr(), n, be.ctypes.data, Cm.data_ptr(), n),
                 lambda: None)
        fl = 8.0 * n ** 3
        out["%sgemm" % p] = {"ms": round(t * 1e3, 2), "tflops": round(fl / t / 1e12, 2),
                             "frac_of_real_mfma_peak": round(fl / t / peak, 3)}
        del B, Cm
    A0 = torch.rand(n, n, dtype=torch.complex128, device="cuda", generator=g) * 2 - (1 + 1j)
    Aw = torch.empty_like(A0)
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    t = best(lambda: L.clapack_zgetrf(X.ColMajor, n, n, Aw.data_ptr(), n, ipiv.data_ptr()),
             lambda: Aw.copy_(A0))
    fl = 8.0 * n ** 3 / 3
    out["zgetrf"] = {"ms": round(t * 1e3, 2), "tflops": round(fl / t / 1e12, 3)}
    rhs = torch.rand(64, n, dtype=torch.complex128, device="cuda", generator=g)
    Xw = torch.empty_like(rhs)
    t = best(lambda: L.clapack_zgetrs(X.ColMajor, X.NoTrans, n, 64, Aw.data_ptr(), n,
                                      ipiv.data_ptr(), Xw.data_ptr(), n),
             lambda: Xw.copy_(rhs))
    out["zgetrs_64"] = {"ms": round(t * 1e3, 2), "tflops": round(8.0 * n * n * 64 / t / 1e12, 3)}
    # Cholesky of the Hermitian positive definite B B^H / n + I (4/3 n^3 flops)
    Bm = torch.randn(n, n, dtype=torch.complex128, device="cuda", generator=g)
    S0 = torch.empty_like(Bm)
    one = np.array([1.0 / n, 0.0])
    zero = np.array([0.0, 0.0])
    L.cblas_zgemm(X.ColMajor, X.NoTrans, X.ConjTrans, n, n, n, one.ctypes.data, Bm.data_ptr(), n,
                  Bm.data_ptr(), n, zero.ctypes.data, S0.data_ptr(), n)
    S0 += torch.eye(n, dtype=torch.complex128, device="cuda")
    del Bm
    t = best(lambda: L.clapack_zpotrf(X.ColMajor, X.Lower, n, Aw.data_ptr(), n),
             lambda: Aw.copy_(S0))
    out["zpotrf"] = {"ms": round(t * 1e3, 2), "tflops": round(4.0 * n ** 3 / 3 / t / 1e12, 3)}
    t = best(lambda: L.clapack_zpotrs(X.ColMajor, X.Lower, n, 64, Aw.data_ptr(), n,
                                      Xw.data_ptr(), n),
             lambda: Xw.copy_(rhs))
    out["zpotrs_64"] = {"ms": round(t * 1e3, 2), "tflops": round(8.0 * n * n * 64 / t / 1e12, 3)}
    # the other complex Level-3 routines on the same operands: zherk
    # (Lower, NoTrans, K = n: 4 n^3 flops), zher2k (8 n^3), zhemm (Left,
    # Lower: 8 n^3), ztrmm (Left, Lower, NoTrans, NonUnit: 4 n^3)
    W = torch.empty_like(S0)
    al = np.array([0.5, -0.25])
    be = np.array([0.25, 0.0])
    runs = [
        ("zherk", 4.0, lambda: L.cblas_zherk(X.ColMajor, X.Lower, X.NoTrans, n, n, 0.5,
                                             S0.data_ptr(), n, 0.25, W.data_ptr(), n)),
        ("zher2k", 8.0, lambda: L.cblas_zher2k(X.ColMajor, X.Lower, X.NoTrans, n, n,
                                               al.ctypes.data, S0.data_ptr(), n, Aw.data_ptr(), n,
                                               0.25, W.data_ptr(), n)),
        ("zhemm", 8.0, lambda: L.cblas_zhemm(X.ColMajor, X.Left, X.Lower, n, n, al.ctypes.data,
                                             S0.data_ptr(), n, Aw.data_ptr(), n, be.ctypes.data,
                                             W.data_ptr(), n)),
        ("ztrmm", 4.0, lambda: L.cblas_ztrmm(X.ColMajor, X.Left, X.Lower, X.NoTrans, X.NonUnit,
                                             n, n, al.ctypes.data, Aw.data_ptr(), n,
                                             W.data_ptr(), n)),
    ]
    for name, c3, fn in runs:
        t = best(fn, lambda: W.copy_(S0))
        out[name] = {"ms": round(t * 1e3, 2), "tflops": round(c3 * n ** 3 / t / 1e12, 2),
                     "frac_of_real_mfma_peak": round(c3 * n ** 3 / t / PEAK_FP64_MFMA, 3)}
    log("complex n=%d: %s" % (n, json.dumps(out)))
    return out


def bench_sweep(n, seed, A6, b6, torch, X, log):
    """The dense-solve configuration's other cases, one solve each: kappa
    1e2 and 1e4 (and the headline's 1e6) with the FP32 tournament LU and
    with half-precision trailing updates (bfloat16 at kappa 1e2 only: beyond
    it its factors stall and the solver falls back), and 16 right-hand sides
    at kappa 1e6.  Backward error, forward error against the FP64 LU's
    solution of the same system, refinement and GMRES steps, and the method
    that produced the result."""
    L = X.lib()
    out = []
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")

    def fp64_solution(A, b):
        Aw, x = A.clone(), b.clone()
        L.atlmx_dgetrf(n, n, X.dptr(Aw), n, X.dptr(ipiv))
        L.atlmx_dgetrs(X.NoTrans, n, b.shape[0], X.dptr(Aw), n, X.dptr(ipiv), X.dptr(x), n)
        torch.cuda.synchronize()
        return x

    def one(A, b, x64, kappa, method, label):
        x = torch.empty_like(b)
        X.set_mixed_method(method)
        try:
            torch.cuda.synchronize()              # workspaces exist from the headline runs
            t0 = time.perf_counter()
            info, st = X.dsgesv_dev(A, b, x)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) * 1e3
        finally:
            X.set_mixed_method(X.MIXED_AUTO)
        nr = b.shape[0]
        be = L.atlmx_dbackward_error(n, nr, X.dptr(A), n, X.dptr(b), n, X.dptr(x), n)
        fe = float(((x - x64).abs().max() / x64.abs().max()).item())
        used = {1: "fp32", 2: "fp32 gepp", 3: "fp64", 5: "bf16 updates", 6: "fp16 updates"}
        row = {"kappa": kappa, "nrhs": nr, "requested": label, "used": used.get(st.method, "?"),
               "ms": round(ms, 2), "backward_error": be, "forward_error_vs_fp64": fe,
               "iterations": st.iterations, "gmres": st.gmres,
               "gmres_steps": st.inner_iterations, "fallback_to_fp64": st.fallback}
        log("sweep kappa=%g nrhs=%d %s -> %s: %.1f ms, be %.2e, fe %.2e, iters %d, gmres %d" %
            (kappa, nr, label, row["used"], ms, be, fe, st.iterations, st.inner_iterations))
        out.append(row)

    for kappa in (1e2, 1e4):
        A, b = make_system(n, kappa, 1, seed + int(kappa), torch, X)
        x64 = fp64_solution(A, b)
        one(A, b, x64, kappa, X.MIXED_AUTO, "fp32")
        one(A, b, x64, kappa, X.MIXED_FP16, "fp16 updates")
        if kappa == 1e2:
            one(A, b, x64, kappa, X.MIXED_BF16, "bf16 updates")
        del A, b, x64
    x64 = fp64_solution(A6, b6)
    one(A6, b6, x64, 1e6, X.MIXED_FP16, "fp16 updates")
    g = torch.Generator(device="cuda")
    g.manual_seed(seed + 16)
    b16 = torch.rand(16, n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    x64 = fp64_solution(A6, b16)
    one(A6, b16, x64, 1e6, X.MIXED_AUTO, "fp32")
    return out


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
    ap.add_argument("--cpu-n", type=int, default=10240)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched and adjoint timings")
    ap.add_argument("--batch-count", type=int, default=20000)
    ap.add_argument("--no-sweep", action="store_true",
                    help="skip the other kappa / precision / 16-RHS cases")
    ap.add_argument("--no-complex", action="store_true",
                    help="skip the complex GEMM / LU timings")
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
    # the headline stops refinement with LAPACK dsgesv's test,
    # ||r||_inf <= ||x||_inf ||A||_inf 2^-53 sqrt(n) (stricter than n eps64)
    X.set_refine_stop(True)

    def mp_step():
        info[0], st[0] = X.dsgesv_dev(A, b, xmp)

    tmp = time_steps(mp_step, args.steps, args.warmup, torch)
    # the same steps again with a HIP event pair around every library launch
    # (on the library's stream): per-kernel-class times and the rooflines
    X.prof_enable(-1)
    tmp_prof = time_steps(mp_step, args.steps, 0, torch)
    classes, raw = {}, {}
    for cls, name in enumerate(["gemm", "panel", "trsm_inv", "laswp", "potf2", "misc", "trsv",
                                "update"]):
        ms, cnt, work, nbytes = X.prof_query(cls, with_bytes=True)
        raw[name] = (ms, cnt, work, nbytes)
        if cnt:
            classes[name] = {"ms_per_step": round(ms / args.steps, 3),
                             "launches_per_step": cnt // args.steps}
    X.prof_enable(0)
    X.set_refine_stop(False)
    bemp = L.atlmx_dbackward_error(n, nrhs, X.dptr(A), n, X.dptr(b), n, X.dptr(xmp), n)
    # checked with torch (rocBLAS), independently of the library: LAPACK's
    # stopping ratio ||r|| / (||x|| ||A|| 2^-53 sqrt(n)) (pass <= 1) and the
    # forward error against the FP64 LU's solution
    Ainf = float(A.abs().sum(dim=0).max().item())    # rows of the matrix = columns of A's storage
    r = b.T - A.T @ xmp.T
    lapack_ratio = float((r.abs().amax(dim=0) /
                          (xmp.T.abs().amax(dim=0) * Ainf * 2.0 ** -53 * n ** 0.5)).max().item())
    fwd = float(((xmp - x64).abs().amax(dim=1) / x64.abs().amax(dim=1)).max().item())
    del r
    gmp = dgesv_flops(n, nrhs) / tmp / 1e9
    log("mixed dsgesv: %.1f ms, %.1f GFLOP/s, iters %d, converged %d, fallback %d, "
        "backward error %.3g, LAPACK stop ratio %.3g, forward error vs FP64 %.3g" %
        (tmp * 1e3, gmp, st[0].iterations, st[0].converged, st[0].fallback, bemp,
         lapack_ratio, fwd))
    ok_mp = (bemp <= n * EPS64) and not st[0].fallback
    value = gmp if ok_mp else g64

    def kernel_roof(name, kernel, bound, peak, unit, traffic_key, per_step_ms):
        ms, cnt, work, nbytes = raw[name]
        if not cnt:
            return None
        avg = ms / cnt
        if bound == "hbm":
            per_launch = work / cnt                      # algorithmic bytes
            achieved = per_launch / (avg * 1e-3) / 1e9
        else:
            per_launch = work / cnt                      # flops
            achieved = per_launch / (avg * 1e-3) / 1e12
        return {"bound": bound, "kernel": kernel, "achieved": round(achieved, 3),
                "peak": peak, "unit": unit, "frac": round(achieved / peak, 5),
                "traffic": measured_traffic(kernel.split(" ")[0], traffic_key),
                "launches": cnt, "avg_launch_ms": round(avg, 4),
                ("algorithmic_bytes_per_launch" if bound == "hbm" else "flops_per_launch"):
                    round(per_launch),
                "ms_per_step": round(per_step_ms, 3)}

    # The kernel that dominates the headline step is the tournament pivot
    # choice (tslu_select_kernel, class "panel"): its algorithmic bytes are
    # the panel read once, (n - k) x 64 FP32 values per launch; it is bound
    # by the latency of its 64 dependent elimination steps per tree level,
    # so its HBM fraction is tiny.  The FP32 MFMA trailing update
    # (update_kernel, class "update") is reported beside it.
    roofline = kernel_roof("panel", "tslu_select_kernel (tournament pivoting, FP32 LU panel)",
                           "hbm", PEAK_HBM / 1e9, "GB/s", "tslu_n%d" % n,
                           raw["panel"][0] / args.steps)
    roof_update = kernel_roof("update", "update_kernel<float,32> (FP32 LU trailing updates)",
                              "mfma", PEAK_FP32_MFMA / 1e12, "TFLOP/s", "tslu_n%d" % n,
                              raw["update"][0] / args.steps)
    if raw["update"][1]:
        roof_update["algorithmic_bytes_per_launch"] = round(raw["update"][3] / raw["update"][1])

    lapack = bench_lapack(8192, args.seed + 8192, torch, X, log)

    sweep = None
    if not args.no_sweep:
        sweep = bench_sweep(n, args.seed, A, b, torch, X, log)

    cplx = None
    if not args.no_complex:
        cplx = bench_complex(4096, args.seed + 4096, torch, X, log)

    batched = None
    if not args.no_batched:
        batched = bench_batched(args.batch_count, args.seed, torch, X, log)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_n, args.seed, torch, X)
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
        "mixed_refine_stop": "LAPACK dsgesv: ||r|| <= ||x|| ||A|| 2^-53 sqrt(n)",
        "mixed_lapack_stop_ratio": lapack_ratio, "mixed_forward_error_vs_fp64": fwd,
        "mixed_iterations": st[0].iterations, "mixed_fallback": st[0].fallback,
        "mixed_factorization": {1: "fp32 LU, tournament pivoting",
                                2: "fp32 partial-pivoting LU",
                                3: "fp64 LU (fallback)"}.get(st[0].method, "?"),
        "mixed_vs_fp64": round(gmp / g64, 3),
        "dgemm_n": gn, "dgemm_tflops": round(dgemm_tflops, 3),
        "dgemm_pct_peak": round(100 * dgemm_tflops * 1e12 / PEAK_FP64_MFMA, 2),
        "kernel_time_per_step": classes, "profiled_ms_per_step": round(tmp_prof * 1e3, 3),
        "roofline": roofline, "roofline_update": roof_update, "cpu_baseline": cpu, "lapack_n8192": lapack, "sweep": sweep,
        "complex_n4096": cplx,
        "batched": batched,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
