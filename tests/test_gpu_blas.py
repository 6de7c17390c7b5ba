"""GPU parity of the Level-3 BLAS (GEMM, TRSM, SYRK, iamax) through the C ABI
against the oracle, with ATLAS's own tester bounds: gemmtst's elementwise
ferr (bin/gemmtst.c:395) and l3blastst's TRSM residual < 1
(doc/TestTime.txt:333-345)."""
import itertools

import numpy as np
import pytest

import atlas_amd as X
import oracle as O

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(2024)


def _f(a):
    return np.asfortranarray(a)


def _gen(shape, dt):
    return _f(RNG.uniform(-0.5, 0.5, shape).astype(dt))


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("ta,tb", list(itertools.product([X.NoTrans, X.Trans], repeat=2)))
@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (17, 9, 5), (130, 129, 33), (255, 64, 257),
                                   (64, 300, 16), (3, 200, 1000), (257, 130, 530), (128, 128, 0)])
def test_gemm_shapes_vs_oracle(dt, ta, tb, M, N, K):
    A = _gen((M, K) if ta == X.NoTrans else (K, M), dt)
    B = _gen((K, N) if tb == X.NoTrans else (N, K), dt)
    C0 = _gen((M, N), dt)
    alpha, beta = 1.25, -0.5
    D = X.gemm(ta, tb, alpha, A, B, beta, C0.copy(order="F"))
    T = O.gemm(ta, tb, alpha, A, B, beta, C0.copy(order="F"), exact_order=True)
    bad, mx, fe = O.gemm_ferr_check(max(K, 1), alpha, beta, T, D)
    assert bad == 0, (mx, fe)


def test_gemm_lda_gaps_untouched_and_beta0_ignores_nan():
    M, N, K, ld = 70, 50, 40, 96
    A = O.matgen(M, K, ld, K * 1112)
    B = O.matgen(K, N, ld, N * 2238)
    Cm = O.matgen(M, N, ld, M * N)
    Cm[:M, 7] = np.nan                      # beta == 0: C must not be read
    D = Cm.copy(order="F")
    X.lib().cblas_dgemm(X.ColMajor, X.NoTrans, X.NoTrans, M, N, K, 1.0, A.ctypes.data, ld,
                        B.ctypes.data, ld, 0.0, D.ctypes.data, ld)
    assert np.all(D[M:, :] == -9e27)        # lda gap (bin/gemmtst.c:102) untouched
    T = Cm.copy(order="F")
    T[:M, :] = 0
    O.gemm(X.NoTrans, X.NoTrans, 1.0, A[:M, :K], B[:K, :N], 0.0, T[:M, :N], exact_order=True)
    assert O.gemm_ferr_check(K, 1.0, 0.0, T[:M, :N], D[:M, :N])[0] == 0


def test_gemm_config_2048_gemmtst():
    """Config C1: ColMajor N/N M=N=K=2048, alpha=1, beta=0.5, gemmtst seeds
    (A: K*1112, B: N*2238, C: M*N) and its elementwise bound."""
    n = 2048
    A = O.matgen(n, n, n, n * 1112)
    B = O.matgen(n, n, n, n * 2238)
    C0 = O.matgen(n, n, n, n * n)
    D = X.gemm(X.NoTrans, X.NoTrans, 1.0, A, B, 0.5, C0.copy(order="F"))
    O.set_threads(16)
    try:
        T = O.gemm(X.NoTrans, X.NoTrans, 1.0, A, B, 0.5, C0.copy(order="F"))
    finally:
        O.set_threads(1)
    bad, mx, fe = O.gemm_ferr_check(n, 1.0, 0.5, T, D)
    assert bad == 0, (mx, fe)
    assert mx < fe / 10


def test_gemm_integer_exact():
    """Small integers: every MFMA partial sum is exact, so the GPU must agree
    bit for bit with int64 arithmetic."""
    for dt in (np.float64, np.float32):
        M, N, K = 200, 150, 300
        Ai = RNG.integers(-4, 5, (M, K))
        Bi = RNG.integers(-4, 5, (K, N))
        Ci = RNG.integers(-50, 51, (M, N))
        D = X.gemm(X.NoTrans, X.Trans, 2.0, _f(Ai.astype(dt)), _f(Bi.T.astype(dt)), 3.0,
                   _f(Ci.astype(dt)))
        assert np.array_equal(D, (2 * (Ai @ Bi) + 3 * Ci).astype(dt))


def test_gemm_rowmajor():
    M, N, K = 33, 47, 29
    A = np.ascontiguousarray(RNG.uniform(-1, 1, (M, K)))
    B = np.ascontiguousarray(RNG.uniform(-1, 1, (K, N)))
    Cm = np.ascontiguousarray(RNG.uniform(-1, 1, (M, N)))
    want = 0.5 * A @ B + 2.0 * Cm
    X.lib().cblas_dgemm(X.RowMajor, X.NoTrans, X.NoTrans, M, N, K, 0.5, A.ctypes.data, K,
                        B.ctypes.data, N, 2.0, Cm.ctypes.data, N)
    assert np.allclose(Cm, want, atol=1e-13)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_trsm_all_variants_vs_oracle(dt):
    thr = 1.0
    for side, uplo, ta, diag in itertools.product([X.Left, X.Right], [X.Upper, X.Lower],
                                                  [X.NoTrans, X.Trans], [X.NonUnit, X.Unit]):
        for M, N in ((70, 33), (1, 5), (129, 130)):
            na = M if side == X.Left else N
            A = _gen((na, na), dt)
            A[np.diag_indices(na)] += np.asarray(2.0 if diag == X.NonUnit else 0.0, dt)
            if diag == X.Unit:   # keep unit-triangular systems well conditioned
                A *= np.asarray(0.1, dt)
            B0 = _gen((M, N), dt)
            Xg = X.trsm(side, uplo, ta, diag, 0.75, A, B0.copy(order="F"))
            r = O.trsm_resid(side, uplo, ta, diag, 0.75, A, Xg, B0)
            assert r < thr, (side, uplo, ta, diag, M, N, r)
            Xo = O.trsm(side, uplo, ta, diag, 0.75, A, B0.copy(order="F"))
            tol = 1e-10 if dt == np.float64 else 2e-3
            assert np.allclose(Xg, Xo, rtol=tol, atol=tol)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_trsm_many_rhs_vs_oracle(dt):
    """Left-side solves with 5..256 right-hand sides (the one-launch path of
    64x16 work items on the matrix cores): every uplo/trans/diag, ragged
    block and column-group counts, a gap between M and ldb that must stay
    untouched, and more work items than one wave of the chip holds."""
    for uplo, ta, diag in itertools.product([X.Upper, X.Lower], [X.NoTrans, X.Trans],
                                            [X.NonUnit, X.Unit]):
        for M, N in ((65, 5), (128, 16), (257, 17), (1000, 64), (2100, 256), (12000, 40)):
            if M == 12000 and (diag == X.Unit or dt == np.float32):
                continue
            A = _gen((M, M), dt) * np.asarray(1.0 / M, dt)
            A[np.diag_indices(M)] = np.asarray(1000.0, dt) if diag == X.Unit else \
                np.asarray(1.0, dt) + np.abs(RNG.uniform(-0.5, 0.5, M)).astype(dt)
            B0 = _gen((M, N), dt)
            Bbuf = _f(np.full((M + 3, N), 7.0, dt))
            Bbuf[:M] = B0
            Xg = X.trsm(X.Left, uplo, ta, diag, 0.75, A, Bbuf[:M])
            assert np.all(Bbuf[M:] == 7.0)
            r = O.trsm_resid(X.Left, uplo, ta, diag, 0.75, A, Xg, B0)
            # as for a few right-hand sides: two roundings per block solution
            assert r < 2.0, (uplo, ta, diag, M, N, r)
            if M <= 1000:
                Xo = O.trsm(X.Left, uplo, ta, diag, 0.75, A, B0.copy(order="F"))
                tol = 1e-10 if dt == np.float64 else 2e-3
                assert np.allclose(Xg, Xo, rtol=tol, atol=tol), (uplo, ta, diag, M, N)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_trsm_few_rhs_vs_oracle(dt):
    """Left-side solves with 1..4 right-hand sides (the one-launch
    flag-chained path): every uplo/trans/diag, ragged block counts, a gap
    between M and ldb that must stay untouched, and a block count above the
    number of workgroups one wave of the chip holds."""
    for uplo, ta, diag in itertools.product([X.Upper, X.Lower], [X.NoTrans, X.Trans],
                                            [X.NonUnit, X.Unit]):
        for M, N in ((1, 1), (63, 2), (64, 4), (65, 3), (700, 1), (20000, 2)):
            if M == 20000 and diag == X.Unit:
                continue
            # off-diagonal entries O(1/M) keep the systems well conditioned
            # at every size; a unit diagonal holds a value that must not be read
            A = _gen((M, M), dt) * np.asarray(1.0 / M, dt)
            A[np.diag_indices(M)] = np.asarray(1000.0, dt) if diag == X.Unit else \
                np.asarray(1.0, dt) + np.abs(RNG.uniform(-0.5, 0.5, M)).astype(dt)
            B0 = _gen((M, N), dt)
            Bbuf = _f(np.full((M + 3, N), 7.0, dt))
            Bbuf[:M] = B0
            Xg = X.trsm(X.Left, uplo, ta, diag, 0.75, A, Bbuf[:M])
            assert np.all(Bbuf[M:] == 7.0)
            r = O.trsm_resid(X.Left, uplo, ta, diag, 0.75, A, Xg, B0)
            # each block's solution is the inverted diagonal block times the
            # right-hand side: two roundings even for a 1 x 1 system, so the
            # tester's normalised residual can reach 2 where a division
            # (one rounding) stays below 1
            assert r < 2.0, (uplo, ta, diag, M, N, r)
            if M <= 700:
                Xo = O.trsm(X.Left, uplo, ta, diag, 0.75, A, B0.copy(order="F"))
                tol = 1e-10 if dt == np.float64 else 2e-3
                assert np.allclose(Xg, Xo, rtol=tol, atol=tol)
            # the many-RHS path (GEMM chain) on the same columns agrees
            Bw = _f(np.concatenate([B0, _gen((M, 5), dt)], axis=1))
            Xw = X.trsm(X.Left, uplo, ta, diag, 0.75, A, Bw)
            tol = 1e-9 if dt == np.float64 else 1e-2
            assert np.allclose(Xg, Xw[:, :N], rtol=tol, atol=tol)


def test_syrk_lower_upper_vs_oracle():
    N, K = 150, 70
    A = _gen((N, K), np.float64)
    for uplo in (X.Lower, X.Upper):
        C0 = _gen((N, N), np.float64)
        D = C0.copy(order="F")
        X.lib().cblas_dsyrk(X.ColMajor, uplo, X.NoTrans, N, K, -1.0, A.ctypes.data, N, 0.5,
                            D.ctypes.data, N)
        T = O.syrk(uplo, X.NoTrans, -1.0, A, 0.5, C0.copy(order="F"))
        tri = np.tril if uplo == X.Lower else np.triu
        other = np.triu if uplo == X.Lower else np.tril
        assert np.allclose(tri(D), tri(T), atol=1e-12)
        assert np.array_equal(other(D, 1 if uplo == X.Lower else -1),
                              other(C0, 1 if uplo == X.Lower else -1))   # untouched


def test_idamax_ties_nan():
    L = X.lib()
    for x in (np.array([1.0, -3.0, 3.0, 2.0]), np.array([0.0, 0.0]),
              np.array([np.nan, 1.0, -1.0]), RNG.uniform(-1, 1, 10000)):
        x = np.ascontiguousarray(x)
        assert L.cblas_idamax(x.size, x.ctypes.data, 1) == O.iamax(x)


def test_device_pointer_api_torch():
    import torch
    dev = torch.device("cuda:0")
    M, N, K = 300, 200, 100
    A = torch.rand(K, M, dtype=torch.float64, device=dev)     # column-major M x K = row-major K x M
    B = torch.rand(N, K, dtype=torch.float64, device=dev)
    Cm = torch.rand(N, M, dtype=torch.float64, device=dev)
    want = (A.T @ B.T) * 2.0 + Cm.T * 0.5
    X.use_torch_stream()
    X.lib().atlmx_dgemm(X.NoTrans, X.NoTrans, M, N, K, 2.0, X.dptr(A), M, X.dptr(B), K, 0.5,
                        X.dptr(Cm), M)
    torch.cuda.synchronize()
    assert torch.allclose(Cm.T, want, atol=1e-12)
    # the cblas ABI also accepts device pointers directly
    C2 = torch.zeros(N, M, dtype=torch.float64, device=dev)
    X.lib().cblas_dgemm(X.ColMajor, X.NoTrans, X.NoTrans, M, N, K, 1.0, X.dptr(A), M,
                        X.dptr(B), K, 0.0, X.dptr(C2), M)
    assert torch.allclose(C2.T, A.T @ B.T, atol=1e-12)


@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (130, 129, 64), (300, 200, 17), (256, 384, 100),
                                   (64, 1000, 130)])
def test_short_k_update_kernel_vs_numpy(M, N, K):
    """atlmx_k_supdate (C += alpha A B, the mixed solver's trailing update)
    against a float64 numpy product of the same float32 inputs."""
    import ctypes as C
    import torch
    X.use_torch_stream()
    rng = np.random.default_rng(M + N + K)
    lda, ldb, ldc = M + 3, K + 1, M + 5
    A = rng.uniform(-1, 1, (K, lda)).astype(np.float32)     # column-major M x K, ld lda
    B = rng.uniform(-1, 1, (N, ldb)).astype(np.float32)
    Cm = rng.uniform(-1, 1, (N, ldc)).astype(np.float32)
    dA, dB, dC = (torch.tensor(x, device="cuda") for x in (A, B, Cm))
    fn = X.lib().atlmx_k_supdate
    fn.restype = C.c_int
    for alpha in (-1.0, 0.5):
        dC.copy_(torch.tensor(Cm, device="cuda"))
        r = fn(M, N, K, C.c_float(alpha), X.dptr(dA), C.c_long(lda), X.dptr(dB), C.c_long(ldb),
               X.dptr(dC), C.c_long(ldc), None, None)
        assert r == 0
        torch.cuda.synchronize()
        got = dC.cpu().numpy()
        want = Cm.T[:M].astype(np.float64) + alpha * (A.T[:M].astype(np.float64) @
                                                      B.T[:K].astype(np.float64))
        assert np.abs(got.T[:M] - want).max() < 1e-5 * K
        assert np.array_equal(got.T[M:], Cm.T[M:])       # rows past M untouched


def _round16(x, lp, mode="rne"):
    """float32 -> bfloat16 (lp 1) or IEEE half (lp 2), back as float64; the
    bfloat16 rounding restated on the bit pattern (nearest even, or
    truncation for the contrast check)"""
    x = np.asarray(x, np.float32)
    if lp == 2:
        return x.astype(np.float16).astype(np.float64)
    u = x.view(np.uint32).astype(np.uint64)
    if mode == "rne":
        u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    else:
        u = u >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("lp", [1, 2])
@pytest.mark.parametrize("M,N,K", [(300, 200, 128), (257, 190, 70), (128, 128, 64)])
def test_low_precision_update_kernel_vs_numpy(lp, M, N, K):
    """atlmx_k_supdate_lp: C += alpha round(A) round(B) with the operands
    rounded to bfloat16 / half (nearest even) and FP32 accumulation, against
    the numpy restatement (the same rounding, float64 products); the
    kernel's result must be clearly closer to it than to the truncated
    rounding of the operands (which pins the rounding and the operand
    layout of the MFMA)."""
    import ctypes as C
    import torch
    X.use_torch_stream()
    rng = np.random.default_rng(7 * M + N + K + lp)
    lda, ldb, ldc = M + 3, K + 1, M + 5
    A = rng.uniform(-1, 1, (K, lda)).astype(np.float32)
    B = rng.uniform(-1, 1, (N, ldb)).astype(np.float32)
    Cm = rng.uniform(-1, 1, (N, ldc)).astype(np.float32)
    dA, dB, dC = (torch.tensor(x, device="cuda") for x in (A, B, Cm))
    fn = X.lib().atlmx_k_supdate_lp
    fn.restype = C.c_int
    r = fn(lp, M, N, K, C.c_float(-1.0), X.dptr(dA), C.c_long(lda), X.dptr(dB), C.c_long(ldb),
           X.dptr(dC), C.c_long(ldc), None, None)
    assert r == 0
    torch.cuda.synchronize()
    got = dC.cpu().numpy().T[:M].astype(np.float64)
    Am, Bm = A.T[:M], B.T[:K]
    want = Cm.T[:M] - _round16(Am, lp) @ _round16(Bm, lp)
    scale = 1.0 + np.abs(Am) @ np.abs(Bm)
    err = np.abs(got - want) / scale
    assert err.max() < 4 * K * 2.0 ** -24, err.max()
    if lp == 1:
        trunc = Cm.T[:M] - _round16(Am, 1, "trunc") @ _round16(Bm, 1, "trunc")
        assert (np.abs(got - trunc) / scale).max() > 10 * err.max()
    assert np.array_equal(dC.cpu().numpy().T[M:], Cm.T[M:])
