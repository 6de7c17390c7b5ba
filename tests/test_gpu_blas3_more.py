"""GPU parity of the rest of the Level-3 BLAS -- cblas_?trmm, cblas_?symm,
cblas_?syr2k -- through the C ABI, against the oracle's reference-order
restatements (ATL_dreftrmm*.c, ATL_drefsymm*.c, ATL_drefsyr2k*.c, pinned bit
for bit to ATLAS's reference BLAS) under gemmtst's elementwise bound
(bin/gemmtst.c:395) with K the length of each dot product; the triangle a
routine must not touch is compared exactly."""
import itertools

import numpy as np
import pytest

import atlas_amd as X
import oracle as O

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(5150)


def _f(a):
    return np.asfortranarray(a)


def _gen(shape, dt):
    return _f(RNG.uniform(-0.5, 0.5, shape).astype(dt))


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("side,uplo,ta,diag", list(itertools.product(
    [X.Left, X.Right], [X.Upper, X.Lower], [X.NoTrans, X.Trans], [X.NonUnit, X.Unit])))
@pytest.mark.parametrize("M,N", [(190, 77), (60, 300)])
def test_trmm_all_variants_vs_oracle(dt, side, uplo, ta, diag, M, N):
    """M = 190 (Left) and N = 300 (Right) span several 128-wide GEMM tiles,
    so each tile position's K band is exercised"""
    na = M if side == X.Left else N
    A = _gen((na, na), dt)
    B0 = _gen((M, N), dt)
    D = X.trmm(side, uplo, ta, diag, 1.5, A, B0.copy(order="F"))
    T = O.trmm(side, uplo, ta, diag, 1.5, A, B0.copy(order="F"))
    bad, mx, fe = O.gemm_ferr_check(na, 1.5, 0.0, T, D)
    assert bad == 0, (mx, fe)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_trmm_alpha_zero_rowmajor_and_gaps(dt):
    M, N, ld = 70, 40, 96
    A = _gen((ld, 70), dt)
    B = _f(np.full((ld, N), 7.0, dt))
    B[:M] = _gen((M, N), dt)
    B[:M, 3] = np.nan                          # alpha == 0: B = 0, not 0 * NaN
    getattr(X.lib(), "cblas_%strmm" % ("d" if dt == np.float64 else "s"))(
        X.ColMajor, X.Left, X.Lower, X.NoTrans, X.NonUnit, M, N, 0.0, A.ctypes.data, ld,
        B.ctypes.data, ld)
    assert np.all(B[:M] == 0) and np.all(B[M:] == 7.0)
    # RowMajor: the same numbers held transposed
    Ar = np.ascontiguousarray(_gen((N, N), dt))
    Br = np.ascontiguousarray(_gen((M, N), dt))
    D = Br.copy()
    getattr(X.lib(), "cblas_%strmm" % ("d" if dt == np.float64 else "s"))(
        X.RowMajor, X.Right, X.Upper, X.Trans, X.Unit, M, N, 2.0, Ar.ctypes.data, N,
        D.ctypes.data, N)
    U = np.triu(Ar.astype(np.float64), 1) + np.eye(N)
    want = 2.0 * Br.astype(np.float64) @ U.T
    tol = 1e-12 if dt == np.float64 else 1e-4
    assert np.abs(D - want).max() <= tol * N


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("side,uplo", list(itertools.product([X.Left, X.Right],
                                                             [X.Upper, X.Lower])))
@pytest.mark.parametrize("alpha,beta", [(1.25, -0.5), (-1.0, 0.0), (0.0, 2.0)])
def test_symm_vs_oracle(dt, side, uplo, alpha, beta):
    M, N = 150, 93
    na = M if side == X.Left else N
    A = _gen((na, na), dt)
    B = _gen((M, N), dt)
    C0 = _gen((M, N), dt)
    D = X.symm(side, uplo, alpha, A, B, beta, C0.copy(order="F"))
    T = O.symm(side, uplo, alpha, A, B, beta, C0.copy(order="F"))
    bad, mx, fe = O.gemm_ferr_check(na, alpha, beta, T, D)
    assert bad == 0, (mx, fe)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("uplo,trans", list(itertools.product([X.Upper, X.Lower],
                                                              [X.NoTrans, X.Trans])))
@pytest.mark.parametrize("alpha,beta", [(0.75, 0.5), (-1.0, 0.0), (0.0, 3.0)])
def test_syr2k_vs_oracle(dt, uplo, trans, alpha, beta):
    N, K = 170, 65
    shp = (N, K) if trans == X.NoTrans else (K, N)
    A, B = _gen(shp, dt), _gen(shp, dt)
    C0 = _gen((N, N), dt)
    D = X.syr2k(uplo, trans, alpha, A, B, beta, C0.copy(order="F"))
    T = O.syr2k(uplo, trans, alpha, A, B, beta, C0.copy(order="F"))
    tri = np.tril(np.ones((N, N), bool)) if uplo == X.Lower else np.triu(np.ones((N, N), bool))
    assert np.array_equal(D[~tri], C0[~tri])            # other triangle untouched
    Dm = np.where(tri, D, T)
    bad, mx, fe = O.gemm_ferr_check(2 * K, alpha, beta, T, _f(Dm))
    assert bad == 0, (mx, fe)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_syrk_trans_rowmajor_and_single(dt):
    """cblas_?syrk with Trans, and RowMajor, against the oracle (the single
    precision twin included)."""
    N, K = 140, 90
    pre = "d" if dt == np.float64 else "s"
    for uplo in (X.Lower, X.Upper):
        A = _gen((K, N), dt)
        C0 = _gen((N, N), dt)
        D = C0.copy(order="F")
        getattr(X.lib(), "cblas_%ssyrk" % pre)(X.ColMajor, uplo, X.Trans, N, K, -1.0,
                                               A.ctypes.data, K, 0.5, D.ctypes.data, N)
        T = O.syrk(uplo, X.Trans, -1.0, A, 0.5, C0.copy(order="F"))
        tri = np.tril(np.ones((N, N), bool)) if uplo == X.Lower else np.triu(np.ones((N, N), bool))
        assert np.array_equal(D[~tri], C0[~tri])
        assert O.gemm_ferr_check(K, -1.0, 0.5, T, _f(np.where(tri, D, T)))[0] == 0
        # RowMajor: C^T held row-major; Upper there is Lower of the column-major view
        Ar = np.ascontiguousarray(_gen((N, K), dt))
        Cr = np.ascontiguousarray(_gen((N, N), dt))
        Dr = Cr.copy()
        getattr(X.lib(), "cblas_%ssyrk" % pre)(X.RowMajor, uplo, X.NoTrans, N, K, 1.0,
                                               Ar.ctypes.data, K, 0.0, Dr.ctypes.data, N)
        full = Ar.astype(np.float64) @ Ar.T.astype(np.float64)
        trr = np.triu(np.ones((N, N), bool)) if uplo == X.Upper else np.tril(np.ones((N, N), bool))
        tol = 1e-12 if dt == np.float64 else 2e-5
        assert np.abs(Dr[trr] - full[trr]).max() <= tol * K
        assert np.array_equal(Dr[~trr], Cr[~trr])
