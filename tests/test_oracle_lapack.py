"""Pins the oracle's LAPACK restatement (recursive getrfC, potrfL/U, getrs,
potrs) and the tester residuals (lutst, llttst, slvtst, gemmtst).

Independent checks: LAPACK dgetrf/dpotrf through scipy/numpy, permutation and
diagonally-dominant inputs with known pivots, integer Cholesky factors that
are exact, and planted mistakes (made in the oracle's output, on the CPU)
that each residual must flag.
"""
import json
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
RNG = np.random.default_rng(77)


def _f(a):
    return np.asfortranarray(a)


def test_lu_pivots_match_lapack_golden():
    g = json.load(open(os.path.join(GOLD, "lu_pivots.json")))
    for n, d in g.items():
        n = int(n)
        A = O.gegen(n, n, d["lda"], seed=d["seed"])
        info, ipiv = O.getrf(A)
        assert info == 0
        c = d["checked_columns"]
        assert c >= n // 2
        assert ipiv[:c].tolist() == d["ipiv"], n


@pytest.mark.parametrize("l1,lcm", [(4096, 4), (16, 1), (1, 1), (100000, 8)])
def test_lu_recursion_params_do_not_change_pivots(l1, lcm):
    """Any cut-off must give partial pivoting (compare with scipy on a
    non-near-tie input)."""
    import scipy.linalg as sla
    O.set_recursion(l1, lcm)
    try:
        n = 150
        A0 = O.gegen(n, n)
        A = A0.copy(order="F")
        gaps = np.zeros(n)
        info, ipiv = O.getrf(A, gaps=gaps)
        assert gaps.min() > 1e-9
        _, piv = sla.lu_factor(A0)
        assert np.array_equal(ipiv, piv)
        assert O.lu_resid(A, ipiv, A0) < 100
    finally:
        O.set_recursion()


@pytest.mark.parametrize("M,N,lda", [(300, 300, 300), (500, 400, 512), (200, 350, 200),
                                     (97, 97, 100), (1, 50, 1), (50, 1, 60), (2, 40, 2)])
def test_lutst_residual(M, N, lda):
    r, npiv = O.lutest(M, N, lda)
    assert 0 <= r < 100         # bin/lutst.c:398 threshold


def test_lutst_float():
    r, _ = O.lutest(256, 256, 256, dtype=np.float32)
    assert 0 <= r < 100


def test_lu_known_pivots_permutation_and_diag_dominant():
    n = 64
    perm = RNG.permutation(n)
    P = np.zeros((n, n))
    P[np.arange(n), perm] = 1.0
    A = _f(P * (1.0 + 0.001 * RNG.uniform(size=(n, n))))   # one nonzero per column
    info, ipiv = O.getrf(A)
    # column j's only nonzero sits where perm maps it; after the swaps of
    # earlier columns the pivot rows are the ones LAPACK would pick:
    import scipy.linalg as sla
    _, piv = sla.lu_factor(P * (1.0 + 0.001 * RNG.uniform(size=(n, n))) * 0 + P)
    assert np.array_equal(ipiv, piv)
    D = _f(RNG.uniform(-1, 1, (n, n)) + n * np.eye(n))
    info, ipiv = O.getrf(D)
    assert np.array_equal(ipiv, np.arange(n))                # findnpvt == 0


def test_lu_singular_info_and_completion():
    n = 40
    A0 = _f(RNG.uniform(-1, 1, (n, n)))
    A0[:, 17] = 0.0                                          # exactly zero column
    A = A0.copy(order="F")
    info, ipiv = O.getrf(A)
    assert info == 18                                        # 1-based first zero pivot
    assert A[17, 17] == 0.0
    assert np.isfinite(A).all()                              # factorization completed


def test_lu_planted_mistakes_fail_residual():
    n = 120
    A0 = O.gegen(n, n)
    A = A0.copy(order="F")
    _, ipiv = O.getrf(A)
    assert O.lu_resid(A, ipiv, A0) < 100
    bad = A.copy(order="F")
    bad[70, 30] += 1e-6                                      # a wrong multiplier
    assert O.lu_resid(bad, ipiv, A0) > 100
    bp = ipiv.copy()
    j = int(np.nonzero(bp[:-1] != np.arange(n - 1))[0][0])
    bp[j] = j                                                # a dropped interchange
    assert O.lu_resid(A, bp, A0) > 100


def test_getrs_vs_numpy_and_rowmajor():
    n, nrhs = 90, 5
    A0 = O.gegen(n, n)
    B0 = _f(RNG.uniform(-1, 1, (n, nrhs)))
    A = A0.copy(order="F")
    _, ipiv = O.getrf(A)
    for trans in (O.NOTRANS, O.TRANS):
        X = O.getrs(A, ipiv, B0.copy(order="F"), trans=trans)
        op = A0[:n, :n] if trans == O.NOTRANS else A0[:n, :n].T
        assert np.allclose(op @ X, B0, atol=1e-10)
    # RowMajor: factor the row-major matrix Ar (same numbers as A0), solve
    # Ar x = b with b stored one RHS per row (ldb >= N).
    Ar = np.ascontiguousarray(A0[:n, :n])
    info, ipr = O.getrf(Ar, order=O.ROW)
    Bt = np.ascontiguousarray(B0.T).copy()        # NRHS x N row-major = N x NRHS col-major
    O.lib().orc_dgetrs(O.ROW, O.NOTRANS, n, nrhs, Ar.ctypes.data, n,
                       np.ascontiguousarray(ipr).ctypes.data, Bt.ctypes.data, n)
    assert np.allclose(A0[:n, :n] @ Bt.T, B0, atol=1e-10)


@pytest.mark.parametrize("uplo", [121, 122])
@pytest.mark.parametrize("N,lda", [(200, 200), (301, 310), (5, 5), (4, 4), (3, 9), (1, 1)])
def test_llttst_residual(uplo, N, lda):
    r = O.llttest(uplo, N, lda)
    assert 0 <= r < 100


def test_potrf_vs_numpy_and_integer_exact():
    n = 60
    # integer L with A = L L^T: Cholesky is exact
    Li = np.tril(RNG.integers(-3, 4, (n, n))).astype(np.float64)
    Li[np.diag_indices(n)] = RNG.integers(1, 4, n)
    A = _f(Li @ Li.T)
    for uplo in (O.LOWER, O.UPPER):
        W = A.copy(order="F")
        assert O.potrf(W, uplo) == 0
        got = np.tril(W) if uplo == O.LOWER else np.triu(W).T
        assert np.array_equal(got, Li), uplo
    S = O.lltgen(O.LOWER, 150)
    Sfull = np.tril(S) + np.tril(S, -1).T
    W = S.copy(order="F")
    assert O.potrf(W, O.LOWER) == 0
    assert np.allclose(np.tril(W), np.linalg.cholesky(Sfull), rtol=1e-12, atol=1e-12)


def test_potrf_not_spd_info_stops():
    n = 30
    A = _f(np.eye(n) * 2.0)
    A[11, 11] = -1.0
    W = A.copy(order="F")
    assert O.potrf(W, O.LOWER) == 12
    assert np.all(W[12:, 12:] == A[12:, 12:])                # stopped: right part untouched


def test_potrs_vs_numpy_rowmajor():
    n, nrhs = 80, 3
    S = O.lltgen(O.LOWER, n)
    Sfull = np.tril(S) + np.tril(S, -1).T
    B0 = _f(RNG.uniform(-1, 1, (n, nrhs)))
    for uplo in (O.LOWER, O.UPPER):
        W = _f(Sfull.copy())
        assert O.potrf(W, uplo) == 0
        X = O.potrs(W, B0.copy(order="F"), uplo)
        assert np.allclose(Sfull @ X, B0, atol=1e-10)
        Wr = np.ascontiguousarray(Sfull).copy()
        assert O.potrf(Wr, uplo, order=O.ROW) == 0
        Bt = np.ascontiguousarray(B0.T).copy()
        O.lib().orc_dpotrs(O.ROW, uplo, n, nrhs, Wr.ctypes.data, n, Bt.ctypes.data, n)
        assert np.allclose(Sfull @ Bt.T, B0, atol=1e-10)


def test_slvtst_residual_and_planted_error():
    n, nrhs = 167, 2                  # dsanity_test size (makes/Make.bin:4499)
    A0 = O.gegen(n, n)
    X0 = O.gegen(n, nrhs)
    B = _f(A0 @ X0)
    A = A0.copy(order="F")
    ipiv = np.zeros(n, dtype=np.int32)
    assert O.lib().orc_dgesv(O.COL, n, nrhs, A.ctypes.data, n, ipiv.ctypes.data,
                             B.ctypes.data, n) == 0
    nrm = O.genrm1(A0)
    assert O.slv_resid(nrm, X0, B) < 100
    B[5, 1] += 1e-8
    assert O.slv_resid(nrm, X0, B) > 100


def test_gemm_ferr_check_flags_planted_error():
    M = N = K = 64
    A = O.matgen(M, K, M, K * 1112)
    B = O.matgen(K, N, K, N * 2238)
    C0 = O.matgen(M, N, M, M * N)
    T = C0.copy(order="F")
    O.gemm(O.NOTRANS, O.NOTRANS, 1.0, A, B, 0.5, T, exact_order=True)
    D = _f(A @ B + 0.5 * C0)
    bad, mx, fe = O.gemm_ferr_check(K, 1.0, 0.5, T, D)
    assert bad == 0 and fe == pytest.approx(2 * (2 * K * 1e-15 + 0.5e-15) + 1e-15)
    D[3, 4] += 10 * fe
    bad, mx, _ = O.gemm_ferr_check(K, 1.0, 0.5, T, D)
    assert bad == 1
    D[3, 4] = np.nan
    assert O.gemm_ferr_check(K, 1.0, 0.5, T, D)[0] == 1


def test_trsm_resid_flags_planted_error():
    M, N = 50, 20
    L = _f(np.tril(RNG.uniform(-1, 1, (M, M))) + 3 * np.eye(M))
    B0 = _f(RNG.uniform(-1, 1, (M, N)))
    X = O.trsm(O.LEFT, O.LOWER, O.NOTRANS, O.NONUNIT, 0.5, L, B0.copy(order="F"))
    assert O.trsm_resid(O.LEFT, O.LOWER, O.NOTRANS, O.NONUNIT, 0.5, L, X, B0) < 1
    X[7, 3] *= 1 + 1e-9
    assert O.trsm_resid(O.LEFT, O.LOWER, O.NOTRANS, O.NONUNIT, 0.5, L, X, B0) > 1
