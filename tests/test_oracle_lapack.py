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


def test_lltgen_diagonal_rule_and_spd():
    """bin/llttst.c:225-259: gegen, the unreferenced triangle zeroed, then
    a_jj = |5.3 a_jj| + ||A||_1 of the symmetric matrix.  Checked through
    numpy: the off-diagonal part is gegen's, the diagonal exceeds every
    row's off-diagonal absolute sum (so A is SPD and Cholesky succeeds)."""
    import oracle as Or
    for uplo in (Or.LOWER, Or.UPPER):
        n = 37
        A = Or.lltgen(uplo, n)
        G = Or.gegen(n, n, seed=n * 1029 + n)
        ref_tri = np.tril(np.ones((n, n), bool), -1) if uplo == Or.LOWER else \
            np.triu(np.ones((n, n), bool), 1)
        assert np.array_equal(A[ref_tri], G[ref_tri])
        assert np.all(A[~ref_tri & ~np.eye(n, dtype=bool)] == 0)
        Gt = np.where(ref_tri | np.eye(n, dtype=bool), G, 0.0)
        S0 = Gt + Gt.T - np.diag(np.diag(Gt))
        nrm = np.linalg.norm(S0, 1)
        assert np.allclose(np.diag(A), np.abs(5.3 * np.diag(G)) + nrm, rtol=1e-15, atol=0)
        S = A + A.T - np.diag(np.diag(A))
        off = np.abs(S).sum(axis=1) - np.abs(np.diag(S))
        assert np.all(np.diag(S) > off)
        np.linalg.cholesky(S)


def test_nan_solution_fails_the_solve_check():
    """slvtst's difference norm skips NaN (bin/slvtst.c:400-415 compares with
    '>'), so the tests' solve check also asserts finiteness: a NaN solution
    must fail it."""
    import oracle as Or
    from tests.checks import solve_ok
    n = 50
    A = Or.gegen(n, n)
    Xc = Or.gegen(n, 2)
    Xn = np.full_like(Xc, np.nan)
    assert Or.slv_resid(Or.genrm1(A), Xc, Xn) == 0.0     # the tester alone passes it
    assert not solve_ok(Or.genrm1(A), Xc, Xn)
    assert solve_ok(Or.genrm1(A), Xc, Xc.copy(order="F"))


def test_inverse_testers_pass_exact_and_fail_planted():
    """The inverse testers (bin/trtritst.c, bin/invtst.c, bin/uumtst.c):
    small residual for numpy's inverse / product, large for a planted error
    of one entry in the 1e-8 range."""
    import oracle as Or
    n = 60
    for uplo in (Or.LOWER, Or.UPPER):
        for diag in (Or.NONUNIT, Or.UNIT):
            A0 = Or.trigen(uplo, diag, n)
            T = np.tril(A0) if uplo == Or.LOWER else np.triu(A0)
            if diag == Or.UNIT:
                np.fill_diagonal(T, 1.0)
            Ti = np.asfortranarray(np.linalg.inv(T))
            assert Or.trtri_resid(uplo, diag, Ti, A0) < 10
            Ti[n - 1 if uplo == Or.LOWER else 0, n // 2] += 1e-8
            assert Or.trtri_resid(uplo, diag, Ti, A0) > 100
    A = Or.gegen(n, n)
    Ai = np.asfortranarray(np.linalg.inv(A))
    assert Or.inv_resid(A, Ai) < 10
    Ai[3, 7] += 1e-8 * np.abs(Ai).max()
    assert Or.inv_resid(A, Ai) > 100
    for uplo in (Or.LOWER, Or.UPPER):
        S = Or.lltgen(uplo, n)
        Sf = np.tril(S) + np.tril(S, -1).T if uplo == Or.LOWER else np.triu(S) + np.triu(S, 1).T
        Si = np.asfortranarray(np.linalg.inv(Sf))
        assert Or.poinv_resid(uplo, S, Si) < 10
        Si[n - 2, 1] += 1e-8 * np.abs(Si).max()
        Si[1, n - 2] += 1e-8 * np.abs(Si).max()
        assert Or.poinv_resid(uplo, S, Si) > 100
        L = np.tril(S) if uplo == Or.LOWER else np.triu(S)
        R = np.asfortranarray(L.T @ L if uplo == Or.LOWER else L @ L.T)
        assert Or.lauum_resid(uplo, R, S) < 10
        R[n - 1, 0] += 1e-6
        R[0, n - 1] += 1e-6
        assert Or.lauum_resid(uplo, R, S) > 100


def test_trinrm1_and_trigen_vs_numpy():
    import oracle as Or
    n = 31
    for uplo in (Or.LOWER, Or.UPPER):
        for diag in (Or.NONUNIT, Or.UNIT):
            A = Or.trigen(uplo, diag, n, padval=-7.0)
            tri = np.tril(np.ones((n, n), bool)) if uplo == Or.LOWER else np.triu(np.ones((n, n), bool))
            ref_part = tri & ~np.eye(n, dtype=bool) if diag == Or.UNIT else tri
            assert np.all(A[~ref_part] == -7.0)              # padding outside what is referenced
            T = np.where(ref_part, A, 0.0)
            assert Or.trinrm1(uplo, diag, A) == pytest.approx(np.linalg.norm(T, 1), rel=1e-14)
            if diag == Or.NONUNIT:                         # diagonally dominant by columns and rows
                off = np.abs(T - np.diag(np.diag(T)))
                assert np.all(np.abs(np.diag(T)) > off.sum(axis=0))
                assert np.all(np.abs(np.diag(T)) > off.sum(axis=1))


@pytest.mark.parametrize("M,N", [(1, 1), (7, 5), (5, 7), (64, 64), (130, 40), (40, 130)])
def test_geqr2_vs_lapack_geqrf(M, N):
    """The oracle's unblocked Householder QR against LAPACK's dgeqrf (raw
    storage and tau through scipy), and qrtst's residual: small for the
    factors, large for a planted error in one reflector entry."""
    import oracle as Or
    import scipy.linalg as sla
    A0 = Or.gegen(M, N)
    A = A0.copy(order="F")
    tau = Or.geqr2(A)
    (qr_l, tau_l), _ = sla.qr(A0, mode="raw")
    assert np.allclose(A, qr_l, atol=1e-12, rtol=0)
    assert np.allclose(tau, tau_l, atol=1e-13, rtol=0)
    assert Or.qr_resid(A, tau, A0) < 10
    if M > 2 and min(M, N) > 0:
        B = A.copy(order="F")
        B[M - 1, 0] += 1e-9
        assert Or.qr_resid(B, tau, A0) > 100


def test_larfg_zero_and_tiny_columns():
    """larfg's special cases: a zero column below the diagonal gives tau = 0
    and leaves the column alone; a column scaled to 1e-300 (below
    safmin/eps) is rescaled first and still factors exactly."""
    import oracle as Or
    A = np.asfortranarray(np.array([[3.0, 1.0], [0.0, 2.0], [0.0, 5.0]]))
    tau = Or.geqr2(A)
    assert tau[0] == 0.0 and A[0, 0] == 3.0
    T = np.asfortranarray(np.array([[3e-300, 1.0], [4e-300, 1.0]]))
    T0 = T.copy(order="F")
    tau = Or.geqr2(T)
    assert T[0, 0] == pytest.approx(-5e-300, rel=1e-14)
    assert Or.qr_resid(T, tau, T0) < 10
