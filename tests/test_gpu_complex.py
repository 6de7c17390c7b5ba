"""GPU parity of the complex Level-3 BLAS (cblas_{z,c}gemm, cblas_{z,c}trsm,
and HERK/HER2K/SYRK/SYR2K/HEMM/SYMM/TRMM) through the C ABI against the
oracle (whose GEMM is bit-exact with ATLAS's ATL_zrefgemm / ATL_crefgemm), with the complex form of gemmtst's elementwise
bound (bin/gemmtst.c:395): |D - T| <= 4 K eps (|alpha| |op A| |op B| +
|beta| |C|) element by element, and l3blastst's TRSM residual (bound 4 for
complex data, where one complex multiply rounds up to ~2*sqrt(2) eps)."""
import ctypes as C
import itertools

import numpy as np
import pytest

import atlas_amd as X
import oracle as O

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(77)
OPS = [X.NoTrans, X.Trans, X.ConjTrans]


def _cplx(shape, dt):
    a = RNG.uniform(-0.5, 0.5, shape) + 1j * RNG.uniform(-0.5, 0.5, shape)
    return np.asfortranarray(a.astype(dt))


def _op(M, t):
    M = M.astype(np.complex128)
    if t == X.NoTrans:
        return M
    if t == X.Trans:
        return M.T
    return M.conj().T


def _eps(dt):
    return np.finfo(np.float64 if dt == np.complex128 else np.float32).eps


def _check_gemm(D, T, ta, tb, alpha, A, B, beta, C0):
    K = A.shape[1] if ta == X.NoTrans else A.shape[0]
    bound = 4 * max(K, 1) * _eps(D.dtype) * (
        abs(alpha) * (np.abs(_op(A, ta)) @ np.abs(_op(B, tb))) + abs(beta) * np.abs(C0)) + 1e-300
    err = np.abs(D.astype(np.complex128) - T.astype(np.complex128))
    assert np.all(err <= bound), (float(np.max(err / bound)), ta, tb, alpha, beta)


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
@pytest.mark.parametrize("ta,tb", list(itertools.product(OPS, repeat=2)))
@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (17, 9, 5), (130, 129, 33), (255, 64, 257),
                                   (3, 200, 300), (128, 128, 0)])
def test_cgemm_shapes_vs_oracle(dt, ta, tb, M, N, K):
    A = _cplx((M, K) if ta == X.NoTrans else (K, M), dt)
    B = _cplx((K, N) if tb == X.NoTrans else (N, K), dt)
    C0 = _cplx((M, N), dt)
    alpha, beta = 1.25 - 0.5j, -0.5 + 0.25j
    D = X.cgemm(ta, tb, alpha, A, B, beta, C0.copy(order="F"))
    T = O.cgemm(ta, tb, alpha, A, B, beta, C0.copy(order="F"))
    _check_gemm(D, T, ta, tb, alpha, A, B, beta, C0)


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
def test_cgemm_alpha_beta_special_cases(dt):
    """alpha = 0 only scales C; beta = 0 never reads C (NaN there stays out);
    real, imaginary and unit betas; imaginary alpha."""
    M, N, K = 70, 45, 38
    for ta, tb in [(X.NoTrans, X.NoTrans), (X.ConjTrans, X.Trans), (X.Trans, X.ConjTrans)]:
        A = _cplx((M, K) if ta == X.NoTrans else (K, M), dt)
        B = _cplx((K, N) if tb == X.NoTrans else (N, K), dt)
        for alpha, beta in [(0.0, 0.5 - 1j), (0.0, 0.0), (0.0, 1.0), (1.0, 0.0), (0.5j, 1.0),
                            (-1.0, 2.0), (0.3 + 0.7j, -1j)]:
            C0 = _cplx((M, N), dt)
            if beta == 0:
                C0[:, 3] = np.nan
            D = X.cgemm(ta, tb, alpha, A, B, beta, C0.copy(order="F"))
            T = O.cgemm(ta, tb, alpha, A, B, beta, C0.copy(order="F"))
            assert np.all(np.isfinite(D)), (alpha, beta)
            if alpha == 0:
                assert np.array_equal(D, T), (alpha, beta)
            _check_gemm(D, T, ta, tb, alpha, A, B, beta, np.nan_to_num(C0))


def test_zgemm_integer_exact():
    """Gaussian integers: every MFMA partial sum is exact, so the GPU must
    agree bit for bit with int64 arithmetic for every op pair."""
    M, N, K = 150, 90, 200
    for ta, tb in itertools.product(OPS, repeat=2):
        Ai = RNG.integers(-4, 5, (M, K)) + 1j * RNG.integers(-4, 5, (M, K))
        Bi = RNG.integers(-4, 5, (K, N)) + 1j * RNG.integers(-4, 5, (K, N))
        # stored so that op(A) == Ai and op(B) == Bi
        A = np.asfortranarray(Ai if ta == X.NoTrans else (Ai.T if ta == X.Trans else Ai.conj().T))
        B = np.asfortranarray(Bi if tb == X.NoTrans else (Bi.T if tb == X.Trans else Bi.conj().T))
        for dt in (np.complex128, np.complex64):
            D = X.cgemm(ta, tb, 1.0, A.astype(dt), B.astype(dt), 0.0,
                        np.zeros((M, N), dtype=dt, order="F"))
            assert np.array_equal(D, (Ai @ Bi).astype(dt)), (ta, tb, dt)


def test_zgemm_rowmajor_and_lda_gaps():
    M, N, K, ld = 61, 47, 35, 80
    L = X.lib()
    for ta, tb in [(X.NoTrans, X.NoTrans), (X.ConjTrans, X.NoTrans), (X.Trans, X.ConjTrans)]:
        # RowMajor: C-ordered arrays
        A = np.ascontiguousarray(_cplx((M, K) if ta == X.NoTrans else (K, M), np.complex128))
        B = np.ascontiguousarray(_cplx((K, N) if tb == X.NoTrans else (N, K), np.complex128))
        C0 = np.ascontiguousarray(_cplx((M, N), np.complex128))
        D = C0.copy()
        al = np.array([0.5, -1.5]); be = np.array([2.0, 0.5])
        L.cblas_zgemm(X.RowMajor, ta, tb, M, N, K, al.ctypes.data, A.ctypes.data, A.shape[1],
                      B.ctypes.data, B.shape[1], be.ctypes.data, D.ctypes.data, N)
        want = (0.5 - 1.5j) * (_op(A, ta) @ _op(B, tb)) + (2.0 + 0.5j) * C0
        assert np.max(np.abs(D - want)) < 1e-12 * K
    # ColMajor with leading-dimension gaps: the gaps stay untouched
    A = np.asfortranarray(np.full((ld, K), 7 + 7j))
    A[:M] = _cplx((M, K), np.complex128)
    B = np.asfortranarray(np.full((ld, N), 7 + 7j))
    B[:K] = _cplx((K, N), np.complex128)
    Cm = np.asfortranarray(np.full((ld, N), -9e27 + 0j))
    Cm[:M] = _cplx((M, N), np.complex128)
    C0 = Cm.copy(order="F")
    al = np.array([1.0, 0.0]); be = np.array([0.0, 1.0])
    L.cblas_zgemm(X.ColMajor, X.NoTrans, X.NoTrans, M, N, K, al.ctypes.data, A.ctypes.data, ld,
                  B.ctypes.data, ld, be.ctypes.data, Cm.ctypes.data, ld)
    assert np.all(Cm[M:] == -9e27 + 0j)
    want = A[:M] @ B[:K] + 1j * C0[:M]
    assert np.max(np.abs(Cm[:M] - want)) < 1e-12 * K


def test_zgemm_2048_vs_numpy():
    n = 2048
    A = _cplx((n, n), np.complex128)
    B = _cplx((n, n), np.complex128)
    C0 = _cplx((n, n), np.complex128)
    D = X.cgemm(X.NoTrans, X.ConjTrans, 1.0 + 0.5j, A, B, 0.5, C0.copy(order="F"))
    want = (1.0 + 0.5j) * (A @ B.conj().T) + 0.5 * C0
    bound = 4 * n * _eps(np.complex128) * (abs(1 + 0.5j) * (np.abs(A) @ np.abs(B).T)
                                           + 0.5 * np.abs(C0))
    assert np.all(np.abs(D - want) <= bound)


def test_cgemm_row_chunks_of_the_expanded_operand():
    """K large enough that the expanded left operand is built in several row
    chunks (1 GiB each): 600 rows of K = 2^18 in complex64 make three."""
    M, N, K = 600, 3, 1 << 18
    A = _cplx((M, K), np.complex64)
    B = _cplx((K, N), np.complex64)
    D = X.cgemm(X.NoTrans, X.NoTrans, 1.0, A, B, 0.0, np.zeros((M, N), np.complex64, order="F"))
    want = A.astype(np.complex128) @ B.astype(np.complex128)
    bound = 4 * K * _eps(np.complex64) * (np.abs(A) @ np.abs(B))
    assert np.all(np.abs(D - want) <= bound)


def _tri(n, uplo, dt, junk=True):
    A = _cplx((n, n), dt)
    A[np.diag_indices(n)] += np.asarray(2.0 + 1.0j, dtype=dt)
    if junk:
        mask = np.tril(np.ones((n, n), bool), -1) if uplo == X.Upper else \
            np.triu(np.ones((n, n), bool), 1)
        A[mask] = np.asarray(1e30 - 1e30j, dtype=dt)
    return A


def _trsm_resid(side, uplo, ta, diag, alpha, A, Xs, B0):
    """l3blastst's TRSM check: ||op(A) X - alpha B|| / (||A|| ||X|| eps n)."""
    n = A.shape[0]
    T = np.tril(A) if uplo == X.Lower else np.triu(A)
    T = T.astype(np.complex128)
    if diag == X.Unit:
        T[np.diag_indices(n)] = 1.0
    E = _op(T, ta)
    Xd = Xs.astype(np.complex128)
    R = E @ Xd if side == X.Left else Xd @ E
    R -= alpha * B0.astype(np.complex128)
    den = np.abs(E).sum(0).max() * max(np.abs(Xd).sum(0).max(), 1e-300) * _eps(Xs.dtype) * n
    return np.abs(R).sum(0).max() / den


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
@pytest.mark.parametrize("side", [X.Left, X.Right])
@pytest.mark.parametrize("uplo", [X.Upper, X.Lower])
@pytest.mark.parametrize("ta", OPS)
@pytest.mark.parametrize("diag", [X.NonUnit, X.Unit])
def test_ctrsm_all_cases_vs_oracle(dt, side, uplo, ta, diag):
    # one RHS over many blocks and 2-4 RHS (tagged TRSV), up to 256 (chained
    # MFMA solve), more (GEMM loop)
    for M, N in [(1, 1), (37, 5), (100, 70), (257, 33), (300, 1), (300, 3), (130, 4), (70, 300)]:
        na = M if side == X.Left else N
        A = _tri(na, uplo, dt)
        B0 = _cplx((M, N), dt)
        alpha = 0.75 - 0.25j
        D = X.ctrsm(side, uplo, ta, diag, alpha, A, B0.copy(order="F"))
        T = O.ctrsm(side, uplo, ta, diag, alpha, A, B0.copy(order="F"))
        # l3blastst's bound is 1 for real data; a complex product or quotient
        # rounds with up to about 2*sqrt(2) eps relative error, so 4 here
        assert _trsm_resid(side, uplo, ta, diag, alpha, A, D, B0) < 4.0
        tol = 1e-9 if dt == np.complex128 else 2e-3
        assert np.max(np.abs(D - T)) <= tol * max(1.0, np.max(np.abs(T))), (M, N)


def test_ztrsm_alpha_zero_rowmajor_and_large():
    L = X.lib()
    # alpha = 0: B = 0 whatever it held
    A = _tri(20, X.Lower, np.complex128)
    B = _cplx((20, 6), np.complex128)
    B[0, 0] = np.nan
    X.ctrsm(X.Left, X.Lower, X.NoTrans, X.NonUnit, 0.0, A, B)
    assert np.all(B == 0)
    # RowMajor: C-ordered arrays, solve X op(A) = alpha B on the right
    M, N = 50, 40
    A = np.ascontiguousarray(_tri(N, X.Upper, np.complex128, junk=False))
    B0 = np.ascontiguousarray(_cplx((M, N), np.complex128))
    Bm = B0.copy()
    al = np.array([1.0, 1.0])
    L.cblas_ztrsm(X.RowMajor, X.Right, X.Upper, X.ConjTrans, X.NonUnit, M, N, al.ctypes.data,
                  A.ctypes.data, N, Bm.ctypes.data, N)
    E = np.triu(A).conj().T
    assert np.max(np.abs(Bm @ E - (1 + 1j) * B0)) < 1e-11
    # 2000 x 300, the blocked path over many 32-row blocks
    n, nr = 2000, 300
    A = _tri(n, X.Lower, np.complex128, junk=False)
    A[np.tril_indices(n, -1)] /= n    # well conditioned
    B0 = _cplx((n, nr), np.complex128)
    D = X.ctrsm(X.Left, X.Lower, X.ConjTrans, X.NonUnit, 1.0, A, B0.copy(order="F"))
    assert _trsm_resid(X.Left, X.Lower, X.ConjTrans, X.NonUnit, 1.0, A, D, B0) < 4.0


# ---------------------------------------------------------------- LU family
def _getrf(A):
    p = "z" if A.dtype == np.complex128 else "c"
    M, N = A.shape
    ipiv = np.zeros(max(min(M, N), 1), dtype=np.int32)
    info = getattr(X.lib(), "clapack_%sgetrf" % p)(X.ColMajor, M, N, A.ctypes.data,
                                                   X._ld(A), ipiv.ctypes.data)
    return info, ipiv[:min(M, N)]


def _lu_resid(LU, ipiv, A0):
    """lutst's ||P A - L U|| / (||A|| eps n) with the factors unpacked."""
    M, N = A0.shape
    k = min(M, N)
    L = np.tril(LU[:, :k].astype(np.complex128), -1) + np.eye(M, k)
    U = np.triu(LU[:k, :].astype(np.complex128))
    P = np.arange(M)
    for j, q in enumerate(ipiv):
        P[[j, q]] = P[[q, j]]
    R = A0[P].astype(np.complex128) - L @ U
    return np.abs(R).sum(0).max() / (np.abs(A0).sum(0).max() * _eps(LU.dtype) * max(M, N))


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
@pytest.mark.parametrize("M,N", [(1, 1), (5, 5), (8, 8), (9, 9), (64, 64), (130, 100),
                                 (100, 130), (300, 300), (517, 517)])
def test_cgetrf_pivots_and_residual_vs_oracle(dt, M, N):
    rng = np.random.default_rng(M * 1000 + N)
    A0 = np.asfortranarray((rng.uniform(-1, 1, (M, N)) + 1j * rng.uniform(-1, 1, (M, N)))
                           .astype(dt))
    A = A0.copy(order="F")
    info, ipiv = _getrf(A)
    T = A0.copy(order="F")
    tinfo, tpiv, gap = O.cgetf2(T)
    assert info == tinfo == 0
    assert _lu_resid(A, ipiv, A0) < 10.0
    if gap > 1e-5:
        assert np.array_equal(ipiv, tpiv)
        tol = 1e-9 if dt == np.complex128 else 5e-3
        assert np.max(np.abs(A - T)) <= tol * max(1.0, np.max(np.abs(T)))


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
def test_cgetrf_zero_pivot_info(dt):
    """info = first exactly zero pivot (1-based), at leaf and panel edges."""
    for col in (0, 7, 8, 63, 127, 128, 200):
        n = 260
        rng = np.random.default_rng(col)
        A = np.asfortranarray((rng.uniform(-1, 1, (n, n)) + 1j * rng.uniform(-1, 1, (n, n)))
                              .astype(dt))
        # rows col.. are zero in columns 0..col: the first col eliminations
        # pivot above them and leave them exactly zero, so U(col,col) = 0
        A[col:, :col + 1] = 0
        T = A.copy(order="F")
        info, ipiv = _getrf(A)
        tinfo, tpiv, _ = O.cgetf2(T)
        assert info == tinfo == col + 1
        assert ipiv[col] == tpiv[col] == col


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
@pytest.mark.parametrize("trans", OPS)
def test_cgetrs_and_gesv(dt, trans):
    p = "z" if dt == np.complex128 else "c"
    L = X.lib()
    for n, nrhs in [(1, 1), (37, 3), (200, 17), (700, 64)]:
        rng = np.random.default_rng(n + nrhs)
        A0 = np.asfortranarray((rng.uniform(-1, 1, (n, n)) + 1j * rng.uniform(-1, 1, (n, n)))
                               .astype(dt))
        B0 = np.asfortranarray((rng.uniform(-1, 1, (n, nrhs)) +
                                1j * rng.uniform(-1, 1, (n, nrhs))).astype(dt))
        A = A0.copy(order="F")
        info, ipiv = _getrf(A)
        assert info == 0
        B = B0.copy(order="F")
        assert getattr(L, "clapack_%sgetrs" % p)(X.ColMajor, trans, n, nrhs, A.ctypes.data, n,
                                                  ipiv.ctypes.data, B.ctypes.data, n) == 0
        E = _op(A0, trans)
        R = E @ B.astype(np.complex128) - B0
        # slvtst: ||A x - b|| / (||A|| ||x|| eps n)
        r = np.abs(R).sum(0).max() / (np.abs(E).sum(0).max() * np.abs(B).sum(0).max() *
                                      _eps(dt) * n)
        assert r < 10.0, (n, nrhs, r)
        if trans == X.NoTrans:
            A = A0.copy(order="F")
            B = B0.copy(order="F")
            piv = np.zeros(n, dtype=np.int32)
            assert getattr(L, "clapack_%sgesv" % p)(X.ColMajor, n, nrhs, A.ctypes.data, n,
                                                     piv.ctypes.data, B.ctypes.data, n) == 0
            assert np.array_equal(piv, ipiv)
            r = np.abs(A0 @ B.astype(np.complex128) - B0).sum(0).max() / (
                np.abs(A0).sum(0).max() * np.abs(B).sum(0).max() * _eps(dt) * n)
            assert r < 10.0


def test_zgesv_rowmajor_and_getrs_rowmajor_conjtrans_as_trans():
    """RowMajor factors A^T; getrs ConjTrans in RowMajor solves with A^T, as
    ATLAS's ATL_getrs.c RowMajor branches do."""
    L = X.lib()
    n, nrhs = 150, 4
    rng = np.random.default_rng(5)
    A0 = np.ascontiguousarray(rng.uniform(-1, 1, (n, n)) + 1j * rng.uniform(-1, 1, (n, n)))
    B0 = np.ascontiguousarray(rng.uniform(-1, 1, (n, nrhs)) + 1j * rng.uniform(-1, 1, (n, nrhs)))
    A = A0.copy()
    B = B0.copy()
    piv = np.zeros(n, dtype=np.int32)
    # for both orders ATLAS's clapack reads B as NRHS vectors of length N
    # stored ldb apart, i.e. the column-major N x NRHS array
    Bc = B0.copy(order="F")
    assert L.clapack_zgesv(X.RowMajor, n, nrhs, A.ctypes.data, n, piv.ctypes.data,
                           Bc.ctypes.data, n) == 0
    assert np.max(np.abs(A0 @ Bc - B0)) < 1e-9
    for trans, E in [(X.NoTrans, A0), (X.Trans, A0.T), (X.ConjTrans, A0.T)]:
        Bt = B0.copy(order="F")
        assert L.clapack_zgetrs(X.RowMajor, trans, n, nrhs, A.ctypes.data, n, piv.ctypes.data,
                                Bt.ctypes.data, n) == 0
        assert np.max(np.abs(E @ Bt - B0)) < 1e-9, trans


def test_zgesv_4096():
    n, nrhs = 4096, 8
    rng = np.random.default_rng(11)
    A0 = np.asfortranarray(rng.uniform(-1, 1, (n, n)) + 1j * rng.uniform(-1, 1, (n, n)))
    B0 = np.asfortranarray(rng.uniform(-1, 1, (n, nrhs)) + 1j * rng.uniform(-1, 1, (n, nrhs)))
    A, B = A0.copy(order="F"), B0.copy(order="F")
    piv = np.zeros(n, dtype=np.int32)
    assert X.lib().clapack_zgesv(X.ColMajor, n, nrhs, A.ctypes.data, n, piv.ctypes.data,
                                 B.ctypes.data, n) == 0
    r = np.abs(A0 @ B - B0).sum(0).max() / (np.abs(A0).sum(0).max() * np.abs(B).sum(0).max() *
                                            _eps(np.complex128) * n)
    assert r < 10.0
    assert _lu_resid(A, piv, A0) < 10.0


def test_cgetrf_1500_pivots_vs_oracle():
    """Order 1500: the leaves hold two rows per thread in registers; pivots
    equal to the oracle's (the smallest margin between a pivot and its
    runner-up is 2e-5 on this input, far above double rounding), residual."""
    n = 1500
    rng = np.random.default_rng(1500)
    A0 = np.asfortranarray(rng.uniform(-1, 1, (n, n)) + 1j * rng.uniform(-1, 1, (n, n)))
    A = A0.copy(order="F")
    info, ipiv = _getrf(A)
    T = A0.copy(order="F")
    tinfo, tpiv, gap = O.cgetf2(T)
    assert info == tinfo == 0
    assert gap > 1e-9
    assert np.array_equal(ipiv, tpiv)
    assert _lu_resid(A, ipiv, A0) < 10.0
    assert np.max(np.abs(A - T)) <= 1e-8 * np.max(np.abs(T))


# ---------------------------------------------------------------- Cholesky
def _hpd(n, dt, seed):
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return np.asfortranarray((B @ B.conj().T / n + np.eye(n)).astype(dt))


def _llt_resid(A0, F, uplo):
    """llttst's ||A - L L^H|| / (||A|| eps n) from the referenced triangle."""
    n = A0.shape[0]
    if uplo == X.Lower:
        L = np.tril(F).astype(np.complex128)
    else:
        L = np.triu(F).astype(np.complex128).conj().T
    R = A0.astype(np.complex128) - L @ L.conj().T
    return np.abs(R).sum(0).max() / (np.abs(A0).sum(0).max() * _eps(F.dtype) * n)


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
@pytest.mark.parametrize("uplo", [X.Lower, X.Upper])
def test_cpotrf_vs_oracle(dt, uplo):
    p = "z" if dt == np.complex128 else "c"
    L = X.lib()
    for n in (1, 5, 64, 65, 130, 300, 700):
        A0 = _hpd(n, dt, n + uplo)
        # junk in the other triangle must stay untouched
        other = np.triu(np.ones((n, n), bool), 1) if uplo == X.Lower else \
            np.tril(np.ones((n, n), bool), -1)
        A0[other] = np.asarray(7.0 - 3.0j, dtype=dt)
        A = A0.copy(order="F")
        assert getattr(L, "clapack_%spotrf" % p)(X.ColMajor, uplo, n, A.ctypes.data, n) == 0
        T = A0.copy(order="F")
        assert O.cpotrf(T, uplo) == 0
        assert np.array_equal(A[other], A0[other]), n
        tri = ~other
        tol = 1e-10 if dt == np.complex128 else 1e-3
        assert np.max(np.abs(A[tri] - T[tri])) <= tol * np.max(np.abs(T[tri])), n
        full = np.where(other, np.conj(A0.T), A0)          # the Hermitian matrix
        assert _llt_resid(full, A, uplo) < 10.0, n


@pytest.mark.parametrize("uplo", [X.Lower, X.Upper])
def test_zpotrf_not_positive_definite_info(uplo):
    L = X.lib()
    for col in (0, 1, 63, 64, 65, 129):
        n = 200
        A = _hpd(n, np.complex128, col)
        A[col, col] = -1.0
        T = A.copy(order="F")
        info = L.clapack_zpotrf(X.ColMajor, uplo, n, A.ctypes.data, n)
        assert info == O.cpotrf(T, uplo), col


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
@pytest.mark.parametrize("uplo", [X.Lower, X.Upper])
def test_cpotrs_posv_and_rowmajor(dt, uplo):
    p = "z" if dt == np.complex128 else "c"
    L = X.lib()
    for n, nrhs in [(1, 1), (50, 3), (300, 17), (700, 64)]:
        A0 = _hpd(n, dt, n)
        rng = np.random.default_rng(n)
        B0 = np.asfortranarray((rng.uniform(-1, 1, (n, nrhs)) +
                                1j * rng.uniform(-1, 1, (n, nrhs))).astype(dt))
        A, B = A0.copy(order="F"), B0.copy(order="F")
        assert getattr(L, "clapack_%sposv" % p)(X.ColMajor, uplo, n, nrhs, A.ctypes.data, n,
                                                 B.ctypes.data, n) == 0
        r = np.abs(A0.astype(np.complex128) @ B - B0).sum(0).max() / (
            np.abs(A0).sum(0).max() * np.abs(B).sum(0).max() * _eps(dt) * n)
        assert r < 10.0, (n, r)
        B2 = B0.copy(order="F")
        assert getattr(L, "clapack_%spotrs" % p)(X.ColMajor, uplo, n, nrhs, A.ctypes.data, n,
                                                  B2.ctypes.data, n) == 0
        assert np.max(np.abs(B2 - B)) <= (1e-12 if dt == np.complex128 else 1e-4) * \
            max(1.0, np.max(np.abs(B)))
        # RowMajor: the stored array is A^T (= conj(A) for Hermitian A); B as
        # NRHS vectors of length n
        Ar = A0.copy(order="C")      # a copy: A0 stays the matrix for the residual
        Br = B0.copy(order="F")
        assert getattr(L, "clapack_%sposv" % p)(X.RowMajor, uplo, n, nrhs, Ar.ctypes.data, n,
                                                 Br.ctypes.data, n) == 0
        r = np.abs(A0.astype(np.complex128) @ Br - B0).sum(0).max() / (
            np.abs(A0).sum(0).max() * np.abs(Br).sum(0).max() * _eps(dt) * n)
        assert r < 10.0, ("rowmajor", n, r)


def test_zposv_4096():
    n, nrhs = 4096, 8
    A0 = _hpd(n, np.complex128, 4096)
    rng = np.random.default_rng(9)
    B0 = np.asfortranarray(rng.uniform(-1, 1, (n, nrhs)) + 1j * rng.uniform(-1, 1, (n, nrhs)))
    A, B = A0.copy(order="F"), B0.copy(order="F")
    assert X.lib().clapack_zposv(X.ColMajor, X.Lower, n, nrhs, A.ctypes.data, n, B.ctypes.data,
                                 n) == 0
    r = np.abs(A0 @ B - B0).sum(0).max() / (np.abs(A0).sum(0).max() * np.abs(B).sum(0).max() *
                                            _eps(np.complex128) * n)
    assert r < 10.0
    assert _llt_resid(A0, A, X.Lower) < 10.0


# --- HERK / HER2K / SYRK / SYR2K / HEMM / SYMM / TRMM through the C ABI
# against the closed forms of tests/checks.py (pinned there against ATLAS's
# own ATL_zref* / ATL_cref* routines), ColMajor and RowMajor, with an
# elementwise bound of gemmtst's form.
from tests.checks import crk_np, chemm_np, ctrmm_np, csym_full   # noqa: E402


def _arr(a, order):
    return a.copy(order="F" if order == X.ColMajor else "C")


def _ldx(a, order):
    return max(1, a.shape[0] if order == X.ColMajor else a.shape[1])


def _pp(a):
    return a.ctypes.data_as(C.c_void_p)


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
@pytest.mark.parametrize("herm,two", [(1, 0), (1, 1), (0, 0), (0, 1)])
@pytest.mark.parametrize("n,K", [(1, 1), (37, 70), (200, 33), (130, 0), (129, 300)])
def test_crk_vs_closed_form(dt, herm, two, n, K):
    p, rt = X._cpre(dt)
    nm = ("her" if herm else "syr") + ("2k" if two else "k")
    f = getattr(X.lib(), "cblas_%s%s" % (p, nm))
    eps = _eps(dt)
    alt = X.ConjTrans if herm else X.Trans
    for order, uplo, trans in itertools.product([X.ColMajor, X.RowMajor], [X.Upper, X.Lower],
                                                [X.NoTrans, alt]):
        shp = (n, K) if trans == X.NoTrans else (K, n)
        A = _arr(_cplx(shp, dt), order)
        B = _arr(_cplx(shp, dt), order) if two else None
        for alpha, beta in [(0.7 - 0.4j, 0.5 + 0.3j), (-1.1 + 0.0j, 0.0), (0.0, -0.5)]:
            if herm and not two:
                alpha = complex(alpha).real
            if herm:
                beta = complex(beta).real
            C0 = _cplx((n, n), dt)
            off = np.triu(np.ones((n, n), bool), 1) if uplo == X.Lower else \
                np.tril(np.ones((n, n), bool), -1)
            C0[off] = np.asarray(3e30 - 1e30j, dtype=dt)
            if beta == 0:
                C0[~off] = np.nan
            got = _arr(C0, order)
            al, be = X._cscalar(alpha, rt), X._cscalar(beta, rt)
            args = [order, uplo, trans, n, K, al[0] if (herm and not two) else _pp(al), _pp(A),
                    _ldx(A, order)]
            if two:
                args += [_pp(B), _ldx(B, order)]
            f(*(args + [be[0] if herm else _pp(be), _pp(got), _ldx(got, order)]))
            want = crk_np(herm, uplo, trans, alpha, A, B, beta, C0)
            opA = _op(A, X.NoTrans if trans == X.NoTrans else alt)
            mag = np.abs(opA) @ np.abs(opA).T
            if two:
                opB = _op(B, X.NoTrans if trans == X.NoTrans else alt)
                mag = np.abs(opA) @ np.abs(opB).T + np.abs(opB) @ np.abs(opA).T
            bound = 4 * (K + 2) * eps * (abs(alpha) * mag +
                                         abs(beta) * np.abs(np.nan_to_num(C0))) + 1e-300
            case = (order, uplo, trans, alpha, beta)
            assert np.array_equal(got[off], C0[off]), case
            err = np.abs(got.astype(np.complex128) - want)[~off]
            assert np.all(err <= bound[~off]), (float(np.max(err / bound[~off])),) + case
            if herm and n:
                assert np.all(np.diag(got).imag == 0), case


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
@pytest.mark.parametrize("herm", [1, 0])
@pytest.mark.parametrize("M,N", [(1, 1), (45, 70), (257, 33)])
def test_chemm_vs_closed_form(dt, herm, M, N):
    p, rt = X._cpre(dt)
    f = getattr(X.lib(), "cblas_%s%s" % (p, "hemm" if herm else "symm"))
    eps = _eps(dt)
    for order, side, uplo in itertools.product([X.ColMajor, X.RowMajor], [X.Left, X.Right],
                                               [X.Upper, X.Lower]):
        na = M if side == X.Left else N
        A = _cplx((na, na), dt)
        off = np.triu(np.ones((na, na), bool), 1) if uplo == X.Lower else \
            np.tril(np.ones((na, na), bool), -1)
        A[off] = np.asarray(1e30 + 1e30j, dtype=dt)       # never read
        if herm:
            A[np.diag_indices(na)] += np.asarray(1e30j, dtype=dt)  # never read
        A = _arr(A, order)
        B = _arr(_cplx((M, N), dt), order)
        for alpha, beta in [(0.6 + 0.2j, -0.3 + 0.9j), (1.0, 0.0)]:
            C0 = _cplx((M, N), dt)
            if beta == 0:
                C0[:] = np.nan
            got = _arr(C0, order)
            al, be = X._cscalar(alpha, rt), X._cscalar(beta, rt)
            f(order, side, uplo, M, N, _pp(al), _pp(A), _ldx(A, order), _pp(B), _ldx(B, order),
              _pp(be), _pp(got), _ldx(got, order))
            want = chemm_np(herm, side, uplo, alpha, A, B, beta, C0)
            F = np.abs(csym_full(herm, uplo, A))
            mag = F @ np.abs(B) if side == X.Left else np.abs(B) @ F
            bound = 4 * (na + 2) * eps * (abs(alpha) * mag +
                                          abs(beta) * np.abs(np.nan_to_num(C0))) + 1e-300
            err = np.abs(got.astype(np.complex128) - want)
            assert np.all(err <= bound), (float(np.max(err / bound)), order, side, uplo, alpha)


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
@pytest.mark.parametrize("M,N", [(1, 1), (45, 70), (257, 33), (40, 300)])
def test_ctrmm_vs_closed_form(dt, M, N):
    """every side/uplo/trans/diag; M = 257 (Left) and N = 300 (Right) span
    several 128-row / 128-column GEMM tiles, so the K band of each tile
    position is exercised"""
    p, rt = X._cpre(dt)
    f = getattr(X.lib(), "cblas_%strmm" % p)
    eps = _eps(dt)
    for order, side, uplo, ta, diag in itertools.product(
            [X.ColMajor, X.RowMajor], [X.Left, X.Right], [X.Upper, X.Lower], OPS,
            [X.NonUnit, X.Unit]):
        na = M if side == X.Left else N
        A = _cplx((na, na), dt)
        off = np.triu(np.ones((na, na), bool), 1) if uplo == X.Lower else \
            np.tril(np.ones((na, na), bool), -1)
        A[off] = np.asarray(1e30 + 1e30j, dtype=dt)       # never read
        if diag == X.Unit:
            A[np.diag_indices(na)] = np.asarray(1e30, dtype=dt)   # never read
        A = _arr(A, order)
        B0 = _cplx((M, N), dt)
        for alpha in (1.0, -0.6 + 0.8j, 0.0):
            got = _arr(B0, order)
            al = X._cscalar(alpha, rt)
            f(order, side, uplo, ta, diag, M, N, _pp(al), _pp(A), _ldx(A, order), _pp(got),
              _ldx(got, order))
            want = ctrmm_np(side, uplo, ta, diag, alpha, A, B0)
            T = np.abs(ctrmm_np(side, uplo, ta, diag, 1.0, np.abs(A), np.eye(na)))  # |op(A)|
            mag = T @ np.abs(B0) if side == X.Left else np.abs(B0) @ T
            bound = 4 * (na + 2) * eps * abs(alpha) * mag + 1e-300
            err = np.abs(got.astype(np.complex128) - want)
            assert np.all(err <= bound), (float(np.max(err / bound)), order, side, uplo, ta, diag)
