"""Pins the oracle's BLAS restatement.

* bit-exact against oracle/_ref (ATLAS's own C reference BLAS, compiled from
  src/blas/reference where /root/reference exists; skipped elsewhere);
* exact on small-integer inputs (every product and partial sum is an integer
  below 2^53, so any summation order gives the same bits) against numpy int64;
* against numpy/scipy within rounding.
"""
import ctypes as C
import itertools

import numpy as np
import pytest

import oracle as O

RNG = np.random.default_rng(1234)


def _ref_or_skip():
    r = O.ref()
    if r is None:
        pytest.skip("oracle/_ref not built (needs /root/reference)")
    return r


def _f(a):
    return np.asfortranarray(a)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("ta,tb", list(itertools.product([O.NOTRANS, O.TRANS], repeat=2)))
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.5), (-0.7, 0.0), (1.3, 1.0), (0.0, 1.3), (0.0, 0.0)])
def test_gemm_ref_bit_exact_vs_atlas_refgemm(dt, ta, tb, alpha, beta):
    R = _ref_or_skip()
    M, N, K = 13, 9, 11
    A = _f(RNG.uniform(-0.5, 0.5, (M, K) if ta == O.NOTRANS else (K, M)).astype(dt))
    B = _f(RNG.uniform(-0.5, 0.5, (K, N) if tb == O.NOTRANS else (N, K)).astype(dt))
    C0 = _f(RNG.uniform(-0.5, 0.5, (M, N)).astype(dt))
    c1 = C0.copy(order="F")
    c2 = C0.copy(order="F")
    O.gemm(ta, tb, alpha, A, B, beta, c1, exact_order=True)
    ct = C.c_double if dt == np.float64 else C.c_float
    fn = R.ATL_drefgemm if dt == np.float64 else R.ATL_srefgemm
    fn.argtypes = [C.c_int] * 5 + [ct, C.c_void_p, C.c_int, C.c_void_p, C.c_int, ct, C.c_void_p, C.c_int]
    fn(ta, tb, M, N, K, alpha, A.ctypes.data, A.shape[0], B.ctypes.data, B.shape[0], beta,
       c2.ctypes.data, M)
    assert np.array_equal(c1, c2)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_trsm_ref_bit_exact_vs_atlas_reftrsm(dt):
    R = _ref_or_skip()
    ct = C.c_double if dt == np.float64 else C.c_float
    fn = R.ATL_dreftrsm if dt == np.float64 else R.ATL_sreftrsm
    fn.argtypes = [C.c_int] * 6 + [ct, C.c_void_p, C.c_int, C.c_void_p, C.c_int]
    for side, uplo, ta, diag in itertools.product([O.LEFT, O.RIGHT], [O.UPPER, O.LOWER],
                                                  [O.NOTRANS, O.TRANS], [O.NONUNIT, O.UNIT]):
        M, N = 11, 7
        na = M if side == O.LEFT else N
        A = _f(RNG.uniform(-0.5, 0.5, (na, na)).astype(dt))
        A[np.diag_indices(na)] += np.asarray(3.0, dtype=dt)
        B0 = _f(RNG.uniform(-0.5, 0.5, (M, N)).astype(dt))
        for alpha in (1.0, -0.6, 0.0):
            b1 = B0.copy(order="F")
            b2 = B0.copy(order="F")
            O.trsm(side, uplo, ta, diag, alpha, A, b1)
            fn(side, uplo, ta, diag, M, N, alpha, A.ctypes.data, na, b2.ctypes.data, M)
            assert np.array_equal(b1, b2), (side, uplo, ta, diag, alpha)


def test_iamax_vs_atlas_refamax_ties_and_nan():
    R = _ref_or_skip()
    R.ATL_idrefamax.argtypes = [C.c_int, C.c_void_p, C.c_int]
    cases = [
        np.array([1.0, -3.0, 3.0, 2.0]),               # tie: first wins
        np.array([0.0, 0.0, 0.0]),                      # all zero -> 0
        np.array([np.nan, 1.0, -1.0, 0.5]),             # NaN never wins
        np.array([1.0, np.nan, 2.0]),
        RNG.uniform(-1, 1, 37),
        np.array([-5.0] * 19),
    ]
    for x in cases:
        x = np.ascontiguousarray(x)
        assert O.iamax(x) == R.ATL_idrefamax(x.size, x.ctypes.data, 1)
    assert O.iamax(np.array([1.0, -3.0, 3.0])) == 1


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_syrk_trmm_vs_atlas_reference(dt):
    R = _ref_or_skip()
    ct = C.c_double if dt == np.float64 else C.c_float
    syrk = R.ATL_drefsyrk if dt == np.float64 else R.ATL_srefsyrk
    syrk.argtypes = [C.c_int] * 4 + [ct, C.c_void_p, C.c_int, ct, C.c_void_p, C.c_int]
    tol = 1e-13 if dt == np.float64 else 2e-5
    for uplo, tr in itertools.product([O.UPPER, O.LOWER], [O.NOTRANS, O.TRANS]):
        N, K = 9, 6
        A = _f(RNG.uniform(-1, 1, (N, K) if tr == O.NOTRANS else (K, N)).astype(dt))
        C0 = _f(RNG.uniform(-1, 1, (N, N)).astype(dt))
        c1, c2 = C0.copy(order="F"), C0.copy(order="F")
        O.syrk(uplo, tr, -1.0, A, 0.5, c1)
        syrk(uplo, tr, N, K, -1.0, A.ctypes.data, A.shape[0], 0.5, c2.ctypes.data, N)
        assert np.allclose(c1, c2, atol=tol, rtol=0)


def test_gemm_integer_kat_exact():
    """Small integers: the exact product is representable, so the oracle must
    reproduce numpy's int64 matmul bit for bit, for every transpose case."""
    for ta, tb in itertools.product([O.NOTRANS, O.TRANS], repeat=2):
        M, N, K = 37, 29, 300
        Ai = RNG.integers(-8, 9, (M, K))
        Bi = RNG.integers(-8, 9, (K, N))
        Ci = RNG.integers(-100, 101, (M, N))
        want = 3 * (Ai @ Bi) + 2 * Ci   # alpha=3, beta=2 (beta*C exact too)
        A = _f((Ai if ta == O.NOTRANS else Ai.T).astype(np.float64))
        B = _f((Bi if tb == O.NOTRANS else Bi.T).astype(np.float64))
        for exact in (True, False):
            Cm = _f(Ci.astype(np.float64))
            O.gemm(ta, tb, 3.0, A, B, 2.0, Cm, exact_order=exact)
            assert np.array_equal(Cm, want.astype(np.float64)), (ta, tb, exact)


def test_gemm_fast_threads_rowmajor_vs_numpy():
    O.set_threads(4)
    try:
        M, N, K = 301, 257, 513
        A = _f(RNG.uniform(-0.5, 0.5, (M, K)))
        B = _f(RNG.uniform(-0.5, 0.5, (K, N)))
        C0 = _f(RNG.uniform(-0.5, 0.5, (M, N)))
        Cm = C0.copy(order="F")
        O.gemm(O.NOTRANS, O.NOTRANS, 1.0, A, B, 0.5, Cm)
        assert np.allclose(Cm, A @ B + 0.5 * C0, atol=1e-12, rtol=0)
        # RowMajor: the same numbers held transposed (cblas_dgemm.c:133)
        Ar, Br = np.ascontiguousarray(A), np.ascontiguousarray(B)
        Cr = np.ascontiguousarray(C0)
        O.lib().orc_dgemm(O.ROW, O.NOTRANS, O.NOTRANS, M, N, K, 1.0, Ar.ctypes.data, K,
                          Br.ctypes.data, N, 0.5, Cr.ctypes.data, N)
        assert np.allclose(Cr, A @ B + 0.5 * C0, atol=1e-12, rtol=0)
    finally:
        O.set_threads(1)


def test_trsm_closed_forms_and_scipy():
    import scipy.linalg as sla
    n, m = 40, 6
    # diagonal A: X = alpha * B / d exactly for powers of two
    d = 2.0 ** RNG.integers(-3, 4, n)
    A = _f(np.diag(d))
    B0 = _f(RNG.uniform(-1, 1, (n, m)))
    X = O.trsm(O.LEFT, O.LOWER, O.NOTRANS, O.NONUNIT, 2.0, A, B0.copy(order="F"))
    assert np.array_equal(X, 2.0 * B0 / d[:, None])
    # general triangles against scipy.linalg.solve_triangular
    L = np.tril(RNG.uniform(-1, 1, (n, n))) + 4 * np.eye(n)
    for uplo, lower in ((O.LOWER, True), (O.UPPER, False)):
        T = _f(L if lower else L.T)
        for ta in (O.NOTRANS, O.TRANS):
            X = O.trsm(O.LEFT, uplo, ta, O.NONUNIT, 1.0, T, B0.copy(order="F"))
            want = sla.solve_triangular(T, B0, lower=lower, trans=0 if ta == O.NOTRANS else 1)
            assert np.allclose(X, want, atol=1e-12)
            Br = _f(RNG.uniform(-1, 1, (m, n)))
            X = O.trsm(O.RIGHT, uplo, ta, O.UNIT, 1.0, T, Br.copy(order="F"))
            Tu = np.tril(T, -1) + np.eye(n) if lower else np.triu(T, 1) + np.eye(n)
            opT = Tu if ta == O.NOTRANS else Tu.T
            assert np.allclose(X @ opT, Br, atol=1e-12)


def test_laswp_forward_backward_inverse():
    A0 = _f(RNG.uniform(-1, 1, (10, 70)))
    piv = np.array([3, 1, 9, 5, 4, 7, 6, 9, 8, 9], dtype=np.int32)
    A = O.laswp(A0.copy(order="F"), 0, 10, piv, 1)
    # forward application equals the explicit permutation product
    perm = np.arange(10)
    for i, p in enumerate(piv):
        perm[[i, p]] = perm[[p, i]]
    assert np.array_equal(A, A0[perm])
    B = O.laswp(A.copy(order="F"), 0, 10, piv, -1)
    assert np.array_equal(B, A0)


def _ref_fn(R, name, dt, nint, nscal_first=True):
    ct = C.c_double if dt == np.float64 else C.c_float
    return getattr(R, ("ATL_d" if dt == np.float64 else "ATL_s") + name), ct


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("alpha", [1.5, -0.75, 0.0])
def test_trmm_ref_bit_exact_vs_atlas_reftrmm(dt, alpha):
    """All 16 side/uplo/trans/diag variants, bit for bit against ATLAS's own
    reference TRMM (src/blas/reference/level3/ATL_dreftrmm*.c)."""
    R = _ref_or_skip()
    fn, ct = _ref_fn(R, "reftrmm", dt, 6)
    fn.argtypes = [C.c_int] * 6 + [ct, C.c_void_p, C.c_int, C.c_void_p, C.c_int]
    for side, uplo, ta, diag in itertools.product([O.LEFT, O.RIGHT], [O.UPPER, O.LOWER],
                                                  [O.NOTRANS, O.TRANS], [O.NONUNIT, O.UNIT]):
        M, N = 13, 9
        na = M if side == O.LEFT else N
        A = _f(RNG.uniform(-1, 1, (na + 3, na)).astype(dt))      # lda = na + 3
        B0 = _f(RNG.uniform(-1, 1, (M + 2, N)).astype(dt))       # ldb = M + 2
        b1, b2 = B0.copy(order="F"), B0.copy(order="F")
        O.trmm(side, uplo, ta, diag, alpha, A[:na, :na], b1[:M, :])
        fn(side, uplo, ta, diag, M, N, alpha, A.ctypes.data, na + 3, b2.ctypes.data, M + 2)
        assert np.array_equal(b1, b2), (side, uplo, ta, diag)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("alpha,beta", [(1.25, 0.5), (-1.0, 0.0), (0.7, 1.0), (0.0, 0.3)])
def test_symm_syr2k_bit_exact_vs_atlas_ref(dt, alpha, beta):
    R = _ref_or_skip()
    ct = C.c_double if dt == np.float64 else C.c_float
    symm = R.ATL_drefsymm if dt == np.float64 else R.ATL_srefsymm
    symm.argtypes = [C.c_int] * 4 + [ct, C.c_void_p, C.c_int, C.c_void_p, C.c_int, ct,
                                     C.c_void_p, C.c_int]
    for side, uplo in itertools.product([O.LEFT, O.RIGHT], [O.UPPER, O.LOWER]):
        M, N = 11, 7
        na = M if side == O.LEFT else N
        A = _f(RNG.uniform(-1, 1, (na, na)).astype(dt))
        B = _f(RNG.uniform(-1, 1, (M, N)).astype(dt))
        C0 = _f(RNG.uniform(-1, 1, (M, N)).astype(dt))
        c1, c2 = C0.copy(order="F"), C0.copy(order="F")
        O.symm(side, uplo, alpha, A, B, beta, c1)
        symm(side, uplo, M, N, alpha, A.ctypes.data, na, B.ctypes.data, M, beta,
             c2.ctypes.data, M)
        assert np.array_equal(c1, c2), (side, uplo)
    syr2k = R.ATL_drefsyr2k if dt == np.float64 else R.ATL_srefsyr2k
    syr2k.argtypes = [C.c_int] * 4 + [ct, C.c_void_p, C.c_int, C.c_void_p, C.c_int, ct,
                                      C.c_void_p, C.c_int]
    for uplo, tr in itertools.product([O.UPPER, O.LOWER], [O.NOTRANS, O.TRANS]):
        N, K = 10, 6
        shp = (N, K) if tr == O.NOTRANS else (K, N)
        A = _f(RNG.uniform(-1, 1, shp).astype(dt))
        B = _f(RNG.uniform(-1, 1, shp).astype(dt))
        C0 = _f(RNG.uniform(-1, 1, (N, N)).astype(dt))
        c1, c2 = C0.copy(order="F"), C0.copy(order="F")
        O.syr2k(uplo, tr, alpha, A, B, beta, c1)
        syr2k(uplo, tr, N, K, alpha, A.ctypes.data, shp[0], B.ctypes.data, shp[0], beta,
              c2.ctypes.data, N)
        assert np.array_equal(c1, c2), (uplo, tr)


def test_trmm_symm_syr2k_closed_forms_vs_numpy():
    """The restatements against the mathematics: dense products of the
    expanded triangle / symmetric matrix (independent of _ref)."""
    n, m = 17, 6
    for side, uplo, ta, diag in itertools.product([O.LEFT, O.RIGHT], [O.UPPER, O.LOWER],
                                                  [O.NOTRANS, O.TRANS], [O.NONUNIT, O.UNIT]):
        A = _f(RNG.uniform(-1, 1, (n, n)))
        Tm = np.tril(A) if uplo == O.LOWER else np.triu(A)
        if diag == O.UNIT:
            np.fill_diagonal(Tm, 1.0)
        opT = Tm if ta == O.NOTRANS else Tm.T
        B0 = _f(RNG.uniform(-1, 1, (n, m) if side == O.LEFT else (m, n)))
        got = O.trmm(side, uplo, ta, diag, 0.5, A, B0.copy(order="F"))
        want = 0.5 * (opT @ B0 if side == O.LEFT else B0 @ opT)
        assert np.allclose(got, want, atol=1e-13, rtol=0), (side, uplo, ta, diag)
    for side, uplo in itertools.product([O.LEFT, O.RIGHT], [O.UPPER, O.LOWER]):
        A = _f(RNG.uniform(-1, 1, (n, n)))
        S = np.tril(A) + np.tril(A, -1).T if uplo == O.LOWER else np.triu(A) + np.triu(A, 1).T
        B = _f(RNG.uniform(-1, 1, (n, m) if side == O.LEFT else (m, n)))
        C0 = _f(RNG.uniform(-1, 1, B.shape))
        got = O.symm(side, uplo, 1.5, A, B, -0.5, C0.copy(order="F"))
        want = 1.5 * (S @ B if side == O.LEFT else B @ S) - 0.5 * C0
        assert np.allclose(got, want, atol=1e-13, rtol=0), (side, uplo)
    for uplo, tr in itertools.product([O.UPPER, O.LOWER], [O.NOTRANS, O.TRANS]):
        A = _f(RNG.uniform(-1, 1, (n, m) if tr == O.NOTRANS else (m, n)))
        B = _f(RNG.uniform(-1, 1, A.shape))
        C0 = _f(RNG.uniform(-1, 1, (n, n)))
        got = O.syr2k(uplo, tr, 0.75, A, B, 2.0, C0.copy(order="F"))
        oA, oB = (A, B) if tr == O.NOTRANS else (A.T, B.T)
        full = 0.75 * (oA @ oB.T + oB @ oA.T) + 2.0 * C0
        tri = np.tril(np.ones((n, n), bool)) if uplo == O.LOWER else np.triu(np.ones((n, n), bool))
        assert np.allclose(got[tri], full[tri], atol=1e-13, rtol=0)
        assert np.array_equal(got[~tri], C0[~tri])          # other triangle untouched


def test_norm_helpers_vs_numpy():
    """genrm1 (max column sum), gediffnrm1, infnrm and lltnrm1 are the
    denominators of every lutst/llttst/slvtst threshold: pin them to numpy's
    norms on non-symmetric inputs, where a row-sum/column-sum mix-up shows."""
    A = _f(RNG.uniform(-1, 1, (23, 17)))
    A[:, 3] *= 9.0                           # one heavy column: norm1 != norm inf
    assert abs(np.linalg.norm(A, 1) - np.linalg.norm(A, np.inf)) > 1.0
    assert O.genrm1(A) == pytest.approx(np.linalg.norm(A, 1), rel=1e-14)
    # a planted mistake (the row-sum norm, i.e. genrm1 of the transpose)
    # must not pass the same check
    assert O.genrm1(_f(A.T)) != pytest.approx(np.linalg.norm(A, 1), rel=1e-3)
    B = _f(RNG.uniform(-1, 1, A.shape))
    assert O.gediffnrm1(A, B) == pytest.approx(np.linalg.norm(A - B, 1), rel=1e-14)
    x = RNG.uniform(-1, 1, 41)
    assert O.infnrm(x) == np.abs(x).max()
    x[7] = np.nan
    assert np.isnan(O.infnrm(x))             # ATL_infnrm.c:30-31 returns the NaN
    for uplo in (O.LOWER, O.UPPER):
        n = 19
        T = _f(RNG.uniform(-1, 1, (n, n)))
        T[:, 2] *= 7.0
        S = np.tril(T) + np.tril(T, -1).T if uplo == O.LOWER else np.triu(T) + np.triu(T, 1).T
        assert O.lltnrm1(uplo, T) == pytest.approx(np.linalg.norm(S, 1), rel=1e-14)
        # reading the other triangle (the planted mistake) gives another norm
        other = O.UPPER if uplo == O.LOWER else O.LOWER
        assert O.lltnrm1(other, T) != pytest.approx(np.linalg.norm(S, 1), rel=1e-3)
