"""Pins the oracle's complex BLAS restatement (oracle/orc_cplx.h).

* GEMM bit-exact against oracle/_ref's ATL_zrefgemm / ATL_crefgemm (ATLAS's
  own C reference BLAS, compiled from src/blas/reference where
  /root/reference exists; skipped elsewhere), every op pair with conjugates
  and the alpha/beta special cases;
* GEMM exact on small Gaussian-integer inputs (every product and partial sum
  is an integer below 2^24, so any summation order gives the same bits)
  against numpy int64 arithmetic -- this pins the conjugation and transpose
  conventions without the reference;
* TRSM against ATL_ztrsm's reference (ATL_zreftrsm / ATL_creftrsm) within
  rounding, all 24 cases, and against numpy's dense solve of op(A);
* the closed forms of HERK, HER2K, SYRK, SYR2K, HEMM, SYMM and TRMM that
  the GPU tests expect (tests/checks.py) against ATL_zref* / ATL_cref*;
* the unblocked complex LU (ATL_getf2) against scipy's LAPACK ?getrf, which
  chooses pivots by the same |Re| + |Im| rule: pivots equal away from near
  ties, factors within rounding, P A = L U always; zero-pivot info and the
  below-SAFMIN division branch by hand.
"""
import ctypes as C
import itertools

import numpy as np
import pytest

import oracle as O

RNG = np.random.default_rng(4321)
OPS = [O.NOTRANS, O.TRANS, O.CONJTRANS]


def _ref_or_skip():
    r = O.ref()
    if r is None:
        pytest.skip("oracle/_ref not built (needs /root/reference)")
    return r


def _cplx(shape, dt, lo=-0.5, hi=0.5):
    a = RNG.uniform(lo, hi, shape) + 1j * RNG.uniform(lo, hi, shape)
    return np.asfortranarray(a.astype(dt))


def _op(M, t):
    if t == O.NOTRANS:
        return M
    if t == O.TRANS:
        return M.T
    return M.conj().T


def _scal(x, rt):
    return np.array([complex(x).real, complex(x).imag], dtype=rt)


ALBE = [(1.0, 0.5 - 0.25j), (-0.7 + 0.3j, 0.0), (1.3, 1.0), (0.0, 1.3 + 0.2j), (0.0, 0.0),
        (0.5j, 1.0), (0.25 - 1.5j, -0.5j)]


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
@pytest.mark.parametrize("ta,tb", list(itertools.product(OPS, repeat=2)))
def test_cgemm_ref_bit_exact_vs_atlas_refgemm(dt, ta, tb):
    R = _ref_or_skip()
    rt = np.float64 if dt == np.complex128 else np.float32
    fn = R.ATL_zrefgemm if dt == np.complex128 else R.ATL_crefgemm
    fn.argtypes = [C.c_int] * 5 + [C.c_void_p] * 2 + [C.c_int, C.c_void_p, C.c_int,
                                                      C.c_void_p, C.c_void_p, C.c_int]
    M, N, K = 13, 9, 11
    A = _cplx((M, K) if ta == O.NOTRANS else (K, M), dt)
    B = _cplx((K, N) if tb == O.NOTRANS else (N, K), dt)
    C0 = _cplx((M, N), dt)
    for alpha, beta in ALBE:
        c1, c2 = C0.copy(order="F"), C0.copy(order="F")
        O.cgemm(ta, tb, alpha, A, B, beta, c1)
        al, be = _scal(alpha, rt), _scal(beta, rt)
        fn(ta, tb, M, N, K, al.ctypes.data, A.ctypes.data, A.shape[0], B.ctypes.data,
           B.shape[0], be.ctypes.data, c2.ctypes.data, M)
        assert np.array_equal(c1, c2), (alpha, beta)


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
def test_cgemm_ref_integer_exact_vs_numpy(dt):
    """Gaussian integers: the result is exact in any order, so the oracle must
    equal numpy's int64 product bit for bit for every op pair."""
    M, N, K = 17, 12, 23
    for ta, tb in itertools.product(OPS, repeat=2):
        Ai = RNG.integers(-4, 5, (M, K)) + 1j * RNG.integers(-4, 5, (M, K))
        Bi = RNG.integers(-4, 5, (K, N)) + 1j * RNG.integers(-4, 5, (K, N))
        Ci = RNG.integers(-4, 5, (M, N)) + 1j * RNG.integers(-4, 5, (M, N))
        # A and B are stored so that op(A) == Ai and op(B) == Bi
        A = np.asfortranarray((Ai if ta == O.NOTRANS else
                               (Ai.T if ta == O.TRANS else Ai.conj().T)).astype(dt))
        B = np.asfortranarray((Bi if tb == O.NOTRANS else
                               (Bi.T if tb == O.TRANS else Bi.conj().T)).astype(dt))
        alpha, beta = 2 - 1j, 1 + 3j
        got = O.cgemm(ta, tb, alpha, A, B, beta, np.asfortranarray(Ci.astype(dt)))
        want = alpha * (Ai @ Bi) + beta * Ci
        assert np.array_equal(got, want.astype(dt)), (ta, tb)


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
def test_cgemm_ref_vs_numpy_rounding(dt):
    tol = 1e-13 if dt == np.complex128 else 2e-5
    M, N, K = 40, 33, 57
    for ta, tb in itertools.product(OPS, repeat=2):
        A = _cplx((M, K) if ta == O.NOTRANS else (K, M), dt)
        B = _cplx((K, N) if tb == O.NOTRANS else (N, K), dt)
        C0 = _cplx((M, N), dt)
        alpha, beta = 0.75 - 0.5j, -0.25 + 1j
        got = O.cgemm(ta, tb, alpha, A, B, beta, C0.copy(order="F"))
        want = alpha * (_op(A.astype(np.complex128), ta) @ _op(B.astype(np.complex128), tb)) \
            + beta * C0.astype(np.complex128)
        assert np.max(np.abs(got - want)) <= tol * K, (ta, tb)


def _tri(n, uplo, dt):
    A = _cplx((n, n), dt)
    A[np.diag_indices(n)] += np.asarray(3.0 + 1.0j, dtype=dt)
    # the other triangle holds junk the solve must not read
    mask = np.tril(np.ones((n, n), bool), -1) if uplo == O.UPPER else \
        np.triu(np.ones((n, n), bool), 1)
    A[mask] = np.asarray(1e30 + 1e30j, dtype=dt)
    return A


def _dense_op(A, uplo, ta, diag):
    T = np.tril(A) if uplo == O.LOWER else np.triu(A)
    T = T.astype(np.complex128)
    if diag == O.UNIT:
        T[np.diag_indices(T.shape[0])] = 1.0
    return _op(T, ta)


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
def test_ctrsm_ref_vs_atlas_reftrsm_and_numpy(dt):
    R = O.ref()   # without oracle/_ref only the numpy pin runs
    rt = np.float64 if dt == np.complex128 else np.float32
    tol = 1e-12 if dt == np.complex128 else 5e-5
    fn = None
    if R is not None:
        fn = R.ATL_zreftrsm if dt == np.complex128 else R.ATL_creftrsm
        fn.argtypes = [C.c_int] * 6 + [C.c_void_p] * 2 + [C.c_int, C.c_void_p, C.c_int]
    for side, uplo, ta, diag in itertools.product([O.LEFT, O.RIGHT], [O.UPPER, O.LOWER], OPS,
                                                  [O.NONUNIT, O.UNIT]):
        M, N = 11, 7
        na = M if side == O.LEFT else N
        A = _tri(na, uplo, dt)
        B0 = _cplx((M, N), dt)
        for alpha in (1.0, -0.6 + 0.8j, 0.0):
            b1 = O.ctrsm(side, uplo, ta, diag, alpha, A, B0.copy(order="F"))
            if fn is not None:
                b2 = B0.copy(order="F")
                al = _scal(alpha, rt)
                fn(side, uplo, ta, diag, M, N, al.ctypes.data, A.ctypes.data, na,
                   b2.ctypes.data, M)
                assert np.max(np.abs(b1 - b2)) <= tol, (side, uplo, ta, diag, alpha)
            E = _dense_op(A, uplo, ta, diag)
            if side == O.LEFT:
                want = np.linalg.solve(E, alpha * B0.astype(np.complex128))
            else:
                want = np.linalg.solve(E.T, (alpha * B0.astype(np.complex128)).T).T
            assert np.max(np.abs(b1 - want)) <= tol, (side, uplo, ta, diag, alpha)


def _perm_from_ipiv(ipiv, m):
    p = np.arange(m)
    for k, q in enumerate(ipiv):
        p[[k, q]] = p[[q, k]]
    return p


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
@pytest.mark.parametrize("M,N", [(1, 1), (7, 7), (40, 25), (25, 40), (120, 120)])
def test_cgetf2_vs_scipy_lu_factor(dt, M, N):
    """scipy's LAPACK ?getrf picks pivots by izamax (|Re| + |Im|, first max),
    the rule ATL_getf2 uses: pivots agree where no near tie exists, and the
    factors agree within rounding."""
    import scipy.linalg as sl
    tol = 1e-12 if dt == np.complex128 else 2e-4
    for seed in range(3):
        rng = np.random.default_rng(seed + 10 * M + N)
        A0 = np.asfortranarray((rng.uniform(-1, 1, (M, N)) +
                                1j * rng.uniform(-1, 1, (M, N))).astype(dt))
        A = A0.copy(order="F")
        info, ipiv, gap = O.cgetf2(A)
        assert info == 0
        lu, piv = sl.lu_factor(A0.astype(np.complex128), check_finite=False)
        if gap > 1e-6:
            assert np.array_equal(ipiv, piv[:min(M, N)]), seed
            assert np.max(np.abs(A - lu)) <= tol * 10, seed
        # P A = L U whatever the ties
        k = min(M, N)
        L = np.tril(A[:, :k].astype(np.complex128), -1) + np.eye(M, k)
        U = np.triu(A[:k, :].astype(np.complex128))
        P = _perm_from_ipiv(ipiv, M)
        assert np.max(np.abs(A0[P].astype(np.complex128) - L @ U)) <= tol * max(M, N)


def test_cgetf2_zero_pivot_info_and_safmin_division():
    # an exactly zero column 3: info = 4, the elimination carries on
    A = np.asfortranarray(RNG.uniform(-1, 1, (6, 6)) + 1j * RNG.uniform(-1, 1, (6, 6)))
    A[:, 3] = 0
    info, ipiv, _ = O.cgetf2(A)
    assert info == 4 and ipiv[3] == 3
    # a pivot below SAFMIN is divided into the column, not inverted
    tiny = np.finfo(np.float64).tiny
    B = np.asfortranarray(np.array([[tiny * 0.25 + 0j, 1.0], [tiny * 0.125j, 2.0]]))
    info, ipiv, _ = O.cgetf2(B)
    assert info == 0 and ipiv[0] == 0
    assert abs(B[1, 0] - 0.5j) < 1e-15


def _hpd(n, dt, seed):
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return np.asfortranarray((B @ B.conj().T / n + np.eye(n)).astype(dt))


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
@pytest.mark.parametrize("uplo", [O.LOWER, O.UPPER])
def test_cpotrf_vs_numpy_cholesky(dt, uplo):
    tol = 1e-12 if dt == np.complex128 else 2e-5
    for n in (1, 2, 7, 40, 130):
        A0 = _hpd(n, dt, n)
        A = A0.copy(order="F")
        assert O.cpotrf(A, uplo) == 0
        L = np.linalg.cholesky(A0.astype(np.complex128))
        if uplo == O.LOWER:
            got = np.tril(A)
            assert np.array_equal(np.triu(A, 1), np.triu(A0, 1))    # other triangle untouched
        else:
            got = np.triu(A).conj().T
            assert np.array_equal(np.tril(A, -1), np.tril(A0, -1))
        assert np.max(np.abs(got - L)) <= tol * n


def test_cpotrf_not_positive_definite_info():
    A = _hpd(12, np.complex128, 3)
    A[5, 5] = -1.0                       # leading minor of order 6 fails
    assert O.cpotrf(A.copy(order="F"), O.LOWER) == 6
    assert O.cpotrf(A.copy(order="F"), O.UPPER) == 6
    A = _hpd(4, np.complex128, 4)
    A[0, 0] = 0.0
    assert O.cpotrf(A, O.LOWER) == 1


# --- the closed forms the complex rank-k / HEMM / SYMM / TRMM GPU tests
# expect (tests/checks.py) against ATLAS's own reference routines: the
# conjugate on her2k's second alpha, the real Hermitian diagonal (whose
# imaginary part is never read), beta on one triangle only, the quick
# returns and beta = 0 not reading C.
from tests.checks import crk_np, chemm_np, ctrmm_np   # noqa: E402

CPTR = C.c_void_p


def _rk_ref(R, dt, herm, two):
    z = dt == np.complex128
    name = ("ATL_%sref" % ("z" if z else "c")) + ("her" if herm else "syr") + \
        ("2k" if two else "k")
    f = getattr(R, name)
    rt = C.c_double if z else C.c_float
    al = rt if (herm and not two) else CPTR
    be = rt if herm else CPTR
    f.argtypes = [C.c_int] * 4 + [al, CPTR, C.c_int] + ([CPTR, C.c_int] if two else []) + \
        [be, CPTR, C.c_int]
    return f


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
@pytest.mark.parametrize("herm,two", [(1, 0), (1, 1), (0, 0), (0, 1)])
def test_crk_closed_form_vs_atlas_ref(dt, herm, two):
    R = _ref_or_skip()
    f = _rk_ref(R, dt, herm, two)
    rt = np.float64 if dt == np.complex128 else np.float32
    tol = 1e-12 if dt == np.complex128 else 2e-5
    alt = O.CONJTRANS if herm else O.TRANS
    n, K = 9, 6
    for uplo, trans in itertools.product([O.UPPER, O.LOWER], [O.NOTRANS, alt]):
        shp = (n, K) if trans == O.NOTRANS else (K, n)
        A, B = _cplx(shp, dt), (_cplx(shp, dt) if two else None)
        for alpha, beta in [(0.7 - 0.4j, 0.5 + 0.3j), (1.0, 1.0), (0.0, 1.0), (0.0, -0.5),
                            (0.0, 0.0), (-1.2 + 0.0j, 0.0), (0.25j, 1.0)]:
            if herm and not two:
                alpha = alpha.real if isinstance(alpha, complex) else alpha
            if herm:
                beta = complex(beta).real
            C0 = _cplx((n, n), dt)
            off = np.triu(np.ones((n, n), bool), 1) if uplo == O.LOWER else \
                np.tril(np.ones((n, n), bool), -1)
            C0[off] = np.asarray(1e30 - 2e30j, dtype=dt)
            if beta == 0:
                C0[~off] = np.nan
            want = crk_np(herm, uplo, trans, alpha, A, B, beta, C0)
            got = C0.copy(order="F")
            al_keep = _scal(alpha, rt)      # held until the call returns
            a = rt(complex(alpha).real) if (herm and not two) else al_keep.ctypes.data
            be_keep = _scal(beta, rt)
            b = rt(complex(beta).real) if herm else be_keep.ctypes.data
            args = [uplo, trans, n, K, a, A.ctypes.data, A.shape[0]]
            if two:
                args += [B.ctypes.data, B.shape[0]]
            f(*(args + [b, got.ctypes.data, n]))
            assert np.array_equal(got[off], C0[off]), (uplo, trans, alpha, beta)
            assert np.all(np.isfinite(got[~off])), (uplo, trans, alpha, beta)
            assert np.max(np.abs(got - want)[~off]) <= tol, (uplo, trans, alpha, beta)
            if herm:
                assert np.all(np.diag(got).imag == 0) or (alpha == 0 and beta == 1)


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
@pytest.mark.parametrize("herm", [1, 0])
def test_chemm_closed_form_vs_atlas_ref(dt, herm):
    R = _ref_or_skip()
    z = dt == np.complex128
    f = getattr(R, "ATL_%sref%s" % ("z" if z else "c", "hemm" if herm else "symm"))
    f.argtypes = [C.c_int] * 4 + [CPTR, CPTR, C.c_int, CPTR, C.c_int, CPTR, CPTR, C.c_int]
    rt = np.float64 if z else np.float32
    tol = 1e-12 if z else 2e-5
    M, N = 8, 5
    for side, uplo in itertools.product([O.LEFT, O.RIGHT], [O.UPPER, O.LOWER]):
        na = M if side == O.LEFT else N
        A = _cplx((na, na), dt)
        off = np.triu(np.ones((na, na), bool), 1) if uplo == O.LOWER else \
            np.tril(np.ones((na, na), bool), -1)
        A[off] = np.asarray(1e30 + 1e30j, dtype=dt)     # never read
        if herm:
            A[np.diag_indices(na)] += np.asarray(1e30j, dtype=dt)   # never read
        B = _cplx((M, N), dt)
        for alpha, beta in [(0.6 + 0.2j, -0.3 + 0.9j), (1.0, 0.0), (0.0, 0.5j)]:
            C0 = _cplx((M, N), dt)
            if beta == 0:
                C0[:] = np.nan
            want = chemm_np(herm, side, uplo, alpha, A, B, beta, C0)
            got = C0.copy(order="F")
            al, be = _scal(alpha, rt), _scal(beta, rt)
            f(side, uplo, M, N, al.ctypes.data, A.ctypes.data, na, B.ctypes.data, M,
              be.ctypes.data, got.ctypes.data, M)
            assert np.max(np.abs(got - want)) <= tol, (side, uplo, alpha, beta)


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
def test_ctrmm_closed_form_vs_atlas_ref(dt):
    R = _ref_or_skip()
    z = dt == np.complex128
    f = R.ATL_zreftrmm if z else R.ATL_creftrmm
    f.argtypes = [C.c_int] * 6 + [CPTR, CPTR, C.c_int, CPTR, C.c_int]
    rt = np.float64 if z else np.float32
    tol = 1e-12 if z else 2e-5
    M, N = 7, 6
    for side, uplo, ta, diag in itertools.product([O.LEFT, O.RIGHT], [O.UPPER, O.LOWER], OPS,
                                                  [O.NONUNIT, O.UNIT]):
        na = M if side == O.LEFT else N
        A = _tri(na, uplo, dt)
        B0 = _cplx((M, N), dt)
        for alpha in (1.0, -0.6 + 0.8j):
            want = ctrmm_np(side, uplo, ta, diag, alpha, A, B0)
            got = B0.copy(order="F")
            al = _scal(alpha, rt)
            f(side, uplo, ta, diag, M, N, al.ctypes.data, A.ctypes.data, na, got.ctypes.data, M)
            assert np.max(np.abs(got - want)) <= tol * 10, (side, uplo, ta, diag, alpha)
