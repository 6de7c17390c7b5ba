"""The Fortran 77 ABI (include/atlas_f77.h) through ctypes with the
gfortran calling convention: dgemm_/sgemm_, dtrsm_, dtrmm_, dsyrk_,
dsyr2k_, dsymm_, idamax_ against the oracle; dgetrf_ returning 1-based
pivots equal to clapack_dgetrf's plus one
(interfaces/lapack/F77/src/f77wrap/ATL_f77wrap_getrf.c:27), dgetrs_ /
dgesv_ / dpotrf_ / dpotrs_ / dposv_ under slvtst, and the inverses."""
import ctypes as C

import numpy as np
import pytest

import atlas_amd as X
import oracle as O
from tests.checks import solve_ok

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(77)
I = C.c_int


def _f(a):
    return np.asfortranarray(a)


def _i(v):
    return C.byref(I(v))


def _d(v, dt=np.float64):
    return C.byref(C.c_double(v) if dt == np.float64 else C.c_float(v))


def _p(a):
    return C.c_void_p(a.ctypes.data)


def _c(ch):
    return C.c_char_p(ch.encode())


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("t", "N"), ("N", "T"), ("C", "t")])
def test_gemm_f77(dt, ta, tb):
    M, N, K = 130, 70, 90
    A = _f(RNG.uniform(-0.5, 0.5, (M, K) if ta == "N" else (K, M)).astype(dt))
    B = _f(RNG.uniform(-0.5, 0.5, (K, N) if tb == "N" else (N, K)).astype(dt))
    C0 = _f(RNG.uniform(-0.5, 0.5, (M, N)).astype(dt))
    D = C0.copy(order="F")
    f = X.lib().dgemm_ if dt == np.float64 else X.lib().sgemm_
    f(_c(ta), _c(tb), _i(M), _i(N), _i(K), _d(1.5, dt), _p(A), _i(A.shape[0]), _p(B),
      _i(B.shape[0]), _d(-0.5, dt), _p(D), _i(M), C.c_size_t(1), C.c_size_t(1))
    T = O.gemm(X.NoTrans if ta == "N" else X.Trans, X.NoTrans if tb == "N" else X.Trans, 1.5, A,
               B, -0.5, C0.copy(order="F"), exact_order=True)
    assert O.gemm_ferr_check(K, 1.5, -0.5, T, D)[0] == 0


def test_level3_f77_vs_oracle():
    L = X.lib()
    n, m = 120, 50
    A = _f(RNG.uniform(-1, 1, (n, n)) + 4 * np.eye(n))
    B0 = _f(RNG.uniform(-1, 1, (n, m)))
    D = B0.copy(order="F")
    L.dtrsm_(_c("L"), _c("U"), _c("T"), _c("N"), _i(n), _i(m), _d(2.0), _p(A), _i(n), _p(D),
             _i(n), C.c_size_t(1), C.c_size_t(1), C.c_size_t(1), C.c_size_t(1))
    assert O.trsm_resid(X.Left, X.Upper, X.Trans, X.NonUnit, 2.0, A, D, B0) < 1
    D = B0.copy(order="F")
    L.dtrmm_(_c("l"), _c("L"), _c("n"), _c("u"), _i(n), _i(m), _d(0.5), _p(A), _i(n), _p(D),
             _i(n), C.c_size_t(1), C.c_size_t(1), C.c_size_t(1), C.c_size_t(1))
    T = O.trmm(X.Left, X.Lower, X.NoTrans, X.Unit, 0.5, A, B0.copy(order="F"))
    assert O.gemm_ferr_check(n, 0.5, 0.0, T, D)[0] == 0
    Cs = _f(RNG.uniform(-1, 1, (n, n)))
    D = Cs.copy(order="F")
    L.dsyrk_(_c("U"), _c("N"), _i(n), _i(m), _d(1.0), _p(B0), _i(n), _d(0.5), _p(D), _i(n),
             C.c_size_t(1), C.c_size_t(1))
    T = O.syrk(X.Upper, X.NoTrans, 1.0, B0, 0.5, Cs.copy(order="F"))
    assert np.allclose(np.triu(D), np.triu(T), atol=1e-12)
    B2 = _f(RNG.uniform(-1, 1, (n, m)))
    D = Cs.copy(order="F")
    L.dsyr2k_(_c("L"), _c("N"), _i(n), _i(m), _d(1.0), _p(B0), _i(n), _p(B2), _i(n), _d(0.0),
              _p(D), _i(n), C.c_size_t(1), C.c_size_t(1))
    T = O.syr2k(X.Lower, X.NoTrans, 1.0, B0, B2, 0.0, Cs.copy(order="F"))
    assert np.allclose(np.tril(D), np.tril(T), atol=1e-12)
    Cm = _f(RNG.uniform(-1, 1, (n, m)))
    D = Cm.copy(order="F")
    L.dsymm_(_c("L"), _c("U"), _i(n), _i(m), _d(1.0), _p(A), _i(n), _p(B0), _i(n), _d(1.0),
             _p(D), _i(n), C.c_size_t(1), C.c_size_t(1))
    T = O.symm(X.Left, X.Upper, 1.0, A, B0, 1.0, Cm.copy(order="F"))
    assert O.gemm_ferr_check(n, 1.0, 1.0, T, D)[0] == 0
    x = np.ascontiguousarray(RNG.uniform(-1, 1, 999))
    L.idamax_.restype = I
    assert L.idamax_(_i(999), _p(x), _i(1)) == O.iamax(x) + 1
    assert L.idamax_(_i(0), _p(x), _i(1)) == 0


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_lapack_f77_pivots_solves_inverses(dt):
    L = X.lib()
    pre = "d" if dt == np.float64 else "s"
    fn = lambda name: getattr(L, pre + name + "_")   # noqa: E731
    n, nrhs = 300, 3
    A0 = O.gegen(n, n, dtype=dt)
    A = A0.copy(order="F")
    ipiv = np.zeros(n, dtype=np.int32)
    info = I(-9)
    fn("getrf")(_i(n), _i(n), _p(A), _i(n), _p(ipiv), C.byref(info))
    assert info.value == 0
    A2 = A0.copy(order="F")
    inf2, ipiv0 = X.getrf(A2)
    assert np.array_equal(ipiv, ipiv0 + 1) and np.array_equal(A, A2)
    Xc = O.gegen(n, nrhs, dtype=dt)
    B = _f((A0.astype(np.float64) @ Xc).astype(dt))
    fn("getrs")(_c("N"), _i(n), _i(nrhs), _p(A), _i(n), _p(ipiv), _p(B), _i(n), C.byref(info),
                C.c_size_t(1))
    assert info.value == 0 and solve_ok(O.genrm1(A0), Xc, B)
    G = A0.copy(order="F")
    B = _f((A0.astype(np.float64) @ Xc).astype(dt))
    piv = np.zeros(n, dtype=np.int32)
    fn("gesv")(_i(n), _i(nrhs), _p(G), _i(n), _p(piv), _p(B), _i(n), C.byref(info))
    assert info.value == 0 and np.array_equal(piv, ipiv) and solve_ok(O.genrm1(A0), Xc, B)
    work = np.zeros(1, dtype=dt)
    fn("getri")(_i(n), _p(A), _i(n), _p(ipiv), _p(work), _i(-1), C.byref(info))
    assert info.value == 0 and work[0] >= n
    work = np.zeros(int(work[0]), dtype=dt)
    fn("getri")(_i(n), _p(A), _i(n), _p(ipiv), _p(work), _i(work.size), C.byref(info))
    assert info.value == 0 and O.inv_resid(A0, A) < 100
    for u, uplo in (("L", X.Lower), ("u", X.Upper)):
        S = O.lltgen(uplo, n, dtype=dt)
        Sf = np.tril(S) + np.tril(S, -1).T if uplo == X.Lower else np.triu(S) + np.triu(S, 1).T
        F = S.copy(order="F")
        fn("potrf")(_c(u), _i(n), _p(F), _i(n), C.byref(info), C.c_size_t(1))
        assert info.value == 0 and O.llt_resid(uplo, F, S) < 100
        B = _f((Sf.astype(np.float64) @ Xc).astype(dt))
        fn("potrs")(_c(u), _i(n), _i(nrhs), _p(F), _i(n), _p(B), _i(n), C.byref(info),
                    C.c_size_t(1))
        assert info.value == 0 and solve_ok(O.genrm1(_f(Sf)), Xc, B)
        F2 = S.copy(order="F")
        B = _f((Sf.astype(np.float64) @ Xc).astype(dt))
        fn("posv")(_c(u), _i(n), _i(nrhs), _p(F2), _i(n), _p(B), _i(n), C.byref(info),
                   C.c_size_t(1))
        assert info.value == 0 and solve_ok(O.genrm1(_f(Sf)), Xc, B)
        fn("potri")(_c(u), _i(n), _p(F), _i(n), C.byref(info), C.c_size_t(1))
        assert info.value == 0 and O.poinv_resid(uplo, S, F) < 100
        R = S.copy(order="F")
        fn("lauum")(_c(u), _i(n), _p(R), _i(n), C.byref(info), C.c_size_t(1))
        assert info.value == 0 and O.lauum_resid(uplo, R, S) < 100
        Tm = O.trigen(uplo, X.NonUnit, n, dtype=dt)
        Ti = Tm.copy(order="F")
        fn("trtri")(_c(u), _c("N"), _i(n), _p(Ti), _i(n), C.byref(info), C.c_size_t(1),
                    C.c_size_t(1))
        assert info.value == 0 and O.trtri_resid(uplo, X.NonUnit, Ti, Tm) < 100


def _z(v, dt):
    """a complex scalar by reference, as {real, imaginary}"""
    return np.array([complex(v).real, complex(v).imag],
                    dtype=np.float64 if dt == np.complex128 else np.float32)


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
def test_complex_f77_gemm_trsm_lu_cholesky(dt):
    """zgemm_/ztrsm_ (and c) against the oracle; zgetrf_ pivots = clapack's
    plus one; zgetrs_ ('C'), zgesv_, zpotrf_/zpotrs_/zposv_ under slvtst."""
    p = "z" if dt == np.complex128 else "c"
    L = X.lib()
    fn = lambda name: getattr(L, p + name + "_")  # noqa: E731
    tol = 1e-10 if dt == np.complex128 else 2e-3
    cz = lambda shape: _f((RNG.uniform(-0.5, 0.5, shape) +  # noqa: E731
                           1j * RNG.uniform(-0.5, 0.5, shape)).astype(dt))
    M, N, K = 70, 50, 40
    A, B, C0 = cz((K, M)), cz((K, N)), cz((M, N))
    D = C0.copy(order="F")
    al, be = _z(1 - 0.5j, dt), _z(0.25j, dt)
    fn("gemm")(_c("C"), _c("N"), _i(M), _i(N), _i(K), _p(al), _p(A), _i(K), _p(B), _i(K),
               _p(be), _p(D), _i(M), C.c_size_t(1), C.c_size_t(1))
    T = O.cgemm(O.CONJTRANS, O.NOTRANS, 1 - 0.5j, A, B, 0.25j, C0.copy(order="F"))
    assert np.max(np.abs(D - T)) <= tol
    Tr = cz((M, M))
    Tr[np.diag_indices(M)] += 2
    Bs = cz((M, 5))
    D = Bs.copy(order="F")
    fn("trsm")(_c("L"), _c("L"), _c("C"), _c("N"), _i(M), _i(5), _p(_z(1.0, dt)), _p(Tr), _i(M),
               _p(D), _i(M), C.c_size_t(1), C.c_size_t(1), C.c_size_t(1), C.c_size_t(1))
    T = O.ctrsm(O.LEFT, O.LOWER, O.CONJTRANS, O.NONUNIT, 1.0, Tr, Bs.copy(order="F"))
    assert np.max(np.abs(D - T)) <= tol * max(1.0, np.max(np.abs(T)))
    n, nrhs = 120, 3
    A0 = cz((n, n))
    A0[np.diag_indices(n)] += 1
    B0 = cz((n, nrhs))
    A1 = A0.copy(order="F")
    ip = np.zeros(n, dtype=np.int32)
    info = C.c_int(-7)
    fn("getrf")(_i(n), _i(n), _p(A1), _i(n), _p(ip), C.byref(info))
    A2 = A0.copy(order="F")
    ip0 = np.zeros(n, dtype=np.int32)
    assert getattr(L, "clapack_%sgetrf" % p)(X.ColMajor, n, n, A2.ctypes.data, n,
                                              ip0.ctypes.data) == 0
    assert info.value == 0 and np.array_equal(ip, ip0 + 1)
    Bx = B0.copy(order="F")
    fn("getrs")(_c("C"), _i(n), _i(nrhs), _p(A1), _i(n), _p(ip), _p(Bx), _i(n), C.byref(info),
                C.c_size_t(1))
    assert info.value == 0
    assert np.max(np.abs(A0.conj().T.astype(np.complex128) @ Bx - B0)) <= tol * n
    A3, Bx = A0.copy(order="F"), B0.copy(order="F")
    fn("gesv")(_i(n), _i(nrhs), _p(A3), _i(n), _p(ip), _p(Bx), _i(n), C.byref(info))
    assert info.value == 0 and np.array_equal(ip, ip0 + 1)
    assert np.max(np.abs(A0.astype(np.complex128) @ Bx - B0)) <= tol * n
    G = cz((n, n)).astype(np.complex128)
    H0 = _f((G @ G.conj().T / n + np.eye(n)).astype(dt))
    for uplo in ("L", "U"):
        H, Bx = H0.copy(order="F"), B0.copy(order="F")
        fn("posv")(_c(uplo), _i(n), _i(nrhs), _p(H), _i(n), _p(Bx), _i(n), C.byref(info),
                   C.c_size_t(1))
        assert info.value == 0
        assert np.max(np.abs(H0.astype(np.complex128) @ Bx - B0)) <= tol * n
        H2, By = H0.copy(order="F"), B0.copy(order="F")
        fn("potrf")(_c(uplo), _i(n), _p(H2), _i(n), C.byref(info), C.c_size_t(1))
        assert info.value == 0
        fn("potrs")(_c(uplo), _i(n), _i(nrhs), _p(H2), _i(n), _p(By), _i(n), C.byref(info),
                    C.c_size_t(1))
        assert np.max(np.abs(By - Bx)) <= tol * max(1.0, np.max(np.abs(Bx)))


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
def test_complex_f77_rank_k_hemm_symm_trmm(dt):
    """ztrmm_, zsyrk_, zsyr2k_, zsymm_, zhemm_, zherk_, zher2k_ (and c)
    against cblas_z* through the C ABI: the F77 wrappers map their CHARACTER
    arguments and scalars onto the same calls, so the results agree bit for
    bit; HERK's alpha/beta and HER2K's beta go by reference as one real."""
    p, rt = X._cpre(dt)
    L = X.lib()
    fn = lambda name: getattr(L, p + name + "_")  # noqa: E731
    cb = lambda name: getattr(L, "cblas_" + p + name)  # noqa: E731
    cz = lambda shape: _f((RNG.uniform(-0.5, 0.5, shape) +  # noqa: E731
                           1j * RNG.uniform(-0.5, 0.5, shape)).astype(dt))
    rs = lambda v: C.byref(C.c_double(v) if dt == np.complex128 else C.c_float(v))  # noqa: E731
    n, K, M = 60, 37, 45
    al, be = _z(0.8 - 0.3j, dt), _z(-0.4 + 0.2j, dt)
    s1 = C.c_size_t(1)
    for uplo, tr, herm in [("U", "N", 1), ("L", "C", 1), ("L", "T", 0), ("U", "N", 0)]:
        A = cz((n, K) if tr == "N" else (K, n))
        B = cz(A.shape)
        lda = A.shape[0]
        U = X.Upper if uplo == "U" else X.Lower
        T = {"N": X.NoTrans, "T": X.Trans, "C": X.ConjTrans}[tr]
        C0 = cz((n, n))
        g1, g2 = C0.copy(order="F"), C0.copy(order="F")
        if herm:
            fn("herk")(_c(uplo), _c(tr), _i(n), _i(K), rs(0.7), _p(A), _i(lda), rs(-0.5), _p(g1),
                       _i(n), s1, s1)
            cb("herk")(X.ColMajor, U, T, n, K, rt(0.7), _p(A), lda, rt(-0.5), _p(g2), n)
        else:
            fn("syrk")(_c(uplo), _c(tr), _i(n), _i(K), _p(al), _p(A), _i(lda), _p(be), _p(g1),
                       _i(n), s1, s1)
            cb("syrk")(X.ColMajor, U, T, n, K, _p(al), _p(A), lda, _p(be), _p(g2), n)
        assert np.array_equal(g1, g2), ("rk", uplo, tr, herm)
        g1, g2 = C0.copy(order="F"), C0.copy(order="F")
        if herm:
            fn("her2k")(_c(uplo), _c(tr), _i(n), _i(K), _p(al), _p(A), _i(lda), _p(B), _i(lda),
                        rs(0.25), _p(g1), _i(n), s1, s1)
            cb("her2k")(X.ColMajor, U, T, n, K, _p(al), _p(A), lda, _p(B), lda, rt(0.25), _p(g2),
                        n)
        else:
            fn("syr2k")(_c(uplo), _c(tr), _i(n), _i(K), _p(al), _p(A), _i(lda), _p(B), _i(lda),
                        _p(be), _p(g1), _i(n), s1, s1)
            cb("syr2k")(X.ColMajor, U, T, n, K, _p(al), _p(A), lda, _p(B), lda, _p(be), _p(g2), n)
        assert np.array_equal(g1, g2), ("r2k", uplo, tr, herm)
    for side, uplo in [("L", "U"), ("R", "L")]:
        na = M if side == "L" else n
        A, B, C0 = cz((na, na)), cz((M, n)), cz((M, n))
        S = X.Left if side == "L" else X.Right
        U = X.Upper if uplo == "U" else X.Lower
        for name in ("hemm", "symm"):
            g1, g2 = C0.copy(order="F"), C0.copy(order="F")
            fn(name)(_c(side), _c(uplo), _i(M), _i(n), _p(al), _p(A), _i(na), _p(B), _i(M),
                     _p(be), _p(g1), _i(M), s1, s1)
            cb(name)(X.ColMajor, S, U, M, n, _p(al), _p(A), na, _p(B), M, _p(be), _p(g2), M)
            assert np.array_equal(g1, g2), (name, side, uplo)
        for ta, dg in [("N", "N"), ("C", "U"), ("T", "N")]:
            g1, g2 = B.copy(order="F"), B.copy(order="F")
            fn("trmm")(_c(side), _c(uplo), _c(ta), _c(dg), _i(M), _i(n), _p(al), _p(A), _i(na),
                       _p(g1), _i(M), s1, s1, s1, s1)
            cb("trmm")(X.ColMajor, S, U, {"N": X.NoTrans, "T": X.Trans, "C": X.ConjTrans}[ta],
                       X.Unit if dg == "U" else X.NonUnit, M, n, _p(al), _p(A), na, _p(g2), M)
            assert np.array_equal(g1, g2), ("trmm", side, uplo, ta, dg)
            assert np.all(np.isfinite(g1))
