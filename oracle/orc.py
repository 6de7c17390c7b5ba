"""ctypes bindings for the oracle (test infrastructure only).

Arrays are numpy, column-major (Fortran order) unless a function says
otherwise.  Enum values are ATLAS's (include/cblas.h:5-10).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROW, COL = 101, 102
NOTRANS, TRANS, CONJTRANS = 111, 112, 113
UPPER, LOWER = 121, 122
NONUNIT, UNIT = 131, 132
LEFT, RIGHT = 141, 142

_lib = None
_ref = None


def lib():
    """The C restatement (oracle/liboracle.so)."""
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        _lib = C.CDLL(path)
        _declare(_lib)
    return _lib


def ref():
    """ATLAS's own C reference BLAS, or None where it was not built."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libatlref.so")
        if not os.path.exists(path):
            return None
        _ref = C.CDLL(path)
    return _ref


_I = C.c_int
_D = C.c_double
_F = C.c_float
_P = C.c_void_p


def _declare(L):
    L.orc_lcg_seed.argtypes = [_P, _I]
    L.orc_lcg_next.restype = _I
    L.orc_rand.restype = _I
    L.orc_dumb_rand.restype = _D
    L.orc_rand_stream.argtypes = [_I, _I, _P]
    L.orc_set_recursion_params.argtypes = [_I, _I]
    L.orc_set_threads.argtypes = [_I]
    L.orc_get_threads.restype = _I
    L.orc_track_pivot_gaps.argtypes = [_P, _I]
    for p, t in (("orc_d", _D), ("orc_s", _F)):
        g = lambda n: getattr(L, p + n)  # noqa: E731
        g("gegen").argtypes = [_I, _I, _P, _I, _I]
        g("matgen").argtypes = [_I, _I, _P, _I, _I]
        g("gechkgap").argtypes = [_I, _I, _P, _I]
        g("gechkgap").restype = _I
        g("epsilon").restype = t
        g("genrm1").argtypes = [_I, _I, _P, _I]
        g("genrm1").restype = t
        g("gediffnrm1").argtypes = [_I, _I, _P, _I, _P, _I]
        g("gediffnrm1").restype = t
        g("infnrm").argtypes = [_I, _P, _I]
        g("infnrm").restype = t
        g("iamax").argtypes = [_I, _P, _I]
        g("iamax").restype = _I
        for n in ("gemm_ref", "gemm_fast"):
            g(n).argtypes = [_I, _I, _I, _I, _I, t, _P, _I, _P, _I, t, _P, _I]
        g("gemm").argtypes = [_I, _I, _I, _I, _I, _I, t, _P, _I, _P, _I, t, _P, _I]
        for n in ("trsm_ref", "trmm_ref"):
            g(n).argtypes = [_I, _I, _I, _I, _I, _I, t, _P, _I, _P, _I]
        g("trsm").argtypes = [_I, _I, _I, _I, _I, _I, _I, t, _P, _I, _P, _I]
        g("syrk_ref").argtypes = [_I, _I, _I, _I, t, _P, _I, t, _P, _I]
        g("trmm_fast").argtypes = [_I, _I, _I, _I, _I, _I, t, _P, _I, _P, _I]
        g("symm_ref").argtypes = [_I, _I, _I, _I, t, _P, _I, _P, _I, t, _P, _I]
        g("syr2k_ref").argtypes = [_I, _I, _I, _I, t, _P, _I, _P, _I, t, _P, _I]
        g("laswp").argtypes = [_I, _P, _I, _I, _I, _P, _I]
        g("getf2").argtypes = [_I, _I, _P, _I, _P]
        g("getf2").restype = _I
        g("getrf").argtypes = [_I, _I, _I, _P, _I, _P]
        g("getrf").restype = _I
        g("getrs").argtypes = [_I, _I, _I, _I, _P, _I, _P, _P, _I]
        g("potrf").argtypes = [_I, _I, _I, _P, _I]
        g("potrf").restype = _I
        g("potrs").argtypes = [_I, _I, _I, _I, _P, _I, _P, _I]
        g("gesv").argtypes = [_I, _I, _I, _P, _I, _P, _P, _I]
        g("gesv").restype = _I
        g("posv").argtypes = [_I, _I, _I, _I, _P, _I, _P, _I]
        g("posv").restype = _I
        g("lu_resid").argtypes = [_I, _I, _P, _I, _P, _P, _I]
        g("lu_resid").restype = t
        g("lutest").argtypes = [_I, _I, _I, _P]
        g("lutest").restype = t
        g("lltgen").argtypes = [_I, _I, _P, _I, _I]
        g("lltnrm1").argtypes = [_I, _I, _P, _I]
        g("lltnrm1").restype = t
        g("llt_resid").argtypes = [_I, _I, _P, _I, _P, _I]
        g("llt_resid").restype = t
        g("llttest").argtypes = [_I, _I, _I]
        g("llttest").restype = t
        g("slv_resid").argtypes = [_I, _I, t, _P, _I, _P, _I]
        g("slv_resid").restype = t
        g("gemm_ferr_check").argtypes = [_I, _I, _I, t, t, _P, _I, _P, _I, _P, _P]
        g("gemm_ferr_check").restype = _I
        g("trsm_resid").argtypes = [_I, _I, _I, _I, _I, _I, t, _P, _I, _P, _I, _P, _I]
        g("trsm_resid").restype = t
        g("trinrm1").argtypes = [_I, _I, _I, _P, _I]
        g("trinrm1").restype = t
        g("trigen").argtypes = [_I, _I, _I, _P, _I, t, _I]
        g("trtri_resid").argtypes = [_I, _I, _I, _P, _I, _P, _I]
        g("trtri_resid").restype = t
        g("inv_resid").argtypes = [_I, _P, _I, _P, _I]
        g("inv_resid").restype = t
        g("poinv_resid").argtypes = [_I, _I, _P, _I, _P, _I]
        g("poinv_resid").restype = t
        g("lauum_resid").argtypes = [_I, _I, _P, _I, _P, _I]
        g("lauum_resid").restype = t
        g("geqr2").argtypes = [_I, _I, _P, _I, _P]
        g("qr_resid").argtypes = [_I, _I, _P, _I, _P, _P, _I]
        g("qr_resid").restype = t
    for p in ("orc_z", "orc_c"):
        getattr(L, p + "gemm_ref").argtypes = [_I] * 5 + [_P, _P, _I, _P, _I, _P, _P, _I]
        getattr(L, p + "trsm_ref").argtypes = [_I] * 6 + [_P, _P, _I, _P, _I]
        getattr(L, p + "getf2").argtypes = [_I, _I, _P, _I, _P, _P]
        getattr(L, p + "getf2").restype = _I
        getattr(L, p + "potrf").argtypes = [_I, _I, _P, _I]
        getattr(L, p + "potrf").restype = _I


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def _pre(dtype):
    return "orc_d" if np.dtype(dtype) == np.float64 else "orc_s"


def _fn(dtype, name):
    return getattr(lib(), _pre(dtype) + name)


def _ld(a):
    """Leading dimension of a Fortran-ordered 2-D array (or view)."""
    assert a.ndim == 2
    if a.flags.f_contiguous:
        return max(a.shape[0], 1)
    if a.strides[0] == a.itemsize:          # Fortran-ordered view
        return a.strides[1] // a.itemsize
    assert a.strides[1] == a.itemsize        # C-ordered: row-major leading dim
    return a.strides[0] // a.itemsize


# ---------------------------------------------------------------- generators
def rand_stream(seed, n):
    out = np.zeros(n, dtype=np.int32)
    lib().orc_rand_stream(seed, n, _p(out))
    return out


def gegen(M, N, lda=None, seed=None, dtype=np.float64):
    """src/testing/ATL_gegen.c: returns the full lda x N buffer (Fortran)."""
    lda = lda or max(M, 1)
    seed = M * N + lda if seed is None else seed
    A = np.zeros((lda, N), dtype=dtype, order="F")
    _fn(dtype, "gegen")(M, N, _p(A), lda, seed)
    return A


def matgen(M, N, lda, seed, dtype=np.float64):
    A = np.zeros((lda, N), dtype=dtype, order="F")
    _fn(dtype, "matgen")(M, N, _p(A), lda, seed)
    return A


def epsilon(dtype=np.float64):
    return _fn(dtype, "epsilon")()


def iamax(x):
    x = np.ascontiguousarray(x)
    return _fn(x.dtype, "iamax")(x.size, _p(x), 1)


def set_threads(n):
    lib().orc_set_threads(int(n))


def set_recursion(l1c_elts=4096, lcm=4):
    lib().orc_set_recursion_params(int(l1c_elts), int(lcm))


# ---------------------------------------------------------------- level 3
def gemm(ta, tb, alpha, A, B, beta, C, exact_order=False, order=COL):
    """C (Fortran array, modified in place) = alpha op(A) op(B) + beta C."""
    dt = C.dtype
    if order == COL:
        M, N = C.shape
        K = A.shape[1] if ta == NOTRANS else A.shape[0]
    else:
        M, N = C.shape[1], C.shape[0]
        K = A.shape[0] if ta == NOTRANS else A.shape[1]
    if exact_order:
        assert order == COL
        _fn(dt, "gemm_ref")(ta, tb, M, N, K, alpha, _p(A), _ld(A), _p(B), _ld(B),
                            beta, _p(C), _ld(C))
    else:
        _fn(dt, "gemm")(order, ta, tb, M, N, K, alpha, _p(A), _ld(A), _p(B), _ld(B),
                        beta, _p(C), _ld(C))
    return C


def _cpre(dtype):
    dt = np.dtype(dtype)
    if dt == np.complex128:
        return "orc_z", np.float64
    if dt == np.complex64:
        return "orc_c", np.float32
    raise TypeError("complex128 or complex64 expected, not %s" % dt)


def _cscalar(x, rt):
    """A complex scalar as the {real, imaginary} pair the C side reads."""
    return np.array([complex(x).real, complex(x).imag], dtype=rt)


def cgemm(ta, tb, alpha, A, B, beta, Cm):
    """Complex Cm (Fortran, in place) = alpha op(A) op(B) + beta Cm in the
    reference loop order (ATL_zrefgemm*.c); op may be CONJTRANS."""
    p, rt = _cpre(Cm.dtype)
    M, N = Cm.shape
    K = A.shape[1] if ta == NOTRANS else A.shape[0]
    al, be = _cscalar(alpha, rt), _cscalar(beta, rt)
    getattr(lib(), p + "gemm_ref")(ta, tb, M, N, K, _p(al), _p(A), _ld(A), _p(B), _ld(B),
                                   _p(be), _p(Cm), _ld(Cm))
    return Cm


def ctrsm(side, uplo, ta, diag, alpha, A, B):
    """Complex B (Fortran, in place) = alpha inv(op(A)) B (Left) or
    alpha B inv(op(A)) (Right); op may be CONJTRANS."""
    p, rt = _cpre(B.dtype)
    M, N = B.shape
    al = _cscalar(alpha, rt)
    getattr(lib(), p + "trsm_ref")(side, uplo, ta, diag, M, N, _p(al), _p(A), _ld(A), _p(B),
                                   _ld(B))
    return B


def cgetf2(A):
    """Complex LU of the Fortran array A in place (ATL_getf2's column loop).
    Returns (info, ipiv (0-based), mingap): mingap is the smallest relative
    margin between a chosen pivot and the runner-up over all columns."""
    p, rt = _cpre(A.dtype)
    M, N = A.shape
    ipiv = np.zeros(max(min(M, N), 1), dtype=np.int32)
    gap = np.zeros(1, dtype=rt)
    info = getattr(lib(), p + "getf2")(M, N, _p(A), _ld(A), _p(ipiv), _p(gap))
    return info, ipiv[:min(M, N)], float(gap[0])


def cpotrf(A, uplo=LOWER):
    """Complex Cholesky of the Fortran array A in place (the referenced
    triangle); returns info."""
    p, _ = _cpre(A.dtype)
    return getattr(lib(), p + "potrf")(uplo, A.shape[0], _p(A), _ld(A))


def trsm(side, uplo, ta, diag, alpha, A, B, order=COL):
    M, N = B.shape if order == COL else (B.shape[1], B.shape[0])
    _fn(B.dtype, "trsm")(order, side, uplo, ta, diag, M, N, alpha, _p(A), _ld(A),
                         _p(B), _ld(B))
    return B


def syrk(uplo, trans, alpha, A, beta, C):
    N = C.shape[0]
    K = A.shape[1] if trans == NOTRANS else A.shape[0]
    _fn(C.dtype, "syrk_ref")(uplo, trans, N, K, alpha, _p(A), _ld(A), beta, _p(C), _ld(C))
    return C


def trmm(side, uplo, ta, diag, alpha, A, B):
    """B (Fortran, in place) = alpha op(A) B (Left) or alpha B op(A) (Right),
    in the reference loop order (ATL_dreftrmm*.c)."""
    M, N = B.shape
    _fn(B.dtype, "trmm_ref")(side, uplo, ta, diag, M, N, alpha, _p(A), _ld(A), _p(B), _ld(B))
    return B


def symm(side, uplo, alpha, A, B, beta, Cm):
    M, N = Cm.shape
    _fn(Cm.dtype, "symm_ref")(side, uplo, M, N, alpha, _p(A), _ld(A), _p(B), _ld(B), beta,
                              _p(Cm), _ld(Cm))
    return Cm


def syr2k(uplo, trans, alpha, A, B, beta, Cm):
    N = Cm.shape[0]
    K = A.shape[1] if trans == NOTRANS else A.shape[0]
    _fn(Cm.dtype, "syr2k_ref")(uplo, trans, N, K, alpha, _p(A), _ld(A), _p(B), _ld(B), beta,
                               _p(Cm), _ld(Cm))
    return Cm


def infnrm(x):
    x = np.ascontiguousarray(x)
    return float(_fn(x.dtype, "infnrm")(x.size, _p(x), 1))


def gediffnrm1(A, B):
    return float(_fn(A.dtype, "gediffnrm1")(A.shape[0], A.shape[1], _p(A), _ld(A), _p(B),
                                            _ld(B)))


def lltnrm1(uplo, A):
    N = A.shape[1]
    return float(_fn(A.dtype, "lltnrm1")(uplo, N, _p(A), _ld(A)))


# ---------------------------------------------------------------- LAPACK
def getrf(A, order=COL, gaps=None):
    """In-place recursive LU (ATL_getrfC).  Returns (info, ipiv)."""
    if order == COL:
        M, N = A.shape
    else:
        M, N = A.shape[1], A.shape[0]
    ipiv = np.zeros(max(min(M, N), 1), dtype=np.int32)
    if gaps is not None:
        lib().orc_track_pivot_gaps(_p(gaps), gaps.size)
    try:
        info = _fn(A.dtype, "getrf")(order, M, N, _p(A), _ld(A), _p(ipiv))
    finally:
        if gaps is not None:
            lib().orc_track_pivot_gaps(None, 0)
    return info, ipiv[: min(M, N)]


def getrs(LU, ipiv, B, trans=NOTRANS, order=COL):
    N = LU.shape[0]
    NRHS = B.shape[1] if order == COL else B.shape[0]
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32)
    _fn(B.dtype, "getrs")(order, trans, N, NRHS, _p(LU), _ld(LU), _p(ipiv), _p(B), _ld(B))
    return B


def potrf(A, uplo=LOWER, order=COL):
    return _fn(A.dtype, "potrf")(order, uplo, A.shape[0], _p(A), _ld(A))


def potrs(L, B, uplo=LOWER, order=COL):
    N = L.shape[0]
    NRHS = B.shape[1] if order == COL else B.shape[0]
    _fn(B.dtype, "potrs")(order, uplo, N, NRHS, _p(L), _ld(L), _p(B), _ld(B))
    return B


def laswp(A, K1, K2, ipiv, inci=1):
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32)
    _fn(A.dtype, "laswp")(A.shape[1], _p(A), _ld(A), K1, K2, _p(ipiv), inci)
    return A


# ---------------------------------------------------------------- testers
def lu_resid(LU, ipiv, A0):
    M, N = A0.shape
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32)
    return float(_fn(A0.dtype, "lu_resid")(M, N, _p(LU), _ld(LU), _p(ipiv), _p(A0), _ld(A0)))


def lutest(M, N, lda, dtype=np.float64):
    npiv = C.c_int(0)
    r = _fn(dtype, "lutest")(M, N, lda, C.byref(npiv))
    return float(r), npiv.value


def lltgen(uplo, N, lda=None, seed=None, dtype=np.float64):
    lda = lda or N
    seed = N * 1029 + lda if seed is None else seed
    A = np.zeros((lda, N), dtype=dtype, order="F")
    _fn(dtype, "lltgen")(uplo, N, _p(A), lda, seed)
    return A


def llt_resid(uplo, L, A0):
    N = A0.shape[1]
    return float(_fn(A0.dtype, "llt_resid")(uplo, N, _p(L), _ld(L), _p(A0), _ld(A0)))


def llttest(uplo, N, lda, dtype=np.float64):
    return float(_fn(dtype, "llttest")(uplo, N, lda))


def genrm1(A):
    return float(_fn(A.dtype, "genrm1")(A.shape[0], A.shape[1], _p(A), _ld(A)))


def slv_resid(nrmA, XC, X):
    N, NRHS = XC.shape
    return float(_fn(XC.dtype, "slv_resid")(N, NRHS, nrmA, _p(XC), _ld(XC), _p(X), _ld(X)))


def gemm_ferr_check(K, alpha, beta, Ctrusted, D):
    M, N = Ctrusted.shape
    mx = C.c_double(0)
    fe = C.c_double(0)
    bad = _fn(D.dtype, "gemm_ferr_check")(M, N, K, alpha, beta, _p(Ctrusted), _ld(Ctrusted),
                                          _p(D), _ld(D), C.byref(mx), C.byref(fe))
    return bad, mx.value, fe.value


def trsm_resid(side, uplo, ta, diag, alpha, A, X, B0):
    M, N = X.shape
    return float(_fn(X.dtype, "trsm_resid")(side, uplo, ta, diag, M, N, alpha, _p(A), _ld(A),
                                            _p(X), _ld(X), _p(B0), _ld(B0)))


def trigen(uplo, diag, N, lda=None, padval=0.0, seed=None, dtype=np.float64):
    """bin/trtritst.c trigen (column-major), seed N*1029+lda by default."""
    lda = lda or max(N, 1)
    seed = N * 1029 + lda if seed is None else seed
    A = np.zeros((lda, N), dtype=dtype, order="F")
    _fn(dtype, "trigen")(uplo, diag, N, _p(A), lda, padval, seed)
    return A


def trinrm1(uplo, diag, A):
    return float(_fn(A.dtype, "trinrm1")(uplo, diag, A.shape[1], _p(A), _ld(A)))


def trtri_resid(uplo, diag, Ainv, A0):
    N = A0.shape[1]
    return float(_fn(A0.dtype, "trtri_resid")(uplo, diag, N, _p(Ainv), _ld(Ainv), _p(A0),
                                              _ld(A0)))


def inv_resid(A, AI):
    return float(_fn(A.dtype, "inv_resid")(A.shape[1], _p(A), _ld(A), _p(AI), _ld(AI)))


def poinv_resid(uplo, A, AI):
    return float(_fn(A.dtype, "poinv_resid")(uplo, A.shape[1], _p(A), _ld(A), _p(AI), _ld(AI)))


def lauum_resid(uplo, R, T0):
    return float(_fn(R.dtype, "lauum_resid")(uplo, T0.shape[1], _p(R), _ld(R), _p(T0),
                                             _ld(T0)))


def geqr2(A):
    """In-place unblocked Householder QR (ATL_geqr2.c / ATL_larfg.c); tau."""
    M, N = A.shape
    tau = np.zeros(max(min(M, N), 1), dtype=A.dtype)
    _fn(A.dtype, "geqr2")(M, N, _p(A), _ld(A), _p(tau))
    return tau[:min(M, N)]


def qr_resid(QR, tau, A0):
    M, N = A0.shape
    tau = np.ascontiguousarray(tau, dtype=A0.dtype)
    return float(_fn(A0.dtype, "qr_resid")(M, N, _p(QR), _ld(QR), _p(tau), _p(A0), _ld(A0)))
