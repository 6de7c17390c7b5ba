"""atlas_amd -- Python binding of libatlas_mi355x (the MI355X Level-3 library).

The product is the C-ABI shared library ``atlas_amd/libatlas_mi355x.so``
(cblas_/clapack_ entry points plus the device-pointer atlmx_ extensions in
include/atlmx.h).  This module only loads it with ctypes and offers thin
numpy (host pointers, through the cblas/clapack ABI) and torch (device
pointers, through atlmx_*) wrappers.  There is no fallback: if the library
is missing, every call raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libatlas_mi355x.so")

RowMajor, ColMajor = 101, 102
NoTrans, Trans, ConjTrans = 111, 112, 113
Upper, Lower = 121, 122
NonUnit, Unit = 131, 132
Left, Right = 141, 142

PROF_GEMM, PROF_PANEL, PROF_TRSM_INV, PROF_LASWP, PROF_POTF2, PROF_MISC, PROF_TRSV, \
    PROF_UPDATE = range(8)

_lib = None


def lib():
    """Load libatlas_mi355x.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError("libatlas_mi355x.so is not built: run `python -m atlas_amd.build`"
                               " (or __graft_entry__.build())")
        L = C.CDLL(LIBPATH, mode=C.RTLD_GLOBAL)
        _declare(L)
        _lib = L
    return _lib


_I, _D, _F, _P, _L = C.c_int, C.c_double, C.c_float, C.c_void_p, C.c_long


def _declare(L):
    for p, t in (("d", _D), ("s", _F)):
        f = getattr(L, "cblas_%sgemm" % p)
        f.argtypes = [_I, _I, _I, _I, _I, _I, t, _P, _I, _P, _I, t, _P, _I]
        f.restype = None
        f = getattr(L, "cblas_%strsm" % p)
        f.argtypes = [_I, _I, _I, _I, _I, _I, _I, t, _P, _I, _P, _I]
        f.restype = None
        f = getattr(L, "cblas_%ssyrk" % p)
        f.argtypes = [_I, _I, _I, _I, _I, t, _P, _I, t, _P, _I]
        f.restype = None
        f = getattr(L, "cblas_i%samax" % p)
        f.argtypes = [_I, _P, _I]
        f.restype = _I
        f = getattr(L, "clapack_%sgetrf" % p)
        f.argtypes = [_I, _I, _I, _P, _I, _P]
        f.restype = _I
        f = getattr(L, "clapack_%sgetrs" % p)
        f.argtypes = [_I, _I, _I, _I, _P, _I, _P, _P, _I]
        f.restype = _I
        f = getattr(L, "clapack_%spotrf" % p)
        f.argtypes = [_I, _I, _I, _P, _I]
        f.restype = _I
        f = getattr(L, "clapack_%spotrs" % p)
        f.argtypes = [_I, _I, _I, _I, _P, _I, _P, _I]
        f.restype = _I
        f = getattr(L, "clapack_%sgesv" % p)
        f.argtypes = [_I, _I, _I, _P, _I, _P, _P, _I]
        f.restype = _I
        f = getattr(L, "clapack_%sposv" % p)
        f.argtypes = [_I, _I, _I, _I, _P, _I, _P, _I]
        f.restype = _I
        f = getattr(L, "cblas_%strmm" % p)
        f.argtypes = [_I, _I, _I, _I, _I, _I, _I, t, _P, _I, _P, _I]
        f.restype = None
        f = getattr(L, "cblas_%ssymm" % p)
        f.argtypes = [_I, _I, _I, _I, _I, t, _P, _I, _P, _I, t, _P, _I]
        f.restype = None
        f = getattr(L, "cblas_%ssyr2k" % p)
        f.argtypes = [_I, _I, _I, _I, _I, t, _P, _I, _P, _I, t, _P, _I]
        f.restype = None
        f = getattr(L, "clapack_%sgetri" % p)
        f.argtypes = [_I, _I, _P, _I, _P]
        f.restype = _I
        for nm in ("potri", "lauum"):
            f = getattr(L, "clapack_%s%s" % (p, nm))
            f.argtypes = [_I, _I, _I, _P, _I]
            f.restype = _I
        f = getattr(L, "clapack_%sgeqrf" % p)
        f.argtypes = [_I, _I, _I, _P, _I, _P]
        f.restype = _I
        f = getattr(L, "clapack_%strtri" % p)
        f.argtypes = [_I, _I, _I, _I, _P, _I]
        f.restype = _I
        # device-pointer API
        f = getattr(L, "atlmx_%sgemm" % p)
        f.argtypes = [_I, _I, _I, _I, _I, t, _P, _I, _P, _I, t, _P, _I]
        f.restype = None
        f = getattr(L, "atlmx_%strsm" % p)
        f.argtypes = [_I, _I, _I, _I, _I, _I, t, _P, _I, _P, _I]
        f.restype = None
        f = getattr(L, "atlmx_%ssyrk" % p)
        f.argtypes = [_I, _I, _I, _I, t, _P, _I, t, _P, _I]
        f.restype = None
        f = getattr(L, "atlmx_%sgetrf" % p)
        f.argtypes = [_I, _I, _P, _I, _P]
        f.restype = _I
        f = getattr(L, "atlmx_%sgetrs" % p)
        f.argtypes = [_I, _I, _I, _P, _I, _P, _P, _I]
        f.restype = None
        f = getattr(L, "atlmx_%spotrf" % p)
        f.argtypes = [_I, _I, _P, _I]
        f.restype = _I
        f = getattr(L, "atlmx_%spotrs" % p)
        f.argtypes = [_I, _I, _I, _P, _I, _P, _I]
        f.restype = None
    for p in ("z", "c"):
        f = getattr(L, "cblas_%sgemm" % p)
        f.argtypes = [_I, _I, _I, _I, _I, _I, _P, _P, _I, _P, _I, _P, _P, _I]
        f.restype = None
        f = getattr(L, "cblas_%strsm" % p)
        f.argtypes = [_I, _I, _I, _I, _I, _I, _I, _P, _P, _I, _P, _I]
        f.restype = None
        f = getattr(L, "clapack_%sgetrf" % p)
        f.argtypes = [_I, _I, _I, _P, _I, _P]
        f.restype = _I
        f = getattr(L, "clapack_%sgetrs" % p)
        f.argtypes = [_I, _I, _I, _I, _P, _I, _P, _P, _I]
        f.restype = _I
        f = getattr(L, "clapack_%sgesv" % p)
        f.argtypes = [_I, _I, _I, _P, _I, _P, _P, _I]
        f.restype = _I
        f = getattr(L, "clapack_%spotrf" % p)
        f.argtypes = [_I, _I, _I, _P, _I]
        f.restype = _I
        for n in ("potrs", "posv"):
            f = getattr(L, "clapack_%s%s" % (p, n))
            f.argtypes = [_I, _I, _I, _I, _P, _I, _P, _I]
            f.restype = _I
        rt = _D if p == "z" else _F
        f = getattr(L, "cblas_%sherk" % p)
        f.argtypes = [_I, _I, _I, _I, _I, rt, _P, _I, rt, _P, _I]
        f.restype = None
        f = getattr(L, "cblas_%sher2k" % p)
        f.argtypes = [_I, _I, _I, _I, _I, _P, _P, _I, _P, _I, rt, _P, _I]
        f.restype = None
        f = getattr(L, "cblas_%ssyrk" % p)
        f.argtypes = [_I, _I, _I, _I, _I, _P, _P, _I, _P, _P, _I]
        f.restype = None
        f = getattr(L, "cblas_%ssyr2k" % p)
        f.argtypes = [_I, _I, _I, _I, _I, _P, _P, _I, _P, _I, _P, _P, _I]
        f.restype = None
        for n in ("hemm", "symm"):
            f = getattr(L, "cblas_%s%s" % (p, n))
            f.argtypes = [_I, _I, _I, _I, _I, _P, _P, _I, _P, _I, _P, _P, _I]
            f.restype = None
        f = getattr(L, "cblas_%strmm" % p)
        f.argtypes = [_I, _I, _I, _I, _I, _I, _I, _P, _P, _I, _P, _I]
        f.restype = None
    L.atlmx_init.restype = _I
    L.atlmx_set_stream.argtypes = [_P]
    L.atlmx_get_stream.restype = _P
    L.atlmx_prof_enable.argtypes = [_I]
    L.atlmx_prof_query.argtypes = [_I, _P, _P, _P]
    L.atlmx_prof_query.restype = _I
    L.atlmx_prof_query2.argtypes = [_I, _P, _P, _P, _P]
    L.atlmx_prof_query2.restype = _I
    L.atlmx_dsgesv.argtypes = [_I, _I, _P, _I, _P, _I, _P, _I, _I, _P]
    L.atlmx_dsgesv.restype = _I
    L.atlmx_dbackward_error.argtypes = [_I, _I, _P, _I, _P, _I, _P, _I]
    L.atlmx_set_mixed_method.argtypes = [_I]
    L.atlmx_set_mixed_method.restype = None
    L.atlmx_dbackward_error.restype = _D
    L.atlmx_dgetrs_noalloc.argtypes = [_I, _I, _I, _P, _I, _P, _P, _I]
    L.atlmx_dgetrs_noalloc.restype = _I
    L.atlmx_noalloc_scratch_bytes.argtypes = []
    L.atlmx_noalloc_scratch_bytes.restype = C.c_size_t


def _pre(dt):
    dt = np.dtype(dt)
    if dt == np.float64:
        return "d"
    if dt == np.float32:
        return "s"
    raise TypeError("unsupported dtype %s" % dt)


def _ptr(a):
    return C.c_void_p(a.ctypes.data)


def _ld(a):
    if a.flags.f_contiguous:
        return max(a.shape[0], 1)
    assert a.strides[0] == a.itemsize, "need column-major storage"
    return a.strides[1] // a.itemsize


# --------------------------------------------------------------------------
# host (numpy, column-major) wrappers of the ATLAS C ABI
# --------------------------------------------------------------------------
def _cpre(dt):
    dt = np.dtype(dt)
    if dt == np.complex128:
        return "z", np.float64
    if dt == np.complex64:
        return "c", np.float32
    raise TypeError("unsupported complex dtype %s" % dt)


def _cscalar(x, rt):
    return np.array([complex(x).real, complex(x).imag], dtype=rt)


def cgemm(ta, tb, alpha, A, B, beta, Cm):
    """Complex Cm (Fortran-ordered, in place) = alpha op(A) op(B) + beta Cm
    through cblas_zgemm / cblas_cgemm (op may be ConjTrans)."""
    p, rt = _cpre(Cm.dtype)
    M, N = Cm.shape
    K = A.shape[1] if ta == NoTrans else A.shape[0]
    al, be = _cscalar(alpha, rt), _cscalar(beta, rt)
    getattr(lib(), "cblas_%sgemm" % p)(ColMajor, ta, tb, M, N, K, _ptr(al), _ptr(A), _ld(A),
                                       _ptr(B), _ld(B), _ptr(be), _ptr(Cm), _ld(Cm))
    return Cm


def ctrsm(side, uplo, ta, diag, alpha, A, B):
    """Complex B (Fortran-ordered, in place) through cblas_ztrsm / cblas_ctrsm."""
    p, rt = _cpre(B.dtype)
    M, N = B.shape
    al = _cscalar(alpha, rt)
    getattr(lib(), "cblas_%strsm" % p)(ColMajor, side, uplo, ta, diag, M, N, _ptr(al),
                                       _ptr(A), _ld(A), _ptr(B), _ld(B))
    return B


def gemm(ta, tb, alpha, A, B, beta, Cm):
    """Cm (Fortran-ordered, in place) = alpha op(A) op(B) + beta Cm."""
    M, N = Cm.shape
    K = A.shape[1] if ta == NoTrans else A.shape[0]
    getattr(lib(), "cblas_%sgemm" % _pre(Cm.dtype))(
        ColMajor, ta, tb, M, N, K, alpha, _ptr(A), _ld(A), _ptr(B), _ld(B), beta, _ptr(Cm),
        _ld(Cm))
    return Cm


def trsm(side, uplo, ta, diag, alpha, A, B):
    M, N = B.shape
    getattr(lib(), "cblas_%strsm" % _pre(B.dtype))(
        ColMajor, side, uplo, ta, diag, M, N, alpha, _ptr(A), _ld(A), _ptr(B), _ld(B))
    return B


def getrf(A):
    M, N = A.shape
    ipiv = np.zeros(max(min(M, N), 1), dtype=np.int32)
    info = getattr(lib(), "clapack_%sgetrf" % _pre(A.dtype))(ColMajor, M, N, _ptr(A), _ld(A),
                                                             _ptr(ipiv))
    return info, ipiv[:min(M, N)]


def getrs(LU, ipiv, B, trans=NoTrans):
    N = LU.shape[0]
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32)
    return getattr(lib(), "clapack_%sgetrs" % _pre(B.dtype))(
        ColMajor, trans, N, B.shape[1], _ptr(LU), _ld(LU), _ptr(ipiv), _ptr(B), _ld(B))


def potrf(A, uplo=Lower):
    return getattr(lib(), "clapack_%spotrf" % _pre(A.dtype))(ColMajor, uplo, A.shape[0],
                                                             _ptr(A), _ld(A))


def potrs(L, B, uplo=Lower):
    return getattr(lib(), "clapack_%spotrs" % _pre(B.dtype))(
        ColMajor, uplo, L.shape[0], B.shape[1], _ptr(L), _ld(L), _ptr(B), _ld(B))


def trmm(side, uplo, ta, diag, alpha, A, B):
    M, N = B.shape
    getattr(lib(), "cblas_%strmm" % _pre(B.dtype))(
        ColMajor, side, uplo, ta, diag, M, N, alpha, _ptr(A), _ld(A), _ptr(B), _ld(B))
    return B


def symm(side, uplo, alpha, A, B, beta, Cm):
    M, N = Cm.shape
    getattr(lib(), "cblas_%ssymm" % _pre(Cm.dtype))(
        ColMajor, side, uplo, M, N, alpha, _ptr(A), _ld(A), _ptr(B), _ld(B), beta, _ptr(Cm),
        _ld(Cm))
    return Cm


def syr2k(uplo, trans, alpha, A, B, beta, Cm):
    N = Cm.shape[0]
    K = A.shape[1] if trans == NoTrans else A.shape[0]
    getattr(lib(), "cblas_%ssyr2k" % _pre(Cm.dtype))(
        ColMajor, uplo, trans, N, K, alpha, _ptr(A), _ld(A), _ptr(B), _ld(B), beta, _ptr(Cm),
        _ld(Cm))
    return Cm


def trtri(A, uplo=Lower, diag=NonUnit):
    return getattr(lib(), "clapack_%strtri" % _pre(A.dtype))(ColMajor, uplo, diag, A.shape[0],
                                                             _ptr(A), _ld(A))


def lauum(A, uplo=Lower):
    return getattr(lib(), "clapack_%slauum" % _pre(A.dtype))(ColMajor, uplo, A.shape[0],
                                                             _ptr(A), _ld(A))


def potri(A, uplo=Lower):
    return getattr(lib(), "clapack_%spotri" % _pre(A.dtype))(ColMajor, uplo, A.shape[0],
                                                             _ptr(A), _ld(A))


def getri(LU, ipiv):
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32)
    return getattr(lib(), "clapack_%sgetri" % _pre(LU.dtype))(ColMajor, LU.shape[0], _ptr(LU),
                                                              _ld(LU), _ptr(ipiv))


# --------------------------------------------------------------------------
# device (torch) helpers
# --------------------------------------------------------------------------
def use_torch_stream():
    """Order library work on torch's current CUDA(HIP) stream."""
    import torch
    lib().atlmx_set_stream(C.c_void_p(torch.cuda.current_stream().cuda_stream))


class IRInfo(C.Structure):
    """atlmx_ir_info (include/atlmx.h)"""
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("fallback", C.c_int),
                ("factor_info", C.c_int), ("backward_error", C.c_double),
                ("method", C.c_int), ("gmres", C.c_int), ("inner_iterations", C.c_int)]


MIXED_AUTO, MIXED_TSLU, MIXED_GEPP = 0, 1, 2
MIXED_BF16, MIXED_FP16 = 5, 6


def set_refine_stop(lapack):
    """refinement stopping test: False -> backward error <= n*2^-52 (default),
    True -> LAPACK dsgesv's."""
    L = lib()
    L.atlmx_set_refine_stop.argtypes = [_I]
    L.atlmx_set_refine_stop.restype = None
    L.atlmx_set_refine_stop(int(bool(lapack)))


def set_mixed_method(m):
    """0 tournament-pivoting FP32 LU, then partial-pivoting FP32 LU, then
    FP64; 1 tournament then FP64; 2 partial pivoting then FP64; 5 / 6 the
    tournament LU with its trailing updates in bfloat16 / half (columns
    scaled by powers of two), then as 0."""
    lib().atlmx_set_mixed_method(int(m))


def dsgesv_dev(A, B, Xout, iter_max=30):
    """FP64-accurate solve of A X = B with an FP32 LU + FP64 refinement.
    A (n x n), B and Xout (n x nrhs) are column-major float64 CUDA tensors given
    as their row-major transposes (A.T storage)."""
    n = A.shape[0]
    nrhs = B.shape[0]
    st = IRInfo()
    info = lib().atlmx_dsgesv(n, nrhs, dptr(A), n, dptr(B), n, dptr(Xout), n, iter_max,
                              C.byref(st))
    return info, st


def dsposv_dev(A, B, Xout, iter_max=30):
    """FP64-accurate SPD solve (FP32 Cholesky + FP64 refinement); A holds the
    full symmetric matrix."""
    n = A.shape[0]
    nrhs = B.shape[0]
    st = IRInfo()
    L = lib()
    L.atlmx_dsposv.argtypes = [_I, _I, _P, _I, _P, _I, _P, _I, _I, _P]
    info = L.atlmx_dsposv(n, nrhs, dptr(A), n, dptr(B), n, dptr(Xout), n, iter_max, C.byref(st))
    return info, st


def dptr(t):
    return C.c_void_p(t.data_ptr())


def prof_enable(mask=-1):
    """mask: bit c records launch class c (True/-1 = all, 0/False = off)."""
    if mask is True:
        mask = -1
    lib().atlmx_prof_enable(int(mask))


def prof_query(cls, with_bytes=False):
    """(ms, launches, work[, algorithmic bytes]) summed over one launch class."""
    ms = C.c_double(0)
    n = C.c_long(0)
    w = C.c_double(0)
    by = C.c_double(0)
    lib().atlmx_prof_query2(cls, C.byref(ms), C.byref(n), C.byref(w), C.byref(by))
    if with_bytes:
        return ms.value, n.value, w.value, by.value
    return ms.value, n.value, w.value


# ---- batched plans (include/atlmx.h) ----------------------------------------
class Batch:
    """A batched plan over lists of CUDA tensors (each a column-major matrix
    given as its row-major transpose, like the rest of this module).  Keeps
    the device pointer arrays alive; run() replays the captured launch on
    the library's stream; close() (or the context manager) frees the plan."""

    def __init__(self, handle, keep):
        self.h = handle
        self._keep = keep

    def run(self):
        lib().atlmx_batch_run(self.h)
        return self

    def config(self):
        nb, grid = C.c_int(0), C.c_int(0)
        lib().atlmx_batch_config(self.h, C.byref(nb), C.byref(grid))
        return nb.value, grid.value

    def close(self):
        if self.h:
            lib().atlmx_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _ptr_array(tensors):
    import torch
    return torch.tensor([t.data_ptr() for t in tensors], dtype=torch.int64, device="cuda")


def _ints(vals):
    arr = (C.c_int * max(len(vals), 1))(*vals)
    return arr


def _setup_batch(L):
    if getattr(L, "_batch_ready", False):
        return
    P = C.c_void_p
    for pre in ("d", "s"):
        f = getattr(L, "atlmx_%sposv_batch_plan" % pre)
        f.argtypes = [C.c_int, P, P, P, C.c_int, P, P, P]
        f.restype = P
        f = getattr(L, "atlmx_%sgesv_batch_plan" % pre)
        f.argtypes = [C.c_int, P, P, P, P, C.c_int, P, P, P]
        f.restype = P
        f = getattr(L, "atlmx_%sgemm_batch_plan" % pre)
        t = C.c_double if pre == "d" else C.c_float
        f.argtypes = [C.c_int, C.c_int, C.c_int, P, P, P, t, P, P, P, P, t, P, P]
        f.restype = P
    L.atlmx_batch_run.argtypes = [P]
    L.atlmx_batch_destroy.argtypes = [P]
    L.atlmx_batch_config.argtypes = [P, P, P]
    L._batch_ready = True


def posv_batch(As, Bs, info):
    """Plan Cholesky factor + solve of A_i X_i = B_i (A_i: n x n tensors,
    B_i: nrhs x n tensors = column-major n x nrhs, info: int32 tensor)."""
    L = lib()
    _setup_batch(L)
    pre = "d" if As[0].dtype.itemsize == 8 else "s"
    n = [a.shape[0] for a in As]
    nrhs = Bs[0].shape[0] if Bs else 0
    pa, pb = _ptr_array(As), (_ptr_array(Bs) if Bs else None)
    keep = (pa, pb, _ints(n), _ints(n))
    h = getattr(L, "atlmx_%sposv_batch_plan" % pre)(
        len(As), keep[2], dptr(pa), keep[3], nrhs, dptr(pb) if pb is not None else None,
        keep[3], dptr(info))
    return Batch(h, keep + (info,))


def gesv_batch(As, ipivs, Bs, info):
    """Plan LU (partial pivoting) factor + solve, ipivs: int32 tensors (n,)."""
    L = lib()
    _setup_batch(L)
    pre = "d" if As[0].dtype.itemsize == 8 else "s"
    n = [a.shape[0] for a in As]
    nrhs = Bs[0].shape[0] if Bs else 0
    pa, pp, pb = _ptr_array(As), _ptr_array(ipivs), (_ptr_array(Bs) if Bs else None)
    keep = (pa, pp, pb, _ints(n), _ints(n))
    h = getattr(L, "atlmx_%sgesv_batch_plan" % pre)(
        len(As), keep[3], dptr(pa), keep[4], dptr(pp), nrhs,
        dptr(pb) if pb is not None else None, keep[4], dptr(info))
    return Batch(h, keep + (info,))


def gemm_batch(ta, tb, alpha, As, Bs, beta, Cs):
    """Plan C_i = alpha op(A_i) op(B_i) + beta C_i; tensors are column-major
    matrices given as row-major transposes (shape (cols, rows))."""
    L = lib()
    _setup_batch(L)
    pre = "d" if Cs[0].dtype.itemsize == 8 else "s"
    m = [c.shape[1] for c in Cs]
    nn = [c.shape[0] for c in Cs]
    k = [(a.shape[0] if ta == NoTrans else a.shape[1]) for a in As]
    lda = [a.shape[1] for a in As]
    ldb = [b.shape[1] for b in Bs]
    ldc = [c.shape[1] for c in Cs]
    pa, pb, pc = _ptr_array(As), _ptr_array(Bs), _ptr_array(Cs)
    keep = (pa, pb, pc, _ints(m), _ints(nn), _ints(k), _ints(lda), _ints(ldb), _ints(ldc))
    h = getattr(L, "atlmx_%sgemm_batch_plan" % pre)(
        ta, tb, len(Cs), keep[3], keep[4], keep[5], alpha, dptr(pa), keep[6], dptr(pb),
        keep[7], beta, dptr(pc), keep[8])
    return Batch(h, keep)


def _setup_adjoint(L):
    if getattr(L, "_adj_ready", False):
        return
    _setup_batch(L)
    P = C.c_void_p
    for pre in ("d", "s"):
        f = getattr(L, "atlmx_%sgesv_t_batch_plan" % pre)
        f.argtypes = [C.c_int, P, P, P, P, C.c_int, P, P, P]
        f.restype = P
        f = getattr(L, "atlmx_%sposv_adjoint_batch_plan" % pre)
        f.argtypes = [C.c_int, P, P, P, C.c_int, P, P, P, P, P, P]
        f.restype = P
    L._adj_ready = True


def gesv_t_batch(As, ipivs, Bs, info):
    """Plan LU factor + transposed solve A_i^T X_i = B_i."""
    L = lib()
    _setup_adjoint(L)
    pre = "d" if As[0].dtype.itemsize == 8 else "s"
    n = [a.shape[0] for a in As]
    pa, pp, pb = _ptr_array(As), _ptr_array(ipivs), _ptr_array(Bs)
    keep = (pa, pp, pb, _ints(n))
    h = getattr(L, "atlmx_%sgesv_t_batch_plan" % pre)(
        len(As), keep[3], dptr(pa), keep[3], dptr(pp), Bs[0].shape[0], dptr(pb), keep[3],
        dptr(info))
    return Batch(h, keep + (info,))


def posv_adjoint_batch(Ls, Xs, Bbs, Gs, Ws):
    """Plan the Cholesky adjoint: G_i = d<Xb, X>/d tril(A_i) from L_i
    (factored A_i), X_i and Bb_i = A_i^{-1} Xb_i."""
    L = lib()
    _setup_adjoint(L)
    pre = "d" if Ls[0].dtype.itemsize == 8 else "s"
    n = [a.shape[0] for a in Ls]
    arrs = [_ptr_array(t) for t in (Ls, Xs, Bbs, Gs, Ws)]
    keep = tuple(arrs) + (_ints(n),)
    h = getattr(L, "atlmx_%sposv_adjoint_batch_plan" % pre)(
        len(Ls), keep[5], dptr(arrs[0]), keep[5], Xs[0].shape[0], dptr(arrs[1]), dptr(arrs[2]),
        keep[5], dptr(arrs[3]), keep[5], dptr(arrs[4]))
    return Batch(h, keep)
