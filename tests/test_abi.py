"""CPU checks of the C-ABI boundary: the library loads, and exports every
function that include/*.h declares (no GPU calls)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "atlas_amd", "libatlas_mi355x.so")


def _declared():
    names = set()
    for h in ("cblas.h", "clapack.h", "atlmx.h", "atlas_f77.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b((?:cblas|clapack|atlmx)_\w+|[a-z]\w*_)\s*\(", txt):
            names.add(m.group(1))
    return names


def test_headers_declare_the_boundary():
    d = _declared()
    for s in ("cblas_sgemm", "cblas_dgemm", "cblas_strsm", "cblas_dtrsm", "clapack_sgetrf",
              "clapack_dgetrf", "clapack_sgetrs", "clapack_dgetrs", "clapack_spotrf",
              "clapack_dpotrf", "clapack_spotrs", "clapack_dpotrs", "clapack_dgesv",
              "clapack_dposv", "cblas_dsyrk", "cblas_idamax", "cblas_xerbla"):
        assert s in d, s


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("libatlas_mi355x.so not built (run __graft_entry__.build())")
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], stdout=subprocess.PIPE,
                         text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if l.strip()}
    missing = sorted(_declared() - exported)
    assert not missing, missing


def test_library_loads_without_gpu():
    import atlas_amd
    L = atlas_amd.lib()
    assert L.cblas_dgemm is not None and L.clapack_dgetrf is not None


def test_xerbla_exits_on_bad_argument():
    """Illegal argument -> 'Parameter %d to routine %s was incorrect', exit(-1)
    (interfaces/blas/C/src/cblas_xerbla.c:13-28), before any GPU work."""
    code = ("import ctypes as C, sys; sys.path.insert(0, %r); import atlas_amd as a; L = a.lib();"
            "L.cblas_dgemm(102, 111, 111, -1, 2, 2, 1.0, None, 1, None, 1, 0.0, None, 1)"
            % ROOT)
    p = subprocess.run(["python", "-c", code], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True)
    assert p.returncode == 255
    assert "Parameter 4 to routine cblas_dgemm was incorrect" in p.stderr
    code = ("import sys; sys.path.insert(0, %r); import atlas_amd as a; L = a.lib();"
            "L.clapack_dgetrf(102, 5, 5, None, 3, None)" % ROOT)
    p = subprocess.run(["python", "-c", code], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True)
    assert p.returncode == 255
    assert "Parameter 6 to routine clapack_dgetrf was incorrect" in p.stderr


def test_complex_level3_xerbla_without_gpu():
    """The complex rank-k, HEMM/SYMM and TRMM argument checks report ATLAS's
    parameter numbers (interfaces/blas/C/src/cblas_zherk.c, cblas_zsyrk.c,
    cblas_zher2k.c, cblas_zhemm.c, cblas_ztrmm.c) before any GPU work."""
    cases = [
        ("L.cblas_zherk(102, 121, 112, 2, 2, C.c_double(1.0), None, 2, C.c_double(0.0), None, 2)",
         "Parameter 3 to routine cblas_zherk"),
        ("L.cblas_csyrk(102, 121, 113, 2, 2, None, None, 2, None, None, 2)",
         "Parameter 3 to routine cblas_csyrk"),
        ("L.cblas_zher2k(102, 122, 111, 4, 2, None, None, 4, None, 3, C.c_double(0.0), None, 4)",
         "Parameter 10 to routine cblas_zher2k"),
        ("L.cblas_csyr2k(101, 121, 111, 3, 2, None, None, 2, None, 2, None, None, 2)",
         "Parameter 13 to routine cblas_csyr2k"),
        ("L.cblas_zhemm(102, 141, 121, 5, 2, None, None, 4, None, 5, None, None, 5)",
         "Parameter 8 to routine cblas_zhemm"),
        ("L.cblas_csymm(102, 141, 123, 2, 2, None, None, 2, None, 2, None, None, 2)",
         "Parameter 3 to routine cblas_csymm"),
        ("L.cblas_ztrmm(102, 141, 121, 111, 133, 2, 2, None, None, 2, None, 2)",
         "Parameter 5 to routine cblas_ztrmm"),
    ]
    for call, msg in cases:
        code = ("import ctypes as C, sys; sys.path.insert(0, %r); import atlas_amd as a;"
                "L = a.lib(); %s; print('RETURNED')" % (ROOT, call))
        p = subprocess.run(["python", "-c", code], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                           text=True, timeout=120)
        assert p.returncode == 255, (call, p.stderr)
        assert msg + " was incorrect" in p.stderr, (call, p.stderr)
        assert "RETURNED" not in p.stdout


def test_batch_plan_order_limits_without_gpu():
    """The batched plans refuse orders the factor kernels cannot hold in
    LDS before any device call: a NULL plan and a message."""
    code = ("import ctypes as C, sys; sys.path.insert(0, %r); import atlas_amd as a; L = a.lib();"
            "L.atlmx_batch_max_order.restype = C.c_int;"
            "lim = L.atlmx_batch_max_order(1, 1);"
            "n = (C.c_int * 2)(5, lim + 1);"
            "L.atlmx_dgesv_batch_plan.restype = C.c_void_p;"
            "h = L.atlmx_dgesv_batch_plan(2, n, None, n, None, 1, None, n, None);"
            "print('LIM', lim, 'PLAN', h)" % ROOT)
    p = subprocess.run(["python", "-c", code], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert "PLAN None" in p.stdout and "LIM 768" in p.stdout
    assert "instance 1 has order 769" in p.stderr



def test_noalloc_refuses_too_large_order_without_gpu():
    """The allocation-free solve checks its arguments before any device call:
    an order beyond its static scratch array is refused with a message."""
    code = ("import sys; sys.path.insert(0, %r); import atlas_amd as a; L = a.lib();"
            "print('RC', L.atlmx_dgetrs_noalloc(111, 16385, 1, None, 16385, None, None, 16385),"
            " L.atlmx_dgetrs_noalloc(111, 10, 1, None, 10, None, None, 9))" % ROOT)
    p = subprocess.run(["python", "-c", code], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert "RC -2 -8" in p.stdout
    assert "N = 16385 is outside [0, 16384]" in p.stderr


def test_f77_xerbla_messages_without_gpu():
    """F77 argument errors are caught before any device call: XERBLA's
    message with the Fortran argument position (interfaces/blas/F77/src/
    xerbla.f), then STOP (exit status 0)."""
    cases = [
        ("L.dgemm_(b'X', b'N', i(2), i(2), i(2), d(1.0), None, i(2), None, i(2), d(0.0), None,"
         " i(2), 1, 1)", "DGEMM  parameter number  1"),
        ("L.dtrsm_(b'L', b'U', b'N', b'N', i(5), i(2), d(1.0), None, i(4), None, i(5), 1, 1, 1, 1)",
         "DTRSM  parameter number  9"),
        ("L.dgetrf_(i(5), i(5), None, i(3), None, C.byref(C.c_int(0)))",
         "DGETRF parameter number  4"),
        ("L.spotrf_(b'Q', i(5), None, i(5), C.byref(C.c_int(0)), 1)",
         "SPOTRF parameter number  1"),
        ("L.dgetri_(i(5), None, i(5), None, None, i(2), C.byref(C.c_int(0)))",
         "DGETRI parameter number  6"),
        ("L.zgemm_(b'N', b'C', i(3), i(2), i(2), None, None, i(2), None, i(2), None, None,"
         " i(3), 1, 1)", "ZGEMM  parameter number  8"),
        ("L.ctrsm_(b'L', b'U', b'Q', b'N', i(5), i(2), None, None, i(5), None, i(5), 1, 1, 1, 1)",
         "CTRSM  parameter number  3"),
        ("L.zgesv_(i(4), i(-1), None, i(4), None, None, i(4), C.byref(C.c_int(0)))",
         "ZGESV  parameter number  2"),
        ("L.cposv_(b'U', i(6), i(1), None, i(6), None, i(5), C.byref(C.c_int(0)), 1)",
         "CPOSV  parameter number  7"),
        ("L.zherk_(b'U', b'T', i(2), i(2), None, None, i(2), None, None, i(2), 1, 1)",
         "ZHERK  parameter number  2"),
        ("L.csyrk_(b'L', b'C', i(2), i(2), None, None, i(2), None, None, i(2), 1, 1)",
         "CSYRK  parameter number  2"),
        ("L.zher2k_(b'L', b'N', i(4), i(2), None, None, i(4), None, i(3), None, None, i(4),"
         " 1, 1)", "ZHER2K parameter number  9"),
        ("L.chemm_(b'R', b'U', i(3), i(5), None, None, i(4), None, i(3), None, None, i(3),"
         " 1, 1)", "CHEMM  parameter number  7"),
        ("L.ztrmm_(b'L', b'U', b'N', b'X', i(2), i(2), None, None, i(2), None, i(2), 1, 1, 1, 1)",
         "ZTRMM  parameter number  4"),
    ]
    for call, msg in cases:
        code = ("import ctypes as C, sys; sys.path.insert(0, %r); import atlas_amd as a;"
                "L = a.lib(); i = lambda v: C.byref(C.c_int(v));"
                "d = lambda v: C.byref(C.c_double(v)); %s; print('RETURNED')" % (ROOT, call))
        p = subprocess.run(["python", "-c", code], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                           text=True, timeout=120)
        assert p.returncode == 0, p.stderr
        assert "** On entry to %s had an illegal value" % msg in p.stdout, p.stdout
        assert "RETURNED" not in p.stdout
