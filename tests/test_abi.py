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
    for h in ("cblas.h", "clapack.h", "atlmx.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b((?:cblas|clapack|atlmx)_\w+)\s*\(", txt):
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
