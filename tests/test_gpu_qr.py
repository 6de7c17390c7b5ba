"""GPU parity of clapack_?geqrf through the C ABI: qrtst's residual
||A - QR||_1 / (||A||_1 eps min(M,N)) < 100 (bin/qrtst.c:1291-1295) from
the returned storage and TAU, R and TAU against the oracle's unblocked
Householder QR (ATL_geqr2.c, pinned to LAPACK) within a rounding bound,
and RowMajor, which holds the same numbers transposed."""
import numpy as np
import pytest

import atlas_amd as X
import oracle as O

pytestmark = pytest.mark.gpu


def _geqrf(A, order=X.ColMajor):
    M, N = A.shape if order == X.ColMajor else (A.shape[0], A.shape[1])
    K = min(M, N)
    tau = np.zeros(max(K, 1), dtype=A.dtype)
    lda = A.shape[0] if order == X.ColMajor else A.shape[1]
    f = X.lib().clapack_dgeqrf if A.dtype == np.float64 else X.lib().clapack_sgeqrf
    info = f(order, M, N, A.ctypes.data, lda, tau.ctypes.data)
    return info, tau[:K]


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("M,N", [(1, 1), (33, 33), (300, 100), (100, 300), (1000, 1000),
                                 (4100, 257)])
def test_geqrf_qrtst_and_oracle(dt, M, N):
    A0 = O.gegen(M, N, dtype=dt)
    A = A0.copy(order="F")
    info, tau = _geqrf(A)
    assert info == 0
    assert np.isfinite(A).all() and np.isfinite(tau).all()
    r = O.qr_resid(A, tau, A0)
    assert r < 100, r
    Ao = A0.copy(order="F")
    to = O.geqr2(Ao)
    eps = np.finfo(dt).eps
    scale = np.abs(np.triu(Ao)).max()
    assert np.abs(np.triu(A) - np.triu(Ao)).max() <= 100 * max(M, N) * eps * scale
    assert np.abs(tau - to).max() <= 100 * max(M, N) * eps


def test_geqrf_rowmajor():
    M, N = 200, 120
    A0 = O.gegen(M, N)
    Ar = np.ascontiguousarray(A0)                 # row-major storage of A
    info, tau = _geqrf(Ar, X.RowMajor)
    assert info == 0
    r = O.qr_resid(np.asfortranarray(Ar), tau, A0)
    assert r < 100, r
