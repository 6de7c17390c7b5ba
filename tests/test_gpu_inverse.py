"""GPU parity of the LAPACK inverses -- clapack_?trtri, ?lauum, ?potri,
?getri -- through the C ABI, against the oracle's restatements of ATLAS's
testers (threshold 100): trtritst's ||inv(A) A - I|| / (||A|| eps N)
(bin/trtritst.c:300-397) on trigen's inputs, uumtst's LAUUM residual
(bin/uumtst.c:300-351) on lltgen's, invtst's geresid / poresid
(bin/invtst.c:388-463).  Plus info on an exactly singular diagonal and the
untouched triangle / unit diagonal."""
import numpy as np
import pytest

import atlas_amd as X
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("uplo", [X.Lower, X.Upper])
@pytest.mark.parametrize("diag", [X.NonUnit, X.Unit])
@pytest.mark.parametrize("N,lda", [(1, 1), (63, 70), (200, 200), (777, 777)])
def test_trtri_trtritst(dt, uplo, diag, N, lda):
    A0 = O.trigen(uplo, diag, N, lda, padval=-9.0, dtype=dt)
    A = A0.copy(order="F")
    assert X.trtri(A[:N, :N], uplo, diag) == 0
    ref = np.tril(np.ones((N, N), bool)) if uplo == X.Lower else np.triu(np.ones((N, N), bool))
    if diag == X.Unit:
        ref &= ~np.eye(N, dtype=bool)
    assert np.array_equal(A[:N, :N][~ref], A0[:N, :N][~ref])     # padding and unit diag kept
    assert np.all(A[N:, :] == A0[N:, :])
    if N == 1 and diag == X.Unit:
        # trtritst's norm leaves a unit diagonal out, so for N = 1 it is 0 and
        # its residual 0/0 (bin/trtritst.c:338,394); the 1 x 1 unit triangle
        # is its own inverse and nothing in it may change (checked above)
        return
    r = O.trtri_resid(uplo, diag, A[:N, :N], A0[:N, :N])
    assert np.isfinite(r) and r < 100, r


def test_trtri_singular_info():
    n = 300
    A0 = O.trigen(X.Upper, X.NonUnit, n)
    A0[137, 137] = 0.0
    A = A0.copy(order="F")
    assert X.trtri(A, X.Upper, X.NonUnit) == 138
    assert np.array_equal(A, A0)
    A[137, 137] = 0.0
    assert X.trtri(A, X.Upper, X.Unit) == 0          # the unit diagonal is not read


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("uplo", [X.Lower, X.Upper])
@pytest.mark.parametrize("N,lda", [(1, 1), (100, 105), (513, 513)])
def test_lauum_uumtst(dt, uplo, N, lda):
    S = O.lltgen(uplo, N, lda, dtype=dt)
    A = S.copy(order="F")
    assert X.lauum(A[:N, :N], uplo) == 0
    unref = np.triu(np.ones((N, N), bool), 1) if uplo == X.Lower else np.tril(np.ones((N, N), bool), -1)
    assert np.array_equal(A[:N, :N][unref], S[:N, :N][unref])
    r = O.lauum_resid(uplo, A[:N, :N], S[:N, :N])
    assert np.isfinite(r) and r < 100, r


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("uplo", [X.Lower, X.Upper])
def test_potri_invtst(dt, uplo):
    n = 400
    S = O.lltgen(uplo, n, dtype=dt)
    A = S.copy(order="F")
    assert X.potrf(A, uplo) == 0
    assert X.potri(A, uplo) == 0
    r = O.poinv_resid(uplo, S, A)
    assert np.isfinite(r) and r < 100, r


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("n", [1, 64, 300, 1000])
def test_getri_invtst(dt, n):
    A0 = O.gegen(n, n, dtype=dt)
    A = A0.copy(order="F")
    info, ipiv = X.getrf(A)
    assert info == 0
    assert X.getri(A, ipiv) == 0
    r = O.inv_resid(A0, A)
    assert np.isfinite(r) and r < 100, r


def test_getri_singular_info_and_rowmajor():
    n = 50
    A0 = O.gegen(n, n)
    A = A0.copy(order="F")
    info, ipiv = X.getrf(A)
    F = A.copy(order="F")
    F[20, 20] = 0.0
    G = F.copy(order="F")
    assert X.getri(G, ipiv) == 21
    assert np.array_equal(G, F)
    # RowMajor: factors of A^T in row-major storage give inv(A) row-major
    Ar = np.ascontiguousarray(O.gegen(n, n, seed=7))
    W = Ar.copy()
    piv = np.zeros(n, dtype=np.int32)
    L = X.lib()
    assert L.clapack_dgetrf(X.RowMajor, n, n, W.ctypes.data, n, piv.ctypes.data) == 0
    assert L.clapack_dgetri(X.RowMajor, n, W.ctypes.data, n, piv.ctypes.data) == 0
    assert np.abs(Ar @ W - np.eye(n)).max() < 1e-10
