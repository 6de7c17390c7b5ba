"""GPU parity of LU / Cholesky / solves through the clapack ABI against the
oracle: pivot vectors bit-identical to the recursive ATL_getrfC restatement
up to the first near-tied column, and ATLAS's tester residuals (lutst,
llttst, slvtst: threshold 100, bin/lutst.c:398, bin/slvtst.c:75)."""
import numpy as np
import pytest

import atlas_amd as X
import oracle as O
from tests.checks import solve_ok

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(99)
TIE = 1e-9


def _f(a):
    return np.asfortranarray(a)


def _pivots_agree(gpu, orc, gaps):
    """Equal up to the first column whose top-2 candidates are near-tied."""
    n = len(orc)
    stop = n
    near = np.nonzero(gaps[:n] < TIE)[0]
    if near.size:
        stop = int(near[0])
    return np.array_equal(gpu[:stop], orc[:stop]), stop


@pytest.mark.parametrize("M,N,lda", [(1, 1, 1), (5, 5, 5), (31, 31, 40), (33, 33, 33),
                                     (100, 100, 100), (257, 257, 260), (300, 200, 300),
                                     (200, 300, 200), (600, 600, 600), (1030, 1030, 1030)])
def test_getrf_lutst_and_pivots(M, N, lda):
    A0 = O.gegen(M, N, lda)                 # lutst input, seed M*N+lda
    A = A0.copy(order="F")
    info, ipiv = X.getrf(A[:M, :N])           # M x N view with leading dimension lda
    assert info == 0
    assert np.all(A[M:, :] == -2.56e9)        # lda gap untouched
    Ao = A0.copy(order="F")
    gaps = np.ones(min(M, N))
    _, opiv = O.getrf(Ao[:M, :N], gaps=gaps)
    ok, stop = _pivots_agree(ipiv, opiv, gaps)
    assert ok, (stop, np.nonzero(ipiv != opiv)[0][:5])
    r = O.lu_resid(A[:M, :N], ipiv, A0[:M, :N])
    assert r < 100, r


@pytest.mark.parametrize("M,N,lda", [(5, 5, 5), (100, 100, 100), (257, 257, 260),
                                     (300, 200, 300), (200, 300, 200), (500, 500, 500)])
def test_sgetrf_lutst_and_pivots(M, N, lda):
    """clapack_sgetrf against the oracle's FP32 getrfC: pivots up to the
    first near-tied column (relative gap 1e-4 in FP32), lutst residual."""
    A0 = O.gegen(M, N, lda, dtype=np.float32)
    A = A0.copy(order="F")
    info, ipiv = X.getrf(A[:M, :N])
    assert info == 0
    Ao = A0.copy(order="F")
    gaps = np.ones(min(M, N))
    _, opiv = O.getrf(Ao[:M, :N], gaps=gaps)
    near = np.nonzero(gaps < 1e-4)[0]
    stop = int(near[0]) if near.size else min(M, N)
    assert np.array_equal(ipiv[:stop], opiv[:stop]), (stop, np.nonzero(ipiv != opiv)[0][:5])
    assert O.lu_resid(A[:M, :N], ipiv, A0[:M, :N]) < 100


def test_getrf_singular_column_info():
    n = 200
    A = _f(RNG.uniform(-1, 1, (n, n)))
    A[:, 77] = 0.0
    info, ipiv = X.getrf(A)
    assert info == 78
    assert np.isfinite(A).all()


@pytest.mark.parametrize("nrhs", [1, 64])
@pytest.mark.parametrize("trans", [X.NoTrans, X.Trans])
def test_getrs_slvtst(nrhs, trans):
    n = 700
    A0 = O.gegen(n, n)
    Xc = O.gegen(n, nrhs)
    B = _f((A0 if trans == X.NoTrans else A0.T) @ Xc)
    A = A0.copy(order="F")
    info, ipiv = X.getrf(A)
    assert info == 0
    X.getrs(A, ipiv, B, trans=trans)
    assert solve_ok(O.genrm1(A0), Xc, B)


def _replay_swaps(ipiv, B, forward):
    """ATLAS's laswp on the rows of B: interchange k with ipiv[k] in order
    (forward) or in reverse order."""
    B = B.copy()
    ks = range(len(ipiv)) if forward else range(len(ipiv) - 1, -1, -1)
    for k in ks:
        p = int(ipiv[k])
        if p != k:
            B[[k, p]] = B[[p, k]]
    return B


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("kind", ["lu_like", "repeated_targets", "below_diagonal"])
def test_getrs_row_interchanges_match_replay(dt, kind):
    """getrs with the identity as its LU factor is its row interchanges
    alone: P B (NoTrans, forward order) and P^T B (Trans, reverse order),
    compared exactly with replaying the interchanges one by one.  LU-like
    ipiv (ipiv[k] >= k) takes the parallel permutation; targets hit many
    times stress its chains; an interchange with a row above (allowed by
    laswp, never made by getrf) takes the replay inside the same kernel."""
    rng = np.random.default_rng(11)
    for n in (1, 2, 65, 700, 4097):
        if kind == "lu_like":
            ipiv = np.array([rng.integers(k, n) for k in range(n)], np.int32)
        elif kind == "repeated_targets":
            ipiv = np.array([max(k, n - 1 - (k % 3)) if rng.random() < 0.7 else k
                             for k in range(n)], np.int32)
        else:
            ipiv = np.array([rng.integers(0, n) for k in range(n)], np.int32)
        B0 = rng.uniform(-1, 1, (n, 5)).astype(dt)
        eye = np.asfortranarray(np.eye(n, dtype=dt))
        for trans, forward in ((X.NoTrans, True), (X.Trans, False)):
            B = _f(B0.copy())
            X.getrs(eye, ipiv, B, trans=trans)
            assert np.array_equal(B, _replay_swaps(ipiv, B0, forward)), (n, trans)


def test_gesv_rowmajor():
    n, nrhs = 300, 3
    A0 = np.ascontiguousarray(RNG.uniform(-1, 1, (n, n)))
    Xc = RNG.uniform(-1, 1, (n, nrhs))
    Bt = np.ascontiguousarray((A0 @ Xc).T)           # one RHS per row, ldb = n
    A = A0.copy()
    ipiv = np.zeros(n, dtype=np.int32)
    info = X.lib().clapack_dgesv(X.RowMajor, n, nrhs, A.ctypes.data, n, ipiv.ctypes.data,
                                 Bt.ctypes.data, n)
    assert info == 0
    assert np.allclose(Bt.T, Xc, atol=1e-9)


@pytest.mark.parametrize("uplo", [X.Lower, X.Upper])
@pytest.mark.parametrize("N,lda", [(1, 1), (4, 4), (127, 130), (128, 128), (129, 129),
                                   (300, 310), (1000, 1000)])
def test_potrf_llttst(uplo, N, lda):
    A0 = O.lltgen(uplo, N, lda)            # llttst input, seed N*1029+lda
    A = A0.copy(order="F")
    assert X.potrf(A[:N, :N], uplo) == 0
    unref = np.triu(np.ones((N, N), bool), 1) if uplo == X.Lower else np.tril(np.ones((N, N), bool), -1)
    assert np.all(A[:N, :N][unref] == A0[:N, :N][unref])       # other triangle untouched
    assert np.all(A[N:, :] == A0[N:, :])
    assert O.llt_resid(uplo, A[:N, :N], A0[:N, :N]) < 100


@pytest.mark.parametrize("uplo", [X.Lower, X.Upper])
@pytest.mark.parametrize("bad", [0, 127, 128, 250, 399])
def test_potrf_not_spd_info(uplo, bad):
    n = 400
    S = O.lltgen(uplo, n)
    S[bad, bad] = -1.0
    A = S.copy(order="F")
    assert X.potrf(A, uplo) == bad + 1
    assert O.potrf(S.copy(order="F"), O.LOWER if uplo == X.Lower else O.UPPER) == bad + 1


@pytest.mark.parametrize("uplo", [X.Lower, X.Upper])
def test_potrs_posv(uplo):
    n, nrhs = 500, 4
    S = O.lltgen(uplo, n)
    Sf = np.tril(S) + np.tril(S, -1).T if uplo == X.Lower else np.triu(S) + np.triu(S, 1).T
    Xc = O.gegen(n, nrhs)
    B = _f(Sf @ Xc)
    A = S.copy(order="F")
    assert X.potrf(A, uplo) == 0
    X.potrs(A, B, uplo)
    assert solve_ok(O.genrm1(_f(Sf)), Xc, B)
    A = S.copy(order="F")
    B = _f(Sf @ Xc)
    info = X.lib().clapack_dposv(X.ColMajor, uplo, n, nrhs, A.ctypes.data, n, B.ctypes.data, n)
    assert info == 0
    assert solve_ok(O.genrm1(_f(Sf)), Xc, B)


@pytest.mark.parametrize("nrhs", [1, 9])
@pytest.mark.parametrize("trans", [X.NoTrans, X.Trans])
def test_sgetrs_slvtst(nrhs, trans):
    n = 400
    A0 = O.gegen(n, n, dtype=np.float32)
    Xc = O.gegen(n, nrhs, dtype=np.float32)
    B = _f(((A0 if trans == X.NoTrans else A0.T).astype(np.float64) @ Xc).astype(np.float32))
    A = A0.copy(order="F")
    info, ipiv = X.getrf(A)
    assert info == 0
    assert X.getrs(A, ipiv, B, trans=trans) == 0
    assert solve_ok(O.genrm1(A0), Xc, B)


@pytest.mark.parametrize("uplo", [X.Lower, X.Upper])
@pytest.mark.parametrize("N,lda", [(1, 1), (127, 130), (129, 129), (600, 600)])
def test_spotrf_llttst(uplo, N, lda):
    A0 = O.lltgen(uplo, N, lda, dtype=np.float32)
    A = A0.copy(order="F")
    assert X.potrf(A[:N, :N], uplo) == 0
    unref = np.triu(np.ones((N, N), bool), 1) if uplo == X.Lower else np.tril(np.ones((N, N), bool), -1)
    assert np.all(A[:N, :N][unref] == A0[:N, :N][unref])
    assert O.llt_resid(uplo, A[:N, :N], A0[:N, :N]) < 100


@pytest.mark.parametrize("uplo", [X.Lower, X.Upper])
def test_spotrs_sposv_sgesv(uplo):
    n, nrhs = 300, 5
    S = O.lltgen(uplo, n, dtype=np.float32)
    Sf = np.tril(S) + np.tril(S, -1).T if uplo == X.Lower else np.triu(S) + np.triu(S, 1).T
    Xc = O.gegen(n, nrhs, dtype=np.float32)
    Bw = _f((Sf.astype(np.float64) @ Xc).astype(np.float32))
    A = S.copy(order="F")
    B = Bw.copy(order="F")
    assert X.potrf(A, uplo) == 0
    assert X.potrs(A, B, uplo) == 0
    assert solve_ok(O.genrm1(_f(Sf)), Xc, B)
    A = S.copy(order="F")
    B = Bw.copy(order="F")
    assert X.lib().clapack_sposv(X.ColMajor, uplo, n, nrhs, A.ctypes.data, n, B.ctypes.data,
                                 n) == 0
    assert solve_ok(O.genrm1(_f(Sf)), Xc, B)
    G = O.gegen(n, n, dtype=np.float32)
    B = _f((G.astype(np.float64) @ Xc).astype(np.float32))
    Gw = G.copy(order="F")
    ipiv = np.zeros(n, dtype=np.int32)
    assert X.lib().clapack_sgesv(X.ColMajor, n, nrhs, Gw.ctypes.data, n, ipiv.ctypes.data,
                                 B.ctypes.data, n) == 0
    assert solve_ok(O.genrm1(G), Xc, B)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_getrf_tall_panel_beyond_131072_rows(dt):
    """A 200,000 x 64 LU: taller than one cooperative panel launch holds;
    pivots against the oracle and the lutst residual."""
    M, N = 200000, 64
    A0 = _f(RNG.uniform(-1, 1, (M, N)).astype(dt))
    A = A0.copy(order="F")
    info, ipiv = X.getrf(A)
    assert info == 0
    Ao = A0.copy(order="F")
    gaps = np.ones(N)
    _, opiv = O.getrf(Ao, gaps=gaps)
    near = np.nonzero(gaps < (1e-9 if dt == np.float64 else 1e-4))[0]
    stop = int(near[0]) if near.size else N
    assert np.array_equal(ipiv[:stop], opiv[:stop])
    assert O.lu_resid(A, ipiv, A0) < 100


def test_two_host_threads():
    """cblas_dgemm and clapack_dgetrf called from two host threads at once
    (each thread has its own stream and workspaces)."""
    import threading
    res = {}

    def gemm_job():
        A = _f(RNG.uniform(-0.5, 0.5, (300, 200)))
        B = _f(RNG.uniform(-0.5, 0.5, (200, 250)))
        for _ in range(20):
            Cm = np.zeros((300, 250), order="F")
            X.gemm(X.NoTrans, X.NoTrans, 1.0, A, B, 0.0, Cm)
        res["gemm"] = np.abs(Cm - A @ B).max()

    def lu_job():
        worst = 0.0
        for k in range(10):
            A0 = O.gegen(400, 400, seed=k + 5)
            A = A0.copy(order="F")
            info, ipiv = X.getrf(A)
            worst = max(worst, O.lu_resid(A, ipiv, A0) if info == 0 else 1e9)
        res["lu"] = worst

    th = [threading.Thread(target=gemm_job), threading.Thread(target=lu_job)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert res["gemm"] < 1e-12
    assert res["lu"] < 100
