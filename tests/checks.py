"""Shared pass/fail checks of the GPU parity tests."""
import numpy as np

import oracle as O

THRESH = 100.0      # lutst/llttst/slvtst threshold (bin/lutst.c:398, bin/slvtst.c:75)


def solve_ok(nrmA, Xc, X, thresh=THRESH):
    """slvtst's residual (bin/slvtst.c:419-441) below the threshold AND every
    entry finite: the tester's difference norm skips NaN (it compares with
    '>'), so a NaN solution would otherwise pass."""
    if not np.isfinite(X).all():
        return False
    return O.slv_resid(nrmA, Xc, X) < thresh


# --- closed forms of the complex rank-k, HEMM/SYMM and TRMM (column-major,
# ATLAS enum values).  Pinned against ATLAS's own reference routines
# (src/blas/reference/level3/ATL_zref{herk,her2k,syrk,syr2k,hemm,symm,trmm}.c)
# by tests/test_oracle_complex.py, then used as the GPU tests' expectation.

def _cop(M, t):
    M = M.astype(np.complex128)
    if t == O.NOTRANS:
        return M
    if t == O.TRANS:
        return M.T
    return M.conj().T


def crk_np(herm, uplo, trans, alpha, A, B, beta, C):
    """C <- alpha P + beta C on C's uplo triangle, the other triangle kept:
    P = op(A) op(A)^H (herk), op(A) op(B)^H + op(B) op(A)^H with conj(alpha)
    on the second (her2k), or the same with ^T and no conjugate (syrk,
    syr2k); op = N or C (herm) / T.  Hermitian forms keep the diagonal real;
    beta = 0 never reads C; (alpha = 0 or K = 0) and beta = 1 returns C as is
    (ATL_zrefherk.c's quick return, diagonal untouched)."""
    C = C.astype(np.complex128)
    n = C.shape[0]
    K = A.shape[1] if trans == O.NOTRANS else A.shape[0]
    if n == 0 or ((alpha == 0 or K == 0) and beta == 1):
        return C
    h = (lambda M: M.conj().T) if herm else (lambda M: M.T)
    opA = A.astype(np.complex128) if trans == O.NOTRANS else h(A.astype(np.complex128))
    if B is None:
        P = alpha * (opA @ h(opA))
    else:
        opB = B.astype(np.complex128) if trans == O.NOTRANS else h(B.astype(np.complex128))
        a2 = np.conj(alpha) if herm else alpha
        P = alpha * (opA @ h(opB)) + a2 * (opB @ h(opA))
    Cs = np.zeros_like(C) if beta == 0 else beta * C
    if herm:
        d = np.diag_indices(n)
        Cs[d] = 0 if beta == 0 else beta * C[d].real
    out = Cs + P
    if herm:
        out[np.diag_indices(n)] = out[np.diag_indices(n)].real
    tri = np.tril(np.ones((n, n), bool)) if uplo == O.LOWER else np.triu(np.ones((n, n), bool))
    return np.where(tri, out, C)


def csym_full(herm, uplo, A):
    """the full matrix of a Hermitian (herm) / symmetric A held in its uplo
    triangle; a Hermitian diagonal's imaginary part is not referenced"""
    A = A.astype(np.complex128)
    T = np.tril(A) if uplo == O.LOWER else np.triu(A)
    S = np.tril(A, -1) if uplo == O.LOWER else np.triu(A, 1)
    F = T + (S.conj().T if herm else S.T)
    if herm:
        F[np.diag_indices(A.shape[0])] = F[np.diag_indices(A.shape[0])].real
    return F


def chemm_np(herm, side, uplo, alpha, A, B, beta, C):
    F = csym_full(herm, uplo, A)
    P = F @ B.astype(np.complex128) if side == O.LEFT else B.astype(np.complex128) @ F
    return alpha * P + (0 if beta == 0 else beta * C.astype(np.complex128))


def ctrmm_np(side, uplo, ta, diag, alpha, A, B):
    T = np.tril(A) if uplo == O.LOWER else np.triu(A)
    T = T.astype(np.complex128)
    if diag == O.UNIT:
        T[np.diag_indices(T.shape[0])] = 1.0
    T = _cop(T, ta)
    B = B.astype(np.complex128)
    return alpha * (T @ B if side == O.LEFT else B @ T)
