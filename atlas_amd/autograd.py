"""torch.autograd.Function wrappers of the batched solves and their
hand-written adjoint kernels (no ATLAS counterpart).

    X = CholeskySolve.apply(A, B)   # A: (batch, n, n) SPD, lower triangle read
    X = LUSolve.apply(A, B)         # A: (batch, n, n) general
    B: (batch, n, nrhs), nrhs <= 8; float64 or float32 CUDA tensors.

Forward runs the batched factor+solve (potrf+potrs or getrf+getrs) on a
copy of A.  Backward recomputes the factorization from the saved A rather
than keeping it, then
  LU:       B-bar = A^{-T} X-bar (getrs with the transpose),
            A-bar = -B-bar X^T   (batched GEMM);
  Cholesky: B-bar = A^{-1} X-bar (potrs), then the Cholesky adjoint kernel:
            L-bar = -(B-bar X^T + X B-bar^T) L, P = Phi(L^T L-bar),
            A-bar = L^{-T} P L^{-1} symmetrised, returned for the lower
            triangle only (the entries potrf reads; each off-diagonal entry
            carries both of its positions, the upper triangle gets zero).
Matrices are handed to the library column-major (as transposed
contiguous copies)."""
import torch

from . import posv_batch, gesv_batch, gesv_t_batch, gemm_batch, posv_adjoint_batch
from . import NoTrans, Trans, use_torch_stream


def _cm(t):
    """column-major copy: storage of t^T, returned as (..., cols, rows)"""
    return t.transpose(-1, -2).contiguous()


def _check(A, B):
    if A.dim() != 3 or B.dim() != 3 or A.shape[1] != A.shape[2] or B.shape[1] != A.shape[1]:
        raise ValueError("A must be (batch, n, n) and B (batch, n, nrhs)")
    if A.dtype not in (torch.float64, torch.float32) or B.dtype != A.dtype:
        raise ValueError("float64 or float32 tensors of one dtype")
    if not A.is_cuda or not B.is_cuda:
        raise ValueError("CUDA tensors")


def _posv(Aw, Xw):
    info = torch.zeros(Aw.shape[0], dtype=torch.int32, device=Aw.device)
    with posv_batch(list(Aw), list(Xw), info) as plan:
        plan.run()
    return info


def _gesv(Aw, Xw, trans=False):
    n = Aw.shape[1]
    piv = torch.zeros(Aw.shape[0], n, dtype=torch.int32, device=Aw.device)
    info = torch.zeros(Aw.shape[0], dtype=torch.int32, device=Aw.device)
    mk = gesv_t_batch if trans else gesv_batch
    with mk(list(Aw), list(piv), list(Xw), info) as plan:
        plan.run()
    return info


class CholeskySolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A, B):
        _check(A, B)
        use_torch_stream()
        Aw, Xw = _cm(A).clone(), _cm(B).clone()          # (batch, n, n), (batch, nrhs, n)
        info = _posv(Aw, Xw)
        if int(info.abs().sum().item()):
            raise RuntimeError("CholeskySolve: a matrix is not positive definite")
        X = Xw.transpose(-1, -2)
        ctx.save_for_backward(A, X)
        return X.contiguous()

    @staticmethod
    def backward(ctx, Xb):
        A, X = ctx.saved_tensors
        use_torch_stream()
        Lw, Bb = _cm(A).clone(), _cm(Xb).clone()
        _posv(Lw, Bb)                                   # L recomputed, Bb = A^{-1} Xb
        Xc = _cm(X)
        G = torch.empty_like(Lw)
        W = torch.empty_like(Lw)
        with posv_adjoint_batch(list(Lw), list(Xc), list(Bb), list(G), list(W)) as plan:
            plan.run()
        return G.transpose(-1, -2), Bb.transpose(-1, -2)


class LUSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A, B):
        _check(A, B)
        use_torch_stream()
        Aw, Xw = _cm(A).clone(), _cm(B).clone()
        info = _gesv(Aw, Xw)
        if int(info.abs().sum().item()):
            raise RuntimeError("LUSolve: a matrix is singular")
        X = Xw.transpose(-1, -2)
        ctx.save_for_backward(A, X)
        return X.contiguous()

    @staticmethod
    def backward(ctx, Xb):
        A, X = ctx.saved_tensors
        use_torch_stream()
        Lw, Bb = _cm(A).clone(), _cm(Xb).clone()
        _gesv(Lw, Bb, trans=True)                       # LU recomputed, Bb = A^{-T} Xb
        Ga = torch.empty_like(Lw)                       # column-major n x n per instance
        Xc = _cm(X)
        # A-bar = -Bb X^T: column-major Bb (n x nrhs), X (n x nrhs) -> op(B) = X^T
        with gemm_batch(NoTrans, Trans, -1.0, list(Bb), list(Xc), 0.0, list(Ga)) as plan:
            plan.run()
        return Ga.transpose(-1, -2), Bb.transpose(-1, -2)
