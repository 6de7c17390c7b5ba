"""Seeded synthetic inputs of the benchmark configurations (shared by
bench.py and the GPU tests).  Matrices are column-major float64 CUDA
tensors held as their row-major transposes, like the rest of atlas_amd.

* haar_system: A = U diag(sigma) V^T, U and V Haar-random orthogonal (QR of
  a seeded Gaussian, columns sign-fixed by diag(R)), sigma geometric from 1
  down to 1/kappa; right-hand sides U[-1,1).  The product runs on the
  library's own DGEMM.
* uniform_matrix: U[-1,1) entries from numpy's PCG64 stream (identical on
  every machine), the LU configuration's input.
* spd_matrix: B B^T / n + I with B standard normal (numpy stream), formed
  on the GPU, the Cholesky configuration's input.
"""
import numpy as np


def haar_system(n, kappa, nrhs, seed, torch, X):
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(seed)

    def haar():
        G = torch.randn(n, n, dtype=torch.float64, device=dev, generator=g)
        Q, R = torch.linalg.qr(G)
        Q *= torch.sign(torch.diagonal(R)).unsqueeze(0)
        return Q

    U = haar()
    V = haar()
    sigma = kappa ** (-torch.arange(n, dtype=torch.float64, device=dev) / max(n - 1, 1))
    U *= sigma.unsqueeze(0)                       # U diag(sigma)
    A = torch.empty(n, n, dtype=torch.float64, device=dev)
    X.use_torch_stream()
    # Column-major views of the row-major tensors: Vc = V^T, Uc = (U diag s)^T,
    # and the matrix the library sees in A's storage is A^T = V diag(s) U^T
    # = Vc^T Uc, whose condition number is kappa as well.
    X.lib().atlmx_dgemm(X.Trans, X.NoTrans, n, n, n, 1.0, X.dptr(V), n, X.dptr(U), n, 0.0,
                        X.dptr(A), n)
    torch.cuda.synchronize()
    del U, V
    b = torch.rand(nrhs, n, dtype=torch.float64, device=dev, generator=g) * 2 - 1
    return A, b


def uniform_matrix(n, seed, m=None):
    """m x n (default n x n) column-major numpy array, entries U[-1,1)."""
    m = n if m is None else m
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.uniform(-1.0, 1.0, (m, n)))


def spd_matrix(n, seed, torch):
    """(host column-major array, device tensor holding the same matrix)."""
    rng = np.random.default_rng(seed)
    B = torch.tensor(rng.standard_normal((n, n)), dtype=torch.float64, device="cuda")
    S = B @ B.T
    S /= n
    S += torch.eye(n, dtype=torch.float64, device="cuda")
    S = (S + S.T) * 0.5                       # exactly symmetric
    del B
    return np.asfortranarray(S.cpu().numpy()), S
