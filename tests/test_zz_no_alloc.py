"""atlmx_dgetrs_noalloc, the allocation-free LU solve, against atlmx_dgetrs.

Both run the same kernels in the same order, so the solutions must be
bit-identical; each case also passes slvtst's residual (bin/slvtst.c:419-441)
against the known solution.  Two calls in a row, the second smaller, check
that the shared static scratch array carries nothing from one call to the
next (flags cleared, inverses rewritten).  Argument errors are refused
before any GPU work.
"""
import ctypes as C

import numpy as np
import pytest

import atlas_amd as X
import oracle as O
from atlas_amd import synth
from tests.checks import solve_ok

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    t.cuda.set_device(0)
    X.use_torch_stream()
    return t


def _factor(torch, n, seed):
    A0 = synth.uniform_matrix(n, seed)
    Ad = torch.tensor(A0.T, device="cuda")            # storage = A, column-major
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    assert X.lib().atlmx_dgetrf(n, n, X.dptr(Ad), n, X.dptr(ipiv)) == 0
    return A0, Ad, ipiv


def _solve_both(torch, A0, Ad, ipiv, nrhs, trans, seed):
    n = A0.shape[0]
    L = X.lib()
    Xc = synth.uniform_matrix(nrhs, seed, m=n)
    op = A0 if trans == X.NoTrans else A0.T
    B0 = np.asfortranarray(op @ Xc)
    Bref = torch.tensor(B0.T, device="cuda")          # rows hold B's columns
    Bna = Bref.clone()
    L.atlmx_dgetrs(trans, n, nrhs, X.dptr(Ad), n, X.dptr(ipiv), X.dptr(Bref), n)
    print("atlmx_dgetrs_noalloc n=%d nrhs=%d trans=%d" % (n, nrhs, trans), flush=True)
    rc = L.atlmx_dgetrs_noalloc(trans, n, nrhs, X.dptr(Ad), n, X.dptr(ipiv), X.dptr(Bna), n)
    torch.cuda.synchronize()
    assert rc == 0
    xr, xn = Bref.cpu().numpy(), Bna.cpu().numpy()
    assert np.array_equal(xr.view(np.uint64), xn.view(np.uint64))
    assert np.isfinite(xn).all()
    assert solve_ok(O.genrm1(A0), Xc, np.asfortranarray(xn.T))


@pytest.mark.parametrize("nrhs", [1, 3, 16, 70])     # 70: two launches of the chained solve
@pytest.mark.parametrize("trans", [X.NoTrans, X.Trans], ids=["N", "T"])
def test_noalloc_matches_getrs_two_calls(torch, nrhs, trans):
    for i, n in enumerate((1500, 333)):              # the second call is smaller
        A0, Ad, ipiv = _factor(torch, n, 9100 + n)
        _solve_both(torch, A0, Ad, ipiv, nrhs, trans, 17 + 31 * i + nrhs)


def test_noalloc_at_its_maximum_order(torch):
    n = 16384                                          # ATLMX_NOALLOC_MAX_N
    A0, Ad, ipiv = _factor(torch, n, 4711)
    _solve_both(torch, A0, Ad, ipiv, 2, X.NoTrans, 5)
    del Ad
    A0, Ad, ipiv = _factor(torch, 200, 4712)           # then a small one again
    _solve_both(torch, A0, Ad, ipiv, 1, X.NoTrans, 6)


def test_noalloc_refuses_bad_arguments(torch, capfd):
    L = X.lib()
    B = torch.zeros(4, dtype=torch.float64, device="cuda")
    assert L.atlmx_dgetrs_noalloc(X.NoTrans, 16385, 1, None, 16385, None, X.dptr(B),
                                  16385) == -2
    assert "16384" in capfd.readouterr().err
    assert L.atlmx_dgetrs_noalloc(X.NoTrans, 10, 1, None, 5, None, X.dptr(B), 10) == -5
    assert L.atlmx_dgetrs_noalloc(7, 10, 1, None, 10, None, X.dptr(B), 10) == -1
    assert L.atlmx_noalloc_scratch_bytes() >= 16384 * 64 * 8
