"""Write the golden fixtures under tests/golden/ (run from the repo root).

Nothing here calls the oracle: every expected value comes from an
independent implementation, so the fixtures pin the oracle rather than echo
it.

* lcg_streams.json -- ATLAS's tester random streams (src/testing/ATL_rand.c).
  The first value X(1) comes from the closed form that the two-int arithmetic
  of ATL_srand reduces to: the seed is held as {iseed, iseed>>16} with the low
  word unmasked (ATL_rand.c:254), i.e. X(0) = iseed + (iseed>>16)*2^16, and
  X(1) = (1103515245*X(0) + 12345) mod 2^31.  X(2), X(3), ... come from
  glibc's random() in its TYPE_0 mode (an 8-byte state), which is the same
  recurrence implemented by the C library.
* lu_pivots.json -- partial-pivoting pivots from LAPACK dgetrf (through
  scipy.linalg.lu_factor) for ATLAS lutst inputs (gegen, seed M*N+lda),
  truncated before the first column whose two largest candidates are within
  a relative 1e-9 of each other.
* c2_getrf_pivots_n8192.npz (`python tests/golden/make_golden.py c2`, about
  3 minutes) -- LAPACK's pivots for the n = 8192 LU configuration's
  U[-1,1) input, with the near-tie screen of a blocked numpy elimination.
"""
import ctypes
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SEEDS = [
    1, 2, 12345, 65535, 65536, 65537, 2**31 - 1, 1000000007,
    2048 * 2048 + 2048,      # lutst / GetGE at n=2048
    2048 * 1112,             # gemmtst A, K=2048
    2048 * 2238,             # gemmtst B, N=2048
    2048 * 2048,             # gemmtst C, M=N=2048
    8192 * 8192 + 8192,      # lutst at n=8192
    8192 * 1029 + 8192,      # llttst at n=8192
    167 * 167 + 167,         # dsanity_test slvtst n=167
    300 * 300 + 300,
]


def x1_closed_form(iseed):
    x0 = (iseed + ((iseed >> 16) << 16)) % 2**31
    return (1103515245 * x0 + 12345) % 2**31


_BUF = ctypes.create_string_buffer(8)


def glibc_type0_stream(x1, n):
    libc = ctypes.CDLL("libc.so.6")
    libc.initstate.restype = ctypes.c_void_p
    libc.initstate.argtypes = [ctypes.c_uint, ctypes.c_void_p, ctypes.c_size_t]
    libc.setstate.restype = ctypes.c_void_p
    libc.setstate.argtypes = [ctypes.c_void_p]
    libc.random.restype = ctypes.c_long
    prev = libc.initstate(ctypes.c_uint(x1), _BUF, 8)   # TYPE_0: state[0] = x1
    out = [x1]
    for _ in range(n - 1):
        out.append(int(libc.random()))
    libc.setstate(prev)                                 # give libc its own state back
    return out


def lcg_fixture(n=64):
    return {str(s): glibc_type0_stream(x1_closed_form(s), n) for s in SEEDS}


def gegen_from_stream(stream, M, N):
    v = 0.5 - np.asarray(stream, dtype=np.float64) / 2147483648.0
    return v[: M * N].reshape((N, M)).T.copy(order="F")


def lu_fixture():
    import scipy.linalg as sla
    out = {}
    for n in (5, 17, 64, 100, 257):
        seed = n * n + n
        stream = glibc_type0_stream(x1_closed_form(seed), n * n)
        A = gegen_from_stream(stream, n, n)
        lu, piv = sla.lu_factor(A)
        # near-tie screening: redo the elimination in extended precision-free
        # float64 numpy to measure the top-2 gap of every pivot column
        W = A.copy()
        stop = n
        for j in range(n):
            col = np.abs(W[j:, j])
            order = np.argsort(-col, kind="stable")
            if col.size > 1 and col[order[0]] > 0:
                gap = (col[order[0]] - col[order[1]]) / col[order[0]]
                if gap < 1e-9:
                    stop = j
                    break
            p = j + int(np.argmax(col))
            W[[j, p]] = W[[p, j]]
            if W[j, j] != 0:
                W[j + 1:, j] /= W[j, j]
                W[j + 1:, j + 1:] -= np.outer(W[j + 1:, j], W[j, j + 1:])
        out[str(n)] = {"seed": seed, "lda": n, "ipiv": [int(x) for x in piv[:stop]],
                       "checked_columns": stop}
    return out


if __name__ == "__main__" and len(__import__("sys").argv) == 1:
    with open(os.path.join(HERE, "lcg_streams.json"), "w") as f:
        json.dump(lcg_fixture(), f, indent=0)
    with open(os.path.join(HERE, "lu_pivots.json"), "w") as f:
        json.dump(lu_fixture(), f, indent=0)
    print("wrote", os.listdir(HERE))


def c2_pivot_fixture(n=8192, seed=20260801, nb=64, tie=1e-9):
    """Partial-pivoting pivots of the LU configuration (n = 8192, A ~
    U[-1,1) from numpy's PCG64 stream with `seed`, as
    atlas_amd.synth.uniform_matrix makes it), from LAPACK dgetrf through
    scipy, cross-checked against a blocked numpy elimination that also
    records the relative gap between the two largest candidates of every
    pivot column; columns from the first gap below `tie` on are not
    compared.  Returns a dict of arrays for np.savez_compressed."""
    import scipy.linalg as sla
    rng = np.random.default_rng(seed)
    A = np.asfortranarray(rng.uniform(-1.0, 1.0, (n, n)))
    _, piv = sla.lu_factor(A.copy(order="F"), check_finite=False)
    W = A.copy(order="F")
    gaps = np.ones(n)
    mypiv = np.zeros(n, dtype=np.int64)
    for j0 in range(0, n, nb):
        j1 = min(n, j0 + nb)
        for j in range(j0, j1):             # panel, right-looking inside
            col = np.abs(W[j:, j])
            p = int(np.argmax(col))
            top = col[p]
            if col.size > 1 and top > 0:
                rest = col.copy()
                rest[p] = -1.0
                gaps[j] = (top - rest.max()) / top
            p += j
            mypiv[j] = p
            if p != j:
                W[[j, p], :] = W[[p, j], :]
            if W[j, j] != 0:
                W[j + 1:, j] /= W[j, j]
                W[j + 1:, j + 1:j1] -= np.outer(W[j + 1:, j], W[j, j + 1:j1])
        if j1 < n:                          # U12 and the trailing update
            L11 = np.tril(W[j0:j1, j0:j1], -1) + np.eye(j1 - j0)
            W[j0:j1, j1:] = sla.solve_triangular(L11, W[j0:j1, j1:], lower=True,
                                                 unit_diagonal=True)
            W[j1:, j1:] -= W[j1:, j0:j1] @ W[j0:j1, j1:]
    near = np.nonzero(gaps < tie)[0]
    stop = int(near[0]) if near.size else n
    assert np.array_equal(mypiv[:stop], piv[:stop]), "numpy elimination disagrees with LAPACK"
    return {"n": np.int64(n), "seed": np.int64(seed), "ipiv": piv.astype(np.int32),
            "checked_columns": np.int64(stop), "min_gap": np.float64(gaps[:stop].min()),
            "a_checksum": np.float64(A.sum()), "a_corner": A[:4, :4].copy()}


if __name__ == "__main__" and len(__import__("sys").argv) > 1 and \
        __import__("sys").argv[1] == "c2":
    fx = c2_pivot_fixture()
    np.savez_compressed(os.path.join(HERE, "c2_getrf_pivots_n8192.npz"), **fx)
    print("c2 fixture: checked columns", int(fx["checked_columns"]), "min gap",
          float(fx["min_gap"]))
