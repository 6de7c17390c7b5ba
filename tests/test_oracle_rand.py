"""Pins the oracle's LCG (restating src/testing/ATL_rand.c) and generators.

The fixture streams come from glibc's TYPE_0 random() seeded with a closed
form of ATL_srand (tests/golden/make_golden.py), not from the oracle.
"""
import json
import os

import numpy as np

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_lcg_matches_glibc_type0_and_closed_form(orc):
    g = json.load(open(os.path.join(GOLD, "lcg_streams.json")))
    assert len(g) >= 16
    for seed, want in g.items():
        got = orc.rand_stream(int(seed), len(want)).tolist()
        assert got == want, seed


def test_lcg_recurrence_big_int():
    """X(n+1) = (1103515245*X(n) + 12345) mod 2^31 on the oracle's own output."""
    import oracle
    s = oracle.rand_stream(987654321, 2000).astype(np.int64)
    nxt = (1103515245 * s[:-1] + 12345) % (1 << 31)
    assert np.array_equal(nxt, s[1:])
    assert s.min() >= 0 and s.max() < (1 << 31)


def test_unmasked_seed_word():
    """ATL_rand.c:254 keeps jseed[0]=iseed unmasked: seeds 0x10000+k and k differ
    in X(0) by 2*2^16, not 2^16 (a masked restatement would fail this)."""
    import oracle
    for iseed in (65536 + 7, 3 * 65536 + 11, 2**30 + 5):
        x0 = (iseed + ((iseed >> 16) << 16)) % 2**31
        x1 = (1103515245 * x0 + 12345) % 2**31
        assert int(oracle.rand_stream(iseed, 1)[0]) == x1
        x0_masked = iseed % 2**31
        assert (1103515245 * x0_masked + 12345) % 2**31 != x1


def test_gegen_layout_and_gap(orc):
    M, N, lda = 7, 5, 10
    A = orc.gegen(M, N, lda)
    s = orc.rand_stream(M * N + lda, M * N)
    want = (0.5 - s / 2147483648.0).reshape(N, M).T
    assert np.array_equal(A[:M, :], want)
    assert np.all(A[M:, :] == -2.56e9)                 # ATL_gegen.c:6 FILLCONST
    assert orc.lib().orc_dgechkgap(M, N, A.ctypes.data, lda) == 0
    A[M + 1, 2] = 0.0
    assert orc.lib().orc_dgechkgap(M, N, A.ctypes.data, lda) == 1
    vals = A[:M, :]
    assert vals.min() > -0.5 and vals.max() <= 0.5


def test_gegen_float_is_rounded_double(orc):
    A = orc.gegen(9, 4, dtype=np.float32)
    D = orc.gegen(9, 4, dtype=np.float64)
    assert np.array_equal(A, D.astype(np.float32))


def test_matgen_pad(orc):
    A = orc.matgen(3, 2, 5, 1112)
    assert np.all(A[3:, :] == -9e27)


def test_epsilon(orc):
    assert orc.epsilon() == 2.0 ** -52
    assert orc.epsilon(np.float32) == np.float32(2.0 ** -23)
