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
