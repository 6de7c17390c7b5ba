"""GPU checks of the mixed-precision solver (FP32 LU + FP64 refinement) and
of its kernels: the tournament's pivot choice against a numpy restatement of
the tournament, its move list and row moves, the root's diagonal-block
factor against LU = (chosen rows) and X inv(X) = I, and every solve against the
FP64 accuracy bar: normwise backward error <= n * eps64 and forward error
against the known solution within kappa * n * eps64."""
import ctypes as C

import numpy as np
import pytest

import atlas_amd as X

pytestmark = pytest.mark.gpu
EPS64 = 2.0 ** -52


@pytest.fixture(scope="module")
def torch():
    import torch as t
    X.use_torch_stream()
    return t


def _dev(t, a, dtype):
    """column-major device copy of a numpy matrix (stored as its transpose)"""
    return t.tensor(np.ascontiguousarray(np.asarray(a).T), dtype=dtype, device="cuda")


def _host(tt):
    return tt.cpu().numpy().T


def _k(name, *args):
    L = X.lib()
    fn = getattr(L, name)
    fn.restype = C.c_int
    r = fn(*args)
    assert r == 0, (name, r)


def _tournament(A, k, w, leaf=256, fan=4):
    """numpy restatement of the tournament (the checker of the GPU kernel):
    partial-pivoting elimination in FP32 on 256-row leaves of the panel
    A[k:, k:k+w], then on the original rows of up to 4 children's choices,
    level by level; returns the root's chosen rows in pivot order."""
    def gepp(rows):
        P = A[rows, k:k + w].astype(np.float32).copy()
        act = np.ones(len(rows), bool)
        sel = []
        for c in range(w):
            if not act.any():
                break
            v = np.where(act, np.abs(P[:, c]), -1.0)
            v = np.where(np.isnan(v), 0.0, v)
            p = int(np.argmax(v))                   # first maximum
            sel.append(rows[p])
            act[p] = False
            if P[p, c] != 0:
                lcol = P[:, c] / P[p, c]
                P[act, c + 1:] -= np.outer(lcol[act], P[p, c + 1:])
        return sel
    n = A.shape[0]
    level = [gepp(list(range(s, min(s + leaf, n)))) for s in range(k, n, leaf)]
    while len(level) > 1:
        level = [gepp([r for ch in level[i:i + fan] for r in ch]) for i in range(0, len(level), fan)]
    return level[0]


def _tslu_select(torch, A, k, w):
    n = A.shape[0]
    lds = (n + 3) // 4 * 4
    S = torch.zeros(n, lds, dtype=torch.float32, device="cuda")     # column j of S = S[j]
    S[:, :n] = torch.tensor(A.astype(np.float32), device="cuda")    # S = A^T: S[j][i] = A[i][j]
    L = X.lib()
    L.atlmx_k_tslu_ws_ints.restype = C.c_size_t
    nws = L.atlmx_k_tslu_ws_ints(n)
    ws = torch.zeros(nws, dtype=torch.int32, device="cuda")
    plan = torch.zeros(2 + 4 * 64, dtype=torch.int32, device="cuda")
    perm = torch.arange(n, dtype=torch.int32, device="cuda")
    blk = torch.zeros(3, 64, 64, dtype=torch.float32, device="cuda")   # lu11, invL, invU
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    _k("atlmx_k_tslu_select", X.dptr(S), C.c_long(lds), n, k, w, X.dptr(ws), X.dptr(plan),
       X.dptr(perm), X.dptr(blk[0]), X.dptr(blk[1]), X.dptr(blk[2]), X.dptr(info), None, None,
       None)
    torch.cuda.synchronize()
    return S, ws, plan, perm, blk, info


@pytest.mark.parametrize("n,k,w", [(64, 0, 64), (200, 0, 64), (300, 64, 64), (1500, 0, 64),
                                   (5000, 128, 64), (70, 64, 6), (20000, 0, 64)])
def test_tournament_choice_and_moves(torch, n, k, w):
    rng = np.random.default_rng(n + k)
    A = rng.uniform(-1, 1, (n, n))
    S, ws, dplan, dperm, blk, info = _tslu_select(torch, A, k, w)
    plan, perm = dplan.cpu().numpy(), dperm.cpu().numpy()
    want = _tournament(A, k, min(w, n - k))
    ws_ = len(want)
    assert int(info.item()) == 0
    # perm: rows k.. in the tournament's pivot order, the rest a permutation
    assert list(perm[k:k + ws_]) == want
    assert sorted(perm) == list(range(n)) and list(perm[:k]) == list(range(k))
    # the move list is the same permutation of rows (dst <- src)
    cur = np.arange(n)
    mv = plan[2:2 + 2 * plan[0]].reshape(-1, 2)
    new = cur.copy()
    new[mv[:, 0]] = cur[mv[:, 1]]
    assert np.array_equal(new, perm) and plan[0] <= 2 * ws_
    # the counters (one per tree node of the k = 0 tree) are left zero
    nodes = sz = -(-n // 256)
    while sz > 1:
        sz = -(-sz // 4)
        nodes += sz
    assert int(ws[:nodes].abs().sum().item()) == 0
    # the diagonal block: L11 U11 = the chosen rows' block, and the inverses
    lu = blk[0].cpu().numpy().T.astype(np.float64)[:ws_, :ws_]
    L11 = np.tril(lu, -1) + np.eye(ws_)
    U11 = np.triu(lu)
    D = A[want][:, k:k + ws_]
    assert np.abs(L11 @ U11 - D).max() <= 1e-4 * max(1.0, np.abs(D).max())
    assert np.abs(L11).max() <= 1.0 + 1e-6            # partial pivoting at the root
    iL = blk[1].cpu().numpy().T.astype(np.float64)
    iU = blk[2].cpu().numpy().T.astype(np.float64)
    assert np.abs(iL[:ws_, :ws_] @ L11 - np.eye(ws_)).max() < 1e-4
    assert np.abs(iU[:ws_, :ws_] @ U11 - np.eye(ws_)).max() < 1e-4 * np.linalg.cond(U11)
    assert np.array_equal(iL[ws_:, ws_:], np.eye(64 - ws_))
    # applying the moves to S moves whole rows of A and writes the factored block
    # in two pieces, as the lookahead applies them: the pair's rows with the
    # factored block first, then the other rows around that gap
    k2 = min(n, k + 2 * 64)
    _k("atlmx_k_tslu_swap", X.dptr(S), C.c_long(S.shape[1]), k, k2, 0, 0, X.dptr(dplan), k, ws_,
       X.dptr(blk[0]), X.dptr(info), None)
    _k("atlmx_k_tslu_swap", X.dptr(S), C.c_long(S.shape[1]), 0, n, k, k2, X.dptr(dplan), k, ws_,
       None, X.dptr(info), None)
    torch.cuda.synchronize()
    Sh = S[:, :n].cpu().numpy()
    want_rows = A[perm].astype(np.float32)
    want_rows[k:k + ws_, k:k + ws_] = blk[0].cpu().numpy().T[:ws_, :ws_]
    assert np.array_equal(Sh, want_rows)


@pytest.mark.parametrize("n,k,e0,e1", [(700, 128, 128, 256), (700, 192, 128, 256),
                                       (300, 256, 256, 300), (5000, 0, 0, 128)])
def test_select_with_moves_matches_select_then_swap(torch, n, k, e0, e1):
    """The tournament launch that also applies its moves to S's rows [e0, e1)
    and writes the factored block leaves S, the plan, perm and the block
    exactly as the tournament followed by atlmx_k_tslu_swap does; where S
    rows [k + w, e1) remain, it also forms U12 = Ut inv(L11)^T there (FP32
    sums, checked against the product of the unfused path's Ut and inverse)."""
    rng = np.random.default_rng(7 * n + k)
    A = rng.uniform(-1, 1, (n, n))
    w = min(64, n - k)
    outs = []
    for fused in (False, True):
        lds = (n + 3) // 4 * 4
        S = torch.zeros(n, lds, dtype=torch.float32, device="cuda")
        S[:, :n] = torch.tensor(A.astype(np.float32), device="cuda")
        L = X.lib()
        L.atlmx_k_tslu_ws_ints.restype = C.c_size_t
        ws = torch.zeros(L.atlmx_k_tslu_ws_ints(n), dtype=torch.int32, device="cuda")
        plan = torch.zeros(2 + 4 * 64, dtype=torch.int32, device="cuda")
        perm = torch.arange(n, dtype=torch.int32, device="cuda")
        blk = torch.zeros(3, 64, 64, dtype=torch.float32, device="cuda")
        info = torch.zeros(1, dtype=torch.int32, device="cuda")
        args = (X.dptr(S), C.c_long(lds), n, k, w, X.dptr(ws), X.dptr(plan), X.dptr(perm),
                X.dptr(blk[0]), X.dptr(blk[1]), X.dptr(blk[2]), X.dptr(info), None, None)
        if fused:
            _k("atlmx_k_tslu_select_swap", *args, e0, e1, e1, None)
        else:
            _k("atlmx_k_tslu_select", *args, None)
            _k("atlmx_k_tslu_swap", X.dptr(S), C.c_long(lds), e0, e1, 0, 0, X.dptr(plan), k, w,
               X.dptr(blk[0]), X.dptr(info), None)
        torch.cuda.synchronize()
        outs.append([t.cpu().numpy() for t in (S, plan, perm, blk, info)])
    S0, S1 = outs[0][0], outs[1][0]
    r0 = k + w
    if e1 > r0:
        # S rows [r0, e1) at S columns k..k+w-1 (A's rows k..k+w-1): U12
        Ut = S0[k:k + w, r0:e1].T.astype(np.float64)          # (rows, t)
        iL = outs[0][3][1].T.astype(np.float64)[:w, :w]       # inv(L11)(j, t)
        want = Ut @ iL.T
        got = S1[k:k + w, r0:e1].T.astype(np.float64)
        scale = np.abs(Ut) @ np.abs(iL).T
        assert np.all(np.abs(got - want) <= 64 * 2.0**-23 * scale + 1e-30)
        S0 = S0.copy(); S1 = S1.copy()
        S0[k:k + w, r0:e1] = 0; S1[k:k + w, r0:e1] = 0
    assert np.array_equal(S0, S1)
    for a, b in zip(outs[0][1:], outs[1][1:]):
        assert np.array_equal(a, b)
    assert int(outs[0][4][0]) == 0


def _as_lp(x, lp):
    """FP32 values as the 16-bit update operands they become (round to
    nearest even): lp 1 bfloat16, lp 2 IEEE half."""
    x = np.asarray(x, np.float32)
    if lp == 1:
        u = x.view(np.uint32).astype(np.uint64)
        u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
        return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32)
    if lp == 2:
        return x.astype(np.float16).astype(np.float32)
    return x


@pytest.mark.parametrize("n,k,wq,lp", [(300, 64, 64, 0), (200, 64, 8, 0), (130, 0, 64, 0),
                                       (300, 64, 64, 1), (300, 64, 40, 2)])
def test_l21_and_second_panel_update_vs_numpy(torch, n, k, wq, lp):
    """One launch forms P's L21 = inv(U11)^T Lt over every S column from
    k + 64 and subtracts U12 L21 from the second panel's wq S rows, with the
    update's operands rounded to 16 bits for lp = 1, 2 (checked against
    numpy on the kernel's own L21, so a rounding boundary cannot flip)."""
    rng = np.random.default_rng(n + k + wq + 10 * lp)
    lds = (n + 3) // 4 * 4
    S0 = np.zeros((n, lds), np.float32)              # S0[col][row]: column-major S
    S0[:, :n] = rng.uniform(-1, 1, (n, n)).astype(np.float32)
    iU = np.triu(rng.uniform(-1, 1, (64, 64))).astype(np.float32) + 2 * np.eye(64, dtype=np.float32)
    S = torch.tensor(S0, device="cuda")
    invu = torch.tensor(np.ascontiguousarray(iU.T), device="cuda")   # column-major iU
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    _k("atlmx_k_tslu_l21up", X.dptr(S), C.c_long(lds), n, k, wq, X.dptr(invu), lp,
       X.dptr(info), None)
    torch.cuda.synchronize()
    S1 = S.cpu().numpy()
    kq = k + 64
    Lt = S0[kq:n, k:kq].T.astype(np.float64)           # (t, column)
    want = iU.T.astype(np.float64) @ Lt
    got = S1[kq:n, k:kq].T.astype(np.float64)
    scale = np.abs(iU.T).astype(np.float64) @ np.abs(Lt)
    assert np.all(np.abs(got - want) <= 64 * 2.0**-23 * scale + 1e-30)
    U12 = _as_lp(S0[k:kq, kq:kq + wq].T, lp).astype(np.float64)     # (i, t)
    L21 = _as_lp(S1[kq:n, k:kq].T, lp).astype(np.float64)           # (t, column)
    Cw = S0[kq:n, kq:kq + wq].T.astype(np.float64) - U12 @ L21
    Cg = S1[kq:n, kq:kq + wq].T.astype(np.float64)
    sc = np.abs(U12) @ np.abs(L21) + np.abs(S0[kq:n, kq:kq + wq].T)
    assert np.all(np.abs(Cg - Cw) <= 64 * 2.0**-23 * sc + 1e-30)
    # nothing else moved
    mask = np.ones_like(S0, bool)
    mask[kq:n, k:kq] = False
    mask[kq:n, kq:kq + wq] = False
    assert np.array_equal(S1[mask], S0[mask])


def test_tournament_zero_pivot_sets_info(torch):
    n = 300
    A = np.random.default_rng(1).uniform(-1, 1, (n, n))
    A[:, 5] = 0.0                                  # column 5 of the panel is zero
    S, ws, dplan, dperm, blk, info = _tslu_select(torch, A, 0, 64)
    assert int(info.item()) == 6
    # with info set, the next call does nothing (perm and counters untouched)
    before = dperm.clone()
    _k("atlmx_k_tslu_select", X.dptr(S), C.c_long(S.shape[1]), n, 64, 64, X.dptr(ws),
       X.dptr(dplan), X.dptr(dperm), X.dptr(blk[0]), X.dptr(blk[1]), X.dptr(blk[2]),
       X.dptr(info), None, None, None)
    torch.cuda.synchronize()
    assert int(ws[:3].abs().sum().item()) == 0 and int(info.item()) == 6   # 2 leaves + root
    assert torch.equal(before, dperm)


def _apply_moves(cols, plan):
    """column moves dst <- src of a move list, every source read first"""
    mv = plan[2:2 + 2 * plan[0]].reshape(-1, 2)
    out = cols.copy()
    out[..., mv[:, 0]] = cols[..., mv[:, 1]]
    return out


@pytest.mark.parametrize("n", [300, 700, 1100])
def test_pair_tail_vs_numpy(torch, n):
    """Two tournaments (P at 0, Q at 64, the second composing both move
    lists), then the pair's tail kernel on a random S: the composed list
    equals P's moves followed by Q's, and the kernel's result equals the
    numpy restatement -- both lists' moves on the S rows outside 0..127,
    then U_P <- U_P inv(L_P)^T, U_Q <- U_Q - U_P B, U_Q <- U_Q inv(L_Q)^T on
    the rows from 128 (B = S rows 0..63, columns 64..127)."""
    rng = np.random.default_rng(n)
    A = rng.uniform(-1, 1, (n, n))
    S, ws, planP, perm, blkP, info = _tslu_select(torch, A, 0, 64)
    planQ = torch.zeros(2 + 4 * 64, dtype=torch.int32, device="cuda")
    planPQ = torch.full((2 + 8 * 64,), -5, dtype=torch.int32, device="cuda")
    blkQ = torch.zeros(3, 64, 64, dtype=torch.float32, device="cuda")
    _k("atlmx_k_tslu_select", X.dptr(S), C.c_long(S.shape[1]), n, 64, 64, X.dptr(ws),
       X.dptr(planQ), X.dptr(perm), X.dptr(blkQ[0]), X.dptr(blkQ[1]), X.dptr(blkQ[2]),
       X.dptr(info), X.dptr(planP), X.dptr(planPQ), None)
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    pP, pQ, pPQ = planP.cpu().numpy(), planQ.cpu().numpy(), planPQ.cpu().numpy()
    assert pPQ[0] <= pP[0] + pQ[0]
    d = pPQ[2:2 + 2 * pPQ[0]:2]
    assert len(set(d.tolist())) == len(d)            # each column written once
    iota = np.arange(n)
    assert np.array_equal(_apply_moves(_apply_moves(iota, pP), pQ), _apply_moves(iota, pPQ))
    # the tail on fresh random data (S rows = entries, torch rows = S columns)
    St = torch.tensor(rng.uniform(-1, 1, S.shape).astype(np.float32), device="cuda")
    iLP = np.tril(rng.uniform(-1, 1, (64, 64)), -1) + np.eye(64)
    iLQ = np.tril(rng.uniform(-1, 1, (64, 64)), -1) + np.eye(64)
    dLP = torch.tensor(iLP.T.astype(np.float32).copy(), device="cuda")   # column-major
    dLQ = torch.tensor(iLQ.T.astype(np.float32).copy(), device="cuda")
    S0 = St.cpu().numpy().astype(np.float64)          # [S column, entry]
    _k("atlmx_k_tslu_tail", X.dptr(St), C.c_long(St.shape[1]), n, 0, X.dptr(planPQ),
       X.dptr(dLP), X.dptr(dLQ), X.dptr(info), None)
    torch.cuda.synchronize()
    got = St.cpu().numpy().astype(np.float64)
    M = S0[:n, :n].T.copy()                           # M[entry, column] = S(entry, column)
    want = M.copy()
    rows = np.r_[0:0, 128:n]
    want[rows] = _apply_moves(M[rows], pPQ)
    UP = want[128:, 0:64] @ iLP.astype(np.float32).T
    UQ = want[128:, 64:128] - UP @ M[0:64, 64:128]
    UQ = UQ @ iLQ.astype(np.float32).T
    want[128:, 0:64], want[128:, 64:128] = UP, UQ
    gotM = got[:n, :n].T
    assert np.array_equal(gotM[:128], M[:128])        # the pair's own rows untouched
    err = np.abs(gotM - want).max() / np.abs(want).max()
    assert err < 1e-5, err


def _system(n, kappa, nrhs, seed):
    """A = U diag(s) V^T (Haar U, V), x known, b = A x"""
    rng = np.random.default_rng(seed)
    Uq, _ = np.linalg.qr(rng.standard_normal((n, n)))
    Vq, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = kappa ** (-np.arange(n) / max(n - 1, 1))
    A = (Uq * s) @ Vq.T
    x = rng.uniform(-1, 1, (n, nrhs))
    return A, x, A @ x


@pytest.mark.parametrize("method,want", [(X.MIXED_AUTO, 1), (X.MIXED_TSLU, 1),
                                         (X.MIXED_GEPP, 2)])
@pytest.mark.parametrize("n,nrhs,kappa", [(1, 1, 1.0), (5, 2, 10.0), (130, 1, 1e4),
                                          (1001, 3, 1e6), (1024, 17, 1e5)])
def test_dsgesv_fp64_accuracy(torch, method, want, n, nrhs, kappa):
    A, x, b = _system(n, kappa, nrhs, seed=n + nrhs)
    dA = _dev(torch, A, torch.float64)
    dB = _dev(torch, b, torch.float64)
    dX = torch.zeros_like(dB)
    X.set_mixed_method(method)
    try:
        info, st = X.dsgesv_dev(dA, dB, dX)
    finally:
        X.set_mixed_method(X.MIXED_AUTO)
    assert info == 0
    assert st.converged == 1 and st.fallback == 0 and st.method == want
    be = X.lib().atlmx_dbackward_error(n, nrhs, X.dptr(dA), n, X.dptr(dB), n, X.dptr(dX), n)
    assert be <= n * EPS64, be
    assert abs(st.backward_error - be) <= 0.5 * be + 1e-300 or be < 1e-17
    xs = _host(dX)
    ferr = np.abs(xs - x).max() / np.abs(x).max()
    assert ferr <= 10 * kappa * n * EPS64, ferr
    assert np.array_equal(_host(dA), A) and np.array_equal(_host(dB), b)   # inputs untouched


@pytest.mark.parametrize("method,n,kappa", [(X.MIXED_BF16, 1024, 1e2), (X.MIXED_FP16, 1024, 1e2),
                                            (X.MIXED_FP16, 1024, 1e4), (X.MIXED_FP16, 4096, 1e2)])
def test_dsgesv_low_precision_updates(torch, method, n, kappa):
    """the tournament LU with its trailing updates in bfloat16 / half: FP64
    accuracy from its factors alone (plain refinement or GMRES-IR), no
    fallback"""
    A, x, b = _system(n, kappa, 1, seed=n)
    dA, dB = _dev(torch, A, torch.float64), _dev(torch, b, torch.float64)
    dX = torch.zeros_like(dB)
    X.set_mixed_method(method)
    try:
        info, st = X.dsgesv_dev(dA, dB, dX)
    finally:
        X.set_mixed_method(X.MIXED_AUTO)
    assert info == 0 and st.converged == 1 and st.fallback == 0 and st.method == method
    be = X.lib().atlmx_dbackward_error(n, 1, X.dptr(dA), n, X.dptr(dB), n, X.dptr(dX), n)
    assert be <= n * EPS64, be
    ferr = np.abs(_host(dX) - x).max() / np.abs(x).max()
    assert ferr <= 10 * kappa * n * EPS64, ferr


def test_dsgesv_half_updates_with_badly_scaled_columns(torch):
    """columns scaled by 10^-6 .. 10^6: without the power-of-two column
    scaling the half-precision operands would overflow (or vanish); with it
    the half-update factorization still serves as the refinement's factors"""
    n = 1024
    A0, x0, _ = _system(n, 1e2, 1, seed=n)
    cs = 10.0 ** np.random.default_rng(9).uniform(-6, 6, n)
    A = A0 * cs                                   # column j times cs[j]
    x = x0 / cs[:, None]
    b = A @ x
    dA, dB = _dev(torch, A, torch.float64), _dev(torch, b, torch.float64)
    dX = torch.zeros_like(dB)
    X.set_mixed_method(X.MIXED_FP16)
    try:
        info, st = X.dsgesv_dev(dA, dB, dX)
    finally:
        X.set_mixed_method(X.MIXED_AUTO)
    assert info == 0 and st.converged == 1 and st.fallback == 0 and st.method == X.MIXED_FP16
    be = X.lib().atlmx_dbackward_error(n, 1, X.dptr(dA), n, X.dptr(dB), n, X.dptr(dX), n)
    assert be <= n * EPS64, be


def test_dsgesv_bfloat16_too_weak_falls_back_to_fp32(torch):
    """at kappa 1e6 the bfloat16-update factors cannot drive refinement (or
    GMRES-IR, which stops when a restart cycle fails to halve the backward
    error); the FP32 tournament LU then solves it"""
    n = 1024
    A, x, b = _system(n, 1e6, 1, seed=n)
    dA, dB = _dev(torch, A, torch.float64), _dev(torch, b, torch.float64)
    dX = torch.zeros_like(dB)
    X.set_mixed_method(X.MIXED_BF16)
    try:
        info, st = X.dsgesv_dev(dA, dB, dX)
    finally:
        X.set_mixed_method(X.MIXED_AUTO)
    assert info == 0 and st.converged == 1 and st.fallback == 0 and st.method == X.MIXED_TSLU
    assert st.inner_iterations <= 300                # the stalled GMRES stopped early
    be = X.lib().atlmx_dbackward_error(n, 1, X.dptr(dA), n, X.dptr(dB), n, X.dptr(dX), n)
    assert be <= n * EPS64, be


def test_dsgesv_falls_back_to_fp64_when_fp32_factor_is_singular(torch):
    """column 0 lies below FP32's range (it rounds to zero): both FP32
    factorizations meet an exactly zero pivot; the FP64 LU solves it"""
    n = 300
    rng = np.random.default_rng(3)
    A = rng.uniform(-1, 1, (n, n)) + 4 * np.eye(n)
    A[:, 0] *= 1e-50
    x = rng.uniform(-1, 1, (n, 1))
    b = A @ x
    dA, dB = _dev(torch, A, torch.float64), _dev(torch, b, torch.float64)
    dX = torch.zeros_like(dB)
    info, st = X.dsgesv_dev(dA, dB, dX)
    assert info == 0 and st.fallback == 1 and st.method == 3 and st.factor_info > 0
    be = X.lib().atlmx_dbackward_error(n, 1, X.dptr(dA), n, X.dptr(dB), n, X.dptr(dX), n)
    assert be <= n * EPS64


@pytest.mark.parametrize("method,want", [(X.MIXED_TSLU, 1), (X.MIXED_GEPP, 2)])
@pytest.mark.parametrize("kappa", [1e8, 3e8])
def test_dsgesv_gmres_ir_beyond_fp32(torch, method, want, kappa):
    """kappa * eps32 >> 1: plain refinement with the FP32 factors stalls, GMRES
    preconditioned by them reaches the FP64 backward-error bar (near 1e9
    the tournament factor needs tens of restarts, so the test stays below)"""
    n = 500
    A, x, b = _system(n, kappa, 2, seed=9)
    dA, dB = _dev(torch, A, torch.float64), _dev(torch, b, torch.float64)
    dX = torch.zeros_like(dB)
    X.set_mixed_method(method)
    try:
        info, st = X.dsgesv_dev(dA, dB, dX)
    finally:
        X.set_mixed_method(X.MIXED_AUTO)
    assert info == 0 and st.converged == 1 and st.fallback == 0
    assert st.method == want and st.gmres == 1 and st.inner_iterations > 0
    be = X.lib().atlmx_dbackward_error(n, 2, X.dptr(dA), n, X.dptr(dB), n, X.dptr(dX), n)
    assert be <= n * EPS64, be


def _spd_system(n, kappa, nrhs, seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = kappa ** (-np.arange(n) / max(n - 1, 1))
    A = (Q * s) @ Q.T
    A = (A + A.T) / 2
    x = rng.uniform(-1, 1, (n, nrhs))
    return A, x, A @ x


@pytest.mark.parametrize("n,nrhs,kappa", [(1, 1, 1.0), (130, 2, 1e4), (1001, 3, 1e6),
                                          (700, 17, 1e5)])
def test_dsposv_fp64_accuracy(torch, n, nrhs, kappa):
    A, x, b = _spd_system(n, kappa, nrhs, seed=n)
    dA, dB = _dev(torch, A, torch.float64), _dev(torch, b, torch.float64)
    dX = torch.zeros_like(dB)
    info, st = X.dsposv_dev(dA, dB, dX)
    assert info == 0
    assert st.converged == 1 and st.fallback == 0 and st.method == 4
    be = X.lib().atlmx_dbackward_error(n, nrhs, X.dptr(dA), n, X.dptr(dB), n, X.dptr(dX), n)
    assert be <= n * EPS64, be
    ferr = np.abs(_host(dX) - x).max() / np.abs(x).max()
    assert ferr <= 10 * kappa * n * EPS64, ferr


def test_dsposv_falls_back_to_fp64(torch):
    n = 300
    A, x, b = _spd_system(n, 1e12, 1, seed=4)
    dA, dB = _dev(torch, A, torch.float64), _dev(torch, b, torch.float64)
    dX = torch.zeros_like(dB)
    info, st = X.dsposv_dev(dA, dB, dX)
    assert info == 0 and st.fallback == 1 and st.method == 3
    be = X.lib().atlmx_dbackward_error(n, 1, X.dptr(dA), n, X.dptr(dB), n, X.dptr(dX), n)
    assert be <= n * EPS64


def test_dsgesv_lapack_stopping_test(torch):
    """LAPACK dsgesv's stricter stopping test also converges, to a residual
    below ||x|| ||A|| 2^-53 sqrt(n)"""
    n = 600
    A, x, b = _system(n, 1e6, 1, seed=12)
    dA, dB = _dev(torch, A, torch.float64), _dev(torch, b, torch.float64)
    dX = torch.zeros_like(dB)
    X.set_refine_stop(True)
    try:
        info, st = X.dsgesv_dev(dA, dB, dX)
    finally:
        X.set_refine_stop(False)
    assert info == 0 and st.converged == 1 and st.fallback == 0
    xs = _host(dX)
    r = np.abs(A @ xs - b).max()
    # the GPU met the test on its own FP64 residual; recomputing A x - b here
    # in another order adds up to ~n eps ||A|| ||x|| of rounding
    assert r <= np.abs(xs).max() * np.abs(A).sum(axis=1).max() * 2.0 ** -53 * (np.sqrt(n) + n)
