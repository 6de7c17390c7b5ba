"""Adjoints of the batched solves (atlas_amd.autograd): torch.autograd.gradcheck
(FP64 central finite differences), a float64 numpy reference gradient from
the closed forms (b-bar = A^{-T} x-bar, A-bar = -b-bar x^T; for Cholesky the
symmetric -(b-bar x^T + x b-bar^T)/2 folded onto the lower triangle), and
the configuration's size: 512 SPD and 512 general systems of order 256 with
8 right-hand sides, loss = sum(w * x), in FP64 and FP32."""
import numpy as np
import pytest

import atlas_amd as X

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    X.use_torch_stream()
    return t


def _spd_batch(rng, b, n):
    B = rng.standard_normal((b, n, n))
    return B @ B.transpose(0, 2, 1) / n + np.eye(n)


def _ref_grads(A, B, W, spd):
    """numpy float64 reference: x = A^{-1} b, loss = sum(w x)"""
    Xs = np.linalg.solve(A, B)
    Bb = np.linalg.solve(A.transpose(0, 2, 1), W)
    if spd:
        S = -(Bb @ Xs.transpose(0, 2, 1) + Xs @ Bb.transpose(0, 2, 1)) / 2   # symmetric grad
        G = np.tril(2 * S)
        idx = np.arange(A.shape[1])
        G[:, idx, idx] = S[:, idx, idx]
    else:
        G = -Bb @ Xs.transpose(0, 2, 1)
    return Xs, G, Bb


@pytest.mark.parametrize("spd", [True, False])
def test_gradcheck_fp64(torch, spd):
    from atlas_amd.autograd import CholeskySolve, LUSolve
    rng = np.random.default_rng(1)
    b, n, r = 3, 9, 2
    A = _spd_batch(rng, b, n) if spd else rng.uniform(-1, 1, (b, n, n)) + 4 * np.eye(n)
    At = torch.tensor(A, device="cuda", requires_grad=True)
    Bt = torch.tensor(rng.uniform(-1, 1, (b, n, r)), device="cuda", requires_grad=True)
    f = CholeskySolve.apply if spd else LUSolve.apply
    if spd:
        # the function reads the lower triangle: perturb only that (upper part symmetric copy)
        def g(Al, Bx):
            return f(torch.tril(Al) + torch.tril(Al, -1).transpose(-1, -2), Bx)
        assert torch.autograd.gradcheck(g, (At, Bt), eps=1e-6, atol=1e-7, rtol=1e-5)
    else:
        assert torch.autograd.gradcheck(f, (At, Bt), eps=1e-6, atol=1e-7, rtol=1e-5)


@pytest.mark.parametrize("spd", [True, False])
def test_finite_differences_lower_triangle(torch, spd):
    """central differences (h = 1e-6) of loss = sum(w x) against single
    entries of A and B; for Cholesky only lower-triangle entries matter"""
    from atlas_amd.autograd import CholeskySolve, LUSolve
    rng = np.random.default_rng(2)
    b, n, r = 2, 16, 3
    A = _spd_batch(rng, b, n) if spd else rng.uniform(-1, 1, (b, n, n)) + 4 * np.eye(n)
    B = rng.uniform(-1, 1, (b, n, r))
    W = rng.standard_normal((b, n, r))
    f = CholeskySolve.apply if spd else LUSolve.apply
    At = torch.tensor(A, device="cuda", requires_grad=True)
    Bt = torch.tensor(B, device="cuda", requires_grad=True)
    Wt = torch.tensor(W, device="cuda")
    (f(At, Bt) * Wt).sum().backward()
    gA, gB = At.grad.cpu().numpy(), Bt.grad.cpu().numpy()

    def loss(Ap, Bp):
        with torch.no_grad():
            return float((f(torch.tensor(Ap, device="cuda"), torch.tensor(Bp, device="cuda"))
                          * Wt).sum().item())
    h = 1e-6
    for (k, i, j) in [(0, 3, 1), (1, 7, 7), (0, 15, 0), (1, 2, 9)]:
        if spd and i < j:
            assert gA[k, i, j] == 0.0
            continue
        Ap, Am = A.copy(), A.copy()
        Ap[k, i, j] += h
        Am[k, i, j] -= h
        if spd and i != j:            # the solver reads the lower triangle only
            Ap[k, j, i] = A[k, j, i]
            Am[k, j, i] = A[k, j, i]
        fd = (loss(Ap, B) - loss(Am, B)) / (2 * h)
        assert abs(fd - gA[k, i, j]) <= 1e-6 * max(1.0, abs(fd)), (k, i, j, fd, gA[k, i, j])
    for (k, i, j) in [(0, 0, 0), (1, 5, 2)]:
        Bp, Bm = B.copy(), B.copy()
        Bp[k, i, j] += h
        Bm[k, i, j] -= h
        fd = (loss(A, Bp) - loss(A, Bm)) / (2 * h)
        assert abs(fd - gB[k, i, j]) <= 1e-6 * max(1.0, abs(fd))


@pytest.mark.parametrize("dt", ["f64", "f32"])
@pytest.mark.parametrize("spd", [True, False])
def test_config_size_vs_reference(torch, dt, spd):
    """512 systems of order 256, 8 right-hand sides: forward solution and
    both gradients against the float64 numpy reference."""
    from atlas_amd.autograd import CholeskySolve, LUSolve
    rng = np.random.default_rng(3)
    b, n, r = 512, 256, 8
    A = _spd_batch(rng, b, n) if spd else rng.uniform(-1, 1, (b, n, n)) + 4 * np.eye(n)
    B = rng.standard_normal((b, n, r))
    W = rng.standard_normal((b, n, r))
    tdt = torch.float64 if dt == "f64" else torch.float32
    A = A.astype(np.float32).astype(np.float64) if dt == "f32" else A
    B = B.astype(np.float32).astype(np.float64) if dt == "f32" else B
    At = torch.tensor(A, dtype=tdt, device="cuda", requires_grad=True)
    Bt = torch.tensor(B, dtype=tdt, device="cuda", requires_grad=True)
    f = CholeskySolve.apply if spd else LUSolve.apply
    Xt = f(At, Bt)
    (Xt * torch.tensor(W, dtype=tdt, device="cuda")).sum().backward()
    Xr, Gr, Bbr = _ref_grads(A, B, W, spd)
    # forward errors grow with the condition number (A + 4I with uniform
    # entries is not well conditioned at n = 256): per-instance bound
    # c * kappa * n * eps of the precision, relative to the instance's scale
    eps = 2.0 ** -52 if dt == "f64" else 2.0 ** -23
    kappa = np.linalg.cond(A)
    got = (Xt.detach().cpu().numpy(), Bt.grad.cpu().numpy(), At.grad.cpu().numpy())
    for name, g, r_ in zip(("x", "b-bar", "A-bar"), got, (Xr, Bbr, Gr)):
        err = np.abs(g - r_).max(axis=(1, 2)) / np.maximum(np.abs(r_).max(axis=(1, 2)), 1e-300)
        bound = 4.0 * kappa * n * eps
        assert np.all(err <= np.maximum(bound, 1e3 * eps)), (name, float((err / bound).max()))


@pytest.mark.parametrize("dt", ["f64", "f32"])
def test_cholesky_solve_order_300(torch, dt):
    """Orders above 256 take the adjoint kernel's global-scratch branch:
    forward and both gradients against the numpy reference at n = 300."""
    from atlas_amd.autograd import CholeskySolve
    rng = np.random.default_rng(31)
    b, n, r = 4, 300, 3
    A = _spd_batch(rng, b, n)
    B = rng.standard_normal((b, n, r))
    W = rng.standard_normal((b, n, r))
    tdt = torch.float64 if dt == "f64" else torch.float32
    if dt == "f32":
        A, B = A.astype(np.float32).astype(np.float64), B.astype(np.float32).astype(np.float64)
    At = torch.tensor(A, dtype=tdt, device="cuda", requires_grad=True)
    Bt = torch.tensor(B, dtype=tdt, device="cuda", requires_grad=True)
    Xt = CholeskySolve.apply(At, Bt)
    (Xt * torch.tensor(W, dtype=tdt, device="cuda")).sum().backward()
    Xr, Gr, Bbr = _ref_grads(A, B, W, True)
    eps = 2.0 ** -52 if dt == "f64" else 2.0 ** -23
    kappa = np.linalg.cond(A)
    got = (Xt.detach().cpu().numpy(), Bt.grad.cpu().numpy(), At.grad.cpu().numpy())
    for name, g, r_ in zip(("x", "b-bar", "A-bar"), got, (Xr, Bbr, Gr)):
        err = np.abs(g - r_).max(axis=(1, 2)) / np.abs(r_).max(axis=(1, 2))
        assert np.all(err <= 4.0 * kappa * n * eps), (name, err)


@pytest.mark.parametrize("dt", ["f64", "f32"])
def test_cholesky_adjoint_uneven_batch(torch, dt):
    """posv_adjoint_batch on instances of orders 9, 200 and 300 in one
    plan, from L, X and Bb computed in numpy: the gradient of <Xb, X> on
    the lower triangle against the closed form tril(2S) - diag(S),
    S = -(Bb X^T + X Bb^T) / 2."""
    rng = np.random.default_rng(33)
    tdt = torch.float64 if dt == "f64" else torch.float32
    npdt = np.float64 if dt == "f64" else np.float32
    sizes, r = [9, 200, 300], 2
    Ls, Xs, Bbs, Gs, Ws, refs = [], [], [], [], [], []
    for n in sizes:
        A = _spd_batch(rng, 1, n)[0].astype(npdt).astype(np.float64)
        Xm = np.linalg.solve(A, rng.standard_normal((n, r)))
        Bb = np.linalg.solve(A, rng.standard_normal((n, r)))
        L = np.linalg.cholesky(A)
        S = -(Bb @ Xm.T + Xm @ Bb.T) / 2
        G = np.tril(2 * S)
        G[np.arange(n), np.arange(n)] = np.diag(S)
        refs.append(G)
        Ls.append(torch.tensor(np.ascontiguousarray(L.T), dtype=tdt, device="cuda"))
        Xs.append(torch.tensor(np.ascontiguousarray(Xm.T), dtype=tdt, device="cuda"))
        Bbs.append(torch.tensor(np.ascontiguousarray(Bb.T), dtype=tdt, device="cuda"))
        Gs.append(torch.zeros(n, n, dtype=tdt, device="cuda"))
        Ws.append(torch.zeros(n, n, dtype=tdt, device="cuda"))
    with X.posv_adjoint_batch(Ls, Xs, Bbs, Gs, Ws) as plan:
        plan.run()
        torch.cuda.synchronize()
    tol = 1e-11 if dt == "f64" else 2e-4
    for n, g, ref in zip(sizes, Gs, refs):
        got = np.tril(g.cpu().numpy().T.astype(np.float64))
        assert np.abs(got - ref).max() <= tol * np.abs(ref).max(), n


def test_batch_plan_rejects_orders_beyond_lds():
    """A gesv order above the LDS limit gives a NULL plan (no launch)."""
    import torch as t
    lim = X.lib().atlmx_batch_max_order(1, 1)
    A = [t.zeros(lim + 1, lim + 1, dtype=t.float64, device="cuda")]
    piv = [t.zeros(lim + 1, dtype=t.int32, device="cuda")]
    info = t.zeros(1, dtype=t.int32, device="cuda")
    plan = X.gesv_batch(A, piv, [], info)
    assert not plan.h
