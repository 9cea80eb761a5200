"""GPU tests of the Cholesky / solve / logdet adjoints behind torch.autograd
(libddla_amd.autograd).  References, all FP64 on the CPU:
  * torch's own autograd of torch.linalg.cholesky / cholesky_solve / logdet
    (the same symmetric-input convention), relative error <= 1e-10;
  * the Gaussian-process likelihood (n = 4096 points uniform in
    [0, 1]^3, seed 4, RBF l = 0.2, s = 1, jitter 1e-2): dLoss/dK against the
    closed form (K^-1 - alpha alpha^T) / 2 and dLoss/dl against
    sum((K^-1 - alpha alpha^T) / 2 * dK/dl) (Rasmussen & Williams 2006,
    eq. 5.9), relative error <= 1e-10; and central finite differences with
    step 1e-6 on l and on 64 random symmetric entry pairs of K.
"""
import ctypes

import numpy as np
import pytest
import torch

import libddla_amd as dd
from libddla_amd import autograd as ag

pytestmark = pytest.mark.gpu

DEV = "cuda"


def spd(n, seed):
    g = np.random.RandomState(seed).standard_normal((n, n))
    return g @ g.T / n + 0.5 * np.eye(n)


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(1e-300, np.abs(b).max()))


@pytest.mark.parametrize("n", [1, 64, 100, 257, 600, 1100, 2100])
def test_cholesky_backward_matches_torch(handle, n):
    """Orders above 1024 take the recursive solves (split orders, ragged
    64-row leaves; 2100 splits twice) and every order the k-blocked
    product."""
    print(f"case cholesky backward n={n}", flush=True)
    a = spd(n, n)
    lbar = np.tril(np.random.RandomState(n + 1).standard_normal((n, n)))
    A = torch.tensor(a, device=DEV, requires_grad=True)
    L = ag.cholesky(A, handle)
    (L * torch.tensor(lbar, device=DEV)).sum().backward()
    Ac = torch.tensor(a, requires_grad=True)
    Lc = torch.linalg.cholesky(Ac)
    (Lc * torch.tensor(lbar)).sum().backward()
    assert rel(L.detach().cpu(), Lc.detach()) < 1e-12
    assert rel(A.grad.cpu(), Ac.grad) < 1e-10
    assert np.allclose(A.grad.cpu().numpy(), A.grad.cpu().numpy().T, rtol=0, atol=0)


@pytest.mark.parametrize("n,nrhs", [(64, 1), (100, 3), (257, 5), (257, 8), (130, 9), (600, 300)])
def test_solve_and_logdet_backward_match_torch(handle, n, nrhs):
    """Up to 8 columns the solve adjoint forms L^T [X | Bbar] in one pass over
    the lower triangle; more columns take the GEMM route (and, from 256
    columns on an order above 512, the recursive solves)."""
    print(f"case solve backward n={n} nrhs={nrhs}", flush=True)
    a = spd(n, 7 * n)
    b = np.random.RandomState(n).standard_normal((n, nrhs))
    xbar = np.random.RandomState(n + 2).standard_normal((n, nrhs))
    lc = torch.linalg.cholesky(torch.tensor(a))
    L = lc.to(DEV).clone().requires_grad_()
    B = torch.tensor(b, device=DEV, requires_grad=True)
    X = ag.cholesky_solve(B, L, handle)
    ((X * torch.tensor(xbar, device=DEV)).sum() + 0.75 * ag.cholesky_logdet(L, handle)).backward()
    Lr = lc.clone().requires_grad_()
    Br = torch.tensor(b, requires_grad=True)
    Xr = torch.cholesky_solve(Br, Lr)
    ((Xr * torch.tensor(xbar)).sum() + 0.75 * 2 * torch.log(torch.diagonal(Lr)).sum()).backward()
    assert rel(X.detach().cpu(), Xr.detach()) < 1e-10
    assert rel(B.grad.cpu(), Br.grad) < 1e-10
    assert rel(torch.tril(L.grad).cpu(), torch.tril(Lr.grad)) < 1e-10
    # vector right-hand side
    y = torch.tensor(b[:, 0], device=DEV, requires_grad=True)
    x = ag.cholesky_solve(y, L.detach(), handle)
    assert x.shape == (n,)
    assert rel(x.detach().cpu(), Xr.detach()[:, 0]) < 1e-10


def gp_problem(n, seed=4):
    rng = np.random.RandomState(seed)
    x = rng.uniform(size=(n, 3))
    y = rng.standard_normal(n)
    return x, y


def gp_kernel(xt, l, s=1.0, jitter=1e-2):
    d2 = ((xt[:, None, :] - xt[None, :, :]) ** 2).sum(-1)
    n = xt.shape[0]
    return s * s * torch.exp(-d2 / (2 * l * l)) + jitter * torch.eye(n, dtype=xt.dtype, device=xt.device), d2


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("n", [1, 64, 100, 257, 1100, 2100])
def test_potri_matches_inverse(handle, n, dtype):
    """A^-1 from the factor (recursive triangular inverse, then tril(X^T X)
    mirrored) against numpy's inverse in FP64; exactly symmetric."""
    print(f"case potri n={n} {dtype}", flush=True)
    a = spd(n, 3 * n + 1)
    L = torch.linalg.cholesky(torch.tensor(a)).to(dtype).to(DEV)
    Ainv = ag.potri(L, handle).cpu().double().numpy()
    ref = np.linalg.inv(a)
    assert rel(Ainv, ref) < (1e-11 if dtype == torch.float64 else 1e-4)
    assert np.array_equal(Ainv, Ainv.T)


GP_LOSSES = {"fused": ag.gp_negative_log_likelihood, "composed": ag.gp_negative_log_likelihood_composed}


@pytest.mark.parametrize("which", sorted(GP_LOSSES))
def test_gp_likelihood_gradient_n4096(handle, which):
    """Both the fused Function (closed-form backward through potri) and the
    composed Cholesky / solve / logdet Functions."""
    print(f"case gp likelihood {which}", flush=True)
    gp_loss = GP_LOSSES[which]
    n = 4096
    x, y = gp_problem(n)
    xt = torch.tensor(x, device=DEV)
    yt = torch.tensor(y, device=DEV)
    l = torch.tensor(0.2, dtype=torch.float64, device=DEV, requires_grad=True)
    K, d2 = gp_kernel(xt, l)
    K.retain_grad()
    loss = gp_loss(K, yt, handle)
    loss.backward()
    kbar = K.grad.detach().cpu().numpy()
    dl = float(l.grad)
    # closed form in FP64 on the CPU
    k = K.detach().cpu().numpy()
    kinv = np.linalg.inv(k)
    alpha = kinv @ y
    kbar_ref = 0.5 * (kinv - np.outer(alpha, alpha))
    rbf = k - 1e-2 * np.eye(n)
    dk_dl = rbf * d2.cpu().numpy() / 0.2 ** 3
    dl_ref = float((kbar_ref * dk_dl).sum())
    assert np.linalg.norm(kbar - kbar_ref) / np.linalg.norm(kbar_ref) < 1e-10
    assert abs(dl - dl_ref) / abs(dl_ref) < 1e-10
    sign, logdet = np.linalg.slogdet(k)
    assert abs(float(loss.detach()) - (0.5 * y @ alpha + 0.5 * logdet)) < 1e-9 * abs(logdet)

    # central differences, step 1e-6
    h = 1e-6

    def f_of_l(lv):
        with torch.no_grad():
            Kp, _ = gp_kernel(xt, torch.tensor(lv, dtype=torch.float64, device=DEV))
            return float(gp_loss(Kp, yt, handle))

    fd_l = (f_of_l(0.2 + h) - f_of_l(0.2 - h)) / (2 * h)
    # The loss carries rounding noise of about cond(K) eps |y^T alpha| ~ 1e-8
    # (cond(K) ~ 1e5 with the 1e-2 jitter), so a step of 1e-6 leaves ~1e-2 of
    # noise in the difference quotient, ~1e-8 of dl; the 1e-10 checks above are
    # the exact ones.
    assert abs(fd_l - dl) <= 1e-7 * abs(dl)
    rng = np.random.RandomState(5)
    Kd = K.detach()
    scale = np.abs(kbar).max()
    for _ in range(64):
        i, j = (int(v) for v in rng.randint(0, n, size=2))
        with torch.no_grad():
            Kp = Kd.clone()
            Kp[i, j] += h
            if i != j:
                Kp[j, i] += h
            fp = float(gp_loss(Kp, yt, handle))
            Kp[i, j] -= 2 * h
            if i != j:
                Kp[j, i] -= 2 * h
            fm = float(gp_loss(Kp, yt, handle))
        fd = (fp - fm) / (2 * h)
        want = kbar[i, j] * (1 if i == j else 2)
        assert abs(fd - want) <= 1e-7 * scale, (i, j, fd, want)


@pytest.mark.parametrize("workers", [1, 3])
def test_gp_backward_concurrently_matches_sequential(handle, workers):
    """Five independent GP problems (n = 700) on `workers` handles, each from
    its own host thread and stream: losses and dK bit for bit those of one
    problem after another on one handle."""
    import libddla_amd as dd

    print(f"case gp concurrent workers={workers}", flush=True)
    n, count = 700, 5
    rng = np.random.RandomState(9)
    ks, ys = [], []
    for q in range(count):
        x = torch.tensor(rng.uniform(size=(n, 3)), device=DEV)
        l = torch.tensor(0.2 + 0.01 * q, dtype=torch.float64, device=DEV)
        ks.append(gp_kernel(x, l)[0].detach())
        ys.append(torch.tensor(rng.standard_normal(n), device=DEV))
    seq = []
    for K, y in zip(ks, ys):
        Kd = K.clone().requires_grad_()
        loss = ag.gp_negative_log_likelihood(Kd, y, handle)
        loss.backward()
        seq.append((float(loss.detach()), Kd.grad.cpu().numpy()))
    handles = [dd.Handle(dd.Backend.GPU) for _ in range(workers)]
    try:
        kds = [K.clone().requires_grad_() for K in ks]
        losses = ag.backward_concurrently(list(zip(kds, ys)), ag.gp_negative_log_likelihood, handles)
        for (l0, g0), l1, Kd in zip(seq, losses, kds):
            assert float(l1) == l0
            assert np.array_equal(Kd.grad.cpu().numpy(), g0)
    finally:
        for h in handles:
            h.destroy()


def test_autograd_rejects_host_tensors_and_non_pd(handle):
    with pytest.raises(ValueError):
        ag.cholesky(torch.eye(3, dtype=torch.float64), handle)
    a = torch.tensor([[1.0, 2.0], [2.0, 1.0]], dtype=torch.float64, device=DEV)
    with pytest.raises(RuntimeError):
        ag.cholesky(a, handle)


def gp_closed_form(k, y):
    """loss, dloss/dK, dloss/dy of 0.5 y^T K^-1 y + 0.5 log det K in FP64 on the CPU."""
    kinv = np.linalg.inv(k)
    alpha = kinv @ y
    sign, logdet = np.linalg.slogdet(k)
    assert sign > 0
    return 0.5 * y @ alpha + 0.5 * logdet, 0.5 * (kinv - np.outer(alpha, alpha)), alpha


@pytest.mark.parametrize("n,batch", [(1, 2), (64, 3), (100, 1), (700, 3), (1000, 2), (4288, 2)])
def test_gp_batched_matches_closed_form(handle, n, batch):
    """The batched Function on `batch` RBF problems of order n (orders off the
    64 grid, a short last block, a last unpaired block in the inverse's merges
    at 4288 = 67 blocks): every loss, dK and dy against the FP64 closed form
    with distinct upstream weights, to 1e-10 relative; dK exactly symmetric."""
    print(f"case gp batched n={n} batch={batch}", flush=True)
    rng = np.random.RandomState(100 + n + batch)
    ks, ys = [], []
    for q in range(batch):
        x = torch.tensor(rng.uniform(size=(n, 3)), device=DEV)
        ks.append(gp_kernel(x, torch.tensor(0.2 + 0.02 * q, dtype=torch.float64, device=DEV))[0].detach())
        ys.append(rng.standard_normal(n))
    K = torch.stack(ks).requires_grad_()
    Y = torch.tensor(np.stack(ys), device=DEV, requires_grad=True)
    w = torch.tensor(rng.uniform(0.5, 2.0, size=batch), device=DEV)
    losses = ag.gp_negative_log_likelihood_batched(K, Y, handle)
    assert losses.shape == (batch,)
    (w * losses).sum().backward()
    kbar = K.grad.cpu().numpy()
    ybar = Y.grad.cpu().numpy()
    for q in range(batch):
        k = ks[q].cpu().numpy()
        loss_ref, kbar_ref, alpha = gp_closed_form(k, ys[q])
        assert abs(float(losses[q]) - loss_ref) <= 1e-10 * max(1.0, abs(loss_ref)), (q, float(losses[q]), loss_ref)
        wq = float(w[q])
        assert rel(kbar[q], wq * kbar_ref) < 1e-10, q
        assert np.array_equal(kbar[q], kbar[q].T)
        assert rel(ybar[q], wq * alpha) < 1e-10, q


def test_gp_batched_config_n4096(handle):
    """Four problems of the GP configuration's order (n = 4096, l = 0.2, s = 1,
    1e-2 jitter): dK and dl against the FP64 closed form to 1e-10, the batched
    losses and gradients against the per-problem fused Function to 1e-11, and
    a central difference on l (step 1e-6) for problem 0."""
    print("case gp batched n=4096", flush=True)
    n, batch = 4096, 4
    rng = np.random.RandomState(4)
    xs = [torch.tensor(rng.uniform(size=(n, 3)), device=DEV) for _ in range(batch)]
    ys = [rng.standard_normal(n) for _ in range(batch)]
    l = torch.tensor(0.2, dtype=torch.float64, device=DEV, requires_grad=True)
    parts = [gp_kernel(x, l) for x in xs]
    K = torch.stack([p[0] for p in parts])
    K.retain_grad()
    Y = torch.tensor(np.stack(ys), device=DEV)
    losses = ag.gp_negative_log_likelihood_batched(K, Y, handle)
    losses.sum().backward()
    kbar = K.grad.detach().cpu().numpy()
    dl = float(l.grad)
    dl_ref = 0.0
    for q in range(batch):
        k = parts[q][0].detach().cpu().numpy()
        loss_ref, kbar_ref, _ = gp_closed_form(k, ys[q])
        assert rel(kbar[q], kbar_ref) < 1e-10, q
        assert abs(float(losses[q]) - loss_ref) <= 1e-10 * abs(loss_ref)
        dl_ref += float((kbar_ref * (k - 1e-2 * np.eye(n)) * parts[q][1].cpu().numpy() / 0.2 ** 3).sum())
        # the per-problem fused Function (ppotrf, ppotrs, diag_logdet, potri)
        Kd = parts[q][0].detach().clone().requires_grad_()
        single = ag.gp_negative_log_likelihood(Kd, torch.tensor(ys[q], device=DEV), handle)
        single.backward()
        assert abs(float(single.detach()) - float(losses[q])) <= 1e-11 * abs(float(single.detach()))
        assert rel(kbar[q], Kd.grad.cpu().numpy()) < 1e-11, q
    assert abs(dl - dl_ref) / abs(dl_ref) < 1e-10
    step = 1e-6

    def f_of_l(lv):
        with torch.no_grad():
            Kp = gp_kernel(xs[0], torch.tensor(lv, dtype=torch.float64, device=DEV))[0].unsqueeze(0)
            return float(ag.gp_negative_log_likelihood_batched(Kp, Y[:1], handle)[0])

    fd = (f_of_l(0.2 + step) - f_of_l(0.2 - step)) / (2 * step)
    dl0 = float((kbar[0] * (parts[0][0].detach().cpu().numpy() - 1e-2 * np.eye(n)) * parts[0][1].cpu().numpy()
                 / 0.2 ** 3).sum())
    assert abs(fd - dl0) <= 1e-7 * abs(dl0)


def test_gp_batched_rejects_bad_input_and_non_pd(handle):
    K = torch.eye(4, dtype=torch.float64, device=DEV).repeat(3, 1, 1)
    Y = torch.ones((3, 4), dtype=torch.float64, device=DEV)
    with pytest.raises(TypeError):
        ag.gp_negative_log_likelihood_batched(K.float(), Y.float(), handle)
    with pytest.raises(ValueError):
        ag.gp_negative_log_likelihood_batched(K.cpu(), Y.cpu(), handle)
    with pytest.raises(ValueError):
        ag.gp_negative_log_likelihood_batched(K, Y[:, :3], handle)
    Kb = K.clone()
    Kb[1, 2, 2] = -1.0  # problem 1 fails at its third pivot; the others are fine
    with pytest.raises(RuntimeError, match="problem 1 .*order 3"):
        ag.gp_negative_log_likelihood_batched(Kb, Y, handle)
    per = ctypes.c_longlong()
    assert dd.lib.ddla_d_gp_batched_workspace_size(4096, ctypes.byref(per)) == 0
    # 2100 problems x 32 pairs of the first merge level exceed one launch: refused before any argument is used
    with pytest.raises(dd.DdlaError, match="65535"):
        dd.check_status(dd.lib.ddla_d_gp_batched_forward(handle.ptr, 2100, 4096, None, 4096 * 4096, None, 4096,
                                                         None, None, None), "ddla_d_gp_batched_forward")
    losses = ag.gp_negative_log_likelihood_batched(K, Y, handle)  # the handle still works
    assert torch.allclose(losses, torch.full((3,), 2.0, dtype=torch.float64, device=DEV))
