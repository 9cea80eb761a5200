"""torch.autograd Functions for the Cholesky path, forward and backward on the
hand-written gfx950 kernels (C ABI of libddla_mi355x.so).

    L = cholesky(A, h)            A = L L^T (FP32/FP64, CUDA tensor), L lower
    X = cholesky_solve(B, L, h)   X = (L L^T)^-1 B
    d = cholesky_logdet(L, h)     log det(L L^T) = 2 sum log L_ii

Backward passes (ddla_{s,d}_*_backward):
    cholesky:       Abar = sym(L^-T Phi(L^T Lbar) L^-1)  -- the symmetric-input
                    gradient, as torch.linalg.cholesky's
    cholesky_solve: Bbar = A^-1 Xbar, Lbar = -tril(Bbar (L^T X)^T + X (L^T Bbar)^T)
    cholesky_logdet: Lbar = 2 g diag(1 / L_ii)
gp_negative_log_likelihood(K, y, h) is one Function whose backward is the
closed form (K^-1 - alpha alpha^T) / 2 with K^-1 from ddla_{s,d}_potri;
gp_negative_log_likelihood_batched(K, y, h) takes a stack of independent
problems (K: batch x n x n, y: batch x n, FP64) and runs every launch over
all of them (ddla_d_gp_batched_forward / _backward);
gp_negative_log_likelihood_composed builds the same loss from the three
Functions above; backward_concurrently runs independent problems' forward
and backward passes on several handles at once.
The library works on column-major storage; tensors are passed through
column-major copies (or views, when they already are), and torch's current
stream is synchronized before each call (every call returns synchronized).
"""
import ctypes

import numpy as np
import torch

from ._capi import check_status, lib
from .matrix import descinit

_PFX = {torch.float64: "d", torch.float32: "s"}


def _pfx(t):
    if t.dtype not in _PFX:
        raise TypeError("float32 and float64 are built")
    if not t.is_cuda:
        raise ValueError("the Cholesky path runs on the GPU: pass CUDA tensors")
    return _PFX[t.dtype]


def _colmajor(t):
    """A column-major (Fortran-ordered) tensor equal to t (2-D); no copy when t already is."""
    if t.dim() == 1:
        t = t.unsqueeze(1)
    if t.stride(0) == 1 and t.stride(1) == max(1, t.shape[0]):
        return t
    return t.t().contiguous().t()


def _empty_colmajor(m, n, like):
    return torch.empty((n, m), dtype=like.dtype, device=like.device).t()


def _zeros_colmajor(m, n, like):
    return torch.zeros((n, m), dtype=like.dtype, device=like.device).t()


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _sync():
    torch.cuda.current_stream().synchronize()


def _desc(h, m, n, ld):
    d = descinit(h, m, n, 64, 64, 0, 0, ld)
    return d, d.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class Cholesky(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A, h):
        p = _pfx(A)
        n = A.shape[0]
        Lc = A.detach().t().contiguous().t().clone() if n else A.detach().clone()
        _sync()
        info = ctypes.c_int(0)
        if n:
            d, dp = _desc(h, n, n, n)
            check_status(getattr(lib, f"ddla_{p}_ppotrf")(h.ptr, b"L", n, _ptr(Lc), 1, 1, dp, ctypes.byref(info)),
                         f"ddla_{p}_ppotrf")
        if info.value != 0:
            raise RuntimeError(f"cholesky: the leading minor of order {info.value} is not positive definite")
        L = torch.tril(Lc)
        ctx.h = h
        ctx.save_for_backward(L)
        return L

    @staticmethod
    def backward(ctx, Lbar):
        (L,) = ctx.saved_tensors
        p = _pfx(L)
        n = L.shape[0]
        Lc, Bc = _colmajor(L), _colmajor(Lbar)
        Abar = _empty_colmajor(n, n, L)
        _sync()
        if n:
            check_status(getattr(lib, f"ddla_{p}_potrf_backward")(ctx.h.ptr, b"L", n, _ptr(Lc), n, _ptr(Bc), n,
                                                                  _ptr(Abar), n), f"ddla_{p}_potrf_backward")
        return Abar, None


class CholeskySolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, B, L, h):
        p = _pfx(L)
        n = L.shape[0]
        vec = B.dim() == 1
        Lc = _colmajor(L.detach())
        X = _colmajor(B.detach()).clone()
        nrhs = X.shape[1]
        _sync()
        if n and nrhs:
            da, dap = _desc(h, n, n, n)
            db, dbp = _desc(h, n, nrhs, n)
            check_status(getattr(lib, f"ddla_{p}_ppotrs")(h.ptr, b"L", b"L", b"N", n, nrhs, _ptr(Lc), dap, _ptr(X),
                                                          dbp), f"ddla_{p}_ppotrs")
        ctx.h, ctx.vec = h, vec
        ctx.save_for_backward(Lc, X)
        return X[:, 0].clone() if vec else X.clone()

    @staticmethod
    def backward(ctx, Xbar):
        Lc, X = ctx.saved_tensors
        p = _pfx(Lc)
        n, nrhs = X.shape
        Xb = _colmajor(Xbar)
        Bbar = _empty_colmajor(n, nrhs, X)
        Lbar = _zeros_colmajor(n, n, X)
        _sync()
        if n and nrhs:
            check_status(getattr(lib, f"ddla_{p}_potrs_backward")(ctx.h.ptr, b"L", n, nrhs, _ptr(Lc), n, _ptr(X), n,
                                                                  _ptr(Xb), n, _ptr(Bbar), n, _ptr(Lbar), n),
                         f"ddla_{p}_potrs_backward")
        return (Bbar[:, 0].clone() if ctx.vec else Bbar), Lbar, None


class CholeskyLogdet(torch.autograd.Function):
    @staticmethod
    def forward(ctx, L, h):
        p = _pfx(L)
        n = L.shape[0]
        Lc = _colmajor(L.detach())
        res = np.zeros(2, dtype=np.float64 if L.dtype == torch.float64 else np.float32)
        _sync()
        if n:
            check_status(getattr(lib, f"ddla_{p}_diag_logdet")(h.ptr, n, _ptr(Lc), n, ctypes.c_void_p(res.ctypes.data)),
                         f"ddla_{p}_diag_logdet")
        ctx.h = h
        ctx.save_for_backward(Lc)
        return torch.tensor(2.0 * float(res[0]), dtype=L.dtype, device=L.device)

    @staticmethod
    def backward(ctx, g):
        (Lc,) = ctx.saved_tensors
        p = _pfx(Lc)
        n = Lc.shape[0]
        Lbar = _zeros_colmajor(n, n, Lc)
        gval = float(g.item())
        _sync()
        if n:
            check_status(getattr(lib, f"ddla_{p}_logdet_backward")(ctx.h.ptr, n, _ptr(Lc), n, gval, _ptr(Lbar), n),
                         f"ddla_{p}_logdet_backward")
        return Lbar, None


def cholesky(A, h):
    return Cholesky.apply(A, h)


def cholesky_solve(B, L, h):
    return CholeskySolve.apply(B, L, h)


def cholesky_logdet(L, h):
    return CholeskyLogdet.apply(L, h)


def potri(L, h):
    """A^-1 (full symmetric, column-major) of A = L L^T from the lower factor L."""
    p = _pfx(L)
    n = L.shape[0]
    Lc = _colmajor(L.detach())
    Ainv = _empty_colmajor(n, n, L)
    _sync()
    if n:
        check_status(getattr(lib, f"ddla_{p}_potri")(h.ptr, b"L", n, _ptr(Lc), n, _ptr(Ainv), n), f"ddla_{p}_potri")
    return Ainv


class GPNegativeLogLikelihood(torch.autograd.Function):
    """loss = y^T K^-1 y / 2 + log det K / 2 in one Function: forward runs
    ppotrf, ppotrs and diag_logdet on the factor; backward uses the closed
    form dloss/dK = (K^-1 - alpha alpha^T) / 2, alpha = K^-1 y (Rasmussen &
    Williams 2006, eq. 5.9, the symmetric-input convention), with K^-1 from
    potri (about n^3 flops, against ~1.8 n^3 for the composed adjoints), and
    dloss/dy = alpha."""

    @staticmethod
    def forward(ctx, K, y, h):
        p = _pfx(K)
        n = K.shape[0]
        Lc = K.detach().t().contiguous().t().clone() if n else K.detach().clone()
        alpha = _colmajor(y.detach()).clone()
        res = np.zeros(2, dtype=np.float64 if K.dtype == torch.float64 else np.float32)
        _sync()
        info = ctypes.c_int(0)
        if n:
            d, dp = _desc(h, n, n, n)
            check_status(getattr(lib, f"ddla_{p}_ppotrf")(h.ptr, b"L", n, _ptr(Lc), 1, 1, dp, ctypes.byref(info)),
                         f"ddla_{p}_ppotrf")
        if info.value != 0:
            raise RuntimeError(f"gp likelihood: the leading minor of order {info.value} is not positive definite")
        if n:
            db, dbp = _desc(h, n, 1, n)
            check_status(getattr(lib, f"ddla_{p}_ppotrs")(h.ptr, b"L", b"L", b"N", n, 1, _ptr(Lc), dp, _ptr(alpha),
                                                          dbp), f"ddla_{p}_ppotrs")
            check_status(getattr(lib, f"ddla_{p}_diag_logdet")(h.ptr, n, _ptr(Lc), n,
                                                               ctypes.c_void_p(res.ctypes.data)),
                         f"ddla_{p}_diag_logdet")
        a = alpha[:, 0]
        ctx.h = h
        ctx.save_for_backward(Lc, a)
        return 0.5 * torch.dot(y.detach(), a) + float(res[0])  # log det K = 2 sum log L_ii

    @staticmethod
    def backward(ctx, g):
        Lc, a = ctx.saved_tensors
        Kbar = potri(Lc, ctx.h)
        Kbar.addr_(a, a, alpha=-1.0)
        Kbar.mul_(0.5 * g)
        return Kbar, g * a, None


def gp_negative_log_likelihood(K, y, h):
    """0.5 y^T K^-1 y + 0.5 log det K (the constant n/2 log 2 pi omitted),
    with the fused closed-form backward."""
    return GPNegativeLogLikelihood.apply(K, y, h)


class GPNegativeLogLikelihoodBatched(torch.autograd.Function):
    """loss[b] = y_b^T K_b^-1 y_b / 2 + log det K_b / 2 for a stack of
    independent problems of one order (K: batch x n x n symmetric, y: batch x
    n, float64 CUDA tensors), with each kernel launch serving every problem:
    the blocked factorization (one panel launch and one trailing update per
    step for all of them), L^-1 by merging inverted diagonal blocks, z = L^-1 y
    and loss = z^T z / 2 + sum log L_ii.  Backward: dloss/dK_b = g_b (K_b^-1 -
    a a^T) / 2 with K_b^-1 = L^-T L^-1 and a = K_b^-1 y_b (Rasmussen &
    Williams 2006, eq. 5.9, the symmetric-input convention), dloss/dy_b = g_b a.
    The device workspace made by forward is kept for backward."""

    @staticmethod
    def forward(ctx, K, y, h):
        if K.dtype != torch.float64 or y.dtype != torch.float64:
            raise TypeError("the batched GP likelihood is built for float64")
        if not (K.is_cuda and y.is_cuda):
            raise ValueError("the batched GP likelihood runs on the GPU: pass CUDA tensors")
        if K.dim() != 3 or K.shape[1] != K.shape[2] or y.shape != K.shape[:2]:
            raise ValueError("K must be batch x n x n and y batch x n")
        batch, n = int(K.shape[0]), int(K.shape[1])
        # K_b is symmetric, so torch's row-major K_b is the column-major one
        Kc = K.detach().contiguous()
        yc = y.detach().contiguous()
        per = ctypes.c_longlong()
        check_status(lib.ddla_d_gp_batched_workspace_size(n, ctypes.byref(per)), "ddla_d_gp_batched_workspace_size")
        # device workspace, 256-byte aligned inside a torch allocation
        raw = torch.empty(batch * per.value + 256, dtype=torch.uint8, device=K.device)
        off = (-raw.data_ptr()) % 256
        work = ctypes.c_void_p(raw.data_ptr() + off)
        loss = torch.empty(batch, dtype=torch.float64, device=K.device)
        info = torch.empty(batch, dtype=torch.int32, device=K.device)
        _sync()
        check_status(lib.ddla_d_gp_batched_forward(h.ptr, batch, n, _ptr(Kc), n * n, _ptr(yc), n, work, _ptr(loss),
                                                   _ptr(info)), "ddla_d_gp_batched_forward")
        bad = torch.nonzero(info).flatten().tolist()
        if bad:
            raise RuntimeError(f"gp likelihood: problem {bad[0]} is not positive definite "
                               f"(leading minor of order {int(info[bad[0]])})")
        ctx.h, ctx.raw, ctx.work, ctx.shape = h, raw, work, (batch, n)
        return loss

    @staticmethod
    def backward(ctx, g):
        batch, n = ctx.shape
        gc = g.detach().to(torch.float64).contiguous()
        Kbar = torch.empty((batch, n, n), dtype=torch.float64, device=gc.device)
        ybar = torch.empty((batch, n), dtype=torch.float64, device=gc.device)
        _sync()
        check_status(lib.ddla_d_gp_batched_backward(ctx.h.ptr, batch, n, ctx.work, _ptr(gc), _ptr(Kbar), n * n,
                                                    _ptr(ybar), n), "ddla_d_gp_batched_backward")
        ctx.h.synchronize()
        ctx.raw = None
        return Kbar, ybar, None


def gp_negative_log_likelihood_batched(K, y, h):
    """Per-problem 0.5 y_b^T K_b^-1 y_b + 0.5 log det K_b (the constant n/2 log
    2 pi omitted) for K of shape batch x n x n and y of shape batch x n, as a
    tensor of batch losses; every kernel launch serves all problems."""
    return GPNegativeLogLikelihoodBatched.apply(K, y, h)


def backward_concurrently(problems, loss_fn, handles):
    """Forward and backward of independent losses at once.

    Problem i (an argument tuple for loss_fn(*args, h)) runs on
    handles[i % len(handles)]; each handle is driven by a host thread of its
    own, with torch's work for its problems enqueued on the handle's compute
    stream (torch.cuda.ExternalStream), so different problems' kernels overlap
    on the GPU -- one problem's latency-bound panel factorization beside
    another's products.  Gradients accumulate into the inputs' .grad as
    loss.backward() would.  The caller's current stream is synchronized first
    (the inputs are ready) and every handle's stream before returning.  Returns
    the detached losses in problem order; the first exception raised by a
    worker is re-raised."""
    import threading

    workers = len(handles)
    if workers < 1:
        raise ValueError("backward_concurrently needs at least one handle")
    torch.cuda.current_stream().synchronize()
    losses = [None] * len(problems)
    errors = []

    def work(w):
        h = handles[w]
        try:
            stream = torch.cuda.ExternalStream(h.stream)
            with torch.cuda.stream(stream):
                for i in range(w, len(problems), workers):
                    loss = loss_fn(*problems[i], h)
                    loss.backward()
                    losses[i] = loss.detach()
            stream.synchronize()
        except BaseException as e:  # re-raised on the calling thread
            errors.append(e)

    threads = [threading.Thread(target=work, args=(w,)) for w in range(workers)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if errors:
        raise errors[0]
    return losses


def gp_negative_log_likelihood_composed(K, y, h):
    """The same loss through the separate Cholesky, solve and log-determinant
    Functions and their adjoints."""
    L = cholesky(K, h)
    alpha = cholesky_solve(y, L, h)
    return 0.5 * torch.dot(y, alpha) + 0.5 * cholesky_logdet(L, h)
