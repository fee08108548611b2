"""Template refinement by gradient steps, beside the reference's averaging update.

The reference has no gradients: it refines templates by averaging the PC
features of the spikes assigned to each cluster (``clustering_qr.py:524-527``)
and the matching path subtracts ``amp * (U^T W)`` per spike
(``template_matching.py:235-236``).  This module differentiates the residual
of that model,

    loss = sum_{c,t} (X - sum_s amp_s * T_{id_s}(. - x_s))[c, t]^2,
    T_i[c, k] = sum_p U[i, p, c] * W[p, k],

with spike times and ids held fixed, with respect to X, U, the temporal basis
W (``wPCA``) and the amplitudes.  The forward and backward passes are the HIP
kernels ``ksm_residual`` / ``ksm_residual_backward`` (float32 or float64;
deterministic summation orders); the backward pass recomputes the residual
and reads each spike's lag window from it, so no per-spike windows are kept.

* ``residual_loss(X, U, W, amps, st, nt)`` -- differentiable loss of one
  padded batch (``torch.autograd.Function``).
* ``residual(X, U, W, amps, st, nt)`` -- the residual itself.
* ``refine_step`` / ``refine`` / ``averaging_update`` -- one preconditioned
  gradient step on U, a sequence of them with backtracking, and the
  feature-averaging update they are compared against.
"""
from __future__ import annotations

import torch

from . import _lib

_DTYPES = {torch.float32: _lib.DTYPE_F32, torch.float64: _lib.DTYPE_F64}


def _check(X, U, W, amps, st, nt):
    if X.device.type != "cuda":
        raise RuntimeError("template refinement runs on a ROCm GPU only")
    dt = X.dtype
    if dt not in _DTYPES:
        raise ValueError(f"X must be float32 or float64 (got {dt})")
    for name, t in (("U", U), ("W", W), ("amps", amps)):
        if t.dtype != dt or t.device != X.device:
            raise ValueError(f"{name} must be {dt} on {X.device}")
    if X.ndim != 2 or U.ndim != 3 or W.ndim != 2:
        raise ValueError("X [C, T], U [K, P, C] and W [P, nt] expected")
    K, P, C = U.shape
    if X.shape[0] != C or W.shape != (P, nt):
        raise ValueError(f"shapes disagree: X {tuple(X.shape)}, U {tuple(U.shape)}, "
                         f"W {tuple(W.shape)}, nt {nt}")
    st = st.to(X.device, torch.int64).contiguous()
    if st.ndim != 2 or st.shape[1] != 2 or amps.reshape(-1).shape[0] != st.shape[0]:
        raise ValueError("st [nsp, 2] (x, id) and amps [nsp] expected")
    if st.shape[0] and (int(st[:, 1].min()) < 0 or int(st[:, 1].max()) >= K):
        raise ValueError(f"template ids must lie in [0, {K})")
    return st


def _orders(st, K):
    """Spike orders: by time (stable) and by (template, time); segment
    offsets of each template in the second."""
    x, ids = st[:, 0], st[:, 1]
    by_t = torch.argsort(x, stable=True)
    by_k = by_t[torch.argsort(ids[by_t], stable=True)]
    seg = torch.zeros(K + 1, dtype=torch.int64, device=st.device)
    seg[1:] = torch.cumsum(torch.bincount(ids, minlength=K), 0)
    return by_t.to(torch.int32), by_k.to(torch.int32), seg.to(torch.int32)


def _workspace(nsp, K, P, C, T, nt, dtype, device):
    lib = _lib.load()
    nb = lib.ksm_residual_workspace(nsp, K, P, C, T, nt, _DTYPES[dtype])
    return torch.empty(max(nb, 1), dtype=torch.uint8, device=device), nb


def _residual(X, U, W, amps, st, order, nt, want_loss=True):
    lib = _lib.load()
    K, P, C = U.shape
    T = X.shape[1]
    X = X.contiguous()
    U = U.contiguous()
    W = W.contiguous()
    amps = amps.reshape(-1).contiguous()
    nsp = st.shape[0]
    R = torch.empty_like(X)
    loss = torch.zeros(1, dtype=torch.float64, device=X.device) if want_loss else None
    ws, nb = _workspace(nsp, K, P, C, T, nt, X.dtype, X.device)
    _lib.check(lib.ksm_residual(X.data_ptr(), T, U.data_ptr(), W.data_ptr(), st.data_ptr(),
                                amps.data_ptr(), order.data_ptr(), nsp, K, P, C, T, nt,
                                R.data_ptr(), T, _lib.ptr(loss), _DTYPES[X.dtype],
                                ws.data_ptr(), nb, _lib.stream_of(X.device)), "ksm_residual")
    return R, loss


def residual(X, U, W, amps, st, nt):
    """R = X - sum_s amp_s T_{id_s}(. - x_s) and sum R^2 (float64 scalar)."""
    st = _check(X, U, W, amps, st, nt)
    order, _, _ = _orders(st, U.shape[0])
    R, loss = _residual(X, U, W, amps, st, order, nt)
    return R, loss[0]


class _ResidualLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, X, U, W, amps, st, nt):
        st = _check(X, U, W, amps, st, nt)
        by_t, by_k, seg = _orders(st, U.shape[0])
        _, loss = _residual(X, U, W, amps, st, by_t, nt)
        ctx.save_for_backward(X, U, W, amps, st, by_t, by_k, seg)
        ctx.nt = nt
        return loss[0].to(X.dtype)

    @staticmethod
    def backward(ctx, gloss):
        X, U, W, amps, st, by_t, by_k, seg = ctx.saved_tensors
        nt = ctx.nt
        lib = _lib.load()
        K, P, C = U.shape
        T = X.shape[1]
        nsp = st.shape[0]
        # the residual again (lag windows are read from it)
        R, _ = _residual(X, U, W, amps, st, by_t, nt, want_loss=False)
        scale = 2.0 * float(gloss)
        dU = torch.empty_like(U.contiguous())
        dW = torch.empty_like(W.contiguous())
        damp = torch.empty(nsp, dtype=X.dtype, device=X.device)
        ws, nb = _workspace(nsp, K, P, C, T, nt, X.dtype, X.device)
        _lib.check(lib.ksm_residual_backward(
            R.data_ptr(), T, U.contiguous().data_ptr(), W.contiguous().data_ptr(),
            st.data_ptr(), amps.reshape(-1).contiguous().data_ptr(), by_k.data_ptr(),
            seg.data_ptr(), nsp, K, P, C, T, nt, scale, dU.data_ptr(), dW.data_ptr(),
            damp.data_ptr(), _DTYPES[X.dtype], ws.data_ptr(), nb,
            _lib.stream_of(X.device)), "ksm_residual_backward")
        dX = R * scale if ctx.needs_input_grad[0] else None
        return dX, dU, dW, damp.reshape(amps.shape), None, None


def residual_loss(X, U, W, amps, st, nt):
    """sum_{c,t} R^2 of one padded batch, differentiable in X, U, W and amps
    (spike times st[:, 0] and template ids st[:, 1] fixed)."""
    return _ResidualLoss.apply(X, U, W, amps, st, nt)


def _gradient(batches, U, W, nt):
    """dloss/dU over ``batches``, the loss, and sum_s amp_s^2 per template."""
    K = U.shape[0]
    U = U.detach().requires_grad_(True)
    total = 0.0
    a2 = torch.zeros(K, dtype=U.dtype, device=U.device)
    for X, st, amps in batches:
        loss = residual_loss(X, U, W, amps, st, nt)
        loss.backward()
        total += float(loss.detach())
        ids = st[:, 1].to(U.device)
        a2.index_add_(0, ids, amps.reshape(-1).to(U.dtype) ** 2)
    return U.grad, total, a2


def _loss(batches, U, W, nt):
    return sum(float(residual(X, U, W, amps, st, nt)[1]) for X, st, amps in batches)


def refine_step(batches, U, W, nt, scale=1.0):
    """One gradient step on U over ``batches`` (a list of (X, st, amps)),
    preconditioned per template by scale / (2 sum_s amp_s^2) -- at scale 1
    the inverse of the loss's curvature in U[i] when the spikes of template i
    overlap no other spike and the rows of W are orthonormal, so that the
    first step from a template set reproduces the amplitude-weighted average
    of the spikes' windows.  Returns the new U and the loss before the step."""
    G, total, a2 = _gradient(batches, U, W, nt)
    step = torch.where(a2 > 0, scale / (2.0 * a2), torch.zeros_like(a2))
    return (U.detach() - step[:, None, None] * G).detach(), total


def refine(batches, U, W, nt, steps=20, max_halvings=40):
    """``steps`` preconditioned gradient steps on U, each at the largest scale
    2^-j (from the last accepted scale, which is doubled again up to 1 after
    each step) that does not increase the loss.  Where spikes of different
    templates overlap, the per-template preconditioner ignores their coupling
    and the unscaled step can overshoot; the backtracking keeps the loss
    non-increasing.  Returns U and the losses [before, after step 1, ...]."""
    U = U.detach()
    scale = 1.0
    losses = []
    for _ in range(steps):
        G, total, a2 = _gradient(batches, U, W, nt)
        if not losses:
            losses.append(total)
        base = torch.where(a2 > 0, 1.0 / (2.0 * a2), torch.zeros_like(a2))[:, None, None] * G
        for _ in range(max_halvings):
            Un = U - scale * base
            ln = _loss(batches, Un, W, nt)
            if ln <= total:
                break
            scale *= 0.5
        else:
            losses.append(total)          # no decrease found: U is kept
            break
        U = Un.detach()
        losses.append(ln)
        scale = min(1.0, 2.0 * scale)
    return U, losses


def averaging_update(batches, U, W, nt):
    """The reference's kind of update with spike times fixed: each template
    becomes the amplitude-weighted average of its spikes' windows projected on
    W, with the spike's own contribution added back to the residual,
        U[i, p, c] = sum_s amp_s F_s[c, p] / sum_s amp_s^2,
        F_s[c, p]  = sum_k (R + amp_s T_i)[c, x_s - h + k] W[p, k]
    (windows clipped to the batch); overlaps between spikes are ignored, as
    feature averaging ignores them.  sum_s amp_s F_s(R) is minus half the
    loss gradient; the own-contribution term is amp_s^2 (W_s W_s^T) U[i] with
    W_s the columns of W inside the batch."""
    K = U.shape[0]
    num = torch.zeros_like(U)
    a2 = torch.zeros(K, dtype=U.dtype, device=U.device)
    h = nt // 2
    for X, st, amps in batches:
        Ug = U.detach().requires_grad_(True)
        loss = residual_loss(X, Ug, W, amps, st, nt)
        loss.backward()
        st = st.to(U.device)
        ids = st[:, 1]
        a = amps.reshape(-1).to(U.dtype)
        t = st[:, :1] - h + torch.arange(nt, device=U.device)[None, :]
        inside = ((t >= 0) & (t < X.shape[1])).to(U.dtype)                 # [nsp, nt]
        G = torch.einsum("sk,pk,qk->spq", inside, W, W)                    # [nsp, P, P]
        own = torch.einsum("spq,sqc->spc", G, U[ids]) * (a * a)[:, None, None]
        num += -0.5 * Ug.grad
        num.index_add_(0, ids, own)
        a2.index_add_(0, ids, a * a)
    keep = a2 > 0
    out = U.clone()
    out[keep] = num[keep] / a2[keep][:, None, None]
    return out
