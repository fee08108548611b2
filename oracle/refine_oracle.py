"""CPU restatement of the template-refinement loss (test oracle).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``; never part of the product
path.

The reference has no gradients.  Its matching path subtracts, for each
accepted spike (column x, template id, amplitude a), ``a * sum_p U[id, p, c] *
W[p, k]`` at sample ``x - nt//2 + k`` (``kilosort/template_matching.py:235-236``:
``Xres[:, iX + lag_x] -= amp * einsum('ijk, jl -> kil', U[iY], W)``).  The
refinement loss is the squared norm of the residual of that model over one
padded batch, with every spike's contribution clipped to the batch:

    loss = sum_{c,t} (X - sum_s a_s T_{id_s}(. - x_s))[c, t]^2.

``residual`` builds the model with the same einsum as the reference line and
one accumulating index_put, so torch autograd (in float64) gives the gradients
the HIP kernels are checked against.  The restatement itself is pinned by
finite differences and by closed forms (``tests/test_refine_oracle.py``).
"""
from __future__ import annotations

import torch


def residual(X, U, W, amps, st, nt):
    """R [C, T] and loss = sum R^2 (differentiable in X, U, W, amps)."""
    C, T = X.shape
    h = nt // 2
    x = st[:, 0].long()
    ids = st[:, 1].long()
    a = amps.reshape(-1)
    lag = torch.arange(-h, h + 1)
    # per spike [C, nt]: a * (U[id]^T W), as template_matching.py:235
    wav = torch.einsum("ijk, jl -> ikl", U[ids], W) * a[:, None, None]
    t = x[:, None] + lag[None, :]                                   # [nsp, nt]
    valid = (t >= 0) & (t < T)
    cc = torch.arange(C)[None, :, None].expand(len(x), C, nt)
    tt = t[:, None, :].expand(len(x), C, nt)
    vm = valid[:, None, :].expand(len(x), C, nt)
    model = torch.zeros((C, T), dtype=X.dtype)
    model = model.index_put((cc[vm], tt[vm]), wav[vm], accumulate=True)
    R = X - model
    return R, (R * R).sum()
