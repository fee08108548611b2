"""Brute-force float64 matching pursuit and closed forms (test oracle).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

``brute_matching`` is an independent statement of the peel in
``kilosort/template_matching.py:170-244``: instead of forming B once and
updating it through the template Gram ``ctc``, it recomputes B from the
current residual by a direct sum over channels and taps every iteration,
so it pins the lag direction and scale of ``ctc`` and of the two-phase
update (B is linear in the residual).  The per-iteration decisions (window
max, threshold, first-index argmax) follow ``:197-217``.

``ctc_direct`` is the closed form ctc[i, j, d + nt] =
sum_c sum_k T_i[c, k] * T_j[c, k + d] with T_i = U_i^T wPCA, the response of
template i's filter to template j placed d samples earlier.
"""
from __future__ import annotations

import numpy as np


def full_templates(U, W):
    """T[j, c, k] = sum_p U[j, p, c] * W[p, k]  (float64)."""
    return np.einsum("jpc,pk->jck", np.asarray(U, np.float64),
                     np.asarray(W, np.float64))


def ctc_direct(U, W):
    """ctc [K, K, 2*nt+1] by direct cross-correlation of full templates."""
    T = full_templates(U, W)
    K, C, nt = T.shape
    out = np.zeros((K, K, 2 * nt + 1))
    for d in range(-nt, nt + 1):
        lo, hi = max(0, -d), min(nt, nt - d)        # k and k + d both in range
        if lo >= hi:
            continue
        out[:, :, d + nt] = np.einsum("ick,jck->ij", T[:, :, lo:hi],
                                      T[:, :, lo + d:hi + d])
    return out


def direct_scores(Xres, T):
    """B[j, x] = sum_c sum_k Xres[c, x + k - nt//2] * T[j, c, k], zero padded."""
    K, C, nt = T.shape
    h = nt // 2
    Tn = Xres.shape[1]
    Xp = np.zeros((C, Tn + 2 * h))
    Xp[:, h:h + Tn] = Xres
    # windows [C, Tn, nt]
    win = np.lib.stride_tricks.sliding_window_view(Xp, nt, axis=1)
    return np.einsum("cxk,jck->jx", win, T)


def brute_matching(X, U, W, nt, Th, max_peels):
    """Matching pursuit with B recomputed from the residual each iteration.

    Returns (times, ids, amps, th_amps, Xres) with spikes in the reference's
    output order (iteration-major, increasing time within an iteration).
    """
    X = np.asarray(X, np.float64)
    T = full_templates(U, W)
    nm = (np.asarray(U, np.float64) ** 2).sum((1, 2))
    Xres = X.copy()
    Tn = X.shape[1]
    h = nt // 2
    out_t, out_i, out_a, out_th = [], [], [], []
    for _ in range(max_peels):
        B = direct_scores(Xres, T)
        Cf = np.maximum(B, 0) ** 2 / nm[:, None]
        Cf[:, :nt] = 0
        Cf[:, -nt:] = 0
        imax = np.argmax(Cf, 0)                       # first index on ties
        Cfmax = Cf[imax, np.arange(Tn)]
        pad = np.full(nt, -np.inf)
        win = np.lib.stride_tricks.sliding_window_view(
            np.concatenate([pad, Cfmax, pad]), 2 * nt + 1)
        Cmax = win.max(1)
        xs = np.nonzero((Cmax > Th ** 2) & (Cmax == Cfmax))[0]
        if len(xs) == 0:
            break
        ids = imax[xs]
        amps = B[ids, xs] / nm[ids]
        out_t.append(xs)
        out_i.append(ids)
        out_a.append(amps)
        out_th.append(np.sqrt(Cmax[xs]))
        for x, j, a in zip(xs, ids, amps):
            Xres[:, x - h:x + h + 1] -= a * T[j]
    cat = (lambda v, dt: np.concatenate(v).astype(dt) if v else np.zeros(0, dt))
    return (cat(out_t, np.int64), cat(out_i, np.int64), cat(out_a, np.float64),
            cat(out_th, np.float64), Xres)


def nearest_channels(xc, yc, nC):
    """Sets of the nC nearest channels of every channel (ties resolved by a
    stable sort; compare as sets where distances tie)."""
    xc = np.asarray(xc, np.float64)
    yc = np.asarray(yc, np.float64)
    d2 = (xc[None, :] - xc[:, None]) ** 2 + (yc[None, :] - yc[:, None]) ** 2
    return np.argsort(d2, 0, kind="stable")[:nC], np.sort(d2, 0)[:nC]


def reference_time(iX, ibatch, NT, nt, nt0min, batch_downsampling=1):
    """Absolute spike time of padded-batch column iX (``:102-104``)."""
    return (iX - nt) + ibatch * batch_downsampling * NT - nt // 2 + nt0min
