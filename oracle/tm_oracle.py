"""CPU restatement of Kilosort4's learned-template matching (test oracle).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke`` and
the ``cpu_baseline`` leg of ``bench.py``; never part of the product path.

Each function restates the reference algorithm with the torch operations the
reference uses, in the same order and with the same shapes, so that on the CPU
the float32 path reproduces the reference's rounding sequence on a given torch
build (checked against fixtures written from the reference itself, see
``tests/golden``).  Passing float64 tensors gives the FP64 oracle: every step
then runs in float64, including the amplitude buffers.

Reference: ``kilosort/template_matching.py`` (cited per function).
"""
from __future__ import annotations

import numpy as np
import torch
from torch.nn.functional import conv1d, max_pool1d

CPU = torch.device("cpu")


def prepare_extract(xc, yc, U, nC, position_limit, device=CPU):
    """Nearest channels, position mask, peak channel and peak-channel PCs.

    Follows ``template_matching.py:44-53``: squared distances, the default
    (unstable) ``np.argsort`` along axis 0 for the nC nearest channels, the
    sorted distances compared with ``position_limit**2``, ``iU`` = argmax over
    channels of the sum over PCs of U**2, and ``Ucc[n, i, p] =
    U[i, p, iCC[n, iU[i]]]``.
    """
    d2 = (xc - xc[:, np.newaxis]) ** 2 + (yc - yc[:, np.newaxis]) ** 2
    near = np.argsort(d2, 0)[:nC]
    mask = np.sort(d2, 0)[:nC] < position_limit ** 2
    iCC = torch.from_numpy(near).to(device)
    iCC_mask = torch.from_numpy(mask).to(device)
    iU = torch.argmax((U ** 2).sum(1), -1)
    Ucc = U[torch.arange(U.shape[0]), :, iCC[:, iU]]
    return iCC, iCC_mask, iU, Ucc


def prepare_matching(ops, U):
    """ctc[i, j, l] = sum_pq UtU[i,j,p,q] WtW[p,q,l]  (``template_matching.py:155-167``).

    WtW is the flipped full cross-correlation of the temporal basis with
    itself (2*nt+1 lags), UtU the channel Gram of the spatial components.
    """
    nt = ops["nt"]
    W = ops["wPCA"].contiguous()
    WtW = torch.flip(conv1d(W.reshape(-1, 1, nt), W.reshape(-1, 1, nt),
                            padding=nt), [2])
    UtU = torch.einsum("ikl, jml -> ijkm", U, U)
    return torch.einsum("ijkm, kml -> ijl", UtU, WtW)


def template_norms(U):
    """nm[i] = sum over C then over P of U**2 (``template_matching.py:176``)."""
    return (U ** 2).sum(-1).sum(-1)


def project(ops, X, U):
    """B[i, t] = sum_{p,c} U[i,p,c] * (X[c] (*) wPCA[p])[t]  (``:180-181``)."""
    nt = ops["nt"]
    W = ops["wPCA"].contiguous()
    B = conv1d(X.unsqueeze(1), W.unsqueeze(1), padding=nt // 2)
    return torch.einsum("ijk, kjl -> il", U, B)


def run_matching(ops, X, U, ctc, capacity=100000, record=None):
    """Matching pursuit over one padded batch (``template_matching.py:170-244``):
    the projection B, then ``peel``."""
    nm = template_norms(U)
    B = project(ops, X, U)
    return peel(ops, B, X, U, ctc, nm, capacity=capacity, record=record)


def peel(ops, B, X, U, ctc, nm, capacity=100000, record=None):
    """The peel loop of ``run_matching`` (``template_matching.py:183-244``)
    from given scores B [K, T] (modified in place) and norms nm.

    Per iteration: Cf = relu(B)^2 / nm with the first and last nt columns
    zeroed; best template per column; a window max of width 2*nt+1; accept
    columns that exceed Th^2 and equal the window max; record (time, id),
    amps = B/nm and th_amps = sqrt(Cmax); subtract the accepted spikes from the
    residual and their ctc columns from B in two interleaved phases.  Where
    the windows of one phase overlap (only exact score ties place accepted
    spikes that close) the indexed subtraction keeps the last spike's write
    (``_sub_last_write``).

    ``record`` (a list) receives per-iteration decision data for the margin
    report when given.
    """
    dt = X.dtype
    Th = ops["Th_learned"]
    nt = ops["nt"]
    max_peels = ops["max_peels"]
    W = ops["wPCA"].contiguous()

    lag_b = torch.arange(-nt, nt + 1)
    lag_x = torch.arange(-(nt // 2), nt // 2 + 1)

    st = torch.zeros((capacity, 2), dtype=torch.int64)
    amps = torch.zeros((capacity, 1), dtype=dt)
    th_amps = torch.zeros((capacity, 1), dtype=dt)
    k = 0
    Xres = X.clone()

    for _ in range(max_peels):
        Cf = torch.relu(B) ** 2 / nm.unsqueeze(-1)
        Cf[:, :nt] = 0
        Cf[:, -nt:] = 0
        Cfmax, imax = torch.max(Cf, 0)
        Cmax = max_pool1d(Cfmax.unsqueeze(0).unsqueeze(0), (2 * nt + 1),
                          stride=1, padding=nt)
        ok = (Cmax[0, 0] > Th ** 2) * (torch.abs(Cmax[0, 0] - Cfmax) < 1e-9)
        xs = torch.nonzero(ok)
        if record is not None:
            record.append(_decision_record(Cf, Cfmax, imax, Th, nt, xs))
        if len(xs) == 0:
            break
        iX = xs[:, :1]
        iY = imax[iX]
        nsp = len(iX)
        st[k:k + nsp, 0] = iX[:, 0]
        st[k:k + nsp, 1] = iY[:, 0]
        amps[k:k + nsp] = B[iY, iX] / nm[iY]
        amp = amps[k:k + nsp]
        th_amps[k:k + nsp] = Cmax[0, 0, iX[:, 0], None] ** .5
        k += nsp
        for j in range(2):
            sel = iY[j::2, 0]
            _sub_last_write(Xres, iX[j::2, 0], lag_x,
                            amp[j::2] * torch.einsum("ijk, jl -> kil", U[sel], W))
            _sub_last_write(B, iX[j::2, 0], lag_b, amp[j::2] * ctc[:, sel, :])

    return st[:k], amps[:k], th_amps[:k], Xres


def _sub_last_write(A, centers, lags, vals):
    """``A[:, centers[:, None] + lags] -= vals`` (vals [R, m, L]) as the
    reference's phase update (template_matching.py:236-238) writes it when the
    windows of two spikes overlap: the subtraction is one non-accumulating
    indexed write of (A gathered before the write) - vals, so a column covered
    by several windows keeps the last spike's value.  That is the order of
    torch's single-threaded index_put; with duplicate indices torch leaves the
    order undefined (its multi-threaded CPU and CUDA index_put do not keep it),
    so the rule is written out here: for every column, the last window
    covering it."""
    m, L = len(centers), len(lags)
    if m == 0:
        return
    cols = (centers[:, None] + lags[None, :]).reshape(-1)
    spike = torch.arange(m).repeat_interleave(L)
    last = torch.full((A.shape[1],), -1, dtype=torch.int64)
    last.scatter_reduce_(0, cols, spike, reduce="amax")
    win = cols[last[cols] == spike]                  # each column once, from its last spike
    who = last[win]
    lag = win - centers[who] - lags[0]
    A[:, win] = A[:, win] - vals[:, who, lag]


def features(ops, Xres, st, amps, iCC, iU, Ucc):
    """PC features of the final residual plus the spike's own template.

    ``template_matching.py:85-86``: gather the nC nearest channels of the
    template's peak channel over the spike's 61-sample window, project on
    wPCA, and add amps * Ucc.  Returns [nC, nsp, P].
    """
    nt = ops["nt"]
    lag_x = torch.arange(-(nt // 2), nt // 2 + 1)
    xfeat = Xres[iCC[:, iU[st[:, 1:2]]], st[:, :1] + lag_x] @ ops["wPCA"].T
    xfeat += amps * Ucc[:, st[:, 1]]
    return xfeat


def extract(ops, bfile, U, capacity=100000, per_batch=None):
    """Whole-recording extraction (``template_matching.py:56-126``).

    Returns (st float64 numpy [N, 3], tF float32 torch [N, nC, P], ops).
    ``per_batch`` (a list) receives the raw per-batch run_matching outputs.
    """
    nC = ops["settings"]["nearest_chans"]
    position_limit = ops["settings"]["position_limit"]
    iCC, iCC_mask, iU, Ucc = prepare_extract(ops["xc"], ops["yc"], U, nC,
                                             position_limit)
    ops["iCC"], ops["iCC_mask"], ops["iU"] = iCC, iCC_mask, iU
    nt = ops["nt"]
    ctc = prepare_matching(ops, U)
    st = np.zeros((10 ** 6, 3), "float64")
    tF = torch.zeros((10 ** 6, nC, ops["settings"]["n_pcs"]))
    k = 0
    for ibatch in range(int(bfile.n_batches)):
        X = bfile.padded_batch_to_torch(ibatch, ops).to(U.dtype)
        stt, amps, th_amps, Xres = run_matching(ops, X, U, ctc, capacity)
        if per_batch is not None:
            per_batch.append((stt.clone(), amps.clone(), th_amps.clone()))
        xfeat = features(ops, Xres, stt, amps, iCC, iU, Ucc)
        if ibatch == 0:
            neg = (stt[:, 0] - nt - nt // 2 + ops["nt0min"]) < 0
            stt, xfeat = stt[~neg, :], xfeat[:, ~neg, :]
            amps, th_amps = amps[~neg, :], th_amps[~neg, :]
        nsp = len(stt)
        if k + nsp > st.shape[0]:
            st = np.concatenate((st, np.zeros_like(st)), 0)
            tF = torch.cat((tF, torch.zeros_like(tF)), 0)
        t_shift = ibatch * bfile.batch_downsampling * ops["batch_size"]
        sttd = stt.double()
        st[k:k + nsp, 0] = ((sttd[:, 0] - nt) + t_shift).numpy() - nt // 2 + ops["nt0min"]
        st[k:k + nsp, 1] = sttd[:, 1].numpy()
        st[k:k + nsp, 2] = th_amps.double().numpy().squeeze()
        tF[k:k + nsp] = xfeat.transpose(0, 1).float()
        k += nsp
    isort = np.argsort(st[:k, 0])
    return st[isort], tF[isort], ops


# ----------------------------------------------------------------------------
# decision margins (near-tie report)

def _decision_record(Cf, Cfmax, imax, Th, nt, xs):
    """Relative margins of the three float comparisons of one iteration."""
    th2 = float(Th) ** 2
    cf = Cfmax.double()
    rec = {}
    # threshold: window maxima that sit close to Th^2
    pad = torch.full((nt,), -np.inf, dtype=torch.float64)
    ext = torch.cat([pad, cf, pad])
    win = ext.unfold(0, 2 * nt + 1, 1)                       # [T, 2nt+1]
    others = win.clone()
    others[:, nt] = -np.inf
    wmax_other = others.max(1).values
    is_lmax = cf >= wmax_other
    lm = cf[is_lmax & (cf > 0)]
    rec["threshold"] = float(((lm - th2).abs() / th2).min()) if len(lm) else np.inf
    # local max: columns above threshold, distance to the best other column
    above = cf > th2
    if above.any():
        rel = (cf[above] - wmax_other[above]).abs() / cf[above]
        rec["localmax"] = float(rel.min())
    else:
        rec["localmax"] = np.inf
    # template argmax at accepted columns
    if len(xs):
        cols = Cf[:, xs[:, 0]].double()
        top2 = torch.topk(cols, 2, dim=0).values if cols.shape[0] > 1 else None
        if top2 is not None:
            rec["argmax"] = float(((top2[0] - top2[1]) / top2[0]).min())
        else:
            rec["argmax"] = np.inf
    else:
        rec["argmax"] = np.inf
    rec["n"] = int(len(xs))
    return rec


def decision_margins(ops, X, U, ctc=None):
    """Minimum relative margins over all iterations of one run_matching."""
    if ctc is None:
        ctc = prepare_matching(ops, U)
    rec = []
    run_matching(ops, X, U, ctc, record=rec)
    out = {key: min(r[key] for r in rec) for key in ("threshold", "localmax", "argmax")}
    out["iterations"] = len(rec)
    return out
