"""Template Gram consumers on the GPU: similarity, alignment, merging.

Mirrors the post-processing of learned templates that reuses the template
Gram of the matching path:

* ``similarity(Wall, W, nt)`` -- ``CCG.similarity`` (``CCG.py:94-102``):
  the maximum over lags of the normalised template cross-correlation, for
  every pair (``ksm_similarity``: the float64 Gram kernel of
  ``prepare_matching`` on the normalised templates);
* ``align_U(U, ops, device)`` -- ``template_matching.align_U``
  (``template_matching.py:128-141``): each template rolled in time so that
  its peak correlation with the universal waveforms sits at the centre
  (``ksm_align_templates``);
* ``merging_function(ops, Wall, clu, st, tF, r_thresh, mode, check_dt,
  device)`` -- ``template_matching.merging_function``
  (``template_matching.py:247-358``): the greedy merge of the most similar
  templates, largest clusters first; the similarity row and lag of each step
  come from the GPU Gram of the current templates (recomputed after each
  merge); mode 'ccg' tests cross-refractoriness with GPU correlogram counts
  (``ksm_acg`` for the clusters' refractoriness, ``ksm_ccg`` per candidate
  pair), mode 'mu' the template norms; ``check_dt`` shifts the absorbed
  cluster's features in PC space (``roll_features``, ``ksm_pc_map``);
* ``postprocess_templates(Wall, ops, clu, st, tF, device)``
  (``template_matching.py:144-152``).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from . import ccg as CG


def _f32(t, device):
    return torch.as_tensor(t).to(device, torch.float32).contiguous()


def similarity_matrix(Wall, W, nt, device, lags=False):
    """(S [K, K] float32 numpy, lag [K, K] int32 numpy or None) for Wall
    [K, C, P] and W [P, nt] (both moved to ``device``)."""
    lib = _lib.load()
    Wall = _f32(Wall, device)
    W = _f32(W, device)
    K, C, P = Wall.shape
    S = torch.empty((K, K), dtype=torch.float32, device=device)
    lag = torch.empty((K, K), dtype=torch.int32, device=device) if lags else None
    nb = lib.ksm_similarity_workspace(K, C, P, nt)
    ws = torch.empty(nb, dtype=torch.uint8, device=device)
    _lib.check(lib.ksm_similarity(Wall.data_ptr(), W.data_ptr(), K, C, P, nt, S.data_ptr(),
                                  _lib.ptr(lag), ws.data_ptr(), nb, _lib.stream_of(device)),
               "ksm_similarity")
    return S.cpu().numpy(), (lag.cpu().numpy() if lags else None)


def similarity(Wall, W, nt=61):
    """``CCG.similarity``: [K, K] float32 numpy array, the maximum over lags of
    the correlation between normalised templates i and j."""
    device = Wall.device if torch.is_tensor(Wall) and Wall.is_cuda else torch.device('cuda')
    device = _lib.require_gpu(device)
    return similarity_matrix(Wall, W, nt, device)[0]


def align_U(U, ops, device=torch.device('cuda')):
    """``template_matching.align_U``: U [K, C, P] -> (Unew [K, P, C] on
    ``device``, imax [K] int64)."""
    device = _lib.require_gpu(device)
    lib = _lib.load()
    U = _f32(U, device)
    K, C, P = U.shape
    nt = int(ops['nt'])
    wPCA = _f32(ops['wPCA'], device)
    wTEMP = _f32(ops['wTEMP'], device)
    nk = wTEMP.shape[0]
    Unew = torch.empty((K, P, C), dtype=torch.float32, device=device)
    imax = torch.empty(K, dtype=torch.int32, device=device)
    nb = lib.ksm_align_workspace(K, C, P, nt, nk)
    ws = torch.empty(nb, dtype=torch.uint8, device=device)
    _lib.check(lib.ksm_align_templates(U.data_ptr(), wPCA.data_ptr(), wTEMP.data_ptr(), K, C, P,
                                       nt, nk, Unew.data_ptr(), imax.data_ptr(), ws.data_ptr(),
                                       nb, _lib.stream_of(device)), "ksm_align_templates")
    return Unew, imax.long()


def shift_operator(W, dt):
    """[P, P] float64: x -> (x @ W) shifted by dt samples with the edge value
    repeated, projected back on W (``roll_features``'s torch.roll by dt and
    its edge fill, ``template_matching.py:363-375``, as one linear map)."""
    W = np.asarray(W, np.float64)
    nt = W.shape[1]
    Sh = np.zeros((nt, nt))
    Sh[np.clip(np.arange(nt) - dt, 0, nt - 1), np.arange(nt)] = 1.0
    return W @ Sh @ W.T


def roll_features(wPCA, tF, Wall, spike_idx, clust_idx, dt, device=None):
    """``template_matching.roll_features``: tF rows ``spike_idx`` (bool mask or
    indices) and template ``clust_idx`` of Wall shifted by dt samples in PC
    space, in place (``ksm_pc_map``); returns (tF, Wall)."""
    device = _lib.require_gpu(device if device is not None else
                              (tF.device if tF.is_cuda else torch.device('cuda')))
    lib = _lib.load()
    M = torch.from_numpy(shift_operator(torch.as_tensor(wPCA).double().cpu().numpy(), dt)
                         .astype(np.float32)).to(device)
    rows = np.nonzero(np.asarray(spike_idx))[0] if np.asarray(spike_idx).dtype == bool \
        else np.asarray(spike_idx, np.int64)
    X = tF if (tF.is_cuda and tF.is_contiguous() and tF.dtype == torch.float32) else \
        tF.to(device, torch.float32).contiguous()
    r = torch.from_numpy(rows.astype(np.int64)).to(device)
    n, R, P = X.shape
    _lib.check(lib.ksm_pc_map(X.data_ptr(), r.data_ptr(), len(rows), R, P, M.data_ptr(),
                              _lib.stream_of(device)), "ksm_pc_map")
    Wk = torch.as_tensor(Wall[clust_idx]).to(device, torch.float32).contiguous().unsqueeze(0)
    one = torch.zeros(1, dtype=torch.int64, device=device)
    _lib.check(lib.ksm_pc_map(Wk.data_ptr(), one.data_ptr(), 1, Wk.shape[1], P, M.data_ptr(),
                              _lib.stream_of(device)), "ksm_pc_map")
    if X is not tF:
        tF.copy_(X.to(tF.device, tF.dtype))
    if torch.is_tensor(Wall):
        Wall[clust_idx] = Wk[0].to(Wall.device, Wall.dtype)
    else:
        Wall[clust_idx] = Wk[0].cpu().numpy()
    return tF, Wall


def cross_refractory(st0, st1, ccg_threshold=0.25, device=None):
    """``CCG.check_CCG(st0, st1)``'s cross-refractoriness (``CCG.py:86-101``):
    the cross-correlogram counts of the two trains (seconds) from
    ``ksm_ccg``, the metrics of ``CCG_metrics`` on the host."""
    device = _lib.require_gpu(device if device is not None else torch.device('cuda'))
    lib = _lib.load()
    nbins, tbin = 500, 1 / 1000
    a = np.sort(np.asarray(st0, np.float64))
    b = np.sort(np.asarray(st1, np.float64))
    T = np.maximum(a.max(), b.max()) - np.minimum(a.min(), b.min())
    ta = torch.from_numpy(a).to(device)
    tb = torch.from_numpy(b).to(device)
    counts = torch.empty(2 * nbins + 1, dtype=torch.int32, device=device)
    _lib.check(lib.ksm_ccg(ta.data_ptr(), a.size, tb.data_ptr(), b.size, nbins, tbin,
                           counts.data_ptr(), _lib.stream_of(device)), "ksm_ccg")
    K = counts.cpu().numpy().view(np.uint32).astype(np.float64)
    R12, Q12, _ = CG.ccg_metrics(a.size, b.size, K, T, nbins, tbin)
    return bool(R12 < ccg_threshold and (Q12 < .05))


def merging_function(ops, Wall, clu, st, tF, r_thresh=0.5, mode='ccg', check_dt=True,
                     device=torch.device('cuda')):
    """``template_matching.merging_function``: the greedy merge of the most
    similar templates, largest clusters first, in mode 'ccg' (merge when the
    spike trains are cross-refractory; only refractory clusters absorb) or
    'mu' (template norms within 20 %); with check_dt the absorbed cluster's
    features and template are shifted to the lag of the similarity maximum
    and its spike times in ``st`` are shifted in place, as in the reference.
    Returns (Ww [K', C, P] CPU, clu2, is_ref or None, st, tF) sorted by spike
    time."""
    if mode not in ('ccg', 'mu'):
        raise ValueError(f"merging_function: unknown mode {mode!r}")
    device = _lib.require_gpu(device)
    nt = int(ops['nt'])
    clu2 = clu.copy()
    clu_unq, ns = np.unique(clu2, return_counts=True)
    Ww = torch.as_tensor(Wall).detach().to(torch.float32).cpu().numpy().copy()
    NN = len(Ww)
    isort = np.argsort(ns)[::-1]
    is_merged = np.zeros(NN, 'bool')
    acg_threshold = ops['settings']['acg_threshold']
    ccg_threshold = ops['settings']['ccg_threshold']
    if mode == 'ccg':
        is_ref, est_contam_rate = CG.refract(clu, st[:, 0] / ops['fs'],
                                             acg_threshold=acg_threshold,
                                             ccg_threshold=ccg_threshold, device=device)
    W = ops['wPCA']
    tF_dev = None
    S, lag = similarity_matrix(Ww, W, nt, device, lags=True)
    t = 0
    nmerge = 0
    while t < NN:
        kk = clu_unq[isort[t]]
        if (mode == 'ccg') and is_ref[kk] == 0:
            t += 1
            continue
        if is_merged[kk]:
            t += 1
            continue
        mu = np.sqrt((Ww.astype(np.float32) ** 2).sum((1, 2), dtype=np.float32))
        cmax = S[kk].copy()
        imax = lag[kk]
        cmax[kk] = 0
        jsort = np.argsort(cmax)[::-1]
        if mode == 'ccg':
            st0 = st[:, 0][clu2 == kk] / ops['fs']
        is_ccg = 0
        for j in range(NN):
            jj = jsort[j]
            if cmax[jj] < r_thresh:
                break
            if mode == 'ccg':
                st1 = st[:, 0][clu2 == jj] / ops['fs']
                is_ccg = cross_refractory(st0, st1, ccg_threshold, device)
            else:
                dmu = 2 * (mu[kk] - mu[jj]) / (mu[kk] + mu[jj])
                is_ccg = abs(dmu) < 0.2
            if is_ccg:
                is_merged[jj] = 1
                dt = int(imax[kk]) - int(imax[jj])
                if dt != 0 and check_dt:
                    idx = (clu2 == jj)
                    if tF_dev is None:
                        tF_dev = torch.as_tensor(tF).to(device, torch.float32).contiguous()
                    tF_dev, Ww = roll_features(W, tF_dev, Ww, idx, jj, dt, device=device)
                    st[idx, 0] -= dt
                a = np.float32(ns[kk] / (ns[kk] + ns[jj]))
                b = np.float32(ns[jj] / (ns[kk] + ns[jj]))
                Ww[kk] = a * Ww[kk] + b * Ww[jj]
                Ww[jj] = 0
                ns[kk] += ns[jj]
                ns[jj] = 0
                clu2[clu2 == jj] = kk
                S, lag = similarity_matrix(Ww, W, nt, device, lags=True)
                break
        if is_ccg == 0:
            t += 1
        else:
            nmerge += 1
    if tF_dev is not None:
        tF.copy_(tF_dev.to(tF.device, tF.dtype))             # in place, as the reference
    imap = np.cumsum((~is_merged).astype('int32')) - 1
    if imap.size > 0:
        clu2 = imap[clu2]
    Ww = torch.from_numpy(Ww[~is_merged])
    is_ref = is_ref[~is_merged] if mode == 'ccg' else None
    sorted_idx = np.argsort(st[:, 0])
    st = np.take_along_axis(st, sorted_idx[..., np.newaxis], axis=0)
    clu2 = clu2[sorted_idx]
    tF = tF[torch.from_numpy(sorted_idx).to(tF.device)]
    return Ww, clu2, is_ref, st, tF


def postprocess_templates(Wall, ops, clu, st, tF, device=torch.device('cuda')):
    """``template_matching.postprocess_templates``: align, then merge templates
    more than 0.9 similar with norms within 20 %; returns Wall3 [K', P, C] on
    ``device``."""
    Wall2, _ = align_U(Wall, ops, device=device)
    Wall3, _, _, _, _ = merging_function(ops, Wall2.transpose(1, 2), clu, st, tF, 0.9, 'mu',
                                         check_dt=False, device=device)
    return Wall3.transpose(1, 2).to(device)
