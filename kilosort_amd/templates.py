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
* ``merging_function(ops, Wall, clu, st, tF, r_thresh, mode='mu',
  check_dt, device)`` -- ``template_matching.merging_function``
  (``template_matching.py:247-358``) in its amplitude ('mu') mode: the
  greedy merge of the most similar templates, largest clusters first; the
  similarity row of each step comes from the GPU Gram of the current
  templates (recomputed after each merge);
* ``postprocess_templates(Wall, ops, clu, st, tF, device)``
  (``template_matching.py:144-152``).

The 'ccg' mode of ``merging_function`` also needs the spike-train
cross-correlograms of ``kilosort.CCG`` (numba) and is not provided here.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib


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


def merging_function(ops, Wall, clu, st, tF, r_thresh=0.5, mode='mu', check_dt=True,
                     device=torch.device('cuda')):
    """``template_matching.merging_function`` in mode 'mu' (merge when the
    template norms differ by less than 20 %).  Returns (Ww [K', C, P] CPU,
    clu2, None, st, tF) sorted by spike time, as the reference."""
    if mode != 'mu':
        raise NotImplementedError("merging_function: only mode='mu' is provided here; "
                                  "mode='ccg' needs kilosort.CCG's cross-correlograms")
    device = _lib.require_gpu(device)
    nt = int(ops['nt'])
    clu2 = clu.copy()
    clu_unq, ns = np.unique(clu2, return_counts=True)
    Ww = torch.as_tensor(Wall).detach().to(torch.float32).cpu().numpy().copy()
    NN = len(Ww)
    isort = np.argsort(ns)[::-1]
    is_merged = np.zeros(NN, 'bool')
    W = ops['wPCA']
    if check_dt:
        raise NotImplementedError("merging_function: check_dt=True (feature rolling) is not "
                                  "provided here")
    S, lag = similarity_matrix(Ww, W, nt, device, lags=True)
    t = 0
    nmerge = 0
    while t < NN:
        kk = clu_unq[isort[t]]
        if is_merged[kk]:
            t += 1
            continue
        mu = np.sqrt((Ww.astype(np.float32) ** 2).sum((1, 2), dtype=np.float32))
        cmax = S[kk].copy()
        cmax[kk] = 0
        jsort = np.argsort(cmax)[::-1]
        is_ccg = 0
        for j in range(NN):
            jj = jsort[j]
            if cmax[jj] < r_thresh:
                break
            dmu = 2 * (mu[kk] - mu[jj]) / (mu[kk] + mu[jj])
            is_ccg = abs(dmu) < 0.2
            if is_ccg:
                is_merged[jj] = 1
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
    imap = np.cumsum((~is_merged).astype('int32')) - 1
    if imap.size > 0:
        clu2 = imap[clu2]
    Ww = torch.from_numpy(Ww[~is_merged])
    sorted_idx = np.argsort(st[:, 0])
    st = np.take_along_axis(st, sorted_idx[..., np.newaxis], axis=0)
    clu2 = clu2[sorted_idx]
    tF = tF[torch.from_numpy(sorted_idx)]
    return Ww, clu2, None, st, tF


def postprocess_templates(Wall, ops, clu, st, tF, device=torch.device('cuda')):
    """``template_matching.postprocess_templates``: align, then merge templates
    more than 0.9 similar with norms within 20 %; returns Wall3 [K', P, C] on
    ``device``."""
    Wall2, _ = align_U(Wall, ops, device=device)
    Wall3, _, _, _, _ = merging_function(ops, Wall2.transpose(1, 2), clu, st, tF, 0.9, 'mu',
                                         check_dt=False, device=device)
    return Wall3.transpose(1, 2).to(device)
