"""CPU restatement of Kilosort4's template Gram consumers (test oracle).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``; never part of the product
path.  Same torch / numpy operations in the same order as
``kilosort/CCG.py:94-102`` (similarity) and ``kilosort/template_matching.py``
:128-141 (align_U), :144-152 (postprocess_templates) and :247-358
(merging_function, both modes, with the feature rolling of :361-379).
"""
from __future__ import annotations

import numpy as np
import torch
from torch.nn.functional import conv1d


def similarity(Wall, W, nt=61):
    """``CCG.py:94-102``."""
    WtW = conv1d(W.reshape(-1, 1, nt), W.reshape(-1, 1, nt), padding=nt)
    WtW = torch.flip(WtW, [2, ])
    mu = (Wall ** 2).sum((1, 2), keepdims=True) ** .5
    Wnorm = Wall / (1e-6 + mu)
    UtU = torch.einsum('ilk, jlm -> ijkm', Wnorm, Wnorm)
    similar_templates = torch.einsum('ijkm, kml -> ijl', UtU, WtW).numpy()
    return similar_templates.max(axis=-1)


def align_U(U, ops):
    """``template_matching.py:128-141``."""
    Uex = torch.einsum('xyz, zt -> xty', U, ops['wPCA'])
    X = Uex.reshape(-1, ops['Nchan']).T
    X = conv1d(X.unsqueeze(1), ops['wTEMP'].unsqueeze(1), padding=ops['nt'] // 2)
    Xmax = X.abs().max(0)[0].max(0)[0].reshape(-1, ops['nt'])
    imax = torch.argmax(Xmax, 1)
    Unew = Uex.clone()
    for j in range(ops['nt']):
        ix = imax == j
        Unew[ix] = torch.roll(Unew[ix], ops['nt'] // 2 - j, -2)
    Unew = torch.einsum('xty, zt -> xzy', Unew, ops['wPCA'])
    return Unew, imax


def merging_function(ops, Wall, clu, st, tF, r_thresh=0.5, mode='mu', check_dt=False):
    """``template_matching.py:247-358`` (both modes; ``st`` is shifted in
    place by check_dt, as in the reference)."""
    from oracle import phy_oracle as PH
    clu2 = clu.copy()
    clu_unq, ns = np.unique(clu2, return_counts=True)
    Ww = Wall.clone()
    NN = len(Ww)
    isort = np.argsort(ns)[::-1]
    is_merged = np.zeros(NN, 'bool')
    acg_threshold = ops['settings']['acg_threshold']
    ccg_threshold = ops['settings']['ccg_threshold']
    if mode == 'ccg':
        is_ref, est_contam_rate = PH.refract(clu, st[:, 0] / ops['fs'],
                                             acg_threshold=acg_threshold,
                                             ccg_threshold=ccg_threshold)
    nt = ops['nt']
    W = ops['wPCA'].contiguous()
    WtW = conv1d(W.reshape(-1, 1, nt), W.reshape(-1, 1, nt), padding=nt)
    WtW = torch.flip(WtW, [2, ])
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
        mu = (Ww ** 2).sum((1, 2), keepdims=True) ** .5
        Wnorm = Ww / (1e-6 + mu)
        UtU = torch.einsum('lk, jlm -> jkm', Wnorm[kk], Wnorm)
        ctc = torch.einsum('jkm, kml -> jl', UtU, WtW)
        cmax, imax = ctc.max(1)
        cmax[kk] = 0
        jsort = np.argsort(cmax.numpy())[::-1]
        if mode == 'ccg':
            st0 = st[:, 0][clu2 == kk] / ops['fs']
        is_ccg = 0
        for j in range(NN):
            jj = jsort[j]
            if cmax[jj] < r_thresh:
                break
            if mode == 'ccg':
                st1 = st[:, 0][clu2 == jj] / ops['fs']
                _, is_ccg, _ = PH.check_CCG(st0, st1, acg_threshold=acg_threshold,
                                            ccg_threshold=ccg_threshold)
            else:
                dmu = 2 * (mu[kk] - mu[jj]) / (mu[kk] + mu[jj])
                is_ccg = dmu.abs() < 0.2
            if is_ccg:
                is_merged[jj] = 1
                dt = (imax[kk] - imax[jj]).item()
                if dt != 0 and check_dt:
                    idx = (clu2 == jj)
                    tF, _ = roll_features(W, tF, Ww, idx, jj, dt)
                    st[idx, 0] -= dt
                Ww[kk] = ns[kk] / (ns[kk] + ns[jj]) * Ww[kk] + ns[jj] / (ns[kk] + ns[jj]) * Ww[jj]
                Ww[jj] = 0
                ns[kk] += ns[jj]
                ns[jj] = 0
                clu2[clu2 == jj] = kk
                break
        if is_ccg == 0:
            t += 1
        else:
            nmerge += 1
    imap = np.cumsum((~is_merged).astype('int32')) - 1
    if imap.size > 0:
        clu2 = imap[clu2]
    Ww = Ww[~is_merged]
    is_ref = is_ref[~is_merged] if mode == 'ccg' else None
    sorted_idx = np.argsort(st[:, 0])
    st = np.take_along_axis(st, sorted_idx[..., np.newaxis], axis=0)
    clu2 = clu2[sorted_idx]
    tF = tF[torch.from_numpy(sorted_idx)]
    return Ww, clu2, is_ref, st, tF


def roll_features(wPCA, tF, Wall, spike_idx, clust_idx, dt):
    """``template_matching.py:361-379``."""
    W = wPCA.cpu()
    feats = torch.roll(tF[spike_idx] @ W, shifts=dt, dims=2)
    temps = torch.roll(Wall[clust_idx:clust_idx + 1] @ wPCA, shifts=dt, dims=2)
    if dt > 0:
        feats[:, :, :dt] = feats[:, :, dt].unsqueeze(-1)
        temps[:, :, :dt] = temps[:, :, dt].unsqueeze(-1)
    elif dt < 0:
        feats[:, :, dt:] = feats[:, :, dt - 1].unsqueeze(-1)
        temps[:, :, dt:] = temps[:, :, dt - 1].unsqueeze(-1)
    tF[spike_idx] = feats @ W.T
    Wall[clust_idx] = temps @ wPCA.T
    return tF, Wall


def postprocess_templates(Wall, ops, clu, st, tF):
    """``template_matching.py:144-152``."""
    Wall2, _ = align_U(Wall, ops)
    Wall3, _, _, _, _ = merging_function(ops, Wall2.transpose(1, 2), clu, st, tF, 0.9)
    return Wall3.transpose(1, 2)


def merge_step_row(Ww, W, nt, kk):
    """The similarity row of one merging step, ``template_matching.py:285-293``:
    (cmax [K], imax [K]) of template kk against all templates, before
    cmax[kk] is zeroed."""
    WtW = conv1d(W.reshape(-1, 1, nt), W.reshape(-1, 1, nt), padding=nt)
    WtW = torch.flip(WtW, [2, ])
    mu = (Ww ** 2).sum((1, 2), keepdims=True) ** .5
    Wnorm = Ww / (1e-6 + mu)
    UtU = torch.einsum('lk, jlm -> jkm', Wnorm[kk], Wnorm)
    ctc = torch.einsum('jkm, kml -> jl', UtU, WtW)
    return ctc.max(1)
