"""CPU restatement of the per-spike parts of Kilosort4's Phy export (test
oracle).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``; never part of the product
path.  Follows, with the same numpy / torch operations in the same order
(the numba functions as plain Python loops):

* ``remove_duplicates`` -- ``kilosort/postprocessing.py:9-30``;
* ``compute_CCG`` -- ``kilosort/CCG.py:8-36``; ``CCG_metrics`` -- ``:38-84``;
  ``check_CCG`` -- ``:86-101``; ``refract`` -- ``:105-120``;
* ``xy_templates`` -- ``kilosort/clustering_qr.py:318-326``; ``get_data_cpu``
  -- ``:560-600``; ``make_pc_features`` -- ``kilosort/postprocessing.py:65-130``.
"""
from __future__ import annotations

import math

import numpy as np
import torch


def remove_duplicates(spike_times, spike_clusters, dt=15):
    """``postprocessing.py:9-30``."""
    keep = np.zeros_like(spike_times, bool)
    cluster_t0 = {}
    for i in range(spike_times.size):
        t = spike_times[i]
        c = spike_clusters[i]
        t0 = cluster_t0[c] if c in cluster_t0 else t - dt
        if t >= (t0 + dt):
            cluster_t0[c] = t
            keep[i] = True
    return spike_times[keep], spike_clusters[keep], keep


def compute_CCG(st1, st2, tbin=1 / 1000, nbins=500):
    """``CCG.py:8-36``."""
    st1 = np.sort(st1)
    st2 = np.sort(st2)
    dt = nbins * tbin
    T = np.maximum(st1.max(), st2.max()) - np.minimum(st1.min(), st2.min())
    ilow = 0
    ihigh = 0
    j = 0
    K = np.zeros(2 * nbins + 1,)
    while j < len(st2):
        while (ihigh < len(st1)) and (st1[ihigh] <= st2[j] + dt):
            ihigh += 1
        while (ilow < len(st1)) and (st1[ilow] <= st2[j] - dt):
            ilow += 1
        if ilow >= len(st1):
            break
        if st1[ilow] > st2[j] + dt:
            j += 1
            continue
        for k in range(ilow, ihigh):
            ibin = int(np.round((st2[j] - st1[k]) / tbin))
            K[ibin + nbins] += 1
        j += 1
    return K, T


def CCG_metrics(st1, st2, K, T, nbins=None, tbin=None):
    """``CCG.py:38-84``."""
    irange1 = np.hstack((np.arange(1, nbins // 2), np.arange(3 * nbins // 2, 2 * nbins)))
    irange2 = np.arange(nbins - 50, nbins - 10)
    irange3 = np.arange(nbins + 10, nbins + 50)
    R00 = K[irange1].sum() / (len(irange1) * tbin * len(st1) * len(st2) / T)
    R1 = K[irange2].sum() / (len(irange2) * tbin * len(st1) * len(st2) / T)
    R2 = K[irange3].sum() / (len(irange3) * tbin * len(st1) * len(st2) / T)
    R01 = np.maximum(R1, R2)
    Q00 = np.maximum(K[irange2].mean(), K[irange3].mean())
    Q00 = np.maximum(Q00, K[irange1].mean())
    a = K[nbins]
    K[nbins] = 0
    Ri = np.zeros(10,)
    Qi = np.zeros(10,)
    for i in range(1, 11):
        irange = np.arange(nbins - i, nbins + i + 1)
        Ri0 = K[irange].sum() / (2 * i * tbin * len(st1) * len(st2) / T)
        Ri[i - 1] = Ri0
        n = K[irange].sum() / 2
        lam = Q00 * i
        p = 1 / 2 * (1 + math.erf((n - lam) / (1e-10 + 2 * lam) ** .5))
        Qi[i - 1] = p
    K[nbins] = a
    R12 = np.min(Ri) / (1e-10 + np.maximum(R00, R01))
    Q12 = np.min(Qi)
    return R12, Q12, Q00


def check_CCG(st1, st2=None, nbins=500, tbin=1 / 1000, acg_threshold=0.2, ccg_threshold=0.25):
    """``CCG.py:86-101``."""
    if st2 is None:
        st2 = st1.copy()
    K, T = compute_CCG(st1, st2, nbins=nbins, tbin=tbin)
    R12, Q12, Q00 = CCG_metrics(st1, st2, K, T, nbins=nbins, tbin=tbin)
    is_refractory = R12 < acg_threshold and (Q12 < .2)
    cross_refractory = R12 < ccg_threshold and (Q12 < .05)
    return is_refractory, cross_refractory, R12


def refract(iclust2, st0, acg_threshold=0.2, ccg_threshold=0.25):
    """``CCG.py:105-120``."""
    Nfilt = iclust2.max() + 1
    is_refractory = np.zeros(Nfilt, )
    cross_refractory = np.zeros(Nfilt, )
    R12 = np.zeros(Nfilt, )
    for kk in range(Nfilt):
        ix = iclust2 == kk
        st1 = st0[ix]
        if (len(st1) > 10) and ((st1.max() - st1.min()) != 0):
            is_refractory[kk], cross_refractory[kk], R12[kk] = check_CCG(
                st1, acg_threshold=acg_threshold, ccg_threshold=ccg_threshold)
    return is_refractory, R12


def xy_templates(ops):
    """``clustering_qr.py:318-326``."""
    iU = ops['iU'].cpu().numpy()
    iC = ops['iCC'][:, ops['iU']]
    xcup, ycup = ops['xc'][iU], ops['yc'][iU]
    xy = np.vstack((xcup, ycup))
    xy = torch.from_numpy(xy)
    return xy, iC


def get_data_cpu(ops, xy, iC, PID, tF, ycenter, xcenter, dmin=20, dminx=32, ix=None,
                 merge_dim=True):
    """``clustering_qr.py:560-600``."""
    PID = torch.from_numpy(PID).long()
    y0 = ycenter
    x0 = xcenter
    if ix is None:
        ix = torch.logical_and(torch.abs(xy[1] - y0) < dmin, torch.abs(xy[0] - x0) < dminx)
    igood = ix[PID].nonzero()[:, 0]
    if len(igood) == 0:
        return None, None, None
    pid = PID[igood]
    data = tF[igood]
    nspikes, nchanraw, nfeatures = data.shape
    ichan, imap = torch.unique(iC[:, ix], return_inverse=True)
    nchan = ichan.nelement()
    dd = torch.zeros((nspikes, nchan, nfeatures))
    for k, j in enumerate(ix.nonzero()[:, 0]):
        ij = torch.nonzero(pid == j)[:, 0]
        dd[ij.unsqueeze(-1), imap[:, k]] = data[ij]
    if merge_dim:
        Xd = torch.reshape(dd, (nspikes, -1))
    else:
        Xd = dd
    return Xd, igood, ichan


def make_pc_features(ops, spike_templates, spike_clusters, tF, with_norms=False):
    """``postprocessing.py:65-130`` (updates tF in place, as the reference).
    with_norms: also return the ranked channel norms [n_clusters, n_chans]
    (``sorted_chans``), so that a check can tell exact ties (norm 0, whose
    order is torch's unstable sort's) from ranked channels."""
    xy, iC = xy_templates(ops)
    n_templates = iC.shape[1]
    n_clusters = np.unique(spike_clusters).size
    n_chans = ops['nearest_chans']
    feature_ind = np.zeros((n_clusters, n_chans), dtype=np.uint32)
    norms = np.zeros((n_clusters, n_chans), dtype=np.float32)
    for i in np.unique(spike_clusters):
        iunq = np.unique(spike_templates[spike_clusters == i]).astype(int)
        ix = torch.from_numpy(np.zeros(n_templates, bool))
        ix[iunq] = True
        Xd, igood, ichan = get_data_cpu(ops, xy, iC, spike_templates, tF, None, None,
                                        dmin=ops['dmin'], dminx=ops['dminx'], ix=ix,
                                        merge_dim=False)
        spike_mean = Xd.mean(0)
        chan_norm = torch.linalg.norm(spike_mean, dim=1)
        sorted_chans, ind = torch.sort(chan_norm, descending=True)
        tF[igood, :] = Xd[:, ind[:n_chans], :]
        feature_ind[i, :] = ichan[ind[:n_chans]].cpu().numpy()
        norms[i, :] = sorted_chans[:n_chans].numpy()
    tF = torch.permute(tF, (0, 2, 1))
    if with_norms:
        return tF, feature_ind, norms
    return tF, feature_ind


def save_to_phy_arrays(st, clu, tF, Wall, ops, imin, similarity, positions):
    """The arrays that ``io.save_to_phy`` (``io.py:248-469``) writes, in a dict
    keyed by file name, computed with the oracle functions in the reference's
    order (``similarity`` and ``positions`` are the template and position
    oracles' functions)."""
    out = {}
    spike_times = st[:, 0].astype('int64') + imin
    spike_templates = st[:, 1].astype('int32')
    spike_clusters = clu
    xs, ys = positions(st, tF, ops)
    spike_positions = np.vstack([xs, ys]).T
    amplitudes = torch.norm(tF, dim=[-2, -1]).cpu().numpy()
    spike_times, spike_clusters, kept_spikes = remove_duplicates(
        spike_times, spike_clusters, dt=ops['duplicate_spike_bins'])
    out['spike_times.npy'] = spike_times
    out['spike_templates.npy'] = spike_clusters
    out['spike_clusters.npy'] = spike_clusters
    out['spike_positions.npy'] = spike_positions[kept_spikes]
    out['spike_detection_templates.npy'] = spike_templates[kept_spikes]
    out['amplitudes.npy'] = amplitudes[kept_spikes]
    out['kept_spikes.npy'] = kept_spikes
    out['similar_templates.npy'] = similarity(Wall, ops['wPCA'].contiguous(), nt=ops['nt'])
    template_amplitudes = torch.norm(Wall, dim=[-2, -1]).cpu().numpy()
    templates = (Wall.unsqueeze(-1).cpu() * ops['wPCA'].cpu()).sum(axis=-2).numpy()
    out['templates.npy'] = templates.transpose(0, 2, 1)
    out['templates_ind.npy'] = np.tile(np.arange(Wall.shape[1])[np.newaxis, :],
                                       (templates.shape[0], 1))
    tF = tF[kept_spikes]
    pc_features, pc_feature_ind, norms = make_pc_features(
        ops, spike_templates[kept_spikes], spike_clusters, tF, with_norms=True)
    out['pc_features.npy'] = pc_features.numpy()
    out['pc_feature_ind.npy'] = pc_feature_ind
    out['_pc_norms'] = norms
    is_ref, est_contam_rate = refract(spike_clusters, spike_times / ops['fs'],
                                      acg_threshold=ops['settings']['acg_threshold'],
                                      ccg_threshold=ops['settings']['ccg_threshold'])
    out['_is_ref'] = is_ref
    out['_contam'] = est_contam_rate
    out['_template_amplitudes'] = template_amplitudes
    return out
