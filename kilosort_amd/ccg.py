"""Refractoriness of clusters from their autocorrelograms, on the GPU.

Mirrors ``kilosort/CCG.py``'s ``refract(iclust2, st0, acg_threshold,
ccg_threshold)`` (``:105-120``) with the same outputs: the autocorrelogram
counts of every cluster (``compute_CCG``, ``:8-36``) come from ``ksm_acg``
in one launch over all clusters (exact integer counts); the metrics of
``CCG_metrics`` (``:38-84``) and the thresholds of ``check_CCG``
(``:86-101``) are evaluated on the host from those counts with the
reference's float64 formulas, per cluster.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import _lib
from .postprocessing import MAX_CLUSTERS, group_by_label


def acg_counts(iclust2, st0, nbins=500, tbin=1 / 1000, active=None, min_spikes=None,
               device=None):
    """(K [Nfilt, 2 nbins + 1] float64 numpy, n [Nfilt], T [Nfilt], active
    [Nfilt]): the autocorrelogram counts of ``compute_CCG(st1, st1)`` for every
    active cluster (other rows 0), the spike counts and the spans max - min of
    each cluster's times.  Active: ``active`` if given, else clusters with
    more than ``min_spikes`` spikes and a nonzero span if that is given, else
    all."""
    device = _lib.require_gpu(device if device is not None else torch.device('cuda'))
    lib = _lib.load()
    iclust2 = np.asarray(iclust2)
    st0 = np.asarray(st0, dtype=np.float64)
    if iclust2.shape != st0.shape or st0.ndim != 1:
        raise ValueError("iclust2 and st0 must be 1-D of one length")
    if iclust2.size and (iclust2.min() < 0 or iclust2.max() >= MAX_CLUSTERS):
        raise ValueError(f"cluster ids must lie in [0, {MAX_CLUSTERS})")
    G = int(iclust2.max()) + 1 if iclust2.size else 1
    if st0.size > 1 and np.any(st0[1:] < st0[:-1]):
        perm = np.argsort(st0, kind='stable')           # the counts ignore spike order
        st0, iclust2 = st0[perm], iclust2[perm]
    lab = torch.as_tensor(iclust2.astype(np.int32)).to(device)
    off, order = group_by_label(lab, G, device)
    offs = off.cpu().numpy()
    n = np.diff(offs)
    T = np.zeros(G)
    nz = n > 0
    if nz.any():
        o = order.cpu().numpy()
        T[nz] = st0[o[offs[1:][nz] - 1]] - st0[o[offs[:-1][nz]]]
    if active is None:
        active = (n > min_spikes) & (T != 0) if min_spikes is not None else np.ones(G, bool)
    act = torch.as_tensor(np.asarray(active, np.uint8)).to(device)
    times = torch.as_tensor(st0).to(device)
    L = 2 * nbins + 1
    counts = torch.empty((G, L), dtype=torch.int32, device=device)
    _lib.check(lib.ksm_acg(times.data_ptr(), st0.size, G, off.data_ptr(), _lib.ptr(order),
                           act.data_ptr(), int(nbins), float(tbin), counts.data_ptr(),
                           _lib.stream_of(device)), "ksm_acg")
    K = counts.cpu().numpy().view(np.uint32).astype(np.float64)
    return K, n, T, np.asarray(active, bool)


def ccg_metrics_rows(n1, n2, K, T, nbins, tbin):
    """``CCG.CCG_metrics`` for many correlograms at once (rows of K, with
    per-row train sizes n1, n2 and spans T): (R12, Q12, Q00) arrays.  The
    counts are integers held in float64, so the row sums are exact in any
    order and every value equals the per-pair formula's."""
    K = K.copy()
    n1 = np.asarray(n1, np.float64)
    n2 = np.asarray(n2, np.float64)
    T = np.asarray(T, np.float64)
    irange1 = np.hstack((np.arange(1, nbins // 2), np.arange(3 * nbins // 2, 2 * nbins)))
    irange2 = np.arange(nbins - 50, nbins - 10)
    irange3 = np.arange(nbins + 10, nbins + 50)
    R00 = K[:, irange1].sum(1) / (len(irange1) * tbin * n1 * n2 / T)
    R1 = K[:, irange2].sum(1) / (len(irange2) * tbin * n1 * n2 / T)
    R2 = K[:, irange3].sum(1) / (len(irange3) * tbin * n1 * n2 / T)
    R01 = np.maximum(R1, R2)
    Q00 = np.maximum(K[:, irange2].sum(1) / len(irange2), K[:, irange3].sum(1) / len(irange3))
    Q00 = np.maximum(Q00, K[:, irange1].sum(1) / len(irange1))
    K[:, nbins] = 0
    Ri = np.zeros((len(K), 10))
    Qi = np.zeros((len(K), 10))
    for i in range(1, 11):
        s = K[:, nbins - i:nbins + i + 1].sum(1)
        Ri[:, i - 1] = s / (2 * i * tbin * n1 * n2 / T)
        m = s / 2
        lam = Q00 * i
        z = (m - lam) / (1e-10 + 2 * lam) ** .5
        Qi[:, i - 1] = [1 / 2 * (1 + math.erf(v)) for v in z]
    R12 = Ri.min(1) / (1e-10 + np.maximum(R00, R01))
    Q12 = Qi.min(1)
    return R12, Q12, Q00


def ccg_metrics(n1, n2, K, T, nbins, tbin):
    """``CCG.CCG_metrics`` from the counts K of two trains of n1 and n2
    spikes spanning T seconds: (R12, Q12, Q00)."""
    K = K.copy()
    irange1 = np.hstack((np.arange(1, nbins // 2), np.arange(3 * nbins // 2, 2 * nbins)))
    irange2 = np.arange(nbins - 50, nbins - 10)
    irange3 = np.arange(nbins + 10, nbins + 50)
    R00 = K[irange1].sum() / (len(irange1) * tbin * n1 * n2 / T)
    R1 = K[irange2].sum() / (len(irange2) * tbin * n1 * n2 / T)
    R2 = K[irange3].sum() / (len(irange3) * tbin * n1 * n2 / T)
    R01 = np.maximum(R1, R2)
    Q00 = np.maximum(K[irange2].mean(), K[irange3].mean())
    Q00 = np.maximum(Q00, K[irange1].mean())
    K[nbins] = 0
    Ri = np.zeros(10,)
    Qi = np.zeros(10,)
    for i in range(1, 11):
        irange = np.arange(nbins - i, nbins + i + 1)
        Ri[i - 1] = K[irange].sum() / (2 * i * tbin * n1 * n2 / T)
        m = K[irange].sum() / 2
        lam = Q00 * i
        Qi[i - 1] = 1 / 2 * (1 + math.erf((m - lam) / (1e-10 + 2 * lam) ** .5))
    R12 = np.min(Ri) / (1e-10 + np.maximum(R00, R01))
    Q12 = np.min(Qi)
    return R12, Q12, Q00


def refract(iclust2, st0, acg_threshold=0.2, ccg_threshold=0.25, device=None):
    """``CCG.refract``: (is_refractory [Nfilt] float64 0/1, R12 [Nfilt])
    for cluster ids ``iclust2`` and spike times ``st0`` in seconds; clusters
    with at most 10 spikes or a zero time span stay 0."""
    nbins, tbin = 500, 1 / 1000
    iclust2 = np.asarray(iclust2)
    Nfilt = int(iclust2.max()) + 1
    K, n, T, active = acg_counts(iclust2, st0, nbins, tbin, min_spikes=10, device=device)
    is_refractory = np.zeros(Nfilt, )
    R12 = np.zeros(Nfilt, )
    a = np.nonzero(active)[0]
    if a.size:
        r12, q12, _ = ccg_metrics_rows(n[a], n[a], K[a], T[a], nbins, tbin)
        is_refractory[a] = (r12 < acg_threshold) & (q12 < .2)
        R12[a] = r12
    return is_refractory, R12
