"""Post-processing for the output files, on the GPU.

Mirrors ``kilosort/postprocessing.py`` with the same signatures and numpy
outputs:

* ``remove_duplicates(spike_times, spike_clusters, dt)`` (``:9-30``): keep a
  spike when it is at least ``dt`` samples after the last kept spike of its
  cluster (``ksm_group`` + ``ksm_remove_duplicates``);
* ``compute_spike_positions(st, tF, ops)`` (``:33-62``): each spike's
  position is the mean of the (x, y) of its template's nearest channels
  weighted by the norms of its PC features there (channels outside
  ``position_limit`` weigh 0) (``ksm_spike_positions``, which also gives the
  Phy amplitudes);
* ``make_pc_features(ops, spike_templates, spike_clusters, tF)``
  (``:65-130``): per cluster, the channels of its templates ranked by the
  norm of the mean feature, and each spike's features re-ordered to the
  top ``nearest_chans`` of them (``ksm_template_sums``,
  ``ksm_pc_channel_rank``, ``ksm_gather_features``).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib

MAX_CLUSTERS = 16384


def _device_of(ops, device):
    if device is None:
        device = ops['iU'].device if torch.is_tensor(ops['iU']) and ops['iU'].is_cuda \
            else torch.device('cuda')
    return _lib.require_gpu(device)


def positions_and_amplitudes(st, tF, ops, device=None):
    """(xs, ys, amplitudes) float32 device tensors from one pass over tF:
    the positions of ``compute_spike_positions`` and the Phy amplitudes
    ``torch.norm(tF, dim=[-2, -1])`` (``io.py:330``)."""
    device = _device_of(ops, device)
    lib = _lib.load()
    ids = torch.as_tensor(np.asarray(st[:, 1]).astype(np.int64)).to(device)
    tFd = torch.as_tensor(tF).to(device, torch.float32).contiguous()
    n, nC, P = tFd.shape
    iU = torch.as_tensor(ops['iU']).to(device, torch.int64).contiguous()
    iCC = torch.as_tensor(ops['iCC']).to(device, torch.int64).contiguous()
    mask = torch.as_tensor(ops['iCC_mask']).to(device, torch.uint8).contiguous()
    xc = torch.as_tensor(np.asarray(ops['xc'], np.float32)).to(device)
    yc = torch.as_tensor(np.asarray(ops['yc'], np.float32)).to(device)
    if iCC.shape != mask.shape or iCC.shape[0] != nC:
        raise ValueError(f"iCC {tuple(iCC.shape)} / iCC_mask {tuple(mask.shape)} do not match "
                         f"{nC} feature channels")
    if n and (int(ids.min()) < 0 or int(ids.max()) >= iU.shape[0]):
        raise ValueError("template ids out of range")
    if iU.numel() and (int(iU.min()) < 0 or int(iU.max()) >= iCC.shape[1]):
        raise ValueError("iU out of range of the channels")
    if xc.numel() < iCC.shape[1] or yc.numel() < iCC.shape[1] or (
            iCC.numel() and (int(iCC.min()) < 0 or int(iCC.max()) >= xc.numel())):
        raise ValueError("iCC out of range of the channel coordinates")
    xs = torch.empty(n, dtype=torch.float32, device=device)
    ys = torch.empty(n, dtype=torch.float32, device=device)
    amp = torch.empty(n, dtype=torch.float32, device=device)
    _lib.check(lib.ksm_spike_positions(tFd.data_ptr(), ids.data_ptr(), n, nC, P, iU.data_ptr(),
                                       iCC.data_ptr(), mask.data_ptr(), iCC.shape[1],
                                       xc.data_ptr(), yc.data_ptr(), xs.data_ptr(),
                                       ys.data_ptr(), amp.data_ptr(), _lib.stream_of(device)),
               "ksm_spike_positions")
    return xs, ys, amp


def compute_spike_positions(st, tF, ops, device=None):
    """``postprocessing.compute_spike_positions``: (xs, ys) float32 numpy
    arrays of the spikes of ``st`` (template ids in ``st[:, 1]``) from
    features ``tF`` [n, nC, P] and ``ops['iCC']``, ``ops['iCC_mask']``,
    ``ops['iU']``, ``ops['xc']``, ``ops['yc']``."""
    xs, ys, _ = positions_and_amplitudes(st, tF, ops, device)
    return xs.cpu().numpy(), ys.cpu().numpy()


def _labels(spike_clusters, device):
    lab = np.asarray(spike_clusters)
    if lab.size and (lab.min() < 0 or lab.max() >= MAX_CLUSTERS):
        raise ValueError(f"cluster ids must lie in [0, {MAX_CLUSTERS})")
    G = int(lab.max()) + 1 if lab.size else 1
    return torch.as_tensor(lab.astype(np.int32)).to(device), G


def group_by_label(lab, G, device):
    """(offsets [G + 1], order [n]) int64 device tensors: the spikes of label g
    are order[offsets[g]:offsets[g + 1]], in input order (``ksm_group``)."""
    lib = _lib.load()
    n = lab.numel()
    nb = lib.ksm_group_workspace(n, G)
    if nb == 0:
        raise ValueError(f"ksm_group: unsupported sizes (n={n}, labels={G})")
    ws = torch.empty(nb, dtype=torch.uint8, device=device)
    off = torch.empty(G + 1, dtype=torch.int64, device=device)
    order = torch.empty(max(n, 1), dtype=torch.int64, device=device)
    _lib.check(lib.ksm_group(lab.data_ptr(), n, G, off.data_ptr(), order.data_ptr(),
                             ws.data_ptr(), nb, _lib.stream_of(device)), "ksm_group")
    return off, order[:n]


def remove_duplicates(spike_times, spike_clusters, dt=15, device=torch.device('cuda')):
    """``postprocessing.remove_duplicates``: (spike_times[keep],
    spike_clusters[keep], keep) numpy."""
    device = _lib.require_gpu(device)
    lib = _lib.load()
    spike_times = np.asarray(spike_times)
    spike_clusters = np.asarray(spike_clusters)
    if spike_times.shape != spike_clusters.shape or spike_times.ndim != 1:
        raise ValueError("spike_times and spike_clusters must be 1-D of one length")
    n = spike_times.size
    lab, G = _labels(spike_clusters, device)
    off, order = group_by_label(lab, G, device)
    times = torch.as_tensor(spike_times.astype(np.int64)).to(device)
    keep = torch.zeros(max(n, 1), dtype=torch.uint8, device=device)
    _lib.check(lib.ksm_remove_duplicates(times.data_ptr(), n, G, off.data_ptr(),
                                         _lib.ptr(order), int(dt), keep.data_ptr(),
                                         _lib.stream_of(device)), "ksm_remove_duplicates")
    keep = keep[:n].cpu().numpy().astype(bool)
    return spike_times[keep], spike_clusters[keep], keep


def make_pc_features(ops, spike_templates, spike_clusters, tF, device=None):
    """``postprocessing.make_pc_features``: (tF permuted to [n, P, nC],
    feature_ind [n_clusters, nearest_chans] uint32).  tF is updated in place
    as in the reference."""
    device = _device_of(ops, device)
    lib = _lib.load()
    spike_templates = np.asarray(spike_templates)
    spike_clusters = np.asarray(spike_clusters)
    iU = torch.as_tensor(ops['iU']).to(device, torch.int64)
    iCC = torch.as_tensor(ops['iCC']).to(device, torch.int64)
    iC = iCC[:, iU].contiguous()                             # [nC, K]: xy_templates' iC
    nC, K = iC.shape
    C = iCC.shape[1]
    n_chans = ops['nearest_chans']
    tFd = torch.as_tensor(tF).to(device, torch.float32).contiguous()
    n, nCf, P = tFd.shape
    if nCf != nC or n_chans != nC:
        raise ValueError(f"tF has {nCf} channel rows, iCC {nC}, nearest_chans {n_chans}")
    if spike_templates.size and (spike_templates.min() < 0 or spike_templates.max() >= K):
        raise ValueError("template ids out of range")
    clusters = np.unique(spike_clusters)
    if clusters.size and clusters.max() >= clusters.size:
        # the reference writes feature_ind[i] into n_clusters rows
        raise IndexError(f"index {clusters.max()} is out of bounds for axis 0 with size "
                         f"{clusters.size}")
    lab, G = _labels(spike_clusters, device)
    tmpl = torch.as_tensor(spike_templates.astype(np.int64)).to(device)
    sums = torch.empty((K, nC, P), dtype=torch.float64, device=device)
    counts = torch.empty(K, dtype=torch.int64, device=device)
    presence = torch.empty((G, K), dtype=torch.uint8, device=device)
    st = _lib.stream_of(device)
    nws = lib.ksm_template_sums_workspace(n, K)
    ws = torch.empty(max(nws, 1), dtype=torch.uint8, device=device)
    _lib.check(lib.ksm_template_sums(tFd.data_ptr(), tmpl.data_ptr(), lab.data_ptr(), n, nC, P,
                                     K, G, sums.data_ptr(), counts.data_ptr(),
                                     presence.data_ptr(), ws.data_ptr(), nws, st),
               "ksm_template_sums")
    sigma = torch.empty((K, nC), dtype=torch.int32, device=device)
    sig_tmp = torch.empty_like(sigma)
    fi = torch.zeros((max(clusters.size, 1), nC), dtype=torch.int32, device=device)
    if n:
        _lib.check(lib.ksm_pc_channel_rank(sums.data_ptr(), counts.data_ptr(),
                                           presence.data_ptr(), iC.data_ptr(), K, nC, P, C, G,
                                           sigma.data_ptr(), sig_tmp.data_ptr(), fi.data_ptr(),
                                           st), "ksm_pc_channel_rank")
    out = torch.empty_like(tFd)
    _lib.check(lib.ksm_gather_features(tFd.data_ptr(), tmpl.data_ptr(), n, nC, P,
                                       sigma.data_ptr(), out.data_ptr(), 0, st),
               "ksm_gather_features")
    feature_ind = fi[:clusters.size].cpu().numpy().view(np.uint32)
    if torch.is_tensor(tF):
        tF.copy_(out)
        res = tF
    else:
        res = out.cpu()
    return torch.permute(res, (0, 2, 1)), feature_ind
