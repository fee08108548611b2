"""CPU restatement of Kilosort4's universal-template spike detection (test
oracle).

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` and ``bench.py``'s
``cpu_baseline`` leg; never part of the product path.

Follows ``kilosort/spikedetect.py`` with the same torch / numpy operations in
the same order:

* ``extract_snippets`` (``spikedetect.py:20-51``);
* ``template_centers`` (``spikedetect.py:101-130``), ``nearest_chans``
  (``:181-187``), the template weights of ``run`` (``:217-242``);
* ``template_match`` (``:133-178``);
* ``yweighted`` (``:190-197``) and the per-batch features and spike table of
  ``run`` (``:256-281``).
"""
from __future__ import annotations

import numpy as np
import torch
from torch.nn.functional import avg_pool2d, conv1d, max_pool1d, max_pool2d


def my_max2d(X, dt):
    """``spikedetect.py:20-25``."""
    Xmax = max_pool2d(X.unsqueeze(0), [2 * dt[0] + 1, 2 * dt[1] + 1], stride=[1, 1],
                      padding=[dt[0], dt[1]])
    return Xmax[0]


def my_sum2d(X, dt):
    """``spikedetect.py:27-32``."""
    Xsum = avg_pool2d(X.unsqueeze(0), [2 * dt[0] + 1, 2 * dt[1] + 1], stride=[1, 1],
                      padding=[dt[0], dt[1]])
    Xsum *= (2 * dt[0] + 1) * (2 * dt[1] + 1)
    return Xsum[0]


def extract_snippets(X, nt, twav_min, Th_single_ch, loc_range=(4, 5), long_range=(6, 30)):
    """``spikedetect.py:34-51``."""
    Xabs = X.abs()
    Xmax = my_max2d(Xabs, loc_range)
    ispeak = torch.logical_and(Xmax == Xabs, Xabs > Th_single_ch).float()
    ispeak_sum = my_sum2d(ispeak, long_range)
    is_peak_iso = ((ispeak_sum == 1) * (ispeak == 1))
    is_peak_iso[:, :nt] = 0
    is_peak_iso[:, -nt:] = 0
    xy = is_peak_iso.nonzero()
    clips = X[xy[:, :1], xy[:, 1:2] - twav_min + torch.arange(nt)]
    return clips


def template_centers(ops):
    """``spikedetect.py:101-130``."""
    shank_idx = ops['kcoords']
    xc = ops['xc']
    yc = ops['yc']
    dmin = ops['settings']['dmin']
    if dmin is None:
        y_uniq = np.unique(yc)
        if y_uniq.size == 1:
            dmin = 1
        else:
            dmin = np.median(np.diff(np.unique(y_uniq)))
    ops['dmin'] = dmin
    ops['dminx'] = dminx = ops['settings']['dminx']
    yup = np.array([])
    xup = np.array([])
    for i in np.unique(shank_idx):
        xc_i = xc[shank_idx == i]
        yc_i = yc[shank_idx == i]
        xmin, xmax, ymin, ymax = xc_i.min(), xc_i.max(), yc_i.min(), yc_i.max()
        yup = np.concatenate([yup, np.arange(ymin, ymax + .01, dmin / 2)])
        nx = np.round((xmax - xmin) / (dminx / 2)) + 1
        xup = np.concatenate([xup, np.linspace(xmin, xmax, int(nx))])
    ops['yup'] = np.unique(yup)
    ops['xup'] = np.unique(xup)
    return ops


def nearest_chans(ys, yc, xs, xc, nC):
    """``spikedetect.py:181-187``."""
    ds = (ys - yc[:, np.newaxis]) ** 2 + (xs - xc[:, np.newaxis]) ** 2
    iC = np.argsort(ds, 0)[:nC]
    iC = torch.from_numpy(iC)
    ds = np.sort(ds, 0)[:nC]
    return iC, ds


def setup(ops):
    """The template grid, nearest channels / templates and size weights of
    ``run`` (``spikedetect.py:217-242``); returns (iC, iC2, weigh) and sets
    ops['ycup'], ops['xcup']."""
    ops = template_centers(ops)
    ys, xs = np.meshgrid(ops['yup'], ops['xup'])
    ys, xs = ys.flatten(), xs.flatten()
    xc, yc = ops['xc'], ops['yc']
    nC = ops['settings']['nearest_chans']
    nC2 = ops['settings']['nearest_templates']
    iC, ds = nearest_chans(ys, yc, xs, xc, nC)
    igood = ds[0, :] <= ops['max_channel_distance'] ** 2
    iC = iC[:, igood]
    ds = ds[:, igood]
    ys = ys[igood]
    xs = xs[igood]
    ops['ycup'], ops['xcup'] = ys, xs
    iC2, _ = nearest_chans(ys, ys, xs, xs, nC2)
    sig = ops['settings']['min_template_size']
    nsizes = ops['settings']['template_sizes']
    ds_torch = torch.from_numpy(ds).float()
    template_sizes = sig * (1 + torch.arange(nsizes))
    weigh = torch.exp(-ds_torch.unsqueeze(-1) / template_sizes ** 2)
    weigh = torch.permute(weigh, (2, 0, 1)).contiguous()
    weigh = weigh / (weigh ** 2).sum(1).unsqueeze(1) ** .5
    return iC, iC2, weigh


def template_match(X, ops, iC, iC2, weigh):
    """``spikedetect.py:133-178`` on the CPU."""
    nt = ops['nt']
    nt0 = ops['settings']['nt0min']
    nk = ops['settings']['n_templates']
    NT = X.shape[-1]
    Nfilt = iC.shape[1]
    niter = 40
    nb = (NT - 1) // niter + 1
    W = ops['wTEMP'].unsqueeze(1)
    B = conv1d(X.unsqueeze(1), W, padding=nt // 2)
    As = torch.zeros((Nfilt, NT), dtype=X.dtype)
    Amaxs = torch.zeros((Nfilt, NT), dtype=X.dtype)
    imaxs = torch.zeros((Nfilt, NT), dtype=torch.int64)
    ti = torch.arange(Nfilt)
    tj = torch.arange(nb)
    for t in range(niter):
        A = torch.einsum('ijk, jklm-> iklm', weigh, B[iC, :, nb * t:nb * (t + 1)])
        A = A.transpose(1, 2)
        A = A.reshape(-1, Nfilt, A.shape[-1])
        Aa, imax = torch.max(A.abs(), 0)
        imax = (1 + imax) * A[imax, ti.unsqueeze(-1), tj[:A.shape[-1]]].sign()
        As[:, nb * t:nb * (t + 1)] = Aa
        imaxs[:, nb * t:nb * (t + 1)] = imax
        Amax = torch.max(Aa[iC2], 0)[0]
        Amaxs[:, nb * t:nb * (t + 1)] = Amax
    Amaxs[:, :nt] = 0
    Amaxs[:, -nt:] = 0
    Amaxs = max_pool1d(Amaxs.unsqueeze(0), (2 * nt0 + 1), stride=1, padding=nt0).squeeze(0)
    xy = torch.logical_and(Amaxs == As, As > ops['Th_universal']).nonzero()
    imax = imaxs[xy[:, 0], xy[:, 1]]
    amp = As[xy[:, 0], xy[:, 1]]
    ssign = imax.sign()
    imax = imax.abs() - 1
    adist = B[iC[:, xy[:, 0]], imax % nk, xy[:, 1]] * ssign
    return xy, imax, amp, adist


def yweighted(yc, iC, adist, xy):
    """``spikedetect.py:190-197``."""
    yy = torch.from_numpy(yc)[iC]
    cF0 = torch.nn.functional.relu(adist)
    cF0 = cF0 / cF0.sum(0)
    yct = (cF0 * yy[:, xy[:, 0]]).sum(0)
    return yct


def detect_batch(X, ops, iC, iC2, weigh, ibatch=0, batch_downsampling=1):
    """One batch of ``run`` (``spikedetect.py:256-281``): the spike rows
    (time in s, y, amp, template, batch, position) and the features."""
    nt = ops['nt']
    xy, imax, amp, adist = template_match(X, ops, iC, iC2, weigh)
    yct = yweighted(ops['yc'], iC, adist, xy)
    tarange = torch.arange(-(nt // 2), nt // 2 + 1)
    xsub = X[iC[:, xy[:, :1]], xy[:, 1:2] + tarange]
    xfeat = xsub @ ops['wPCA'].T
    tF = xfeat.transpose(0, 1).numpy()
    nsp = len(xy)
    st = np.zeros((nsp, 6), 'float64')
    t_shift = ibatch * batch_downsampling * (ops['batch_size'] / ops['fs'])
    st[:, 0] = ((xy[:, 1].numpy() - nt) / ops['fs'] + t_shift)
    st[:, 1] = yct.numpy()
    st[:, 2] = amp.numpy()
    st[:, 3] = imax.numpy()
    st[:, 4] = ibatch
    st[:, 5] = xy[:, 0].numpy()
    return st, tF, (xy, imax, amp, adist)
