"""Universal-template spike detection on the GPU (Kilosort4's first pass).

Mirrors ``kilosort/spikedetect.py``: ``template_centers``, ``nearest_chans``,
``template_match(X, ops, iC, iC2, weigh, device)`` with the reference's
outputs ``(xy, imax, amp, adist)``, and ``run(ops, bfile, device,
progress_bar, clear_cache, verbose)`` returning ``(st, tF, ops)`` with the
reference's 6-column spike table (time in s, y position, amplitude,
template, batch, template position) and features.

Every per-batch step runs in the HIP kernels of ``ksm_detect``
(``libksm.so``): the temporal correlation with the universal waveforms, the
size-weighted channel sums and their maxima, the spatio-temporal
neighbourhood maxima, the ordered spike list, the channel-weighted y
positions and the PC features.  The set-up of ``run`` -- the template grid,
nearest channels and templates, the size weights -- is host numpy as in the
reference.  The universal waveforms and PCs come from the data as in the
reference (``extract_wPCA_wTEMP``: isolated-peak snippets found on the GPU by
``ksm_snippets``, then scikit-learn's TruncatedSVD and KMeans on the host),
or from ``ops['wTEMP']`` / ``ops['wPCA']`` when ``templates_from_data`` is
off (the reference then loads its built-in file).
"""
from __future__ import annotations

import logging

import numpy as np
import torch

from . import _lib

logger = logging.getLogger(__name__)


def template_centers(ops):
    """Grid of universal-template positions per shank
    (``spikedetect.py:101-130``): y every dmin/2 from the lowest to the
    highest site, x on a linspace with step about dminx/2."""
    shank_idx = ops['kcoords']
    xc, yc = ops['xc'], ops['yc']
    dmin = ops['settings']['dmin']
    if dmin is None:
        y_uniq = np.unique(yc)
        dmin = 1 if y_uniq.size == 1 else np.median(np.diff(np.unique(y_uniq)))
    ops['dmin'] = dmin
    ops['dminx'] = dminx = ops['settings']['dminx']
    yup, xup = np.array([]), np.array([])
    for i in np.unique(shank_idx):
        xs_i, ys_i = xc[shank_idx == i], yc[shank_idx == i]
        yup = np.concatenate([yup, np.arange(ys_i.min(), ys_i.max() + .01, dmin / 2)])
        nx = np.round((xs_i.max() - xs_i.min()) / (dminx / 2)) + 1
        xup = np.concatenate([xup, np.linspace(xs_i.min(), xs_i.max(), int(nx))])
    ops['yup'] = np.unique(yup)
    ops['xup'] = np.unique(xup)
    return ops


def nearest_chans(ys, yc, xs, xc, nC, device=torch.device('cuda')):
    """The nC nearest sites of each position by squared distance, with the
    reference's ``np.argsort`` (``spikedetect.py:181-187``)."""
    ds = (ys - yc[:, np.newaxis]) ** 2 + (xs - xc[:, np.newaxis]) ** 2
    iC = torch.from_numpy(np.argsort(ds, 0)[:nC]).to(device)
    return iC, np.sort(ds, 0)[:nC]


def prepare(ops, device=torch.device('cuda')):
    """The set-up of ``run`` (``spikedetect.py:217-242``): template grid,
    nearest channels ``iC`` and templates ``iC2``, size weights ``weigh``."""
    ops = template_centers(ops)
    ys, xs = np.meshgrid(ops['yup'], ops['xup'])
    ys, xs = ys.flatten(), xs.flatten()
    xc, yc = ops['xc'], ops['yc']
    iC, ds = nearest_chans(ys, yc, xs, xc, ops['settings']['nearest_chans'], device=device)
    igood = ds[0, :] <= ops['max_channel_distance'] ** 2
    iC = iC[:, torch.from_numpy(igood).to(device)]
    ds = ds[:, igood]
    ys, xs = ys[igood], xs[igood]
    ops['ycup'], ops['xcup'] = ys, xs
    iC2, _ = nearest_chans(ys, ys, xs, xs, ops['settings']['nearest_templates'], device=device)
    sig = ops['settings']['min_template_size']
    nsizes = ops['settings']['template_sizes']
    ds_t = torch.from_numpy(ds).float()
    sizes = sig * (1 + torch.arange(nsizes))
    weigh = torch.exp(-ds_t.unsqueeze(-1) / sizes ** 2)
    weigh = torch.permute(weigh, (2, 0, 1)).contiguous()
    weigh = weigh / (weigh ** 2).sum(1).unsqueeze(1) ** .5
    return iC, iC2, weigh.to(device)


class _Detector:
    """Device copies of the per-recording tables and a workspace."""

    def __init__(self, ops, iC, iC2, weigh, device):
        self.device = device
        self.iC = iC.to(device, torch.int32).contiguous()
        self.iC2 = iC2.to(device, torch.int32).contiguous()
        self.weigh = weigh.to(device, torch.float32).contiguous()
        self.wTEMP = ops['wTEMP'].to(device, torch.float32).contiguous()
        wPCA = ops.get('wPCA')
        self.wPCA = (wPCA if wPCA is not None else self.wTEMP).to(device, torch.float32)
        self.wPCA = self.wPCA.contiguous()
        self.yc = torch.as_tensor(np.asarray(ops['yc'], np.float32)).to(device)
        self.nt = int(ops['nt'])
        self.nt0 = int(ops['settings']['nt0min'])
        self.th = float(ops['Th_universal'])
        self.cap = 4096
        self._ws = None

    def run(self, X):
        lib = _lib.load()
        X = X.to(self.device, torch.float32)
        if X.stride(1) != 1:
            X = X.contiguous()
        C, T = X.shape
        nk = self.wTEMP.shape[0]
        nC, Nfilt = self.iC.shape
        nsizes = self.weigh.shape[0]
        P = self.wPCA.shape[0]
        nb = lib.ksm_detect_workspace(C, T, nk, Nfilt)
        if self._ws is None or self._ws.numel() < nb:
            self._ws = torch.empty(nb, dtype=torch.uint8, device=self.device)
        count = torch.zeros(1, dtype=torch.int32, device=self.device)
        s = _lib.stream_of(self.device)
        while True:
            cap = self.cap
            xy = torch.empty((cap, 2), dtype=torch.int32, device=self.device)
            imax = torch.empty(cap, dtype=torch.int32, device=self.device)
            amp = torch.empty(cap, dtype=torch.float32, device=self.device)
            adist = torch.empty((nC, cap), dtype=torch.float32, device=self.device)
            yct = torch.empty(cap, dtype=torch.float32, device=self.device)
            tF = torch.empty((cap, nC, P), dtype=torch.float32, device=self.device)
            _lib.check(lib.ksm_detect(
                X.data_ptr(), X.stride(0), C, T, self.wTEMP.data_ptr(), nk, self.nt,
                self.iC.data_ptr(), nC, self.iC2.data_ptr(), self.iC2.shape[0],
                self.weigh.data_ptr(), nsizes, Nfilt, self.nt0, self.th, self.yc.data_ptr(),
                self.wPCA.data_ptr(), P, cap, xy.data_ptr(), imax.data_ptr(), amp.data_ptr(),
                adist.data_ptr(), yct.data_ptr(), tF.data_ptr(), count.data_ptr(),
                self._ws.data_ptr(), nb, s), "ksm_detect")
            n = int(count.item())
            if n <= cap:
                break
            self.cap = 1 << int(np.ceil(np.log2(n)))      # run again with room for all
        return (xy[:n], imax[:n], amp[:n], adist[:, :n], yct[:n], tF[:n])


def extract_snippets(X, nt, twav_min, Th_single_ch, loc_range=(4, 5), long_range=(6, 30),
                     device=torch.device('cuda')):
    """``spikedetect.extract_snippets`` (``spikedetect.py:34-51``) on the GPU:
    the [n, nt] waveforms around isolated single-channel peaks of X, in
    (channel, time) order (``ksm_snippets``)."""
    device = _lib.require_gpu(device)
    lib = _lib.load()
    X = X.to(device, torch.float32)
    if X.stride(1) != 1:
        X = X.contiguous()
    C, T = X.shape
    nb = lib.ksm_snippets_workspace(C, T)
    ws = torch.empty(nb, dtype=torch.uint8, device=device)
    count = torch.zeros(1, dtype=torch.int32, device=device)
    cap = 1024
    while True:
        clips = torch.empty((cap, nt), dtype=torch.float32, device=device)
        xy = torch.empty((cap, 2), dtype=torch.int32, device=device)
        _lib.check(lib.ksm_snippets(X.data_ptr(), X.stride(0), C, T, nt, twav_min,
                                    float(Th_single_ch), loc_range[0], loc_range[1],
                                    long_range[0], long_range[1], cap, clips.data_ptr(),
                                    xy.data_ptr(), count.data_ptr(), ws.data_ptr(), nb,
                                    _lib.stream_of(device)), "ksm_snippets")
        n = int(count.item())
        if n <= cap:
            return clips[:n]
        cap = 1 << int(np.ceil(np.log2(n)))


def _snippet_sample(ops, bfile, nt, twav_min, Th_single_ch, nskip, max_clips, device):
    """Isolated single-channel peaks of every ``nskip``-th batch, as one
    device tensor [n, nt].  The sample stops before the first batch whose
    clips would take it past ``max_clips`` (the reference keeps a buffer of
    500 000 clips and stops there, ``spikedetect.py:56-69``)."""
    parts, n = [], 0
    for j in range(0, int(bfile.n_batches), nskip):
        X = bfile.padded_batch_to_torch(j, ops)
        c = extract_snippets(X, nt=nt, twav_min=twav_min, Th_single_ch=Th_single_ch,
                             device=device)
        if n + len(c) > max_clips:
            break
        parts.append(c)
        n += len(c)
    return torch.cat(parts, 0) if parts else torch.zeros((0, nt), device=device)


def extract_wPCA_wTEMP(ops, bfile, nt=61, twav_min=20, Th_single_ch=6, nskip=25,
                       device=torch.device('cuda'), max_clips=500000):
    """Universal PCs and waveforms from the data (``spikedetect.py:53-92``).

    The snippet search runs on the GPU (``ksm_snippets``); the clips are
    scaled to unit norm there and copied to the host once.  The basis and
    the waveform clusters come from scikit-learn as in the reference (both
    are randomised there too): ``TruncatedSVD(n_components=n_pcs)`` gives
    ``wPCA`` = its components, ``KMeans(n_clusters=n_templates, n_init=10)``
    gives ``wTEMP`` = its centres scaled to unit norm."""
    from sklearn.cluster import KMeans
    from sklearn.decomposition import TruncatedSVD
    clips = _snippet_sample(ops, bfile, nt, twav_min, Th_single_ch, nskip, max_clips, device)
    unit = (clips / clips.square().sum(1, keepdim=True).sqrt()).cpu().numpy()
    settings = ops['settings']
    basis = TruncatedSVD(n_components=settings['n_pcs']).fit(unit).components_
    centres = KMeans(n_clusters=settings['n_templates'], n_init=10).fit(unit).cluster_centers_
    wPCA = torch.from_numpy(basis).to(device).float()
    wTEMP = torch.from_numpy(centres).to(device).float()
    wTEMP = wTEMP / wTEMP.norm(dim=1, keepdim=True)
    return wPCA, wTEMP


def template_match(X, ops, iC, iC2, weigh, device=torch.device('cuda')):
    """``spikedetect.template_match`` (``spikedetect.py:133-178``) on the GPU:
    ``xy`` int64 [nsp, 2] (template, column) in (template, column) order,
    ``imax`` int64 (size * n_templates + waveform), ``amp`` float32,
    ``adist`` float32 [nC, nsp]."""
    device = _lib.require_gpu(device)
    det = _Detector(ops, iC, iC2, weigh, device)
    xy, imax, amp, adist, _, _ = det.run(X)
    return xy.long(), imax.long(), amp, adist


def run(ops, bfile, device=torch.device('cuda'), progress_bar=None, clear_cache=False,
        verbose=False):
    """``spikedetect.run`` (``spikedetect.py:199-307``): universal waveforms
    from the data (``settings['templates_from_data']``, the default) or from
    ``ops['wTEMP']`` / ``ops['wPCA']``.  Returns ``(st, tF, ops)``: st float64
    [nsp, 6] = (time in s, y position, amplitude, template, batch, template
    position index), tF float32 [nsp, nC, n_pcs] (numpy, as the reference),
    and ops with ``iC``, ``iC2``, ``weigh``, ``ycup``, ``xcup``."""
    device = _lib.require_gpu(device)
    if ops['settings'].get('templates_from_data', True):
        logger.info('Re-computing universal templates from data.')
        ops['wPCA'], ops['wTEMP'] = extract_wPCA_wTEMP(
            ops, bfile, nt=ops['nt'], twav_min=ops['nt0min'],
            Th_single_ch=ops['settings']['Th_single_ch'], nskip=25, device=device)
    elif 'wTEMP' not in ops or 'wPCA' not in ops:
        raise ValueError("templates_from_data=False needs ops['wTEMP'] and ops['wPCA'] "
                         "(the reference's built-in waveforms, kilosort.spikedetect.get_waves)")
    iC, iC2, weigh = prepare(ops, device=device)
    logger.info(f'Number of universal templates: {iC.shape[1]}')
    det = _Detector(ops, iC, iC2, weigh, device)
    nt = ops['nt']
    parts, feats = [], []
    for ibatch in range(int(bfile.n_batches)):
        X = bfile.padded_batch_to_torch(ibatch, ops)
        xy, imax, amp, _, yct, tF = det.run(X)
        nsp = len(xy)
        st = np.zeros((nsp, 6), 'float64')
        t_shift = ibatch * bfile.batch_downsampling * (ops['batch_size'] / ops['fs'])
        xyc = xy.cpu().numpy().astype(np.int64)
        st[:, 0] = ((xyc[:, 1] - nt) / ops['fs'] + t_shift)
        st[:, 1] = yct.cpu().numpy()
        st[:, 2] = amp.cpu().numpy()
        st[:, 3] = imax.cpu().numpy()
        st[:, 4] = ibatch
        st[:, 5] = xyc[:, 0]
        parts.append(st)
        feats.append(tF.cpu().numpy())
        if progress_bar is not None:
            progress_bar.emit(int((ibatch + 1) / bfile.n_batches * 100))
    st = np.concatenate(parts, 0) if parts else np.zeros((0, 6), 'float64')
    nC = iC.shape[0]
    P = det.wPCA.shape[0]
    tF = np.concatenate(feats, 0) if feats else np.zeros((0, nC, P), 'float32')
    ops['iC'] = iC
    ops['iC2'] = iC2
    ops['weigh'] = weigh
    return st, tF, ops
