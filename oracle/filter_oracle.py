"""CPU restatement of Kilosort4's batch preprocessing (test oracle).

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` and ``bench.py``'s
``cpu_baseline`` leg; never part of the product path.

Follows, with the same torch / numpy / scipy operations in the same order:

* ``padded_batch``: ``BinaryRWFile.padded_batch_to_torch`` and
  ``_get_batch_edges`` (``kilosort/io.py:771-831``): the padded window of
  ``NT + 2 nt`` samples, uint16 shift, scale / shift, edge replication on the
  first and last batch;
* ``filter_batch``: ``BinaryFiltered.filter`` (``kilosort/io.py:990-1026``):
  channel map, sign inversion, channel means, common median reference,
  high-pass filtering in the Fourier domain, the artifact test and the
  whitening (optionally drift) matrix;
* ``get_highpass_filter`` / ``fft_highpass``
  (``kilosort/preprocessing.py:121-154``);
* ``kernel2D`` / ``drift_matrix`` (``kilosort/preprocessing.py:37-69``).
"""
from __future__ import annotations

import numpy as np
import torch
from scipy.interpolate import interp1d
from scipy.signal import butter, filtfilt
from torch.fft import fft, fftshift, ifft

CPU = torch.device("cpu")


def batch_edges(ibatch, NT, nt, imin, imax):
    """``io.py:771-783``."""
    if ibatch == 0:
        return imin, imin + NT + nt
    bstart = np.int64(imin + ibatch * NT - nt)
    return bstart, min(imax, np.int64(bstart + NT + 2 * nt))


def padded_batch(data, ibatch, NT, nt, n_batches, imin=0, imax=None, dtype="int16",
                 scale=None, shift=None):
    """``io.py:789-831`` for a [n_samples, n_chan_bin] array (batch_downsampling
    1): returns the float32 [n_chan_bin, NT + 2 nt] batch and [bstart, bend]."""
    imax = data.shape[0] if imax is None else imax
    bstart, bend = batch_edges(ibatch, NT, nt, imin, imax)
    d = np.asarray(data[bstart:bend]).T
    if dtype == "uint16":
        d = d.astype("float32")
        d = d - 2 ** 15
    if scale is not None:
        d = d * scale
    if shift is not None:
        d = d + shift
    nsamp = d.shape[-1]
    X = torch.zeros((data.shape[1], NT + 2 * nt))
    if ibatch == 0:
        X[:, nt:nt + nsamp] = torch.from_numpy(np.ascontiguousarray(d)).float()
        X[:, :nt] = X[:, nt:nt + 1]
        bstart = imin - nt
    elif ibatch == n_batches - 1:
        X[:, :nsamp] = torch.from_numpy(np.ascontiguousarray(d)).float()
        X[:, nsamp:] = X[:, nsamp - 1:nsamp]
        bend += nt
    else:
        X[:] = torch.from_numpy(np.ascontiguousarray(d)).float()
    return X, [bstart, bend]


def get_highpass_filter(fs=30000, cutoff=300):
    """``preprocessing.py:121-137``: the impulse response (length 30122) of a
    3rd-order Butterworth high-pass applied forwards and backwards."""
    NT = 30122
    b, a = butter(3, cutoff, fs=fs, btype="high")
    x = np.zeros(NT)
    x[NT // 2] = 1
    hp_filter = filtfilt(b, a, x).copy()
    return torch.from_numpy(hp_filter).float()


def fft_highpass(hp_filter, NT=30122):
    """``preprocessing.py:139-154``: the filter padded or cropped to NT, in the
    Fourier domain."""
    ft = hp_filter.shape[0]
    if ft < NT:
        pad = (NT - ft) // 2
        return fft(torch.cat((torch.zeros(pad, dtype=hp_filter.dtype), hp_filter,
                              torch.zeros(pad + (NT - pad * 2 - ft), dtype=hp_filter.dtype))))
    if ft > NT:
        crop = (ft - NT) // 2
        return fft(hp_filter[crop:crop + NT])
    return fft(hp_filter)


def kernel2D(x, y, sig=1):
    """``preprocessing.py:37-41``."""
    ds = ((x.unsqueeze(1) - y) ** 2).sum(-1)
    return torch.exp(-ds / (2 * sig ** 2))


def drift_matrix(ops, dshift):
    """``preprocessing.py:44-69``: the interpolation matrix that shifts the
    channels by the batch's drift estimate."""
    yblk = ops["yblk"]
    if ops["nblocks"] == 1:
        shifts = dshift
    else:
        finterp = interp1d(yblk, dshift, fill_value="extrapolate", kind="linear")
        shifts = finterp(ops["probe"]["yc"])
    xp = np.vstack((ops["probe"]["xc"], ops["probe"]["yc"])).T
    yp = xp.copy()
    yp[:, 1] -= shifts
    xp = torch.from_numpy(xp)
    yp = torch.from_numpy(yp)
    Kyx = kernel2D(yp, xp, ops["settings"]["sig_interp"])
    return Kyx @ ops["iKxx"]


def filter_batch(X, chan_map=None, invert_sign=False, do_CAR=True, hp_filter=None,
                 artifact_threshold=np.inf, whiten_mat=None, M=None, skip_preproc=False):
    """``io.py:990-1026`` on the CPU (M: the batch's drift matrix or None)."""
    if chan_map is not None:
        X = X[chan_map]
    if invert_sign:
        X = X * -1
    X = X - X.mean(1).unsqueeze(1)
    if do_CAR:
        X = X - torch.median(X, 0)[0]
    if skip_preproc:
        return X
    if hp_filter is not None:
        fwav = fft_highpass(hp_filter, NT=X.shape[1])
        X = torch.real(ifft(fft(X) * torch.conj(fwav)))
        X = fftshift(X, dim=-1)
    if artifact_threshold < np.inf:
        if torch.any(torch.abs(X) >= artifact_threshold):
            return torch.zeros_like(X)
    if whiten_mat is not None:
        if M is not None:
            X = (M @ whiten_mat) @ X
        else:
            X = whiten_mat @ X
    return X


def circular_highpass_direct(X, hp_filter):
    """The Fourier-domain filtering written as a sum over taps, in float64:
    out[c, n] = sum_m hpad[m] X[c, (n - T//2 + m) mod T] with hpad the filter
    padded or cropped to T as in ``fft_highpass`` (the product with the
    conjugate spectrum is a circular correlation; ``fftshift`` rolls it by
    T//2).  Used to pin the index conventions of ``filter_batch``."""
    X = np.asarray(X, np.float64)
    T = X.shape[1]
    h = np.asarray(hp_filter, np.float64)
    ft = h.shape[0]
    hpad = np.zeros(T)
    if ft < T:
        pad = (T - ft) // 2
        hpad[pad:pad + ft] = h
    elif ft > T:
        crop = (ft - T) // 2
        hpad[:] = h[crop:crop + T]
    else:
        hpad[:] = h
    out = np.zeros_like(X)
    n = np.arange(T)
    for m in np.nonzero(hpad)[0]:
        out += hpad[m] * X[:, (n - T // 2 + m) % T]
    return out
