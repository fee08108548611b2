"""In-memory batch source with the reference's padded-batch semantics.

``extract`` only needs a duck-typed ``bfile`` with ``n_batches``,
``batch_downsampling`` and ``padded_batch_to_torch(ibatch, ops)``
(``template_matching.py:73,83,102``).  ``ArrayBatches`` serves batches of an
already whitened recording held in host or device memory and reproduces the
batch edges of ``BinaryRWFile`` exactly:

* batch count and last-batch trimming: ``io.py:646-655``;
* batch edges: ``io.py:771-783``;
* padding to ``NT + 2*nt`` with edge replication on the first and last batch:
  ``io.py:789-831``.

Filtering (CAR, high-pass, whitening; ``io.py:990-1026``) is not applied; the
data are taken to be whitened already.
"""
from __future__ import annotations

import numpy as np
import torch


class ArrayBatches:
    """Padded batches of a [n_samples, C] float32 recording."""

    def __init__(self, data, NT: int = 60000, nt: int = 61, device=None,
                 batch_downsampling: int = 1, tmin_samples: int = 0):
        if isinstance(data, np.ndarray):
            data = torch.from_numpy(np.ascontiguousarray(data))
        if data.ndim != 2:
            raise ValueError("data must be [n_samples, n_channels]")
        self.data = data
        self.device = torch.device(device) if device is not None else data.device
        self.NT = np.int64(NT)
        self.nt = int(nt)
        self.n_chan_bin = int(data.shape[1])
        self.total_samples = int(data.shape[0])
        self.imin = max(np.int64(tmin_samples), 0)
        self.imax = np.int64(self.total_samples)
        self.n_batches_raw = np.int64(np.ceil(self.n_samples / self.NT))
        # a last batch shorter than nt samples is dropped (io.py:650-655)
        a, b = self._get_batch_edges(self.n_batches_raw - 1)
        tail = b - a - self.nt
        if tail < self.nt:
            self.n_batches_raw -= 1
            self.imax -= tail
        self.set_downsampling(batch_downsampling)

    @property
    def n_samples(self) -> int:
        return int(self.imax - self.imin)

    def set_downsampling(self, downsampling: int):
        self.batch_downsampling = int(downsampling)
        self.n_batches = np.int64(self.n_batches_raw / self.batch_downsampling)

    def _get_batch_edges(self, ibatch):
        if ibatch == 0:
            bstart = self.imin
            bend = self.imin + self.NT + self.nt
        else:
            ibatch = np.int64(ibatch)
            bstart = np.int64(self.imin + ibatch * self.NT - self.nt)
            bend = min(self.imax, np.int64(bstart + self.NT + 2 * self.nt))
        return int(bstart), int(bend)

    def padded_batch_to_torch(self, ibatch, ops=None, return_inds=False):
        ibatch = int(ibatch) * self.batch_downsampling
        bstart, bend = self._get_batch_edges(ibatch)
        # from pinned host memory the copy is queued on the caller's stream and
        # the host goes on (the recording outlives the copy)
        src = self.data[bstart:bend]
        seg = src.to(self.device, torch.float32,
                     non_blocking=bool(src.device.type == 'cpu' and src.is_pinned())).T
        nsamp = seg.shape[-1]
        NT, nt = int(self.NT), self.nt
        X = torch.zeros((self.n_chan_bin, NT + 2 * nt), device=self.device)
        if ibatch == 0:
            X[:, nt:nt + nsamp] = seg
            X[:, :nt] = X[:, nt:nt + 1]
            bstart = int(self.imin) - nt
        elif ibatch == self.n_batches - 1:
            X[:, :nsamp] = seg
            X[:, nsamp:] = X[:, nsamp - 1:nsamp]
            bend += nt
        else:
            X[:] = seg
        if return_inds:
            return X, [bstart, bend]
        return X
