"""Batch preprocessing on the GPU behind the reference's ``BinaryFiltered``.

``FilteredBatches`` is the batch source that ``extract`` reads
(``bfile.padded_batch_to_torch(ibatch, ops)``): a binary file (or an
in-memory array) of raw samples [n_samples, n_chan_bin] cut into padded
batches of ``NT + 2 nt`` samples and filtered as ``BinaryFiltered.filter``
does (``kilosort/io.py:789-831, 990-1026``):

    channel map -> sign -> minus channel means -> minus the median across
    channels (CAR) -> high-pass filter -> artifact test -> whitening, with
    the batch's drift matrix when ``dshift`` is given.

The raw int16 / uint16 / float32 window of each batch is copied to the
device through a pinned staging buffer and every step runs in the HIP
kernels of ``ksm_filter_batch`` (``libksm.so``).  The reference applies its
high-pass filter in the Fourier domain (``preprocessing.py:139-154``); that
product is a circular correlation with the filter's impulse response centred
at T/2, which the kernels compute as a banded Toeplitz product on MFMA over
the taps where the impulse response has not decayed below float32
resolution (``highpass_taps``).

Host-side set-up, as in the reference: the Butterworth design
(``get_highpass_filter``, scipy) and the drift interpolation of each batch's
shift to the channels (scipy ``interp1d``) and its Gaussian kernel; the
products ``Kyx @ iKxx`` and ``M @ Wrot`` run on the device (``ksm_gemm_f32``).
"""
from __future__ import annotations

import os
import threading

import numpy as np
import torch
from scipy.interpolate import interp1d
from scipy.signal import butter, filtfilt

from . import _lib

_FORMATS = {"int16": _lib.RAW_I16, "uint16": _lib.RAW_U16, "float32": _lib.RAW_F32}
_NP = {"int16": np.int16, "uint16": np.uint16, "float32": np.float32}
TAP_TAIL = 2.0 ** -26      # cut taps whose summed magnitude is below this share


def get_highpass_filter(fs=30000, cutoff=300, device=torch.device("cuda")):
    """Impulse response (30122 samples) of a 3rd-order Butterworth high-pass
    run forwards and backwards (``preprocessing.py:121-137``)."""
    NT = 30122
    b, a = butter(3, cutoff, fs=fs, btype="high")
    x = np.zeros(NT)
    x[NT // 2] = 1
    hp_filter = filtfilt(b, a, x).copy()
    return torch.from_numpy(hp_filter).to(device).float()


def highpass_taps(hp_filter, T, tail=TAP_TAIL):
    """Taps g[-J..J] of the circular correlation that the reference's
    Fourier-domain filter computes on a batch of T samples:
    ``ifft(fft(X) * conj(fft(hpad)))`` rolled by T//2 equals
    out[n] = sum_j hpad[j + T//2] x[(n + j) mod T], with hpad the filter
    zero padded (centred) or cropped to T as in ``fft_highpass``
    (``preprocessing.py:139-154``).  J is the smallest half-width whose
    dropped taps sum to at most ``tail`` of sum |g|.  Returns (float32 taps
    of length 2J+1 as in the filter, J)."""
    h = hp_filter.detach().cpu().float().numpy()
    ft = h.shape[0]
    hpad = np.zeros(T, np.float32)
    if ft < T:
        pad = (T - ft) // 2
        hpad[pad:pad + ft] = h
    elif ft > T:
        crop = (ft - T) // 2
        hpad[:] = h[crop:crop + T]
    else:
        hpad[:] = h
    g = np.roll(hpad, -(T // 2))            # g[j mod T] = hpad[j + T//2]
    a = np.abs(g.astype(np.float64))
    total = a.sum()
    if total == 0:
        return np.zeros(1, np.float32), 0
    J = (T - 1) // 2
    # dropped magnitude for half-width j: everything with |index| > j
    idx = np.arange(T)
    dist = np.minimum(idx, T - idx)         # |j| of entry idx
    order = np.bincount(dist, weights=a)    # magnitude per |j|
    kept = np.cumsum(order)
    ok = np.nonzero(total - kept <= tail * total)[0]
    if len(ok):
        J = int(ok[0])
    J = min(J, (T - 1) // 2)
    taps = np.concatenate([g[T - J:], g[:J + 1]]) if J > 0 else g[:1].copy()
    return taps.astype(np.float32), J


def drift_kernel(ops, dshift):
    """Kyx of ``get_drift_matrix`` (``preprocessing.py:44-64``): the batch's
    shift interpolated to the channels and the Gaussian kernel between the
    shifted and the original sites (float32, as the reference's probe
    coordinates)."""
    yblk = ops["yblk"]
    if ops["nblocks"] == 1:
        shifts = dshift
    else:
        finterp = interp1d(yblk, dshift, fill_value="extrapolate", kind="linear")
        shifts = finterp(ops["probe"]["yc"])
    xp = np.vstack((ops["probe"]["xc"], ops["probe"]["yc"])).T
    yp = xp.copy()
    yp[:, 1] -= shifts
    sig = ops["settings"]["sig_interp"]
    ds = ((yp[:, np.newaxis, :] - xp[np.newaxis, :, :]) ** 2).sum(-1)
    return np.exp(-ds / (2 * sig ** 2)).astype(xp.dtype)


_copy_pool = None
_COPY_CHUNK = 4 << 20            # bytes per host copy task


def _host_copy(dst, src):
    """dst[:] = src (1-D uint8 arrays) in chunks over a small thread pool:
    numpy releases the GIL for the copies, and one host thread copies a
    batch into pinned memory at a fraction of the memory bandwidth."""
    global _copy_pool
    n = src.nbytes
    if n <= 2 * _COPY_CHUNK:
        dst[:] = src
        return
    if _copy_pool is None:
        from concurrent.futures import ThreadPoolExecutor
        _copy_pool = ThreadPoolExecutor(max_workers=max(1, min(8, (os.cpu_count() or 2) // 2)))
    step = max(_COPY_CHUNK, -(-n // 32))
    futs = [_copy_pool.submit(np.copyto, dst[a:a + step], src[a:a + step])
            for a in range(0, n, step)]
    for f in futs:
        f.result()


class _Staging:
    """Two pinned host buffers used in turn; a buffer is rewritten only after
    its previous host-to-device copy has completed."""

    def __init__(self, nbytes):
        self.bufs = [torch.empty(nbytes, dtype=torch.uint8, pin_memory=True) for _ in range(2)]
        self.events = [None, None]
        self.k = 0

    def put(self, arr, dev_buf, stream):
        k = self.k
        self.k ^= 1
        if self.events[k] is not None:
            self.events[k].synchronize()
        src = self.bufs[k][:arr.nbytes]
        _host_copy(src.numpy(), np.ascontiguousarray(arr).view(np.uint8).reshape(-1))
        with torch.cuda.stream(stream):
            dev_buf[:arr.nbytes].copy_(src, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(stream)
        self.events[k] = ev


class FilteredBatches:
    """GPU counterpart of ``BinaryFiltered`` (``io.py:962-1044``) with the
    batch geometry of ``BinaryRWFile`` (``io.py:564-831``): same constructor
    arguments, ``n_batches``, ``padded_batch_to_torch(ibatch, ops=None,
    return_inds=False, skip_preproc=False)`` and ``filter(X, ops, ibatch,
    skip_preproc)``.  Raw dtypes: int16 (default), uint16, float32.  The
    batches are produced on the current torch stream of ``device``; one
    instance is used from one host thread at a time."""

    supported_dtypes = ["int16", "uint16", "float32"]

    def __init__(self, filename=None, n_chan_bin: int = 0, fs: int = 30000, NT: int = 60000,
                 nt: int = 61, nt0min: int = 20, chan_map=None, hp_filter=None,
                 whiten_mat=None, dshift=None, device=None, do_CAR: bool = True,
                 artifact_threshold: float = np.inf, invert_sign: bool = False, dtype=None,
                 tmin: float = 0.0, tmax: float = np.inf, shift=None, scale=None,
                 file_object=None, batch_downsampling=1):
        self.device = _lib.require_gpu(torch.device("cuda") if device is None else device)
        self.fs = np.float64(fs)
        self.n_chan_bin = int(n_chan_bin)
        self.filename = filename
        self.NT = np.int64(NT)
        self.nt = np.int64(nt)
        self.nt0min = np.int64(nt0min)
        self.shift, self.scale = shift, scale
        if file_object is not None:
            dtype = file_object.dtype
        self.dtype = str(np.dtype(dtype if dtype is not None else "int16"))
        if self.dtype not in self.supported_dtypes:
            raise ValueError(f"raw dtype {self.dtype} is not supported "
                             f"(one of {self.supported_dtypes})")
        self.file_object = file_object
        if file_object is None:
            itemsize = np.dtype(self.dtype).itemsize
            nbytes = np.memmap(filename, mode="r", dtype=np.uint8).shape[0]
            self.total_samples = int(nbytes // (itemsize * self.n_chan_bin))
        else:
            n, c = file_object.shape
            if c != self.n_chan_bin:
                raise ValueError(f"file_object has {c} channels, n_chan_bin={n_chan_bin}")
            self.total_samples = int(n)
        self.imin = max(np.int64(np.float64(tmin) * self.fs), 0)
        self.imax = np.int64(self.total_samples) if tmax == np.inf else \
            min(np.int64(np.float64(tmax) * self.fs), np.int64(self.total_samples))
        self.n_batches_raw = np.int64(np.ceil(self.n_samples / self.NT))
        a, b = self._get_batch_edges(self.n_batches_raw - 1)
        tail = b - a - self.nt
        if tail < self.nt:                       # io.py:650-655
            self.n_batches_raw -= 1
            self.imax -= tail
        self.set_downsampling(batch_downsampling)

        self.chan_map = chan_map
        self.invert_sign = invert_sign
        self.do_CAR = do_CAR
        self.artifact_threshold = artifact_threshold
        self.hp_filter = hp_filter
        self.whiten_mat = whiten_mat
        self.dshift = dshift
        self._lock = threading.Lock()
        self._ws = {}
        self._taps = {}
        self._staging = None
        self._raw_dev = None
        cm = None if chan_map is None else np.asarray(chan_map, np.int64)
        if cm is not None and (cm.min() < 0 or cm.max() >= self.n_chan_bin):
            raise ValueError("chan_map entries must lie in [0, n_chan_bin)")
        self._cmap = None if cm is None else torch.from_numpy(cm.astype(np.int32)).to(self.device)
        self.C = self.n_chan_bin if cm is None else len(cm)
        self._whiten = None
        if whiten_mat is not None:
            W = torch.as_tensor(whiten_mat).to(self.device, torch.float32).contiguous()
            if W.shape != (self.C, self.C):
                raise ValueError(f"whiten_mat must be [{self.C}, {self.C}]")
            self._whiten = W

    # ---- batch geometry (io.py:564-787) ----
    @property
    def n_samples(self) -> int:
        return int(self.imax - self.imin)

    @property
    def shape(self):
        return self.n_samples, self.n_chan_bin

    def set_downsampling(self, downsampling):
        self.batch_downsampling = downsampling
        self.n_batches = np.int64(self.n_batches_raw / self.batch_downsampling)

    def _get_batch_edges(self, ibatch):
        if ibatch == 0:
            bstart = self.imin
            bend = self.imin + self.NT + self.nt
        else:
            ibatch = np.int64(ibatch)
            bstart = np.int64(self.imin + (ibatch * self.NT) - self.nt)
            bend = min(self.imax, np.int64(bstart + self.NT + 2 * self.nt))
        return bstart, bend

    @property
    def file(self):
        if self.file_object is not None:
            return self.file_object
        return np.memmap(self.filename, mode="r", dtype=self.dtype,
                         shape=(self.total_samples, self.n_chan_bin))

    # ---- filtering ----
    def _taps_for(self, T):
        if self.hp_filter is None:
            return None, -1
        if T not in self._taps:
            taps, J = highpass_taps(self.hp_filter, T)
            if 2 * J + 1 > T:
                raise ValueError(f"the high-pass filter needs {2 * J + 1} taps, more than "
                                 f"the {T} samples of a batch")
            self._taps[T] = (torch.from_numpy(taps).to(self.device), J)
        return self._taps[T]

    def _workspace(self, C, T, J, flags):
        lib = _lib.load()
        nb = lib.ksm_filter_workspace(C, T, J, flags & _lib.FILTER_WHITEN)
        ws = self._ws.get("buf")
        if ws is None or ws.numel() < nb:
            ws = torch.empty(max(nb, 1), dtype=torch.uint8, device=self.device)
            self._ws["buf"] = ws
        return ws, nb

    def _matrix(self, ops, ibatch):
        """Whitening matrix of batch ibatch, times its drift matrix
        (io.py:1019-1026)."""
        if self._whiten is None:
            return None
        if self.dshift is None or ops is None or ibatch is None:
            return self._whiten
        lib = _lib.load()
        Kyx = torch.from_numpy(np.ascontiguousarray(
            drift_kernel(ops, self.dshift[ibatch]), np.float32)).to(self.device)
        iK = torch.as_tensor(ops["iKxx"]).to(self.device, torch.float32).contiguous()
        C = self.C
        M = torch.empty((C, C), device=self.device)
        MW = torch.empty((C, C), device=self.device)
        s = _lib.stream_of(self.device)
        _lib.check(lib.ksm_gemm_f32(Kyx.data_ptr(), C, iK.data_ptr(), C, M.data_ptr(), C,
                                    C, C, C, s), "ksm_gemm_f32")
        _lib.check(lib.ksm_gemm_f32(M.data_ptr(), C, self._whiten.data_ptr(), C, MW.data_ptr(),
                                    C, C, C, C, s), "ksm_gemm_f32")
        return MW

    def _run(self, raw_ptr, fmt, ldraw, nsamp, off, edge, T, ops, ibatch, skip_preproc):
        lib = _lib.load()
        flags = edge
        if self.invert_sign:
            flags |= _lib.FILTER_INVERT
        if self.do_CAR:
            flags |= _lib.FILTER_CAR
        if skip_preproc:
            flags |= _lib.FILTER_SKIP
        scale = shift = 0.0
        if self.scale is not None:
            flags |= _lib.FILTER_SCALE
            scale = float(self.scale)
        if self.shift is not None:
            flags |= _lib.FILTER_SHIFT
            shift = float(self.shift)
        taps, J = (None, -1) if skip_preproc else self._taps_for(T)
        Wm = None if skip_preproc else self._matrix(ops, ibatch)
        if Wm is not None:
            flags |= _lib.FILTER_WHITEN
        C = self.C
        X = torch.empty((C, T), device=self.device)
        ws, nb = self._workspace(C, T, J, flags)
        thr = float(self.artifact_threshold)
        _lib.check(lib.ksm_filter_batch(
            raw_ptr, fmt, ldraw, nsamp, off, _lib.ptr(self._cmap), C, T, flags, scale, shift,
            _lib.ptr(taps), J, thr, _lib.ptr(Wm), X.data_ptr(), T, ws.data_ptr(), nb,
            _lib.stream_of(self.device)), "ksm_filter_batch")
        return X

    def filter(self, X, ops=None, ibatch=None, skip_preproc=False):
        """``BinaryFiltered.filter`` (io.py:990-1026) of a float32 device batch
        X [n_chan_bin, T]."""
        X = X.to(self.device, torch.float32)
        if X.stride(1) != 1:
            X = X.contiguous()
        if X.shape[0] != self.n_chan_bin:
            raise ValueError(f"X must have {self.n_chan_bin} rows")
        T = int(X.shape[1])
        with self._lock:
            return self._run(X.data_ptr(), _lib.RAW_F32_CM, X.stride(0), T, 0, 0, T, ops,
                             ibatch, skip_preproc)

    def padded_batch_to_torch(self, ibatch, ops=None, return_inds=False, skip_preproc=False):
        """``BinaryFiltered.padded_batch_to_torch`` (io.py:1037-1044): batch
        ibatch padded to NT + 2 nt samples (io.py:789-831) and filtered."""
        ibatch_in = ibatch
        ibatch = ibatch * self.batch_downsampling
        bstart, bend = self._get_batch_edges(ibatch)
        data = self.file[bstart:bend]
        nsamp = int(data.shape[0])
        T = int(self.NT + 2 * self.nt)
        if ibatch == 0:
            off, edge = int(self.nt), _lib.FILTER_EDGE_FIRST
            bstart = self.imin - self.nt
        elif ibatch == self.n_batches - 1:
            off, edge = 0, _lib.FILTER_EDGE_LAST
            bend += self.nt
        else:
            off, edge = 0, 0
        nsamp = min(nsamp, T - off)
        data = data[:nsamp]
        stream = torch.cuda.current_stream(self.device)
        with self._lock:
            nbytes = T * self.n_chan_bin * np.dtype(self.dtype).itemsize
            if self._staging is None:
                self._staging = _Staging(nbytes)
                self._raw_dev = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
            arr = np.asarray(data, dtype=_NP[self.dtype])
            self._staging.put(arr, self._raw_dev, stream)
            X = self._run(self._raw_dev.data_ptr(), _FORMATS[self.dtype], self.n_chan_bin,
                          nsamp, off, edge, T, ops, ibatch_in, skip_preproc)
        if return_inds:
            return X, [bstart, bend]
        return X
