"""Drop-in replacement for Kilosort4's learned-template matching path.

Same Python surface as ``kilosort/template_matching.py``:

* ``prepare_extract(xc, yc, U, nC, position_limit, device)`` (ref :14-53)
* ``extract(ops, bfile, U, device, progress_bar)``             (ref :56-126)
* ``prepare_matching(ops, U)``                                  (ref :155-167)
* ``run_matching(ops, X, U, ctc, device)``                      (ref :170-244)

with the same ``ops`` read set (``nt``, ``Th_learned``, ``max_peels``,
``wPCA``, ``nt0min``, ``batch_size``, ``xc``, ``yc``,
``settings[nearest_chans, position_limit, n_pcs]``), the same ``ops`` writes
(``iCC``, ``iCC_mask``, ``iU``) and the same outputs: ``st`` float64 numpy
[N, 3] (time, template, th_amp), ``tF`` float32 CPU tensor [N, nC, P].

All arithmetic runs in the HIP kernels of ``libksm.so`` on a ROCm GPU; torch
provides device memory and the stream.  The one host computation is the
nearest-channel table, which the reference also computes on the host with
numpy (its unstable ``argsort`` tie order is reproduced by making the same
call).  There is no CPU path: a CPU ``device`` or a missing library raises.

Optional key ``ops['matching_precision']``:

* ``'fp32'`` (default): FP32 correlation and FP32 MFMA projection -- the
  parity form;
* ``'bf16'``: BF16 operands on MFMA with FP32 accumulation;
* ``'bf16x3'``: split BF16 operands (hi*hi + hi*lo + lo*hi) on MFMA;
* ``'fp8'``: OCP e4m3 operands on MFMA with FP32 accumulation, the
  correlation scaled per (channel, PC) by a bound on its magnitude and U per
  template.

In the reduced-precision forms the projection also returns a per-column
bound on its scores' error (proven for the BF16 forms, a 6-sigma rounding
model for FP8), and the peel takes every accept/reject decision that bound
leaves open on scores re-scored in FP32 from the residual, so decisions
match FP32 arithmetic; accepted spikes' amplitudes, th_amps and the
subtraction use their FP32 scores.

Optional key ``ops['matching_streams']`` (default ``MATCHING_STREAMS``):
device streams over which ``extract`` spreads its batches, one host thread
each; results are identical for any value and come back in batch order.
"""
from __future__ import annotations

import ctypes
import logging
import os
import queue
import threading

import numpy as np
import torch

from . import _lib

logger = logging.getLogger(__name__)

ST_CAPACITY = 100000          # per-batch spike buffer of run_matching (ref :186)

# peel iterations that accepted spikes in the last run_matching call (of any
# thread; _iters.n holds the calling thread's)
last_peel_iterations = 0
_iters = threading.local()

def _default_streams() -> int:
    """Device streams (one host thread each) over which extract runs its
    batches: one per HIP hardware queue the process has beside the default
    stream's, up to 4.  HIP gives a process GPU_MAX_HW_QUEUES hardware queues
    (4 unless the environment sets it before HIP starts); streams beyond them
    share queues and serialise.  Measured at 383 channels x 700 templates:
    3 streams on 4 queues 2.42 ms per batch, 4 streams on 8 queues 2.32 ms,
    4 streams on 4 queues 2.64 ms."""
    try:
        queues = int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))
    except ValueError:
        queues = 4
    return max(1, min(4, queues - 1))


MATCHING_STREAMS = _default_streams()


def _lags(nt: int) -> int:
    return 2 * nt + 1


def _ldl(nt: int) -> int:
    return (_lags(nt) + 3) // 4 * 4


def _f32(t, device, name):
    if not torch.is_tensor(t):
        t = torch.as_tensor(np.asarray(t))
    if t.device != device:
        t = t.to(device)
    if t.dtype != torch.float32:
        raise TypeError(f"{name} must be float32 (got {t.dtype})")
    return t.contiguous()


def _check_U(U, device):
    U = _f32(U, device, "U")
    if U.ndim != 3:
        raise ValueError(f"U must be [K, n_pcs, C] (got shape {tuple(U.shape)})")
    return U


class _Workspace:
    """Grow-only device scratch buffers, one set per (device, stream): calls
    on different streams (e.g. batches run concurrently from several host
    threads) never share scratch memory."""

    def __init__(self):
        self._bufs = {}

    def get(self, device, name, nbytes):
        key = (str(device), torch.cuda.current_stream(device).cuda_stream, name)
        buf = self._bufs.get(key)
        if buf is None or buf.numel() < nbytes:
            buf = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)
            self._bufs[key] = buf
        return buf

    def clear(self):
        self._bufs.clear()


_ws = _Workspace()


class _Gram:
    """Peel-side form of one (templates, Gram) pair."""
    __slots__ = ("base", "ldl", "rows", "nrows", "perm", "rank", "lohi", "off", "cc",
                 "ready", "src")


class _GramCache:
    """Per-(U, ctc) derived data: the Gram in firing-template-major storage
    (copied once when the caller passes the reference's contiguous layout),
    its nonzero-row lists, the peel order of the templates (sorted by peak
    channel) and the compact lag-major Gram over that order."""

    MAX_ENTRIES = 16          # recordings interleaved by extract_many stay cached
    MAX_BYTES = 8 << 30       # device memory the cache may hold

    def __init__(self):
        self.entries = {}         # key -> _Gram, in insertion (use) order
        self.lock = threading.Lock()

    def get(self, ctc, U, nt, device):
        K, P, C = U.shape
        key = (ctc.data_ptr(), ctc._version, tuple(ctc.shape), tuple(ctc.stride()),
               U.data_ptr(), U._version, tuple(U.shape))
        with self.lock:
            g = self.entries.pop(key, None)
            if g is not None:
                self.entries[key] = g             # most recently used last
        if g is not None:
            # built on another stream: order this stream after it
            torch.cuda.current_stream(device).wait_event(g.ready)
            return g
        L = _lags(nt)
        if tuple(ctc.shape) != (K, K, L):
            raise ValueError(f"ctc must be [{K}, {K}, {L}] (got {tuple(ctc.shape)})")
        s = ctc.stride()
        if ctc.dtype == torch.float32 and s[2] == 1 and s[1] == K * s[0] and s[0] >= L \
                and s[0] % 4 == 0 and ctc.device == device:
            base, ldl = ctc, s[0]            # firing-template-major storage
        else:
            ldl = _ldl(nt)
            store = torch.zeros((K, K, ldl), dtype=torch.float32, device=device)
            store[:, :, :L] = ctc.to(device, torch.float32).permute(1, 0, 2)
            base = store[:, :, :L].permute(1, 0, 2)
        lib = _lib.load()
        stream = _lib.stream_of(device)
        g = _Gram()
        g.base, g.ldl = base, ldl
        i32 = dict(dtype=torch.int32, device=device)
        g.rows = torch.empty((K, K), **i32)
        g.nrows = torch.empty((K,), **i32)
        _lib.check(lib.ksm_ctc_rows(base.data_ptr(), K, L, ldl, g.rows.data_ptr(),
                                    g.nrows.data_ptr(), stream), "ksm_ctc_rows")
        g.perm = torch.empty((K,), **i32)
        g.rank = torch.empty((K,), **i32)
        _lib.check(lib.ksm_template_order(U.data_ptr(), K, P, C, g.perm.data_ptr(),
                                          g.rank.data_ptr(), stream), "ksm_template_order")
        g.lohi = torch.empty((K, 2), **i32)
        g.off = torch.empty((K + 1,), dtype=torch.int64, device=device)
        _lib.check(lib.ksm_gram_ranges(g.rows.data_ptr(), g.nrows.data_ptr(), g.rank.data_ptr(),
                                       K, L, g.lohi.data_ptr(), g.off.data_ptr(), stream),
                   "ksm_gram_ranges")
        total = int(g.off[K].item())
        g.cc = torch.zeros((max(total, 4),), dtype=torch.float32, device=device)
        _lib.check(lib.ksm_gram_compact(base.data_ptr(), ldl, g.perm.data_ptr(),
                                        g.lohi.data_ptr(), g.off.data_ptr(), K, L,
                                        g.cc.data_ptr(), stream), "ksm_gram_compact")
        g.ready = torch.cuda.Event()
        g.ready.record(torch.cuda.current_stream(device))
        # the key holds addresses: keep the tensors alive while cached, so
        # that no other tensor can take their memory and match the key
        g.src = (ctc, U)
        with self.lock:
            self.entries[key] = g
            # least recently used first out: at most MAX_ENTRIES entries and
            # MAX_BYTES of device memory (the newest entry always stays)
            while len(self.entries) > 1 and (len(self.entries) > self.MAX_ENTRIES or
                                             self._bytes() > self.MAX_BYTES):
                self.entries.pop(next(iter(self.entries)))
        return g

    def _bytes(self):
        """Device bytes the cached entries hold (their copies and tables; a
        caller's own ctc and U are counted too, as the entry keeps them)."""
        n = 0
        for g in self.entries.values():
            st = {t.untyped_storage().data_ptr(): t.untyped_storage().nbytes()
                  for t in (g.base, g.rows, g.nrows, g.perm, g.rank, g.lohi, g.off, g.cc, *g.src)
                  if torch.is_tensor(t)}
            n += sum(st.values())
        return n

    def clear(self):
        with self.lock:
            self.entries.clear()


_grams = _GramCache()


def prepare_extract(xc, yc, U, nC, position_limit, device=torch.device('cuda')):
    """Nearest channels, position mask, peak channel and peak-channel PCs.

    Reference ``template_matching.py:14-53``.  ``iCC``/``iCC_mask`` come from
    the same numpy calls on the host (for the reference's argsort tie order);
    ``iU`` and ``Ucc`` from ``ksm_prepare_extract``.
    """
    device = _lib.require_gpu(device)
    xc = np.asarray(xc)
    yc = np.asarray(yc)
    ds = (xc - xc[:, np.newaxis]) ** 2 + (yc - yc[:, np.newaxis]) ** 2
    iCC = torch.from_numpy(np.argsort(ds, 0)[:nC]).to(device)
    iCC_mask = torch.from_numpy(np.sort(ds, 0)[:nC] < position_limit ** 2).to(device)
    U = _check_U(U, device)
    K, P, C = U.shape
    if iCC.shape[1] != C:
        raise ValueError(f"probe has {iCC.shape[1]} channels but U has {C}")
    iCC = iCC.to(torch.int64).contiguous()
    nC = iCC.shape[0]
    iU = torch.empty((K,), dtype=torch.int64, device=device)
    Ucc = torch.empty((nC, K, P), dtype=torch.float32, device=device)
    lib = _lib.load()
    _lib.check(lib.ksm_prepare_extract(U.data_ptr(), K, P, C, iCC.data_ptr(), nC,
                                       iU.data_ptr(), Ucc.data_ptr(),
                                       _lib.stream_of(device)),
               "ksm_prepare_extract")
    return iCC, iCC_mask, iU, Ucc


def prepare_matching(ops, U):
    """Template Gram ctc [K, K, 2*nt+1] (reference ``template_matching.py:155-167``).

    Formed in float64 by ``ksm_prepare_matching`` and rounded to float32.  The
    returned tensor indexes like the reference's (``ctc[i, j, l]``) and is a
    permuted view of firing-template-major storage, the layout the peel reads.
    """
    W = ops['wPCA']
    device = _lib.require_gpu(W.device if torch.is_tensor(W) else 'cuda')
    W = _f32(W, device, "wPCA")
    U = _check_U(U, device)
    K, P, C = U.shape
    nt = int(ops['nt'])
    if tuple(W.shape) != (P, nt):
        raise ValueError(f"wPCA must be [{P}, {nt}] (got {tuple(W.shape)})")
    L, ldl = _lags(nt), _ldl(nt)
    store = torch.empty((K, K, ldl), dtype=torch.float32, device=device)
    lib = _lib.load()
    nbytes = lib.ksm_prepare_matching_workspace(K, P, C, nt)
    ws = _ws.get(device, "prepare", nbytes)
    _lib.check(lib.ksm_prepare_matching(U.data_ptr(), W.data_ptr(), K, P, C, nt,
                                        store.data_ptr(), ldl, ws.data_ptr(), nbytes,
                                        _lib.stream_of(device)),
               "ksm_prepare_matching")
    return store[:, :, :L].permute(1, 0, 2)


_PRECISIONS = {'fp32': _lib.PREC_FP32, 'bf16': _lib.PREC_BF16, 'bf16x3': _lib.PREC_BF16X3,
               'fp8': _lib.PREC_FP8}


def _precision(ops):
    name = str(ops.get('matching_precision', 'fp32')).lower()
    if name not in _PRECISIONS:
        raise ValueError(f"unknown matching_precision {name!r} (one of {sorted(_PRECISIONS)})")
    return _PRECISIONS[name]


def _tile_m(ops, K):
    """Template rows per tile of the BF16 forms' projection GEMM (64 or 128;
    0: the library's default), an autotuning choice: ``ops['matching_tile_m']``."""
    t = ops.get('matching_tile_m')
    return 0 if t is None else int(t)


def run_matching(ops, X, U, ctc, device=torch.device('cuda')):
    """Matching pursuit over one padded batch (reference ``template_matching.py:170-244``).

    Returns ``st`` int64 [k, 2] (column, template), ``amps`` and ``th_amps``
    float32 [k, 1], and the residual ``Xres`` [C, T], all on the device, in the
    reference's order (by peel iteration, then by time).
    """
    device = _lib.require_gpu(device)
    Th = float(ops['Th_learned'])
    nt = int(ops['nt'])
    max_peels = int(ops['max_peels'])
    W = _f32(ops['wPCA'], device, "wPCA")
    U = _check_U(U, device)
    X = _f32(X, device, "X")
    K, P, C = U.shape
    if X.ndim != 2 or X.shape[0] != C:
        raise ValueError(f"X must be [{C}, T] (got {tuple(X.shape)})")
    T = X.shape[1]
    prec = _precision(ops)
    lib = _lib.load()
    stream = _lib.stream_of(device)

    g = _grams.get(ctc, U, nt, device)
    nm = torch.empty((K,), dtype=torch.float32, device=device)
    _lib.check(lib.ksm_template_norms(U.data_ptr(), K, P, C, nm.data_ptr(), stream),
               "ksm_template_norms")

    # column-major scores over the peel order: Bt[t][r] = B[perm[r], t]
    ldbt = (K + 63) // 64 * 64
    rows_t = (T + 127) // 128 * 128
    Bt = _ws.get(device, "Bt", rows_t * ldbt * 4)
    pbytes = lib.ksm_project_workspace(C, T, K, P, nt, prec)
    pws = _ws.get(device, "project", pbytes)
    if prec == _lib.PREC_FP32:
        bound = None
        _lib.check(lib.ksm_project(X.data_ptr(), T, W.data_ptr(), U.data_ptr(), g.perm.data_ptr(),
                                   C, T, K, P, nt, Bt.data_ptr(), ldbt, prec, pws.data_ptr(),
                                   pbytes, stream), "ksm_project")
    else:
        # reduced-precision scores and their per-column error bound, with
        # which the peel takes every decision the bound leaves open on FP32
        # re-scored values
        bbytes = lib.ksm_bound_bytes(C, T, K, P, nt)
        bound = _ws.get(device, "bound", bbytes)
        _lib.check(lib.ksm_project_bounded(X.data_ptr(), T, W.data_ptr(), U.data_ptr(),
                                           g.perm.data_ptr(), C, T, K, P, nt, Bt.data_ptr(),
                                           ldbt, prec, _tile_m(ops, K), bound.data_ptr(), bbytes, pws.data_ptr(),
                                           pbytes, stream), "ksm_project_bounded")

    Xres = X.clone()
    st = torch.empty((ST_CAPACITY, 2), dtype=torch.int64, device=device)
    amps = torch.empty((ST_CAPACITY, 1), dtype=torch.float32, device=device)
    th_amps = torch.empty((ST_CAPACITY, 1), dtype=torch.float32, device=device)
    info = torch.zeros((4,), dtype=torch.int32, device=device)
    kbytes = lib.ksm_peel_workspace(K, T, max_peels)
    kws = _ws.get(device, "peel", kbytes)
    _lib.check(lib.ksm_peel(Bt.data_ptr(), ldbt, Xres.data_ptr(), T, U.data_ptr(), W.data_ptr(),
                            nm.data_ptr(), g.perm.data_ptr(), g.rank.data_ptr(),
                            g.cc.data_ptr(), g.lohi.data_ptr(), g.off.data_ptr(), K, P, C, T,
                            nt, Th, max_peels,
                            0 if bound is None else _lib.PEEL_RESCORE,
                            None if bound is None else bound.data_ptr(),
                            st.data_ptr(), amps.data_ptr(),
                            th_amps.data_ptr(), ST_CAPACITY, info.data_ptr(), kws.data_ptr(),
                            kbytes, stream),
               "ksm_peel")
    k, status, iters, rescored = (int(v) for v in info.cpu())
    if status == _lib.KSM_ERR_CAPACITY:
        raise RuntimeError(f"more than {ST_CAPACITY} spikes in one batch "
                           "(the reference's per-batch buffer overflows too)")
    _lib.check(status, "ksm_peel (device)")
    global last_peel_iterations
    last_peel_iterations = iters
    _iters.n = iters
    _iters.rescored = rescored
    logger.debug("run_matching: %d spikes in %d peel iterations (%d FP32 re-scores)", k, iters,
                 rescored)
    return st[:k], amps[:k], th_amps[:k], Xres


def _features(ops, Xres, st, amps, iCC, iU, Ucc, device):
    nsp = st.shape[0]
    nC, K, P = Ucc.shape
    C, T = Xres.shape
    tF = torch.empty((nsp, nC, P), dtype=torch.float32, device=device)
    if nsp == 0:
        return tF
    W = _f32(ops['wPCA'], device, "wPCA")
    lib = _lib.load()
    _lib.check(lib.ksm_features(Xres.data_ptr(), T, W.data_ptr(), st.contiguous().data_ptr(),
                                amps.contiguous().data_ptr(), nsp, iCC.data_ptr(),
                                iU.data_ptr(), Ucc.data_ptr(), nC, K, P, C, int(ops['nt']),
                                tF.data_ptr(), _lib.stream_of(device)),
               "ksm_features")
    return tF


class _Ctx:
    """What one recording's batches are matched with."""
    __slots__ = ("ops", "U", "ctc", "iCC", "iU", "Ucc")

    def __init__(self, ops, U, ctc, iCC, iU, Ucc):
        self.ops, self.U, self.ctc = ops, U, ctc
        self.iCC, self.iU, self.Ucc = iCC, iU, Ucc


# idle lanes (host thread + device stream + pinned buffers) per device, reused
# by later pipelines: a new host thread pays for its own HIP set-up, pinned
# progress words and pinned buffers, milliseconds per pipeline
_idle_lanes = {}
_idle_lock = threading.Lock()


def _lane_take(device):
    with _idle_lock:
        free = _idle_lanes.get(str(device))
        if free:
            return free.pop()
    return _Lane(device)


def _lane_give(lane):
    with _idle_lock:
        free = _idle_lanes.setdefault(str(lane.device), [])
        if len(free) < 16:
            free.append(lane)
            return
    lane.jobs.put(None)


class _Graph:
    """An instantiated graph of libksm's recorder."""

    __slots__ = ("lib", "h")

    def __init__(self, lib, h):
        self.lib, self.h = lib, h

    def launch(self, device):
        _lib.check(self.lib.ksm_graph_launch(self.h, _lib.stream_of(device)), "ksm_graph_launch")

    def destroy(self):
        if self.h:
            h, self.h = self.h, None
            _lib.check(self.lib.ksm_graph_destroy(h), "ksm_graph_destroy")

    def __del__(self):
        # a launch still in flight is freed by the runtime when it completes
        if self.h:
            self.lib.ksm_graph_destroy(self.h)
            self.h = None


class _GraphStep:
    """One recording's batches on one lane as launches of recorded graphs.

    The batch's work -- residual copy, projection (BF16 forms: with its error
    bound), the first ``peels`` iterations of the peel (``ksm_peel_range``),
    the features of the spikes the device counted (``ksm_features_counted``)
    and the copy of the peel's info words to pinned host memory -- is
    recorded once by the library into a HIP graph over static buffers
    (``ksm_graph_begin``/``ksm_graph_end``: nodes are added directly, no
    stream capture, so the other lanes' threads are unaffected) and launched
    per batch, one launch instead of a hundred.  Where a batch needs more
    iterations than the graph holds, tail graphs of ``_TAIL`` iterations
    (plus features) continue the same peel.  The first batch runs the same kernels eagerly
    (the peel with its host-polled end) and its iteration count picks
    ``peels`` unless ``ops['matching_graph_peels']`` gives it.  Results are
    those of run_matching + features (same kernels, same order).
    """

    _TAIL = 8

    def __init__(self, ctx, C, T, device):
        ops, U = ctx.ops, _check_U(ctx.U, device)
        self.ctx, self.device = ctx, device
        self.lib = lib = _lib.load()
        self.U = U
        K, P, _ = U.shape
        self.K, self.P, self.C, self.T = K, P, C, T
        self.nt = nt = int(ops['nt'])
        self.Th = float(ops['Th_learned'])
        self.max_peels = int(ops['max_peels'])
        self.W = _f32(ops['wPCA'], device, "wPCA")
        self.prec = _precision(ops)
        self.tile = _tile_m(ops, K)
        self.g = _grams.get(ctx.ctc, U, nt, device)
        stream = _lib.stream_of(device)
        f32 = dict(dtype=torch.float32, device=device)
        self.nm = torch.empty((K,), **f32)
        _lib.check(lib.ksm_template_norms(U.data_ptr(), K, P, C, self.nm.data_ptr(), stream),
                   "ksm_template_norms")
        self.ldbt = (K + 63) // 64 * 64
        rows_t = (T + 127) // 128 * 128
        self.X = torch.empty((C, T), **f32)
        self.Xres = torch.empty((C, T), **f32)
        self.Bt = torch.empty((rows_t, self.ldbt), **f32)
        self.pbytes = lib.ksm_project_workspace(C, T, K, P, nt, self.prec)
        self.pws = torch.empty((max(self.pbytes, 1),), dtype=torch.uint8, device=device)
        self.bbytes = 0 if self.prec == _lib.PREC_FP32 else lib.ksm_bound_bytes(C, T, K, P, nt)
        self.bound = (torch.empty((self.bbytes,), dtype=torch.uint8, device=device)
                      if self.bbytes else None)
        self.kbytes = lib.ksm_peel_workspace(K, T, self.max_peels)
        self.kws = torch.empty((self.kbytes,), dtype=torch.uint8, device=device)
        self.st = torch.empty((ST_CAPACITY, 2), dtype=torch.int64, device=device)
        self.amps = torch.empty((ST_CAPACITY, 1), **f32)
        self.th = torch.empty((ST_CAPACITY, 1), **f32)
        self.info = torch.zeros((4,), dtype=torch.int32, device=device)
        self.info_h = torch.zeros((4,), dtype=torch.int32, pin_memory=True)
        nC = ctx.Ucc.shape[0]
        self.nC = nC
        self.tF = torch.empty((ST_CAPACITY, nC, P), **f32)
        self.graph = None
        self.peels = ops.get('matching_graph_peels')
        self.tails = {}

    # -- the work of one batch, enqueued on the current stream ---------------
    def _project(self):
        lib, g, stream = self.lib, self.g, _lib.stream_of(self.device)
        _lib.check(lib.ksm_copy_async(self.Xres.data_ptr(), self.X.data_ptr(),
                                      self.X.numel() * 4, _lib.COPY_D2D, stream), "ksm_copy_async")
        a = (self.X.data_ptr(), self.T, self.W.data_ptr(), self.U.data_ptr(), g.perm.data_ptr(),
             self.C, self.T, self.K, self.P, self.nt, self.Bt.data_ptr(), self.ldbt, self.prec)
        if self.bound is None:
            _lib.check(lib.ksm_project(*a, self.pws.data_ptr(), self.pbytes, stream),
                       "ksm_project")
        else:
            _lib.check(lib.ksm_project_bounded(*a, self.tile, self.bound.data_ptr(), self.bbytes,
                                               self.pws.data_ptr(), self.pbytes, stream),
                       "ksm_project_bounded")

    def _peel_args(self):
        g = self.g
        flags = 0 if self.bound is None else _lib.PEEL_RESCORE
        return (self.Bt.data_ptr(), self.ldbt, self.Xres.data_ptr(), self.T, self.U.data_ptr(),
                self.W.data_ptr(), self.nm.data_ptr(), g.perm.data_ptr(), g.rank.data_ptr(),
                g.cc.data_ptr(), g.lohi.data_ptr(), g.off.data_ptr(), self.K, self.P, self.C,
                self.T, self.nt, self.Th, self.max_peels, flags,
                None if self.bound is None else self.bound.data_ptr(), self.st.data_ptr(),
                self.amps.data_ptr(), self.th.data_ptr(), ST_CAPACITY, self.info.data_ptr(),
                self.kws.data_ptr(), self.kbytes)

    def _features(self):
        c, stream = self.ctx, _lib.stream_of(self.device)
        _lib.check(self.lib.ksm_features_counted(
            self.Xres.data_ptr(), self.T, self.W.data_ptr(), self.st.data_ptr(),
            self.amps.data_ptr(), self.info.data_ptr(), ST_CAPACITY, c.iCC.data_ptr(),
            c.iU.data_ptr(), c.Ucc.data_ptr(), self.nC, self.K, self.P, self.C, self.nt,
            self.tF.data_ptr(), stream), "ksm_features_counted")
        _lib.check(self.lib.ksm_copy_async(self.info_h.data_ptr(), self.info.data_ptr(), 16,
                                           _lib.COPY_D2H, stream), "ksm_copy_async")

    def _range(self, it0, it1):
        _lib.check(self.lib.ksm_peel_range(*self._peel_args(), it0, it1,
                                           _lib.stream_of(self.device)), "ksm_peel_range")

    def _capture(self, fn):
        """Records fn's library calls on the current stream into a graph."""
        lib, stream = self.lib, _lib.stream_of(self.device)
        _lib.check(lib.ksm_graph_begin(stream), "ksm_graph_begin")
        h = ctypes.c_void_p()
        try:
            fn()
        except BaseException:
            lib.ksm_graph_end(stream, ctypes.byref(h))      # discards the recording
            raise
        _lib.check(lib.ksm_graph_end(stream, ctypes.byref(h)), "ksm_graph_end")
        return _Graph(lib, h.value)

    def release(self):
        """Frees the recorded graphs (no launch of them may be pending)."""
        for g in [self.graph, *self.tails.values()]:
            if g is not None:
                g.destroy()
        self.graph = None
        self.tails = {}

    def _wait(self):
        ev = torch.cuda.Event()
        ev.record()
        ev.synchronize()
        k, status, iters, rescored = (int(v) for v in self.info_h)
        if status == _lib.KSM_ERR_CAPACITY:
            raise RuntimeError(f"more than {ST_CAPACITY} spikes in one batch "
                               "(the reference's per-batch buffer overflows too)")
        _lib.check(status, "ksm_peel (device)")
        return k, iters, rescored

    def run(self, X):
        """Matching + features of batch X (a device tensor [C, T] ready on the
        current stream).  Returns (k, iterations, st, th_amps, tF) with st,
        th_amps and tF views of the static buffers (valid until the next run)."""
        self.X.copy_(X)
        if self.graph is None:
            # first batch: eagerly, with the host-polled end of the peel
            self._project()
            _lib.check(self.lib.ksm_peel(*self._peel_args(), _lib.stream_of(self.device)),
                       "ksm_peel")
            self._features()
            k, iters, _ = self._wait()
            if self.peels is None:
                self.peels = min(self.max_peels, (iters + 8 + 3) // 4 * 4)
            n = int(self.peels)

            def head():
                self._project()
                self._range(0, n)
                self._features()
            self.graph = self._capture(head)
            return k, iters, self.st[:k], self.th[:k], self.tF[:k]
        self.graph.launch(self.device)
        k, iters, _ = self._wait()
        it = int(self.peels)
        # more iterations than the graph holds: continue the same peel
        while iters >= it and it < self.max_peels:
            it1 = min(self.max_peels, it + self._TAIL)
            tail = self.tails.get(it)
            if tail is None:
                def fn(a=it, b=it1):
                    self._range(a, b)
                    self._features()
                tail = self._capture(fn)
                self.tails[it] = tail
            tail.launch(self.device)
            k, iters, _ = self._wait()
            it = it1
        return k, iters, self.st[:k], self.th[:k], self.tF[:k]


_LANE_SLOTS = 3                 # pinned result buffer sets per lane


class _Lane:
    """One device stream of a _BatchPipeline and the host thread that drives
    it: run_matching + features per job, then the host copies of (st,
    th_amps) into one of its pinned buffer sets; the features stay on the
    device."""

    def __init__(self, device):
        self.device = device
        self.stream = torch.cuda.Stream(device=device)
        self.jobs = queue.Queue()
        self.results = queue.Queue()
        self.bufs = [None] * _LANE_SLOTS
        self.graph_steps = {}
        self.thread = threading.Thread(target=self._run, daemon=True)
        self.thread.start()

    def drop_graphs(self):
        """Releases this lane's captured graphs and their static buffers."""
        steps, self.graph_steps = self.graph_steps, {}
        for step in steps.values():
            step.release()

    def _host(self, slot, k):
        b = self.bufs[slot]
        if b is None or b[0].shape[0] < k:
            cap = max(2 * k, 1 << 15)
            b = (torch.empty((cap, 2), dtype=torch.int64, pin_memory=True),
                 torch.empty((cap, 1), dtype=torch.float32, pin_memory=True))
            self.bufs[slot] = b
        return b

    def _run(self):
        dev = self.device
        with torch.cuda.device(dev), torch.cuda.stream(self.stream):
            while True:
                job = self.jobs.get()
                if job is None:
                    return
                slot, X, tag, ready, c, graphed = job
                try:
                    self.stream.wait_event(ready)
                    if graphed:
                        step = self.graph_steps.get(id(c))
                        if step is None:
                            # one recording's graph steps at a time per lane
                            self.drop_graphs()
                            step = _GraphStep(c, X.shape[0], X.shape[1], dev)
                            self.graph_steps = {id(c): step}
                        k, iters, stt, th, xf = step.run(X)
                        xfeat = xf.clone()
                    else:
                        stt, amps, th, Xres = run_matching(c.ops, X, c.U, c.ctc, device=dev)
                        iters = _iters.n
                        xfeat = _features(c.ops, Xres, stt, amps, c.iCC, c.iU, c.Ucc, dev)
                        k = stt.shape[0]
                    hb = self._host(slot, k)
                    if k:
                        hb[0][:k].copy_(stt, non_blocking=True)
                        hb[1][:k].copy_(th, non_blocking=True)
                    done = torch.cuda.Event()
                    done.record(self.stream)
                    done.synchronize()
                    del X, stt, th
                    self.results.put((tag, hb[0][:k], hb[1][:k], xfeat, iters))
                except BaseException as e:      # handed to the submitting thread
                    self.results.put(e)


class _BatchPipeline:
    """run_matching + PC features over a sequence of batches, on ``lanes``
    device streams at once with one host thread each (the peel's host loop
    follows its stream's progress), so that one batch's latency-bound late
    peel iterations and small kernels overlap other batches' work.  Batch i
    runs on lane i mod lanes, two batches deep; each lane copies its results
    to pinned host buffers, three sets per lane.

    ``submit(X, tag, ctx)`` queues batch X (a device tensor made on the
    caller's current stream) of the recording ``ctx`` (default: the one the
    pipeline was made with) and returns the batches finished so far that the
    lanes need back, in submission order, as a list of ``(tag, st int64
    [k, 2], th_amps float32 [k, 1], tF float32 [k, nC, P])`` -- st and
    th_amps views into pinned buffers that stay valid until the next
    ``submit`` or ``flush`` returns, tF a device tensor (completed); ``flush()`` returns all remaining batches and ``close()`` hands
    the lanes (threads, streams, pinned buffers) back to a per-device pool for
    later pipelines.  ``peel_iterations`` sums the peel iterations of the
    batches returned.
    """

    def __init__(self, ops, U, ctc, iCC, iU, Ucc, device, lanes=None):
        self.ctx = _Ctx(ops, U, ctc, iCC, iU, Ucc) if ops is not None else None
        self.device = device
        if self.ctx is not None:
            self.prepare(self.ctx)
        n = MATCHING_STREAMS if lanes is None else int(lanes)
        if n < 1:
            raise ValueError("lanes must be >= 1")
        self.lanes = [_lane_take(device) for _ in range(n)]
        self.n_sub = 0
        self.n_done = 0
        self.peel_iterations = 0

    def prepare(self, ctx):
        """Builds the recording's derived Gram on the caller's stream (lanes
        wait for it) before its batches are submitted."""
        _grams.get(ctx.ctc, _check_U(ctx.U, self.device), int(ctx.ops['nt']), self.device)

    def _take(self):
        lane = self.lanes[self.n_done % len(self.lanes)]
        res = lane.results.get()
        self.n_done += 1
        if isinstance(res, BaseException):
            raise res
        tag, st, th, tf, iters = res
        self.peel_iterations += iters
        return tag, st, th, tf

    def submit(self, X, tag=None, ctx=None, graphed=False):
        """graphed: run the batch as a replay of the lane's captured graphs of
        this recording (_GraphStep), for recordings with several batches per
        lane."""
        ctx = self.ctx if ctx is None else ctx
        n = len(self.lanes)
        i = self.n_sub
        out = []
        # each lane holds up to _LANE_SLOTS - 1 batches (one running, the next
        # queued so that the lane never waits for this thread); batch i takes
        # its lane's host buffer set (i // n) % _LANE_SLOTS, last used by batch
        # i - _LANE_SLOTS n, which was handed out at least one submit ago
        while self.n_done <= i - (_LANE_SLOTS - 1) * n:
            out.append(self._take())
        ready = torch.cuda.Event()
        ready.record(torch.cuda.current_stream(self.device))
        self.lanes[i % n].jobs.put(((i // n) % _LANE_SLOTS, X, tag, ready, ctx, bool(graphed)))
        self.n_sub += 1
        return out

    def flush(self):
        out = []
        while self.n_done < self.n_sub:
            out.append(self._take())
        return out

    def close(self):
        """Waits for the batches still in flight (results dropped) and hands
        the lanes back for reuse."""
        while self.n_done < self.n_sub:
            lane = self.lanes[self.n_done % len(self.lanes)]
            lane.results.get()
            self.n_done += 1
        for lane in self.lanes:
            lane.drop_graphs()           # the lane is idle: its graphs and buffers go
            _lane_give(lane)
        self.lanes = []

    def __del__(self):
        if getattr(self, "lanes", None):
            try:
                self.close()
            except Exception:
                pass


def _extract_batch(ops, X, U, ctc, iCC, iU, Ucc, device):
    """run_matching + features of one padded batch; host copies of
    (st int64 [k, 2], th_amps float32 [k, 1], tF float32 [k, nC, P])."""
    stt, amps, th_amps, Xres = run_matching(ops, X, U, ctc, device=device)
    xfeat = _features(ops, Xres, stt, amps, iCC, iU, Ucc, device)
    return stt.cpu(), th_amps.cpu(), xfeat.cpu()


class _Collector:
    """A recording's outputs as ``extract`` assembles them (reference
    ``template_matching.py:88-124``): batch-0 spikes before the recording
    start dropped, float64 times shifted by the batch offset, and a final sort
    by time with numpy's default argsort over the rows in batch order.  Rows
    are kept per batch and joined at the end instead of written into the
    reference's preallocated 10^6-row buffers (whose zero-filled pages cost
    more than a small recording's matching); the features stay on the device
    and come back in one copy, in the final order; the returned arrays are
    the same."""

    def __init__(self, ops, bfile, nC, P, progress_bar=None):
        self.ops, self.bfile = ops, bfile
        self.nC, self.P = nC, P
        self.st_parts, self.tf_parts = [], []
        self.n_batches = int(bfile.n_batches)
        self.progress_bar = progress_bar
        self.last = None                  # host st of the last batch (error reports)

    def add(self, jb, stt_h, th_h, xfeat_d):
        ops = self.ops
        nt = ops['nt']
        self.last = stt_h
        x = stt_h[:, 0].numpy()
        # the features stay on the device until result(); the caller's stream
        # reads them there
        if xfeat_d.numel():
            xfeat_d.record_stream(torch.cuda.current_stream(xfeat_d.device))
        if jb == 0:
            # times before the recording start (trough alignment) are dropped
            keep = (x - nt - nt // 2 + ops['nt0min']) >= 0
            stt_h, th_h = stt_h[keep], th_h[keep]
            xfeat_d = xfeat_d[torch.from_numpy(keep).to(xfeat_d.device)]
        nsp = len(stt_h)
        t_shift = jb * self.bfile.batch_downsampling * (ops['batch_size'])
        sd = stt_h.numpy().astype(np.float64)
        st = np.empty((nsp, 3), np.float64)
        st[:, 0] = ((sd[:, 0] - nt) + t_shift) - nt // 2 + ops['nt0min']
        st[:, 1] = sd[:, 1]
        st[:, 2] = th_h.numpy().astype(np.float64).reshape(-1)
        self.st_parts.append(st)
        self.tf_parts.append(xfeat_d)
        if self.progress_bar is not None:
            self.progress_bar.emit(int((jb + 1) / self.n_batches * 100))

    def result(self):
        st = (np.concatenate(self.st_parts, 0) if self.st_parts
              else np.zeros((0, 3), np.float64))
        isort = np.argsort(st[:, 0])
        if self.tf_parts:
            dev = self.tf_parts[0].device
            tf_d = torch.cat(self.tf_parts, 0)[torch.from_numpy(isort).to(dev)]
            # into page-locked host memory: a copy to pageable memory runs at
            # a small fraction of the link's rate (1.1 GB: 290 ms against 20)
            tF = torch.empty(tf_d.shape, dtype=tf_d.dtype, pin_memory=True)
            tF.copy_(tf_d)
        else:
            tF = torch.zeros((0, self.nC, self.P))
        self.tf_parts = []
        return np.take(st, isort, axis=0), tF      # st[isort]; take gathers rows faster


def _prepare_recording(ops, U, device):
    """prepare_extract + prepare_matching of one recording; writes
    ops['iCC'], ops['iCC_mask'], ops['iU'] (reference :62-67)."""
    nC = ops['settings']['nearest_chans']
    position_limit = ops['settings']['position_limit']
    U = _check_U(U, device)
    iCC, iCC_mask, iU, Ucc = prepare_extract(ops['xc'], ops['yc'], U, nC, position_limit,
                                             device=device)
    ops['iCC'] = iCC
    ops['iCC_mask'] = iCC_mask
    ops['iU'] = iU
    ctc = prepare_matching(ops, U)
    return _Ctx(ops, U, ctc, iCC, iU, Ucc)


def extract(ops, bfile, U, device=torch.device('cuda'), progress_bar=None):
    """Spike extraction over every batch of ``bfile`` (reference ``template_matching.py:56-126``).

    Returns ``(st, tF, ops)``: ``st`` float64 [N, 3] = (time in samples,
    template id, th_amp) sorted by time with numpy's default argsort, ``tF``
    float32 [N, nC, n_pcs] on the CPU; ``ops`` gains ``iCC``, ``iCC_mask``,
    ``iU``.
    """
    device = _lib.require_gpu(device)
    ctx = _prepare_recording(ops, U, device)
    col = _Collector(ops, bfile, ctx.iCC.shape[0], ops['settings']['n_pcs'], progress_bar)
    ibatch, X = -1, None
    pipe = _BatchPipeline(ctx.ops, ctx.U, ctx.ctc, ctx.iCC, ctx.iU, ctx.Ucc, device,
                          lanes=_streams(ops))
    try:
        for ibatch in range(col.n_batches):
            if ibatch % 100 == 0:
                logger.debug("extract: batch %d of %d", ibatch, col.n_batches)
            X = bfile.padded_batch_to_torch(ibatch, ops)
            for res in pipe.submit(X, tag=ibatch):
                col.add(*res)
            X = None
        for res in pipe.flush():
            col.add(*res)
    except Exception:
        logger.exception(f'Error in template_matching.extract on batch {ibatch}')
        logger.debug(f'X shape: {None if X is None else tuple(X.shape)}')
        logger.debug(f'stt shape: {None if col.last is None else tuple(col.last.shape)}')
        raise
    finally:
        pipe.close()
    st, tF = col.result()
    return st, tF, ops


def _streams(ops):
    n = ops.get('matching_streams', MATCHING_STREAMS)
    return MATCHING_STREAMS if n is None else int(n)


def extract_many(jobs, device=torch.device('cuda'), streams=None, graphs=None):
    """``extract`` over several recordings of any sizes at once.

    ``jobs`` is a list of ``(ops, bfile, U)``; returns a list of ``(st, tF,
    ops)``, each exactly what ``extract(ops, bfile, U)`` returns.  All
    recordings' batches go through one batch pipeline (recording by
    recording, batch by batch), so that small recordings -- whose batches
    leave most of the GPU idle -- share the device streams with their
    neighbours instead of running one after another.  ``streams`` defaults to
    ``MATCHING_STREAMS``.

    ``graphs``: run the batches of a recording as replays of captured HIP
    graphs (``_GraphStep``) where the recording has at least two batches per
    stream (True: always; False: never; None: that rule).  Per-recording
    choices of an autotuner (``autotune.tune_many``) come in each ``ops``:
    ``matching_tile_m`` (GEMM tile of the BF16 forms),
    ``matching_graph_peels`` (peel iterations in the captured graph),
    ``matching_graphs``.
    """
    device = _lib.require_gpu(device)
    cols = [_Collector(ops, bfile, ops['settings']['nearest_chans'], ops['settings']['n_pcs'])
            for ops, bfile, _ in jobs]
    lanes = MATCHING_STREAMS if streams is None else int(streams)
    pipe = _BatchPipeline(None, None, None, None, None, None, device, lanes=lanes)
    r, ibatch = -1, -1
    try:
        for r, (ops, bfile, U) in enumerate(jobs):
            # prepared when its turn comes: only the recordings in flight hold
            # their Gram on the device
            c = _prepare_recording(ops, U, device)
            pipe.prepare(c)
            nb = cols[r].n_batches
            g = ops.get('matching_graphs', graphs)
            graphed = (nb >= 2 * lanes) if g is None else bool(g)
            for ibatch in range(nb):
                X = bfile.padded_batch_to_torch(ibatch, ops)
                for tag, *res in pipe.submit(X, tag=(r, ibatch), ctx=c, graphed=graphed):
                    cols[tag[0]].add(tag[1], *res)
                X = None
        for tag, *res in pipe.flush():
            cols[tag[0]].add(tag[1], *res)
    except Exception:
        logger.exception(f'Error in template_matching.extract_many on recording {r}, '
                         f'batch {ibatch}')
        raise
    finally:
        pipe.close()
    out = []
    for (ops, _, _), col in zip(jobs, cols):
        st, tF = col.result()
        out.append((st, tF, ops))
    return out
