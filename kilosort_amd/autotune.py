"""Picks the matching path's schedule parameters by measurement.

* Streams: how many batches run concurrently (``MATCHING_STREAMS``, one host
  thread and one device stream each).  Large batches keep the GPU busy with a
  few streams; many small recordings need more to hide per-batch latency.
  ``tune_streams`` times ``extract_many`` over a sample of the caller's
  recordings for each candidate count and caches the fastest per workload
  signature.
* Tile shape: template rows per tile of the BF16 forms' projection GEMM (64 or
  128; small template counts waste less of a 64-row tile, and finer tiles skip
  more all-zero reduction steps), timed per recording shape (``tune_tile``).
* Peel schedule: how many peel iterations the captured graph of a
  graph-replayed recording holds (``tune_peels``).

``tune_many`` applies all three to a list of recordings for ``extract_many``.
"""
from __future__ import annotations

import time

import torch

from . import template_matching as tm

_cache: dict = {}


def signature(jobs):
    """Workload key: per recording (channels, templates, batches, precision)."""
    out = []
    for ops, bfile, U in jobs:
        out.append((int(U.shape[2]), int(U.shape[0]), int(bfile.n_batches),
                    str(ops.get('matching_precision', 'fp32'))))
    return tuple(sorted(out))


def tune_streams(jobs, device=torch.device('cuda'), candidates=(1, 2, 3, 4, 6), reps=1):
    """Returns (best stream count, {count: seconds}) for ``extract_many`` over
    ``jobs`` (list of (ops, bfile, U)); results are cached by ``signature``.
    Each candidate runs once untimed (compilation, caches) and ``reps`` times
    timed."""
    key = (signature(jobs), tuple(candidates))
    if key in _cache:
        return _cache[key]
    times = {}
    for n in candidates:
        tm.extract_many([(dict(o), b, U) for o, b, U in jobs], device=device, streams=n)
        torch.cuda.synchronize(device)
        t0 = time.perf_counter()
        for _ in range(reps):
            tm.extract_many([(dict(o), b, U) for o, b, U in jobs], device=device, streams=n)
        torch.cuda.synchronize(device)
        times[n] = (time.perf_counter() - t0) / reps
    best = min(times, key=times.get)
    _cache[key] = (best, times)
    return best, times


# ---------------------------------------------------------------------------
# The batched form's choices: GEMM tile height, peel schedule, streams.

def _first_batch(ops, bfile, device):
    X = bfile.padded_batch_to_torch(0, ops)
    return X if X.device == device else X.to(device)


def tune_tile(ops, X, U, device=torch.device('cuda'), reps=3):
    """Template rows per tile of the BF16 forms' projection GEMM (64 or 128)
    for this recording's shape, by timing the bounded projection of one batch
    with each; cached per (channels, templates, precision).  Returns (best,
    {tile: ms})."""
    from . import _lib
    prec_name = str(ops.get('matching_precision', 'fp32'))
    K, P, C = U.shape
    key = ("tile", C, K, prec_name)
    if key in _cache:
        return _cache[key]
    prec = tm._precision(ops)
    if prec == _lib.PREC_FP32:
        _cache[key] = (0, {})
        return _cache[key]
    lib = _lib.load()
    nt = int(ops['nt'])
    T = X.shape[1]
    W = tm._f32(ops['wPCA'], device, "wPCA")
    ldbt = (K + 63) // 64 * 64
    Bt = torch.empty(((T + 127) // 128 * 128, ldbt), dtype=torch.float32, device=device)
    nb = lib.ksm_project_workspace(C, T, K, P, nt, prec)
    ws = torch.empty((nb,), dtype=torch.uint8, device=device)
    nbb = lib.ksm_bound_bytes(C, T, K, P, nt)
    bb = torch.empty((nbb,), dtype=torch.uint8, device=device)
    stream = _lib.stream_of(device)
    times = {}
    for tile in (128, 64):
        def once():
            _lib.check(lib.ksm_project_bounded(X.data_ptr(), T, W.data_ptr(), U.data_ptr(), None,
                                               C, T, K, P, nt, Bt.data_ptr(), ldbt, prec, tile,
                                               bb.data_ptr(), nbb, ws.data_ptr(), nb, stream),
                       "ksm_project_bounded")
        once()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            once()
        e1.record()
        e1.synchronize()
        times[tile] = e0.elapsed_time(e1) / reps
    best = min(times, key=times.get)
    _cache[key] = (best, times)
    return best, times


def tune_peels(ops, X, U, device=torch.device('cuda'), margin=6):
    """Peel schedule of the graph-replayed batches: the peel iterations the
    captured graph holds, from one batch's measured iteration count plus a
    margin (a batch that needs more continues with tail graphs; every
    iteration the graph holds beyond the end costs two launches that return at
    once).  Returns the count."""
    ctc = tm.prepare_matching(ops, U)
    tm.run_matching(ops, X, U, ctc, device=device)
    iters = int(tm._iters.n)
    return int(min(int(ops['max_peels']), (iters + margin + 1) // 2 * 2))


def tune_many(jobs, device=torch.device('cuda'), stream_candidates=(2, 3, 4),
              sample_recordings=6, sample_batches=2):
    """Autotuner of the batched many-recordings form (``extract_many``).

    Per recording (ops, bfile, U): the GEMM tile height of the BF16 forms
    (``tune_tile``, timed per shape), the peel schedule of its captured graph
    (``tune_peels``, from its first batch) and whether its batches run as
    graph replays (two or more batches per stream); for the whole job list,
    the stream count (``tune_streams`` over a sample: the first
    ``sample_recordings`` recordings, at most ``sample_batches`` batches each).
    The choices are written into each ``ops`` (``matching_tile_m``,
    ``matching_graph_peels``, ``matching_graphs``) and returned with the
    stream count: (streams, [per-recording dict], {streams: seconds})."""
    plans = []
    for ops, bfile, U in jobs:
        U = tm._check_U(U, device)
        X = _first_batch(ops, bfile, device)
        tile, _ = tune_tile(ops, X, U, device)
        if tile:
            ops['matching_tile_m'] = tile
        ops['matching_graph_peels'] = tune_peels(ops, X, U, device)
        plans.append(dict(tile_m=tile, graph_peels=ops['matching_graph_peels']))
        del X
    sample = []
    for ops, bfile, U in jobs[:sample_recordings]:
        sample.append((dict(ops), _Prefix(bfile, sample_batches), U))
    streams, times = tune_streams(sample, device=device, candidates=stream_candidates)
    for (ops, bfile, U), plan in zip(jobs, plans):
        graphed = int(bfile.n_batches) >= 2 * streams
        ops['matching_graphs'] = graphed
        plan['graphs'] = graphed
    return streams, plans, times


class _Prefix:
    """The first n batches of a batch source."""

    def __init__(self, bfile, n):
        self.bfile = bfile
        self.n_batches = min(int(bfile.n_batches), int(n))
        self.batch_downsampling = getattr(bfile, 'batch_downsampling', 1)

    def padded_batch_to_torch(self, ibatch, ops=None):
        return self.bfile.padded_batch_to_torch(ibatch, ops)
