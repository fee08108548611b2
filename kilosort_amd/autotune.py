"""Picks the batch pipeline's stream count by measurement.

The matching path has one schedule parameter that depends on the workload:
how many batches run concurrently (``MATCHING_STREAMS``, one host thread and
one device stream each).  Large batches keep the GPU busy with a few streams;
many small recordings need more to hide per-batch latency.  ``tune_streams``
times ``extract_many`` over a sample of the caller's recordings for each
candidate count and caches the fastest per workload signature.
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
