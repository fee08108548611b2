"""CPU restatement of Kilosort4's spike positions for the output files (test
oracle).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``; never part of the product
path.  Same torch operations in the same order as
``kilosort/postprocessing.py:33-62`` (``compute_spike_positions``).
"""
from __future__ import annotations

import torch


def compute_spike_positions(st, tF, ops):
    """``postprocessing.py:33-62``: (xs, ys) numpy arrays."""
    cpu = torch.device('cpu')
    icc_mask = ops['iCC_mask'].to(cpu)
    iU = ops['iU'].to(cpu)
    iCC = ops['iCC'].to(cpu)
    tmass = torch.norm(tF, 2, dim=-1)
    tmask = icc_mask[:, iU[st[:, 1]]].T
    tmass = tmass * tmask
    tmass = tmass / tmass.sum(1, keepdim=True)
    xc = torch.from_numpy(ops['xc'])
    yc = torch.from_numpy(ops['yc'])
    chs = iCC[:, iU[st[:, 1]]]
    xc0 = xc[chs.T]
    yc0 = yc[chs.T]
    xs = (xc0 * tmass).sum(1).numpy()
    ys = (yc0 * tmass).sum(1).numpy()
    return xs, ys
