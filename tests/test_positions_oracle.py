"""Pins of the spike-position oracle (CPU) by cases the weighted mean fixes:
features on one channel put the spike on that channel, equal norms on two
channels put it half way, masked channels weigh nothing, and a brute-force
loop over spikes and channels on random features."""
import numpy as np
import torch

from oracle import positions_oracle as PO


def layout(K=4, nC=3, C=8, seed=0):
    rng = np.random.default_rng(seed)
    xc = rng.uniform(0, 60, C).astype(np.float32)
    yc = rng.uniform(0, 400, C).astype(np.float32)
    iU = torch.from_numpy(rng.integers(0, C, K))
    iCC = torch.from_numpy(rng.integers(0, C, (nC, C)))
    mask = torch.from_numpy(rng.uniform(size=(nC, C)) < 0.8)
    mask[0] = True
    return dict(xc=xc, yc=yc, iU=iU, iCC=iCC, iCC_mask=mask)


def test_single_channel_and_midpoint():
    ops = layout()
    st = np.array([[0.0, 2], [0.0, 3]])
    tF = torch.zeros((2, 3, 6))
    tF[0, 0, 2] = -5.0                                     # all mass on nearest channel 0
    tF[1, 0, :] = 1.0
    tF[1, 1, :] = -1.0                                     # equal norms on channels 0 and 1
    ops['iCC_mask'][:2] = True
    xs, ys = PO.compute_spike_positions(st, tF, ops)
    c0 = ops['iCC'][0, ops['iU'][2]]
    assert xs[0] == ops['xc'][c0] and ys[0] == ops['yc'][c0]
    u = ops['iU'][3]
    a, b = ops['iCC'][0, u], ops['iCC'][1, u]
    np.testing.assert_allclose(xs[1], (ops['xc'][a] + ops['xc'][b]) / 2, rtol=1e-6)
    np.testing.assert_allclose(ys[1], (ops['yc'][a] + ops['yc'][b]) / 2, rtol=1e-6)


def test_masked_channel_weighs_nothing():
    ops = layout()
    ops['iCC_mask'][:] = True
    u = int(ops['iU'][1])
    ops['iCC_mask'][2, u] = False
    st = np.array([[0.0, 1]])
    tF = torch.zeros((1, 3, 6))
    tF[0, 0, 0] = 3.0
    tF[0, 2, 0] = 1e6                                      # huge but masked out
    xs, ys = PO.compute_spike_positions(st, tF, ops)
    c0 = ops['iCC'][0, u]
    assert xs[0] == ops['xc'][c0] and ys[0] == ops['yc'][c0]


def test_brute_force_loop():
    ops = layout(K=7, nC=5, C=12, seed=3)
    rng = np.random.default_rng(4)
    n = 50
    st = np.zeros((n, 6))
    st[:, 1] = rng.integers(0, 7, n)
    tF = torch.from_numpy(rng.standard_normal((n, 5, 6)).astype(np.float32))
    xs, ys = PO.compute_spike_positions(st, tF, ops)
    for s in range(n):
        u = int(ops['iU'][int(st[s, 1])])
        w = np.array([np.sqrt((tF[s, j].double() ** 2).sum().item()) * bool(ops['iCC_mask'][j, u])
                      for j in range(5)])
        w /= w.sum()
        ch = ops['iCC'][:, u].numpy()
        np.testing.assert_allclose(xs[s], (w * ops['xc'][ch]).sum(), rtol=1e-5)
        np.testing.assert_allclose(ys[s], (w * ops['yc'][ch]).sum(), rtol=1e-5)
