"""Pins of the universal-template detection oracle (CPU): template_match
against a brute-force restatement of its definition, a single planted spike
with closed-form amplitude, position, features and y, and the template grid."""
import numpy as np
import pytest
import torch

from oracle import detect_oracle as DO


def probe(C, pitch=20.0):
    c = np.arange(C)
    xc = np.where(c % 2 == 0, 16.0, 48.0).astype(np.float32)
    yc = (pitch * (c // 2) + 0.5 * pitch * (c % 2)).astype(np.float32)
    return xc, yc


def make_ops(C, nk=6, nt=61, seed=0, nC2=100, dtype=torch.float64):
    xc, yc = probe(C)
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((nk, nt))
    w /= np.sqrt((w ** 2).sum(1, keepdims=True))
    ops = dict(xc=xc, yc=yc, kcoords=np.zeros(C), nt=nt, Th_universal=9.0,
               max_channel_distance=32.0, fs=30000.0, batch_size=1000,
               wTEMP=torch.from_numpy(w).to(dtype),
               wPCA=torch.from_numpy(np.linalg.qr(rng.standard_normal((nt, 6)))[0].T.copy()).to(dtype),
               settings=dict(dmin=None, dminx=32.0, nearest_chans=10, nearest_templates=nC2,
                             min_template_size=10.0, template_sizes=5, nt0min=20,
                             n_templates=nk))
    return ops


def brute_template_match(X, ops, iC, iC2, weigh):
    """The definition written out with loops (float64)."""
    X = np.asarray(X, np.float64)
    W = ops['wTEMP'].double().numpy()
    nk, nt = W.shape
    C, T = X.shape
    h = nt // 2
    nt0 = ops['settings']['nt0min']
    B = np.zeros((C, nk, T))
    for m in range(nt):
        lo, hi = max(0, h - m), min(T, T + h - m)
        B[:, :, lo:hi] += X[:, None, lo + m - h:hi + m - h] * W[None, :, m, None]
    iCn, iC2n, wn = iC.numpy(), iC2.numpy(), weigh.double().numpy()
    nsz, nC, Nf = wn.shape
    A = np.zeros((nsz, nk, Nf, T))
    for i in range(nsz):
        for f in range(Nf):
            A[i, :, f, :] = np.einsum('j,jkt->kt', wn[i, :, f], B[iCn[:, f]])
    Aq = A.reshape(nsz * nk, Nf, T)               # q = i * nk + k
    q = np.abs(Aq).argmax(0)                       # first maximum
    As = np.abs(Aq).max(0)
    sgn = np.sign(np.take_along_axis(Aq, q[None], 0)[0])
    Im = (1 + q) * sgn
    Z = As[iC2n].max(0)
    Z[:, :nt] = 0
    Z[:, -nt:] = 0
    Amaxs = np.zeros_like(Z)
    for t in range(T):
        Amaxs[:, t] = Z[:, max(0, t - nt0):min(T, t + nt0 + 1)].max(1)
    xy = np.argwhere((Amaxs == As) & (As > ops['Th_universal']))
    imax = Im[xy[:, 0], xy[:, 1]]
    amp = As[xy[:, 0], xy[:, 1]]
    ssign = np.sign(imax)
    imax = np.abs(imax).astype(np.int64) - 1
    adist = B[iCn[:, xy[:, 0]], imax % nk, xy[:, 1]] * ssign
    return xy, imax, amp, adist


def test_template_match_matches_brute_force():
    C, T = 16, 720                  # the reference chunks columns by 40: keep the last non-empty
    ops = make_ops(C, seed=1)
    iC, iC2, weigh = DO.setup(ops)
    weigh = weigh.double()
    rng = np.random.default_rng(2)
    X = rng.standard_normal((C, T)) * 4.0
    for t0, c0 in [(150, 3), (400, 10), (520, 6)]:      # a few strong events
        X[c0:c0 + 3, t0 - 30:t0 + 31] += 25.0 * ops['wTEMP'][2].numpy()
    xy, imax, amp, adist = DO.template_match(torch.from_numpy(X), ops, iC, iC2, weigh)
    bxy, bimax, bamp, badist = brute_template_match(X, ops, iC, iC2, weigh)
    assert len(bxy) >= 3
    np.testing.assert_array_equal(xy.numpy(), bxy)
    np.testing.assert_array_equal(imax.numpy(), bimax)
    np.testing.assert_allclose(amp.numpy(), bamp, rtol=1e-12)
    np.testing.assert_allclose(adist.numpy(), badist, rtol=1e-12, atol=1e-12)


def test_single_planted_spike_closed_form():
    """X = a * (weights of template f0, size i0, on its channels) (x)
    wTEMP[k0] at t0: the only spike is (f0, t0) with amp = a, imax = i0*nk+k0,
    adist = a * weights, y = the weight-averaged channel y, features =
    a * w_j * <wTEMP[k0], wPCA[p]>."""
    C, T, a = 16, 600, 50.0
    ops = make_ops(C, seed=3)
    iC, iC2, weigh = DO.setup(ops)
    weigh = weigh.double()
    nk, nt = ops['wTEMP'].shape
    f0, i0, k0, t0 = int(iC.shape[1] // 2), 2, 4, 300
    w = weigh[i0, :, f0].numpy()                   # unit norm over the nC channels
    X = np.zeros((C, T))
    for j in range(iC.shape[0]):
        X[iC[j, f0], t0 - nt // 2:t0 + nt // 2 + 1] += a * w[j] * ops['wTEMP'][k0].numpy()
    st, tF, (xy, imax, amp, adist) = DO.detect_batch(torch.from_numpy(X), ops, iC, iC2, weigh)
    assert xy.tolist() == [[f0, t0]]
    assert int(imax[0]) == i0 * nk + k0
    # sum_j w_j^2 = 1 up to the float32 rounding of the weights
    assert abs(float(amp[0]) - a) <= 1e-6 * a
    np.testing.assert_allclose(adist[:, 0].numpy(), a * w, atol=1e-9)
    yc = ops['yc'][iC[:, f0].numpy()].astype(np.float64)
    assert abs(st[0, 1] - (w * yc).sum() / w.sum()) <= 1e-4
    proj = ops['wTEMP'][k0].numpy() @ ops['wPCA'].numpy().T        # [P]
    np.testing.assert_allclose(tF[0], a * w[:, None] * proj[None, :], atol=1e-9)
    assert st[0, 0] == (t0 - nt) / ops['fs']
    assert st[0, 3] == i0 * nk + k0 and st[0, 5] == f0


def test_template_grid():
    """y every dmin/2 (dmin = the median spacing of the distinct site
    depths), x from the leftmost to the rightmost site in steps of ~dminx/2."""
    ops = make_ops(16)
    DO.template_centers(ops)
    assert ops['dmin'] == 10.0                    # staggered 20 um rows: depths 10 um apart
    np.testing.assert_allclose(ops['yup'], np.arange(0, 150 + .01, 5.0))
    np.testing.assert_allclose(ops['xup'], [16.0, 32.0, 48.0])


def test_snippets_isolated_peaks_closed_form():
    """One isolated peak above threshold gives one clip starting twav_min
    samples before it; a second peak within 6 channels and 30 samples makes
    both non-isolated; a peak below threshold or within nt of the edges is
    not taken."""
    C, T, nt, tw = 20, 400, 61, 20
    X = torch.zeros(C, T, dtype=torch.float64)
    X[5, 200] = -10.0                     # isolated (|X| is used)
    X[15, 150] = 12.0                     # pair within the isolation window
    X[18, 170] = 11.0
    X[2, 300] = 5.0                       # below threshold 6
    X[9, 30] = 20.0                       # within nt of the start
    clips = DO.extract_snippets(X, nt, tw, 6.0)
    assert clips.shape == (1, nt)
    ref = X[5, 200 - tw:200 - tw + nt]
    np.testing.assert_array_equal(clips[0].numpy(), ref.numpy())
    assert clips[0, tw] == -10.0
