"""Pins of the template-consumer oracle (CPU): similarity against explicit
template cross-correlations, align_U against a loop restatement, and the
'mu' merge on planted duplicates."""
import numpy as np
import torch

from oracle import templates_oracle as TO


def basis(nt=61, P=6, seed=0):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((nt, P)))
    return torch.from_numpy(q.T.copy())                      # orthonormal rows [P, nt]


def test_similarity_is_max_lag_correlation_of_normalised_templates():
    rng = np.random.default_rng(1)
    K, C, P, nt = 5, 7, 6, 61
    W = basis(nt, P)
    Wall = torch.from_numpy(rng.standard_normal((K, C, P)))
    S = TO.similarity(Wall, W, nt)
    Wn = Wall.numpy() / (1e-6 + np.sqrt((Wall.numpy() ** 2).sum((1, 2))))[:, None, None]
    T = np.einsum('kcp,pt->kct', Wn, W.numpy())              # [K, C, nt]
    ref = np.zeros((K, K))
    for i in range(K):
        for j in range(K):
            best = -np.inf
            for d in range(-nt, nt + 1):
                a = T[i][:, max(0, -d):nt - max(0, d)]
                b = T[j][:, max(0, d):nt - max(0, -d)]
                best = max(best, float((a * b).sum()) if a.size else 0.0)
            ref[i, j] = best
    np.testing.assert_allclose(S, ref, atol=1e-9)
    np.testing.assert_allclose(np.diag(S), 1.0, atol=1e-6)   # orthonormal W: unit self-similarity
    np.testing.assert_allclose(S, S.T, atol=1e-12)


def test_align_u_matches_loop_restatement():
    rng = np.random.default_rng(2)
    K, C, P, nt = 6, 5, 6, 61
    wPCA = basis(nt, P, 3)
    wTEMP = basis(nt, 4, 4)
    U = torch.from_numpy(rng.standard_normal((K, C, P)))
    ops = dict(wPCA=wPCA, wTEMP=wTEMP, nt=nt, Nchan=C)
    Unew, imax = TO.align_U(U, ops)
    Y = np.einsum('kcp,pt->kct', U.numpy(), wPCA.numpy())     # [K, C, nt]
    seq = np.concatenate([Y[k] for k in range(K)], axis=1)   # [C, K nt]
    h = nt // 2
    pad = np.pad(seq, ((0, 0), (h, h)))
    corr = np.stack([np.stack([np.correlate(pad[c], wTEMP.numpy()[q], 'valid')
                               for c in range(C)]) for q in range(wTEMP.shape[0])])
    peak = np.abs(corr).max((0, 1)).reshape(K, nt)
    im = peak.argmax(1)
    np.testing.assert_array_equal(imax.numpy(), im)
    for k in range(K):
        s = h - im[k]
        rolled = np.roll(Y[k], s, axis=1)                       # out[t] = in[t - s]
        np.testing.assert_allclose(Unew[k].numpy(), wPCA.numpy() @ rolled.T, atol=1e-10)


def test_mu_merge_of_planted_duplicates():
    """Template 1 is 1.1 x template 0 (similarity 1, norms within 20 %): merged
    into 0 with count weights; template 2 (unrelated) stays; labels remap and
    spikes come back sorted by time."""
    rng = np.random.default_rng(5)
    C, P, nt = 8, 6, 61
    W = basis(nt, P, 6)
    w0 = rng.standard_normal((C, P))
    Wall = torch.from_numpy(np.stack([w0, 1.1 * w0, rng.standard_normal((C, P))]))
    clu = np.array([0] * 5 + [1] * 3 + [2] * 4)
    st = np.stack([rng.permutation(12).astype(float), np.zeros(12)], 1)
    tF = torch.arange(12 * 2, dtype=torch.float64).reshape(12, 2)
    ops = dict(nt=nt, wPCA=W, settings=dict(acg_threshold=0.2, ccg_threshold=0.25))
    Ww, clu2, _, st2, tF2 = TO.merging_function(ops, Wall, clu, st, tF, 0.9)
    assert Ww.shape[0] == 2
    np.testing.assert_allclose(Ww[0].numpy(), (5 * w0 + 3 * 1.1 * w0) / 8, atol=1e-12)
    order = np.argsort(st[:, 0])
    np.testing.assert_array_equal(st2[:, 0], np.sort(st[:, 0]))
    np.testing.assert_array_equal(clu2, np.where(clu == 2, 1, 0)[order])
    np.testing.assert_array_equal(tF2.numpy(), tF.numpy()[order])


def test_roll_features_moves_the_peak_by_dt():
    """A waveform in the span of orthonormal PCs, peaked at t0: after the
    roll, the waveform of the new features peaks at t0 + dt (both signs)."""
    nt, P = 61, 12
    t = np.arange(nt)
    w = np.exp(-0.5 * ((t - 30) / 3.0) ** 2)
    rng = np.random.default_rng(8)
    q, _ = np.linalg.qr(np.column_stack([w, np.roll(w, 4), np.roll(w, -4),
                                         rng.standard_normal((nt, P - 3))]))
    W = torch.from_numpy(q.T.copy())                          # [P, nt], w in the span
    x = torch.from_numpy((w @ q)[None, None, :].copy())      # one spike, one channel
    Wall = torch.from_numpy(np.repeat((w @ q)[None, None, :], 2, 0).copy())
    for dt in (4, -4):
        tF, Wl = TO.roll_features(W, x.clone(), Wall.clone(), np.array([True]), 1, dt)
        assert int(np.argmax((tF[0, 0] @ W).numpy())) == 30 + dt
        assert int(np.argmax((Wl[1, 0] @ W).numpy())) == 30 + dt
        np.testing.assert_array_equal(Wl[0].numpy(), Wall[0].numpy())   # other rows untouched


def test_ccg_mode_merges_a_split_refractory_unit():
    """Two clusters with one template, made by splitting one refractory train,
    are cross-refractory and merge; a third cluster with the same template but
    an independent Poisson train stays apart."""
    rng = np.random.default_rng(9)
    C, P, nt, fs = 8, 6, 61, 30000.0
    W = basis(nt, P, 10)
    w0 = rng.standard_normal((C, P))
    Wall = torch.from_numpy(np.stack([w0, w0, w0]))
    unit = np.cumsum(0.004 + rng.exponential(0.03, 3000))    # refractory, ~30 Hz
    half = rng.uniform(size=unit.size) < 0.5
    other = np.sort(rng.uniform(0, unit.max(), 1500))
    times = np.concatenate([unit, other])
    clu = np.concatenate([np.where(half, 0, 1), np.full(other.size, 2)])
    o = np.argsort(times)
    st = np.stack([np.round(times[o] * fs), np.zeros(o.size)], 1)
    clu = clu[o]
    tF = torch.zeros((o.size, 2))
    ops = dict(nt=nt, wPCA=W, fs=fs, settings=dict(acg_threshold=0.2, ccg_threshold=0.25))
    Ww, clu2, is_ref, st2, _ = TO.merging_function(ops, Wall, clu, st, tF, 0.5, 'ccg',
                                                   check_dt=False)
    assert Ww.shape[0] == 2
    assert len(np.unique(clu2)) == 2
    merged = clu2[np.isin(clu[np.argsort(st[:, 0], kind='stable')], [0, 1])]
    assert len(np.unique(merged)) == 1
