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
    ops = dict(nt=nt, wPCA=W)
    Ww, clu2, _, st2, tF2 = TO.merging_function(ops, Wall, clu, st, tF, 0.9)
    assert Ww.shape[0] == 2
    np.testing.assert_allclose(Ww[0].numpy(), (5 * w0 + 3 * 1.1 * w0) / 8, atol=1e-12)
    order = np.argsort(st[:, 0])
    np.testing.assert_array_equal(st2[:, 0], np.sort(st[:, 0]))
    np.testing.assert_array_equal(clu2, np.where(clu == 2, 1, 0)[order])
    np.testing.assert_array_equal(tF2.numpy(), tF.numpy()[order])
