"""Pins of the Phy-export oracle (CPU): duplicate removal on a known answer,
the correlogram loop against brute force over all pairs, refractoriness of
trains with and without a refractory period, and make_pc_features against a
spike-by-spike numpy replay of its in-place updates."""
import numpy as np
import torch

from oracle import phy_oracle as PH


def test_remove_duplicates_known_answer():
    t = np.array([0, 5, 10, 20, 34, 40, 3, 18], np.int64)
    c = np.array([0, 0, 0, 0, 0, 0, 1, 1], np.int32)
    ts, cs, keep = PH.remove_duplicates(t, c, dt=15)
    np.testing.assert_array_equal(keep, [1, 0, 0, 1, 0, 1, 1, 1])
    np.testing.assert_array_equal(ts, [0, 20, 40, 3, 18])
    np.testing.assert_array_equal(cs, [0, 0, 0, 1, 1])


def brute_ccg(st1, st2, tbin, nbins):
    dt = nbins * tbin
    K = np.zeros(2 * nbins + 1)
    for a in st2:
        for b in st1:
            if a - dt < b <= a + dt:
                K[int(np.round((a - b) / tbin)) + nbins] += 1
    return K


def test_compute_ccg_matches_all_pairs():
    rng = np.random.default_rng(0)
    for trial in range(4):
        st1 = np.sort(rng.uniform(0, 3, 120))
        st2 = np.sort(np.concatenate([rng.uniform(0, 3, 90), st1[:5]]))   # shared times
        K, T = PH.compute_CCG(st1, st2, tbin=1 / 1000, nbins=500)
        np.testing.assert_array_equal(K, brute_ccg(st1, st2, 1 / 1000, 500))
        assert T == max(st1.max(), st2.max()) - min(st1.min(), st2.min())
    K, _ = PH.compute_CCG(st1, st1.copy())
    assert K[500] >= len(st1)                                # every spike with itself
    np.testing.assert_array_equal(K, K[::-1])                # an ACG is symmetric


def test_refract_separates_refractory_trains():
    rng = np.random.default_rng(1)
    n = 4000
    poisson = np.sort(rng.uniform(0, 400, n))
    isi = 0.004 + rng.exponential(0.1, n)                    # 4 ms dead time
    refractory = np.cumsum(isi)
    st0 = np.concatenate([poisson, refractory, [1.0] * 5])
    clu = np.concatenate([np.zeros(n, int), np.ones(n, int), [2] * 5])
    is_ref, R12 = PH.refract(clu, st0)
    assert is_ref[1] == 1 and is_ref[0] == 0 and is_ref[2] == 0
    assert R12[1] < 0.05 and 0.5 < R12[0] < 1.5 and R12[2] == 0


def replay_make_pc_features(iC, n_chans, spike_templates, spike_clusters, tF):
    """numpy, spike by spike: for each cluster in ascending order, place each
    spike's current rows at its template's channels, rank the union by the
    norm of the mean, and overwrite the rows of every spike of those
    templates."""
    tF = tF.numpy().copy()
    n, nC, P = tF.shape
    clusters = np.unique(spike_clusters)
    fi = np.zeros((clusters.size, n_chans), np.uint32)
    for i in clusters:
        S = sorted(set(spike_templates[spike_clusters == i].tolist()))
        chans = sorted(set(iC[:, S].ravel().tolist()))
        pos = {c: u for u, c in enumerate(chans)}
        rows = [s for s in range(n) if spike_templates[s] in S]
        dd = np.zeros((len(rows), len(chans), P), np.float32)
        for r, s in enumerate(rows):
            for j in range(nC):
                dd[r, pos[iC[j, spike_templates[s]]]] = tF[s, j]
        norm = np.linalg.norm(dd.astype(np.float64).mean(0), axis=1)
        ind = np.argsort(-norm, kind='stable')[:n_chans]
        for r, s in enumerate(rows):
            tF[s] = dd[r, ind]
        fi[i] = np.array(chans)[ind]
    return tF.transpose(0, 2, 1), fi


def pc_case(seed, n=600, K=9, C=20, nC=4, n_clu=6):
    rng = np.random.default_rng(seed)
    xc = rng.uniform(0, 40, C).astype(np.float32)
    yc = np.arange(C, dtype=np.float32) * 20
    ds = (xc - xc[:, None]) ** 2 + (yc - yc[:, None]) ** 2
    iCC = torch.from_numpy(np.argsort(ds, 0)[:nC])
    iU = torch.from_numpy(rng.integers(0, C, K))
    ops = dict(iU=iU, iCC=iCC, xc=xc, yc=yc, nearest_chans=nC, dmin=20.0, dminx=32.0)
    spike_templates = rng.integers(0, K, n)
    owner = rng.integers(0, n_clu, K)
    spike_clusters = owner[spike_templates]
    split = rng.uniform(size=n) < 0.3                         # templates shared by clusters
    spike_clusters[split] = rng.integers(0, n_clu, split.sum())
    spike_clusters[:n_clu] = np.arange(n_clu)                # every cluster id present
    tF = torch.from_numpy(rng.standard_normal((n, nC, 6)).astype(np.float32))
    tF += torch.from_numpy(rng.uniform(0, 3, (K, nC, 1)).astype(np.float32))[spike_templates]
    return ops, spike_templates, spike_clusters, tF


def test_make_pc_features_matches_replay():
    for seed in range(3):
        ops, stt, clu, tF = pc_case(seed)
        iC = ops['iCC'][:, ops['iU']].numpy()
        want, fi_want = replay_make_pc_features(iC, ops['nearest_chans'], stt, clu, tF)
        got, fi = PH.make_pc_features(ops, stt, clu, tF.clone())
        np.testing.assert_array_equal(fi, fi_want)
        np.testing.assert_allclose(got.numpy(), want, rtol=0, atol=0)
        assert got.shape == (len(stt), 6, ops['nearest_chans'])


def test_vectorised_ccg_metrics_equal_the_per_pair_formula():
    """The refractoriness metrics evaluated for many correlograms at once
    (host part of the GPU refract) equal CCG_metrics row by row, bit for bit."""
    from kilosort_amd.ccg import ccg_metrics_rows
    rng = np.random.default_rng(3)
    G = 40
    K = rng.poisson(rng.uniform(0.1, 30, (G, 1)), (G, 1001)).astype(np.float64)
    K[::7, 490:511] = 0                                       # refractory dips
    n1 = rng.integers(11, 5000, G)
    n2 = n1.copy()
    T = rng.uniform(10, 3600, G)
    R12, Q12, Q00 = ccg_metrics_rows(n1, n2, K, T, 500, 1 / 1000)
    for g in range(G):
        st1 = np.zeros(int(n1[g]))
        r, q, q0 = PH.CCG_metrics(st1, st1, K[g].copy(), T[g], nbins=500, tbin=1 / 1000)
        assert (R12[g], Q12[g], Q00[g]) == (r, q, q0)
