"""Phy-export kernels (ksm_group, ksm_remove_duplicates, ksm_acg,
ksm_template_sums, ksm_gather_features through kilosort_amd.postprocessing
and kilosort_amd.ccg) against the Phy-export oracle.

Bar: groupings, kept-spike masks, correlogram counts, refractory flags,
feature channel indices and re-ordered features identical (integer counts
and copied values); R12 identical (the same float64 formulas on the same
counts)."""
import sys

import numpy as np
import pytest
import torch

from kilosort_amd import _lib
from kilosort_amd import ccg as CG
from kilosort_amd import postprocessing as PP
from oracle import phy_oracle as PH

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda")


def say(msg):
    print(f"[gpu-phy] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _lib.load()
    yield


@pytest.mark.parametrize("n,G", [(0, 1), (1, 1), (5000, 1), (100000, 1000), (300001, 16384)])
def test_group_is_stable_argsort(n, G):
    rng = np.random.default_rng(n + G)
    lab = rng.integers(0, G, n).astype(np.int32)
    say(f"group n={n} G={G}")
    off, order = PP.group_by_label(torch.from_numpy(lab).to(DEV), G, DEV)
    want_off = np.concatenate([[0], np.cumsum(np.bincount(lab, minlength=G))])
    np.testing.assert_array_equal(off.cpu().numpy(), want_off)
    np.testing.assert_array_equal(order.cpu().numpy(), np.argsort(lab, kind='stable'))


@pytest.mark.parametrize("shuffle", [False, True])
def test_remove_duplicates_matches_oracle(shuffle):
    rng = np.random.default_rng(7)
    n = 200000
    t = np.sort(rng.integers(0, 30000 * 600, n)).astype(np.int64)
    t[1::7] = t[0::7][:len(t[1::7])] + rng.integers(0, 30, len(t[1::7]))   # near-duplicates
    clu = rng.integers(0, 300, n).astype(np.int32)
    if shuffle:
        p = rng.permutation(n)
        t, clu = t[p], clu[p]
    say(f"remove_duplicates n={n} shuffle={shuffle}")
    ts, cs, keep = PP.remove_duplicates(t, clu, dt=15, device=DEV)
    rts, rcs, rkeep = PH.remove_duplicates(t, clu, dt=15)
    assert 0 < rkeep.sum() < n
    np.testing.assert_array_equal(keep, rkeep)
    np.testing.assert_array_equal(ts, rts)
    np.testing.assert_array_equal(cs, rcs)


def trains(seed):
    """Poisson and refractory trains, a tiny cluster and a zero-span one."""
    rng = np.random.default_rng(seed)
    parts, ids = [], []
    for k in range(24):
        m = int(rng.integers(200, 2500))
        if k % 3 == 0:
            t = np.cumsum(0.003 + rng.exponential(0.05, m))
        else:
            t = np.sort(rng.uniform(0, 100, m))
        parts.append(t)
        ids.append(np.full(m, k))
    parts += [np.sort(rng.uniform(0, 100, 8)), np.full(30, 12.5)]
    ids += [np.full(8, 24), np.full(30, 25)]
    st0, clu = np.concatenate(parts), np.concatenate(ids)
    o = np.argsort(st0, kind='stable')
    return st0[o], clu[o]


def test_acg_counts_match_oracle():
    st0, clu = trains(3)
    say(f"acg n={len(st0)}")
    K, n, T, _ = CG.acg_counts(clu, st0, device=DEV)
    for kk in range(clu.max() + 1):
        st1 = st0[clu == kk]
        Kr, Tr = PH.compute_CCG(st1, st1.copy())
        np.testing.assert_array_equal(K[kk], Kr)
        assert n[kk] == len(st1) and T[kk] == Tr


def test_refract_matches_oracle():
    st0, clu = trains(4)
    say(f"refract n={len(st0)}")
    is_ref, R12 = CG.refract(clu, st0, device=DEV)
    r_ref, r_R12 = PH.refract(clu, st0)
    np.testing.assert_array_equal(is_ref, r_ref)
    np.testing.assert_array_equal(R12, r_R12)
    assert 0 < is_ref.sum() < len(is_ref)


def pc_case(seed, n, K, C, nC, n_clu, shared=0.3):
    rng = np.random.default_rng(seed)
    xc = rng.uniform(0, 40, C).astype(np.float32)
    yc = np.arange(C, dtype=np.float32) * 20
    ds = (xc - xc[:, None]) ** 2 + (yc - yc[:, None]) ** 2
    iCC = torch.from_numpy(np.argsort(ds, 0)[:nC])
    iU = torch.from_numpy(rng.integers(0, C, K))
    ops = dict(iU=iU, iCC=iCC, xc=xc, yc=yc, nearest_chans=nC, dmin=20.0, dminx=32.0)
    spike_templates = rng.integers(0, K, n)
    spike_templates[:K] = np.arange(K)                       # every template has spikes
    owner = rng.permutation(np.arange(K) % n_clu)            # every cluster owns one
    spike_clusters = owner[spike_templates]
    split = rng.uniform(size=n) < shared
    spike_clusters[split] = rng.integers(0, n_clu, split.sum())
    tF = torch.from_numpy(rng.standard_normal((n, nC, 6)).astype(np.float32))
    tF += torch.from_numpy(rng.uniform(0, 3, (K, nC, 1)).astype(np.float32))[spike_templates]
    return ops, spike_templates, spike_clusters, tF


@pytest.mark.parametrize("n,K,C,nC,n_clu,shared", [(3000, 12, 32, 4, 8, 0.3),
                                                   (20000, 40, 64, 10, 30, 0.3),
                                                   (20000, 40, 64, 10, 30, 0.0),
                                                   (60000, 700, 383, 10, 200, 0.1)])
def test_make_pc_features_matches_oracle(n, K, C, nC, n_clu, shared):
    """Channel slots whose ranked norm is exactly 0 (all-zero rows left by an
    earlier cluster) are ties that the reference orders by torch's unstable
    sort; the kernel path takes the lowest channel.  Those slots hold zeros
    either way, so the features match bit for bit and feature_ind matches
    on every ranked (non-zero) slot."""
    ops, stt, clu, tF = pc_case(n, n, K, C, nC, n_clu, shared)
    say(f"make_pc_features n={n} K={K} nC={nC} clusters={n_clu} shared={shared}")
    tF_gpu = tF.clone()
    got, fi = PP.make_pc_features(ops, stt, clu, tF_gpu, device=DEV)
    want, fi_want, norms = PH.make_pc_features(ops, stt, clu, tF.clone(), with_norms=True)
    ranked = norms > 0
    np.testing.assert_array_equal(fi[ranked], fi_want[ranked])
    if shared < 0.1:
        assert ranked.all()
    assert fi.dtype == np.uint32 and got.shape == want.shape
    np.testing.assert_array_equal(got.numpy(), want.numpy())
    np.testing.assert_array_equal(tF_gpu.numpy(), want.permute(0, 2, 1).numpy())   # in place


def phy_case(seed=11, C=64, K=30, nC=10, n=30000, n_clu=20, fs=30000.0):
    from kilosort_amd import synth
    from kilosort_amd import template_matching as TMg
    rec = synth.make_recording(seed, C, K, 1000, with_data=False)
    U = torch.from_numpy(rec.U)                               # [K, P, C]
    iCC, iCC_mask, iU, _ = TMg.prepare_extract(rec.xc, rec.yc, U.to(DEV), nC, 40.0, device=DEV)
    rng = np.random.default_rng(seed)
    t = np.sort(rng.integers(0, int(60 * fs), n)).astype(np.float64)
    dup = rng.uniform(size=n) < 0.05
    t[1:][dup[1:]] = t[:-1][dup[1:]] + 4                      # artefact duplicates
    t = np.sort(t)
    tmpl = rng.integers(0, K, n)
    tmpl[:K] = np.arange(K)
    owner = rng.permutation(np.arange(K) % n_clu)
    clu = owner[tmpl]
    split = rng.uniform(size=n) < 0.1
    clu[split] = rng.integers(0, n_clu, split.sum())
    st = np.stack([t, tmpl, rng.uniform(8, 30, n)], 1)
    tF = torch.from_numpy(rng.standard_normal((n, nC, 6)).astype(np.float32))
    tF += torch.from_numpy(rng.uniform(0, 3, (K, nC, 1)).astype(np.float32))[tmpl]
    Wall = U.transpose(1, 2).contiguous()
    xc, yc = rec.xc.astype(np.float32), rec.yc.astype(np.float32)
    ops = dict(iU=iU, iCC=iCC, iCC_mask=iCC_mask, xc=xc, yc=yc, nearest_chans=nC, dmin=20.0,
               dminx=32.0, wPCA=torch.from_numpy(rec.wPCA.copy()), nt=rec.nt, fs=fs,
               duplicate_spike_bins=7, Wrot=torch.eye(C) * 2.0,
               settings=dict(acg_threshold=0.2, ccg_threshold=0.25, n_chan_bin=C, fs=fs,
                             filename='/data/rec.bin'))
    probe = dict(chanMap=np.arange(C), xc=xc, yc=yc, kcoords=np.zeros(C))
    return st, clu.astype(np.int32), tF, Wall, probe, ops


@pytest.mark.parametrize("n,K,G", [(200000, 700, 500), (1000, 50, 7), (0, 30, 4)])
def test_template_sums_exact(n, K, G):
    """ksm_template_sums against numpy: features that are multiples of 2^-8
    below 2^7 in magnitude sum exactly in float64 in any order, so the sums
    must equal numpy's bit for bit; counts and the label x template presence
    exactly; templates without spikes give zero rows."""
    lib = _lib.load()
    rng = np.random.default_rng(n + K)
    nC, P = 10, 6
    tF = (rng.integers(-2 ** 15, 2 ** 15, size=(n, nC, P)) / 256.0).astype(np.float32)
    tmpl = rng.integers(0, K - 3, size=n).astype(np.int64)      # the last 3 templates empty
    lab = rng.integers(0, G, size=n).astype(np.int32)
    say(f"template sums n={n} K={K} G={G}")
    tFd = torch.from_numpy(tF).to(DEV)
    td = torch.from_numpy(tmpl).to(DEV)
    ld = torch.from_numpy(lab).to(DEV)
    sums = torch.full((K, nC, P), float("nan"), dtype=torch.float64, device=DEV)
    counts = torch.full((K,), -1, dtype=torch.int64, device=DEV)
    presence = torch.full((G, K), 7, dtype=torch.uint8, device=DEV)
    nws = lib.ksm_template_sums_workspace(n, K)
    ws = torch.empty(max(nws, 1), dtype=torch.uint8, device=DEV)
    _lib.check(lib.ksm_template_sums(tFd.data_ptr(), td.data_ptr(), ld.data_ptr(), n, nC, P, K,
                                     G, sums.data_ptr(), counts.data_ptr(),
                                     presence.data_ptr(), ws.data_ptr(), nws,
                                     _lib.stream_of(DEV)), "ksm_template_sums")
    ref = np.zeros((K, nC, P), np.float64)
    np.add.at(ref, tmpl, tF.astype(np.float64))
    np.testing.assert_array_equal(sums.cpu().numpy(), ref)
    np.testing.assert_array_equal(counts.cpu().numpy(), np.bincount(tmpl, minlength=K))
    pres = np.zeros((G, K), np.uint8)
    pres[lab, tmpl] = 1
    np.testing.assert_array_equal(presence.cpu().numpy(), pres)
    # a workspace one byte short is refused
    if nws > 256:
        st = lib.ksm_template_sums(tFd.data_ptr(), td.data_ptr(), ld.data_ptr(), n, nC, P, K, G,
                                   sums.data_ptr(), counts.data_ptr(), presence.data_ptr(),
                                   ws.data_ptr(), nws - 1, _lib.stream_of(DEV))
        assert st == _lib.KSM_ERR_WORKSPACE


def test_save_to_phy_matches_oracle(tmp_path):
    from kilosort_amd import phy as PHY
    from oracle import positions_oracle as PO
    from oracle import templates_oracle as TO
    st, clu, tF, Wall, probe, ops = phy_case()
    say(f"save_to_phy n={len(st)} K={Wall.shape[0]}")
    res = PHY.save_to_phy(st, clu, tF.clone(), Wall, probe, ops, imin=100, results_dir=tmp_path,
                          device=DEV)
    ops_c = dict(ops, iU=ops['iU'].cpu(), iCC=ops['iCC'].cpu(), iCC_mask=ops['iCC_mask'].cpu())
    want = PH.save_to_phy_arrays(st, clu, tF.clone(), Wall, ops_c, 100, TO.similarity,
                                 PO.compute_spike_positions)
    load = lambda name: np.load(tmp_path / name)
    for name in ('spike_times.npy', 'spike_templates.npy', 'spike_clusters.npy',
                 'spike_detection_templates.npy', 'kept_spikes.npy', 'templates_ind.npy',
                 'pc_features.npy'):
        got = load(name)
        assert got.dtype == want[name].dtype, name
        np.testing.assert_array_equal(got, want[name], err_msg=name)
    assert 0 < (~want['kept_spikes.npy']).sum()
    ranked = want['_pc_norms'] > 0
    np.testing.assert_array_equal(load('pc_feature_ind.npy')[ranked],
                                  want['pc_feature_ind.npy'][ranked])
    scale = float(np.abs(probe['yc']).max())
    np.testing.assert_allclose(load('spike_positions.npy'), want['spike_positions.npy'],
                               atol=1e-5 * scale, rtol=0)
    np.testing.assert_allclose(load('amplitudes.npy'), want['amplitudes.npy'], rtol=1e-6)
    np.testing.assert_allclose(load('similar_templates.npy'), want['similar_templates.npy'],
                               atol=1e-5)
    tw = want['templates.npy']
    np.testing.assert_allclose(load('templates.npy'), tw, atol=1e-5 * np.abs(tw).max(), rtol=0)
    np.testing.assert_array_equal(res[2], want['_is_ref'])
    np.testing.assert_array_equal(res[3], want['_contam'])
    labels = (tmp_path / 'cluster_KSLabel.tsv').read_text().splitlines()
    assert labels[0] == 'cluster_id\tKSLabel' and len(labels) == clu.max() + 2
    assert labels[1:] == [f'{i}\t{["mua", "good"][int(r)]}' for i, r in enumerate(want['_is_ref'])]
    amp_rows = (tmp_path / 'cluster_Amplitude.tsv').read_text().splitlines()[1:]
    assert amp_rows == [f'{i}\t{p:.1f}' for i, p in enumerate(want['_template_amplitudes'])]
    assert (tmp_path / 'params.py').read_text().splitlines()[0] == f"n_channels_dat = {Wall.shape[1]}"


def test_empty_inputs():
    ts, cs, keep = PP.remove_duplicates(np.zeros(0, np.int64), np.zeros(0, np.int32), 7,
                                        device=DEV)
    assert ts.size == 0 and cs.size == 0 and keep.size == 0
    ops, _, _, _ = pc_case(1, 100, 5, 16, 4, 2)
    pcf, fi = PP.make_pc_features(ops, np.zeros(0, np.int64), np.zeros(0, np.int32),
                                  torch.zeros((0, 4, 6)), device=DEV)
    assert pcf.shape == (0, 6, 4) and fi.shape == (0, 4)
