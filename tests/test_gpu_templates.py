"""Template Gram consumers on the GPU (ksm_similarity, ksm_align_templates
through kilosort_amd.templates) against the template oracle.

Bar: similarity and aligned templates within 1e-5 (float64 Gram sums here,
float32 einsums in the oracle); alignment indices, merges, labels and
spike order identical."""
import sys

import numpy as np
import pytest
import torch

from kilosort_amd import _lib, synth
from kilosort_amd import templates as TM
from oracle import templates_oracle as TO

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda")


def say(msg):
    print(f"[gpu-templates] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _lib.load()
    yield


def templates(seed, C, K, dup=0):
    """Synthetic learned templates Wall [K, C, P] (reference layout) plus
    `dup` near-duplicates (scaled copies of the first templates)."""
    rec = synth.make_recording(seed, C, K, 1000, with_data=False)
    U = torch.from_numpy(rec.U)                          # [K, P, C]
    Wall = U.transpose(1, 2).contiguous()
    if dup:
        Wall = torch.cat([Wall, 1.05 * Wall[:dup]], 0)
    wtemp = synth.make_recording(seed + 1, C, 4, 1000, with_data=False).wPCA
    return Wall, torch.from_numpy(rec.wPCA), torch.from_numpy(wtemp), rec.nt


def test_similarity_matches_oracle():
    Wall, W, _, nt = templates(71, 96, 80)
    say(f"similarity K={Wall.shape[0]} C=96")
    S = TM.similarity(Wall.to(DEV), W.to(DEV), nt)
    ref = TO.similarity(Wall, W, nt)
    assert S.dtype == np.float32 and S.shape == ref.shape
    np.testing.assert_allclose(S, ref, atol=1e-5)
    np.testing.assert_allclose(np.diag(S), 1.0, atol=1e-5)


def test_similarity_383_channels_700_templates():
    Wall, W, _, nt = templates(72, 383, 700)
    say("similarity K=700 C=383")
    S = TM.similarity(Wall.to(DEV), W.to(DEV), nt)
    ref = TO.similarity(Wall[:64], W, nt)
    # rows of the first 64 templates against themselves
    np.testing.assert_allclose(S[:64, :64], ref, atol=1e-5)


def test_align_u_matches_oracle():
    Wall, W, wT, nt = templates(73, 64, 48)
    ops = dict(wPCA=W, wTEMP=wT, nt=nt, Nchan=64)
    say("align_U K=48 C=64")
    Unew, imax = TM.align_U(Wall.to(DEV), dict(ops, wPCA=W.to(DEV), wTEMP=wT.to(DEV)), device=DEV)
    Uref, iref = TO.align_U(Wall, ops)
    np.testing.assert_array_equal(imax.cpu().numpy(), iref.numpy())
    scale = np.abs(Uref.numpy()).max()
    assert np.abs(Unew.cpu().numpy() - Uref.numpy()).max() <= 1e-5 * scale


def test_postprocess_templates_merges_like_oracle():
    Wall, W, wT, nt = templates(74, 64, 40, dup=6)
    K = Wall.shape[0]
    rng = np.random.default_rng(0)
    clu = rng.integers(0, K, 2000)
    clu[:K] = np.arange(K)                               # every template has spikes
    st = np.stack([rng.permutation(2000).astype(float), np.zeros(2000)], 1)
    tF = torch.from_numpy(rng.standard_normal((2000, 3)).astype(np.float32))
    ops = dict(wPCA=W, wTEMP=wT, nt=nt, Nchan=64,
               settings=dict(acg_threshold=0.2, ccg_threshold=0.25))
    say(f"postprocess_templates K={K} (6 planted duplicates)")
    W3 = TM.postprocess_templates(Wall.to(DEV), dict(ops, wPCA=W.to(DEV), wTEMP=wT.to(DEV)), clu,
                                  st.copy(), tF.clone(), device=DEV)
    R3 = TO.postprocess_templates(Wall, ops, clu, st.copy(), tF.clone())
    say(f"templates after merging: {W3.shape[0]} (oracle {R3.shape[0]})")
    assert W3.shape == R3.shape and W3.shape[0] < K
    scale = np.abs(R3.numpy()).max()
    assert np.abs(W3.cpu().numpy() - R3.numpy()).max() <= 1e-5 * scale
    # merging_function outputs on the aligned templates
    Wall2, _ = TO.align_U(Wall, ops)
    out = TM.merging_function(dict(ops, wPCA=W.to(DEV)), Wall2.transpose(1, 2), clu, st.copy(),
                              tF.clone(), 0.9, 'mu', check_dt=False, device=DEV)
    ref = TO.merging_function(ops, Wall2.transpose(1, 2), clu, st.copy(), tF.clone(), 0.9)
    np.testing.assert_array_equal(out[1], ref[1])
    np.testing.assert_array_equal(out[3], ref[3])
    np.testing.assert_array_equal(out[4].numpy(), ref[4].numpy())


def test_similarity_lags_match_merge_rows():
    """lag[k][j] is the imax of merging_function's row for template k (the
    lag that check_dt turns into a time shift)."""
    Wall, W, _, nt = templates(75, 48, 24)
    Wall = Wall.double()
    shifted = torch.einsum('kcp,pq->kcq', Wall[:4], torch.from_numpy(TM.shift_operator(W.numpy(), 3)))
    Wall = torch.cat([Wall, shifted]).float()
    S, lag = TM.similarity_matrix(Wall.to(DEV), W.to(DEV), nt, DEV, lags=True)
    for kk in range(Wall.shape[0]):
        cmax, imax = TO.merge_step_row(Wall, W, nt, kk)
        np.testing.assert_allclose(S[kk], cmax.numpy(), atol=1e-5)
        np.testing.assert_array_equal(lag[kk], imax.numpy())
    assert lag[0][0] == nt and lag[0][24] != nt                 # shifted copy peaks off centre


def test_ccg_counts_match_oracle():
    from oracle import phy_oracle as PH
    rng = np.random.default_rng(12)
    lib = _lib.load()
    for n1, n2 in ((0, 50), (1, 1), (3000, 2000), (20000, 500)):
        a = np.sort(rng.uniform(0, 60, n1))
        b = np.sort(np.concatenate([rng.uniform(0, 60, n2 - min(n2, 10)), a[:min(n2, 10, n1)]])
                    if n1 else rng.uniform(0, 60, n2))
        counts = torch.empty(1001, dtype=torch.int32, device=DEV)
        ta, tb = torch.from_numpy(a).to(DEV), torch.from_numpy(b).to(DEV)
        _lib.check(lib.ksm_ccg(_lib.ptr(ta) if n1 else 0, n1, tb.data_ptr(), len(b), 500, 1 / 1000,
                               counts.data_ptr(), _lib.stream_of(DEV)), "ksm_ccg")
        got = counts.cpu().numpy().astype(np.float64)
        if n1 == 0:
            assert not got.any()
            continue
        want, _ = PH.compute_CCG(a, b)
        np.testing.assert_array_equal(got, want)


def merge_case(seed=13, C=32, K=10, fs=30000.0):
    """K templates plus shifted copies of templates 0-2 (3 samples, in PC
    space); each copy shares a refractory train with its original (split at
    random), the others have independent trains (refractory or not)."""
    Wall, W, _, nt = templates(seed, C, K)
    Wall = Wall.double()
    Msh = torch.from_numpy(TM.shift_operator(W.numpy(), 3))
    Wall = torch.cat([Wall, torch.einsum('kcp,pq->kcq', Wall[:3], Msh)]).float()
    rng = np.random.default_rng(seed)
    parts, ids = [], []
    for k in range(K):
        m = int(rng.integers(800, 2000))
        t = np.cumsum(0.004 + rng.exponential(0.04, m)) if k % 2 == 0 else \
            np.sort(rng.uniform(0, 60, m))
        lab = np.full(m, k)
        if k < 3:
            lab[rng.uniform(size=m) < 0.4] = K + k
        parts.append(t)
        ids.append(lab)
    times, clu = np.concatenate(parts), np.concatenate(ids)
    o = np.argsort(times, kind='stable')
    st = np.stack([np.round(times[o] * fs), np.zeros(o.size), np.ones(o.size)], 1)
    clu = clu[o]
    tF = torch.from_numpy(rng.standard_normal((o.size, 5, W.shape[0])).astype(np.float32))
    ops = dict(nt=nt, wPCA=W, fs=fs, settings=dict(acg_threshold=0.2, ccg_threshold=0.25))
    return ops, Wall, clu, st, tF


@pytest.mark.parametrize("check_dt", [False, True])
def test_merging_ccg_mode_matches_oracle(check_dt):
    ops, Wall, clu, st, tF = merge_case()
    say(f"merging_function ccg K={Wall.shape[0]} n={len(st)} check_dt={check_dt}")
    st_g, tF_g = st.copy(), tF.clone()
    out = TM.merging_function(dict(ops, wPCA=ops['wPCA'].to(DEV)), Wall.to(DEV), clu, st_g, tF_g,
                              0.5, 'ccg', check_dt=check_dt, device=DEV)
    st_r, tF_r = st.copy(), tF.clone()
    ref = TO.merging_function(ops, Wall, clu, st_r, tF_r, 0.5, 'ccg', check_dt=check_dt)
    say(f"templates after merging: {out[0].shape[0]} (oracle {ref[0].shape[0]})")
    assert out[0].shape == ref[0].shape and out[0].shape[0] < Wall.shape[0]
    np.testing.assert_array_equal(out[1], ref[1])
    np.testing.assert_array_equal(out[2], ref[2])
    np.testing.assert_array_equal(out[3], ref[3])
    np.testing.assert_array_equal(st_g, st_r)                   # shifted in place alike
    scale = np.abs(ref[0].numpy()).max()
    assert np.abs(out[0].numpy() - ref[0].numpy()).max() <= 1e-5 * scale
    fs_ = np.abs(ref[4].numpy()).max()
    assert np.abs(out[4].numpy() - ref[4].numpy()).max() <= 1e-5 * fs_
    assert np.abs(tF_g.numpy() - tF_r.numpy()).max() <= 1e-5 * fs_
    if check_dt:
        assert (st_r[:, 0] != st[:, 0]).any()                    # some spikes were shifted
