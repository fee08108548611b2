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
    ops = dict(wPCA=W, wTEMP=wT, nt=nt, Nchan=64)
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
