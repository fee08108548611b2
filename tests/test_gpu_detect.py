"""Universal-template detection (ksm_detect through kilosort_amd.spikedetect)
against the detection oracle on the same float32 batches.

Bar: spike positions (template, column) and imax identical; amp, adist, y
positions and features within 1e-5 relative (FP32 sums in another order)."""
import sys

import numpy as np
import pytest
import torch

from kilosort_amd import _lib, synth
from kilosort_amd import spikedetect as SD
from kilosort_amd.batches import ArrayBatches
from oracle import detect_oracle as DO

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda")
CPU = torch.device("cpu")


def say(msg):
    print(f"[gpu-detect] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _lib.load()
    yield


def make_ops(rec, NT, nC2=100):
    return dict(xc=rec.xc, yc=rec.yc, kcoords=np.zeros(len(rec.xc)), nt=rec.nt,
                Th_universal=9.0, max_channel_distance=32.0, fs=30000.0, batch_size=NT,
                wTEMP=torch.from_numpy(rec.wPCA.copy()), wPCA=torch.from_numpy(rec.wPCA.copy()),
                settings=dict(dmin=None, dminx=32.0, nearest_chans=10, nearest_templates=nC2,
                              min_template_size=10.0, template_sizes=5, nt0min=20,
                              n_templates=6, n_pcs=6, templates_from_data=False,
                              Th_single_ch=6.0))


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)) if a.size else 0.0


def check_batch(X, ops, iC, iC2, weigh, name):
    ref = DO.template_match(X.cpu(), ops, iC.cpu(), iC2.cpu(), weigh.cpu())
    ops_d = dict(ops)
    got = SD.template_match(X.to(DEV), ops_d, iC.to(DEV), iC2.to(DEV), weigh.to(DEV), device=DEV)
    xy, imax, amp, adist = (t.cpu() for t in got)
    say(f"{name}: {len(xy)} spikes (oracle {len(ref[0])}), Nfilt {iC.shape[1]}")
    np.testing.assert_array_equal(xy.numpy(), ref[0].numpy())
    np.testing.assert_array_equal(imax.numpy(), ref[1].numpy())
    assert xy.dtype == torch.int64 and imax.dtype == torch.int64
    assert rel(amp, ref[2]) <= 1e-5
    assert rel(adist, ref[3]) <= 1e-5
    return len(xy)


def test_template_match_64_channels():
    rec = synth.make_recording(61, 64, 48, 30000 + 122, rate_range=(5.0, 20.0))
    ops = make_ops(rec, 30000)
    iC, iC2, weigh = DO.setup(dict(ops))
    X = ArrayBatches(rec.data, NT=30000, nt=61, device=CPU).padded_batch_to_torch(0)
    n = check_batch(X, ops, iC, iC2, weigh, "C=64")
    assert n > 100


def test_template_match_other_sizes():
    """Other nearest-channel and template-size counts than the defaults (8
    channels, 3 sizes): the score pass for sizes given at run time, which
    stores the first maximum's index and sign for every position."""
    rec = synth.make_recording(63, 64, 48, 30000 + 122, rate_range=(5.0, 20.0))
    ops = make_ops(rec, 30000)
    ops["settings"] = dict(ops["settings"], nearest_chans=8, template_sizes=3)
    iC, iC2, weigh = DO.setup(dict(ops))
    assert iC.shape[0] == 8 and weigh.shape[0] == 3
    X = ArrayBatches(rec.data, NT=30000, nt=61, device=CPU).padded_batch_to_torch(0)
    n = check_batch(X, ops, iC, iC2, weigh, "C=64, 8 channels, 3 sizes")
    assert n > 100


def test_template_match_383_channels():
    """The bench's geometry: 383 channels, 1149 template positions."""
    rec = synth.make_recording(62, 383, 256, 60000 + 122, rate_range=(2.0, 10.0))
    ops = make_ops(rec, 60000)
    iC, iC2, weigh = DO.setup(dict(ops))
    X = ArrayBatches(rec.data, NT=60000, nt=61, device=CPU).padded_batch_to_torch(0)
    n = check_batch(X, ops, iC, iC2, weigh, "C=383")
    assert n > 1000


def test_run_matches_oracle_batches():
    """run() over three batches: the spike table and features equal the
    oracle's per-batch rows; capacity growth is exercised (first capacity
    below the spikes of one batch)."""
    NT = 20000
    rec = synth.make_recording(63, 48, 32, 3 * NT + 122, rate_range=(10.0, 30.0))
    ops = make_ops(rec, NT)
    bf = ArrayBatches(rec.data, NT=NT, nt=61, device=DEV)
    st, tF, ops_out = SD.run(dict(ops), bf, device=DEV)
    ops_c = dict(ops)
    iC, iC2, weigh = DO.setup(ops_c)
    bf_c = ArrayBatches(rec.data, NT=NT, nt=61, device=CPU)
    parts, feats = [], []
    for ib in range(int(bf_c.n_batches)):
        s, f, _ = DO.detect_batch(bf_c.padded_batch_to_torch(ib), ops_c, iC, iC2, weigh, ibatch=ib)
        parts.append(s)
        feats.append(f)
    st_ref, tF_ref = np.concatenate(parts), np.concatenate(feats)
    say(f"run: {len(st)} spikes (oracle {len(st_ref)})")
    assert len(st_ref) > 300
    np.testing.assert_array_equal(st[:, [0, 3, 4, 5]], st_ref[:, [0, 3, 4, 5]])
    assert rel(st[:, 1], st_ref[:, 1]) <= 1e-5
    assert rel(st[:, 2], st_ref[:, 2]) <= 1e-5
    assert rel(tF, tF_ref) <= 1e-5
    np.testing.assert_array_equal(ops_out["iC"].cpu().numpy(), iC.numpy())
    np.testing.assert_array_equal(ops_out["iC2"].cpu().numpy(), iC2.numpy())


def test_capacity_growth():
    rec = synth.make_recording(64, 32, 24, 10000 + 122, rate_range=(20.0, 40.0))
    ops = make_ops(rec, 10000)
    iC, iC2, weigh = DO.setup(dict(ops))
    X = ArrayBatches(rec.data, NT=10000, nt=61, device=DEV).padded_batch_to_torch(0)
    det = SD._Detector(ops, iC.to(DEV), iC2.to(DEV), weigh.to(DEV), DEV)
    det.cap = 8
    xy, imax, amp, adist, yct, tF = det.run(X)
    assert len(xy) > 8 and det.cap >= len(xy)
    ref = DO.template_match(X.cpu(), ops, iC, iC2, weigh)
    np.testing.assert_array_equal(xy.cpu().numpy(), ref[0].numpy())


@pytest.mark.parametrize("C", [64, 383])
def test_snippets_match_oracle(C):
    """Isolated-peak snippets: the same (channel, time) peaks in the same order
    and the same waveforms (copied samples, so bit-identical)."""
    rec = synth.make_recording(65 + C, C, 32, 60000 + 122, rate_range=(5.0, 20.0))
    X = ArrayBatches(rec.data, NT=60000, nt=61, device=CPU).padded_batch_to_torch(0)
    ref = DO.extract_snippets(X, 61, 20, 6.0)
    got = SD.extract_snippets(X.to(DEV), 61, 20, 6.0, device=DEV).cpu()
    say(f"snippets C={C}: {len(got)} (oracle {len(ref)})")
    assert len(ref) > 50
    np.testing.assert_array_equal(got.numpy(), ref.numpy())


def test_universal_waveforms_from_data():
    """extract_wPCA_wTEMP: snippets of every 25th batch, then scikit-learn's
    SVD and k-means as the reference: orthonormal PCs, unit-norm waveforms."""
    NT = 20000
    rec = synth.make_recording(66, 48, 32, 3 * NT + 122, rate_range=(10.0, 30.0))
    ops = make_ops(rec, NT)
    bf = ArrayBatches(rec.data, NT=NT, nt=61, device=DEV)
    wPCA, wTEMP = SD.extract_wPCA_wTEMP(ops, bf, nt=61, twav_min=20, Th_single_ch=6.0,
                                        nskip=1, device=DEV)
    assert wPCA.shape == (6, 61) and wTEMP.shape == (6, 61)
    G = (wPCA @ wPCA.T).cpu().numpy()
    np.testing.assert_allclose(G, np.eye(6), atol=1e-4)
    np.testing.assert_allclose((wTEMP ** 2).sum(1).cpu().numpy(), 1.0, atol=1e-5)

