"""GPU batch preprocessing (ksm_filter_batch through FilteredBatches) against
the preprocessing oracle: padded batches of int16 / uint16 / float32 raw
data, channel map, sign, CAR, the high-pass filter, the artifact test,
whitening and drift.

Bar: max |X - X_oracle| <= 1e-5 * max |X_oracle| (float32 arithmetic in a
different order: a cut filter tail of 2^-26 of its mass, FFT rounding in
the oracle); the CAR-only stage, whose arithmetic is the same, within
2 ulp-scale (1e-6 relative)."""
import sys

import numpy as np
import pytest
import torch

from kilosort_amd import _lib, synth
from kilosort_amd import preprocess as PP
from oracle import filter_oracle as FO

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda")


def say(msg):
    print(f"[gpu-filter] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _lib.load()
    yield


def raw_recording(seed, C, n, dtype="int16"):
    """Spiking data plus what the filter removes: per-channel offsets, a
    slow drift and a common-mode signal; quantised like a binary file."""
    rec = synth.make_recording(seed, C, 16, n, rate_range=(5.0, 20.0))
    rng = np.random.default_rng(seed)
    t = np.arange(n)[:, None]
    x = 20.0 * rec.data + rng.uniform(-300, 300, C)[None, :] \
        + 80.0 * np.sin(2 * np.pi * t / 7000.0) + 30.0 * rng.standard_normal((n, 1))
    if dtype == "int16":
        return np.clip(np.round(x), -32768, 32767).astype(np.int16)
    if dtype == "uint16":
        return (np.clip(np.round(x), -32768, 32767) + 32768).astype(np.uint16)
    return x.astype(np.float32)


def whitening(C, seed):
    rng = np.random.default_rng(seed)
    A = np.eye(C) + 0.05 * rng.standard_normal((C, C))
    return (A @ A.T / C ** 0.5).astype(np.float32)


def oracle_batch(data, fb, ibatch, W=None, M=None, skip=False, dtype="int16", cmap=None,
                 scale=None, shift=None):
    X, inds = FO.padded_batch(data, ibatch, int(fb.NT), int(fb.nt), int(fb.n_batches),
                              imin=int(fb.imin), imax=int(fb.imax), dtype=dtype, scale=scale,
                              shift=shift)
    hp = None if fb.hp_filter is None else fb.hp_filter.cpu()
    out = FO.filter_batch(X, chan_map=cmap, invert_sign=fb.invert_sign, do_CAR=fb.do_CAR,
                          hp_filter=hp, artifact_threshold=fb.artifact_threshold,
                          whiten_mat=None if W is None else torch.from_numpy(W), M=M,
                          skip_preproc=skip)
    return out.numpy(), inds


def rel_err(got, ref):
    return float(np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-30))


@pytest.mark.parametrize("ibatch", [0, 1, 4])
def test_int16_full_pipeline_matches_oracle(ibatch):
    C, NT, n = 64, 6000, 4 * 6000 + 500
    data = raw_recording(11, C, n)
    W = whitening(C, 1)
    fb = PP.FilteredBatches(n_chan_bin=C, NT=NT, file_object=data, device=DEV,
                            hp_filter=PP.get_highpass_filter(device=DEV), whiten_mat=W)
    assert fb.n_batches == 5                  # the last holds 500 samples + padding
    say(f"int16 C={C} NT={NT} batch {ibatch}")
    X, inds = fb.padded_batch_to_torch(ibatch, return_inds=True)
    ref, rinds = oracle_batch(data, fb, ibatch, W=W)
    assert inds == rinds
    e = rel_err(X.cpu().numpy(), ref)
    say(f"batch {ibatch}: rel err {e:.2e}")
    assert e <= 1e-5


def test_uint16_chan_map_invert_no_car():
    C, NT, n = 48, 4000, 3 * 4000
    data = raw_recording(12, C, n, "uint16")
    cmap = np.random.default_rng(2).permutation(C)[:40]
    fb = PP.FilteredBatches(n_chan_bin=C, NT=NT, file_object=data, device=DEV, chan_map=cmap,
                            hp_filter=PP.get_highpass_filter(device=DEV), invert_sign=True,
                            do_CAR=False)
    say("uint16 chan_map invert no-CAR")
    for ib in range(int(fb.n_batches)):
        X = fb.padded_batch_to_torch(ib)
        ref, _ = oracle_batch(data, fb, ib, dtype="uint16", cmap=cmap)
        assert X.shape == (40, NT + 122)
        assert rel_err(X.cpu().numpy(), ref) <= 1e-5


def test_float32_scale_shift_and_skip_preproc():
    C, NT, n = 33, 3000, 2 * 3000 + 200
    data = raw_recording(13, C, n, "float32")
    fb = PP.FilteredBatches(n_chan_bin=C, NT=NT, file_object=data, device=DEV, scale=0.5,
                            shift=3.0, hp_filter=PP.get_highpass_filter(device=DEV))
    say("float32 scale/shift, skip_preproc")
    for ib in range(int(fb.n_batches)):
        X = fb.padded_batch_to_torch(ib, skip_preproc=True)
        ref, _ = oracle_batch(data, fb, ib, skip=True, dtype="float32", scale=0.5, shift=3.0)
        # the same float32 operations except the channel mean (double sum here)
        assert rel_err(X.cpu().numpy(), ref) <= 1e-6
        X = fb.padded_batch_to_torch(ib)
        ref, _ = oracle_batch(data, fb, ib, dtype="float32", scale=0.5, shift=3.0)
        assert rel_err(X.cpu().numpy(), ref) <= 1e-5


def test_filter_of_a_device_batch_without_highpass():
    """filter(X) of a float device batch; CAR over an even channel count
    (lower median), whitening without a high-pass filter."""
    C, T = 64, 5000
    rng = np.random.default_rng(14)
    X = (rng.standard_normal((C, T)) * 10 + rng.uniform(-5, 5, (C, 1))).astype(np.float32)
    W = whitening(C, 3)
    fb = PP.FilteredBatches(n_chan_bin=C, NT=T - 122, file_object=np.zeros((T, C), np.int16),
                            device=DEV, whiten_mat=W)
    got = fb.filter(torch.from_numpy(X).to(DEV)).cpu().numpy()
    ref = FO.filter_batch(torch.from_numpy(X), do_CAR=True, whiten_mat=torch.from_numpy(W)).numpy()
    assert rel_err(got, ref) <= 1e-5
    fb2 = PP.FilteredBatches(n_chan_bin=C, NT=T - 122, file_object=np.zeros((T, C), np.int16),
                             device=DEV)
    got = fb2.filter(torch.from_numpy(X).to(DEV)).cpu().numpy()
    ref = FO.filter_batch(torch.from_numpy(X), do_CAR=True).numpy()
    assert rel_err(got, ref) <= 1e-6


@pytest.mark.parametrize("thr", [np.inf, 1e9, 50.0])
def test_artifact_threshold(thr):
    C, NT, n = 32, 3000, 3 * 3000
    data = raw_recording(15, C, n)
    data[4000, 5] = 30000                    # a large transient in batch 1
    W = whitening(C, 4)
    fb = PP.FilteredBatches(n_chan_bin=C, NT=NT, file_object=data, device=DEV, whiten_mat=W,
                            hp_filter=PP.get_highpass_filter(device=DEV),
                            artifact_threshold=thr)
    say(f"artifact threshold {thr}")
    for ib in range(int(fb.n_batches)):
        X = fb.padded_batch_to_torch(ib).cpu().numpy()
        ref, _ = oracle_batch(data, fb, ib, W=W)
        if not np.any(ref):
            assert not np.any(X)
        else:
            assert rel_err(X, ref) <= 1e-5
    if thr == 50.0:
        assert not np.any(fb.padded_batch_to_torch(1).cpu().numpy())


def test_drift_matrix_batches():
    C, NT, n = 40, 3000, 3 * 3000
    data = raw_recording(16, C, n)
    W = whitening(C, 5)
    xc = np.tile(np.array([16.0, 48.0], np.float32), C // 2)
    yc = (np.arange(C) // 2 * 20.0).astype(np.float32)
    xp = np.vstack((xc, yc)).T
    Kxx = torch.from_numpy(FO.kernel2D(torch.from_numpy(xp), torch.from_numpy(xp), 20.0).numpy())
    iKxx = torch.linalg.inv(Kxx + 0.01 * torch.eye(C))
    ops = dict(yblk=np.array([200.0]), nblocks=1, probe=dict(xc=xc, yc=yc),
               settings=dict(sig_interp=20.0), iKxx=iKxx.to(DEV))
    dshift = np.array([0.0, 7.5, -12.0])
    fb = PP.FilteredBatches(n_chan_bin=C, NT=NT, file_object=data, device=DEV, whiten_mat=W,
                            hp_filter=PP.get_highpass_filter(device=DEV), dshift=dshift)
    say("drift matrices")
    ops_c = dict(ops, iKxx=iKxx)
    for ib in range(3):
        X = fb.padded_batch_to_torch(ib, ops).cpu().numpy()
        M = FO.drift_matrix(ops_c, dshift[ib])
        ref, _ = oracle_batch(data, fb, ib, W=W, M=M)
        e = rel_err(X, ref)
        say(f"drift batch {ib}: rel err {e:.2e}")
        assert e <= 1e-5


def test_full_size_383_channels():
    """The bench's geometry: 383 channels (odd: a true median), NT=60000."""
    C, NT = 383, 60000
    n = 2 * NT + 1000
    data = raw_recording(17, C, n)
    W = whitening(C, 6)
    fb = PP.FilteredBatches(n_chan_bin=C, NT=NT, file_object=data, device=DEV, whiten_mat=W,
                            hp_filter=PP.get_highpass_filter(device=DEV))
    say("383 channels, NT=60000, batch 1")
    X = fb.padded_batch_to_torch(1).cpu().numpy()
    ref, _ = oracle_batch(data, fb, 1, W=W)
    e = rel_err(X, ref)
    say(f"383ch: rel err {e:.2e}")
    assert e <= 1e-5


def test_gemm_f32_odd_shapes():
    lib = _lib.load()
    g = torch.Generator().manual_seed(0)
    for M, N, K in [(37, 52, 29), (383, 383, 383), (1, 4, 1), (130, 260, 17)]:
        A = torch.randn(M, K, generator=g)
        B = torch.randn(K, N, generator=g)
        Ad, Bd = A.to(DEV), B.to(DEV)
        Cd = torch.full((M, N), float("nan"), device=DEV)
        _lib.check(lib.ksm_gemm_f32(Ad.data_ptr(), K, Bd.data_ptr(), N, Cd.data_ptr(), N, M, N,
                                    K, _lib.stream_of(DEV)), "ksm_gemm_f32")
        ref = (A.double() @ B.double()).numpy()
        got = Cd.cpu().double().numpy()
        assert np.abs(got - ref).max() <= 1e-5 * max(np.abs(ref).max(), 1.0), (M, N, K)


class _OracleFiltered:
    """A bfile whose batches are the oracle's padded + filtered batches."""

    def __init__(self, data, fb, W):
        self.data, self.fb, self.W = data, fb, W
        self.n_batches = fb.n_batches
        self.batch_downsampling = 1

    def padded_batch_to_torch(self, ibatch, ops=None):
        out, _ = oracle_batch(self.data, self.fb, ibatch, W=self.W)
        return torch.from_numpy(out).to(DEV)


def test_extract_reads_filtered_batches():
    """FilteredBatches is the bfile of extract: spikes found from its batches
    agree with those found from the oracle's filtered batches (the inputs
    differ by float32 rounding only)."""
    from kilosort_amd import template_matching as tm
    C, NT = 64, 30000
    rec = synth.config1_recording(seed=3, n_batches=2)
    n = 2 * NT
    rng = np.random.default_rng(3)
    raw = 20.0 * rec.data[:n] + rng.uniform(-300, 300, C)[None, :] \
        + 30.0 * rng.standard_normal((n, 1))
    raw = np.clip(np.round(raw), -32768, 32767).astype(np.int16)
    W = (np.eye(C) / 20.0).astype(np.float32)
    fb = PP.FilteredBatches(n_chan_bin=C, NT=NT, file_object=raw, device=DEV, whiten_mat=W,
                            hp_filter=PP.get_highpass_filter(device=DEV))
    ops = dict(nt=61, nt0min=rec.nt0min, Th_learned=8, max_peels=100, batch_size=NT,
               wPCA=torch.from_numpy(rec.wPCA).to(DEV), xc=rec.xc, yc=rec.yc,
               settings=dict(nearest_chans=10, position_limit=100.0, n_pcs=6))
    U = torch.from_numpy(rec.U).to(DEV)
    say("extract from FilteredBatches vs oracle-filtered batches")
    st, tF, _ = tm.extract(dict(ops), fb, U, device=DEV)
    st_o, tF_o, _ = tm.extract(dict(ops), _OracleFiltered(raw, fb, W), U, device=DEV)
    a = set(map(tuple, st[:, :2].astype(np.int64).tolist()))
    b = set(map(tuple, st_o[:, :2].astype(np.int64).tolist()))
    agree = len(a & b) / max(len(a), len(b))
    say(f"{len(st)} spikes vs {len(st_o)}: agreement {100 * agree:.3f} %")
    assert len(st_o) > 500
    assert agree >= 0.999


def test_extract_reads_filtered_batches_383_channels():
    """The streamed form at the benchmark's shape (383 channels, 700
    templates, two batches of NT=60000 from a raw int16 recording): spikes
    found through FilteredBatches against those found from the oracle's
    filtered batches.  The filtered inputs differ by float32 rounding
    (<= 1e-5 relative, test_full_size_383_channels), so a decision can only
    differ where two scores are that close."""
    from kilosort_amd import template_matching as tm
    C, K, NT = 383, 700, 60000
    rec = synth.make_recording(21, C, K, 2 * NT, rate_range=(1.0, 20.0),
                               amp_range=(8.0, 40.0), overlap_frac=0.2)
    rng = np.random.default_rng(21)
    raw = 20.0 * rec.data + rng.uniform(-300, 300, C)[None, :] \
        + 30.0 * rng.standard_normal((2 * NT, 1))
    raw = np.clip(np.round(raw), -32768, 32767).astype(np.int16)
    W = (np.eye(C) / 20.0).astype(np.float32)
    fb = PP.FilteredBatches(n_chan_bin=C, NT=NT, file_object=raw, device=DEV, whiten_mat=W,
                            hp_filter=PP.get_highpass_filter(device=DEV))
    ops = dict(nt=61, nt0min=rec.nt0min, Th_learned=8, max_peels=100, batch_size=NT,
               wPCA=torch.from_numpy(rec.wPCA).to(DEV), xc=rec.xc, yc=rec.yc,
               settings=dict(nearest_chans=10, position_limit=100.0, n_pcs=6))
    U = torch.from_numpy(rec.U).to(DEV)
    say("383 ch x 700 templates: extract from FilteredBatches vs oracle-filtered batches")
    st, tF, _ = tm.extract(dict(ops), fb, U, device=DEV)
    st_o, tF_o, _ = tm.extract(dict(ops), _OracleFiltered(raw, fb, W), U, device=DEV)
    a = set(map(tuple, st[:, :2].astype(np.int64).tolist()))
    b = set(map(tuple, st_o[:, :2].astype(np.int64).tolist()))
    agree = len(a & b) / max(len(a), len(b))
    say(f"383ch: {len(st)} spikes vs {len(st_o)}: agreement {100 * agree:.4f} %")
    assert len(st_o) > 10000
    assert agree >= 0.999

