"""Pins of the preprocessing oracle (CPU): the Fourier-domain high-pass against
the direct sum over taps, the filter's gain against the Butterworth closed
form, the common median reference against a sort, the padded-batch edges,
and the drift matrix at zero drift."""
import numpy as np
import pytest
import torch

from oracle import filter_oracle as FO


@pytest.mark.parametrize("T,ft", [(60, 31), (61, 61), (64, 70), (200, 57), (201, 263)])
def test_fourier_highpass_is_circular_correlation(T, ft):
    """filter_batch's FFT product, conjugate and fftshift equal
    out[n] = sum_m hpad[m] x[(n - T//2 + m) mod T] (float64); an asymmetric
    filter makes a flipped or shifted index fail."""
    g = torch.Generator().manual_seed(T * 1000 + ft)
    X = torch.randn(5, T, generator=g, dtype=torch.float64)
    h = torch.randn(ft, generator=g, dtype=torch.float64)
    got = FO.filter_batch(X, do_CAR=False, hp_filter=h).numpy()
    Xc = (X - X.mean(1, keepdim=True)).numpy()
    ref = FO.circular_highpass_direct(Xc, h)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-10 * np.abs(ref).max())


@pytest.mark.parametrize("f_hz", [60.0, 150.0, 300.0, 600.0, 3000.0])
def test_highpass_gain_is_squared_butterworth(f_hz):
    """A periodic sinusoid through the default filter (fs 30 kHz, 300 Hz)
    comes out scaled by |H|^2 = 1 / (1 + (tan(pi fc/fs) / tan(pi f/fs))^6),
    the forward-backward gain of the bilinear 3rd-order Butterworth."""
    fs, fc, T = 30000.0, 300.0, 60122
    k = round(f_hz * T / fs)                  # whole periods in the batch
    f = k * fs / T
    t = np.arange(T)
    x = np.sin(2 * np.pi * k * t / T + 0.3)
    X = torch.from_numpy(np.stack([x, 2 * x]))
    hp = FO.get_highpass_filter(fs, fc).double()
    out = FO.filter_batch(X, do_CAR=False, hp_filter=hp).numpy()
    gain = np.dot(out[0], x) / np.dot(x, x)
    ref = 1.0 / (1.0 + (np.tan(np.pi * fc / fs) / np.tan(np.pi * f / fs)) ** 6)
    assert abs(gain - ref) <= 1e-5
    np.testing.assert_allclose(out[1], 2 * out[0], rtol=0, atol=1e-9)


def test_common_median_reference_is_lower_median():
    """With an even channel count the median removed per sample is the lower
    of the two middle values (torch.median), after the channel means."""
    g = torch.Generator().manual_seed(5)
    X = torch.randn(8, 300, generator=g, dtype=torch.float64)
    got = FO.filter_batch(X, do_CAR=True).numpy()
    Y = (X - X.mean(1, keepdim=True)).numpy()
    med = np.sort(Y, axis=0)[(Y.shape[0] - 1) // 2]
    np.testing.assert_array_equal(got, Y - med)


def test_padded_batch_edges():
    """First batch: nt copies of the first sample, then the data; last batch:
    the data, then copies of the last sample; middle batches: the window
    [i NT - nt, i NT + NT + nt); uint16 data are shifted by 2^15."""
    NT, nt, n = 100, 7, 330
    rng = np.random.default_rng(3)
    data = rng.integers(-3000, 3000, size=(n, 6)).astype(np.int16)
    nb = int(np.ceil(n / NT))
    X0, i0 = FO.padded_batch(data, 0, NT, nt, nb)
    assert i0 == [-nt, NT + nt]
    np.testing.assert_array_equal(X0[:, :nt].numpy(), np.repeat(data[:1].T, nt, 1))
    np.testing.assert_array_equal(X0[:, nt:].numpy(), data[:NT + nt].T)
    X1, _ = FO.padded_batch(data, 1, NT, nt, nb)
    np.testing.assert_array_equal(X1.numpy(), data[NT - nt:2 * NT + nt].T)
    X3, i3 = FO.padded_batch(data, 3, NT, nt, nb)
    tail = data[3 * NT - nt:]
    np.testing.assert_array_equal(X3[:, :len(tail)].numpy(), tail.T)
    np.testing.assert_array_equal(X3[:, len(tail):].numpy(),
                                  np.repeat(tail[-1:].T, NT + 2 * nt - len(tail), 1))
    assert i3[1] == n + nt
    u = (data.astype(np.int32) + 2 ** 15).astype(np.uint16)
    Xu, _ = FO.padded_batch(u, 1, NT, nt, nb, dtype="uint16")
    np.testing.assert_array_equal(Xu.numpy(), X1.numpy())


def test_drift_matrix_at_zero_drift():
    """dshift = 0: M = Kxx (Kxx + 0.01 I)^-1, and (M W) X with W = I is M X."""
    rng = np.random.default_rng(4)
    xc = rng.uniform(0, 40, 12)
    yc = np.arange(12) * 20.0
    xp = torch.from_numpy(np.vstack((xc, yc)).T)
    Kxx = FO.kernel2D(xp, xp, 20.0)
    iKxx = torch.linalg.inv(Kxx + 0.01 * torch.eye(12, dtype=torch.float64))
    ops = dict(yblk=np.array([100.0]), nblocks=1, probe=dict(xc=xc, yc=yc),
               settings=dict(sig_interp=20.0), iKxx=iKxx)
    M = FO.drift_matrix(ops, 0.0)
    np.testing.assert_allclose(M.numpy(), (Kxx @ iKxx).numpy(), rtol=0, atol=1e-12)
    # Kxx (Kxx + 0.01 I)^-1 = I - 0.01 (Kxx + 0.01 I)^-1
    np.testing.assert_allclose(M.numpy(), np.eye(12) - 0.01 * iKxx.numpy(), atol=1e-9)
    X = torch.from_numpy(rng.standard_normal((12, 50)))
    out = FO.filter_batch(X, do_CAR=False, whiten_mat=torch.eye(12, dtype=torch.float64), M=M)
    Xc = X - X.mean(1, keepdim=True)
    np.testing.assert_allclose(out.numpy(), (M @ Xc).numpy(), atol=1e-12)


def test_artifact_threshold_zeroes_the_batch():
    X = torch.zeros(3, 50, dtype=torch.float64)
    X[1, 10] = 100.0
    out = FO.filter_batch(X, do_CAR=False, artifact_threshold=50.0,
                          whiten_mat=torch.eye(3, dtype=torch.float64))
    assert (out == 0).all()
    out = FO.filter_batch(X, do_CAR=False, artifact_threshold=200.0)
    assert out.abs().max() > 90
