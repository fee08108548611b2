"""Host logic of the GPU preprocessing: the taps that replace the reference's
Fourier-domain high-pass filter (CPU, no GPU work)."""
import numpy as np
import pytest
import torch

from kilosort_amd import preprocess as PP
from oracle import filter_oracle as FO


def circular(x, taps, J):
    T = x.shape[1]
    n = np.arange(T)
    out = np.zeros_like(x)
    for j in range(-J, J + 1):
        out += taps[J + j] * x[:, (n + j) % T]
    return out


@pytest.mark.parametrize("T", [60122, 3001, 2000])
def test_taps_reproduce_the_fourier_filter(T):
    """Circular correlation with the cut taps equals the oracle's FFT filter
    (float64) to the dropped tail's share."""
    hp = FO.get_highpass_filter()
    taps, J = PP.highpass_taps(hp, T)
    assert taps.dtype == np.float32 and len(taps) == 2 * J + 1
    assert 300 < J < 1000                      # the 300 Hz filter decays within ~20 ms
    rng = np.random.default_rng(T)
    x = rng.standard_normal((3, T))
    x -= x.mean(1, keepdims=True)
    ref = FO.filter_batch(torch.from_numpy(x), do_CAR=False, hp_filter=hp.double()).numpy()
    got = circular(x, taps.astype(np.float64), J)
    bound = PP.TAP_TAIL * np.abs(hp.double().numpy()).sum() * np.abs(x).max()
    assert np.abs(got - ref).max() <= bound + 1e-12


def test_taps_of_a_short_asymmetric_filter():
    """An asymmetric filter shorter than the batch: the taps are the filter
    itself with its element ft//2 ... placed at j = (index - pad) - T//2."""
    T = 101
    h = torch.arange(1, 12, dtype=torch.float32)         # 11 taps, all nonzero
    taps, J = PP.highpass_taps(h, T, tail=0.0)
    x = np.random.default_rng(1).standard_normal((2, T))
    ref = FO.circular_highpass_direct(x, h.double())
    np.testing.assert_allclose(circular(x, taps.astype(np.float64), J), ref, atol=1e-9)


def test_zero_filter_gives_one_zero_tap():
    taps, J = PP.highpass_taps(torch.zeros(31), 100)
    assert J == 0 and taps.tolist() == [0.0]
