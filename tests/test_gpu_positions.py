"""Spike positions and amplitudes for the output files (ksm_spike_positions
through kilosort_amd.postprocessing) against the position oracle and the
Frobenius norm of the features.

Bar: positions within 1e-5 of the largest channel coordinate (float32 sums
in another order); spikes with no feature mass are NaN in both."""
import sys

import numpy as np
import pytest
import torch

from kilosort_amd import _lib, synth
from kilosort_amd import postprocessing as PP
from kilosort_amd import template_matching as TMg
from oracle import positions_oracle as PO

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _lib.load()
    yield


@pytest.mark.parametrize("C,K,n,nC,P", [(64, 64, 5000, 10, 6), (383, 700, 200000, 10, 6),
                                        (64, 30, 2000, 16, 6), (96, 40, 3000, 32, 6),
                                        (96, 40, 3001, 32, 9)])
def test_positions_match_oracle(C, K, n, nC, P):
    """nC P <= 255 stages the rows through LDS (blocks of 256, 128 or 64
    spikes); 32 x 9 = 288 takes the unstaged kernel."""
    rec = synth.make_recording(80 + C, C, K, 1000, with_data=False)
    U = torch.from_numpy(rec.U).to(DEV)
    iCC, iCC_mask, iU, _ = TMg.prepare_extract(rec.xc, rec.yc, U, nC, 60.0, device=DEV)
    ops = dict(xc=rec.xc.astype(np.float32), yc=rec.yc.astype(np.float32), iCC=iCC,
               iCC_mask=iCC_mask, iU=iU)
    rng = np.random.default_rng(C)
    st = np.zeros((n, 6))
    st[:, 1] = rng.integers(0, K, n)
    tF = torch.from_numpy(rng.standard_normal((n, nC, P)).astype(np.float32))
    tF[:3] = 0                                              # no mass: 0/0
    print(f"[gpu-positions] C={C} K={K} n={n} nC={nC} P={P}", flush=True)
    sys.stdout.flush()
    xs, ys = PP.compute_spike_positions(st, tF.to(DEV), ops, device=DEV)
    rx, ry = PO.compute_spike_positions(st, tF, ops)
    _, _, amp = PP.positions_and_amplitudes(st, tF, ops, device=DEV)
    ramp = torch.norm(tF, dim=[-2, -1]).numpy()             # io.py:330
    np.testing.assert_allclose(amp.cpu().numpy(), ramp, rtol=1e-6, atol=1e-7)
    assert xs.dtype == np.float32 and xs.shape == (n,)
    assert np.isnan(xs[:3]).all() and np.isnan(rx[:3]).all()
    scale = max(np.abs(ops['xc']).max(), np.abs(ops['yc']).max())
    np.testing.assert_allclose(xs, rx, atol=1e-5 * scale, rtol=0)
    np.testing.assert_allclose(ys, ry, atol=1e-5 * scale, rtol=0)


def test_positions_empty():
    rec = synth.make_recording(90, 32, 8, 1000, with_data=False)
    U = torch.from_numpy(rec.U).to(DEV)
    iCC, iCC_mask, iU, _ = TMg.prepare_extract(rec.xc, rec.yc, U, 10, 60.0, device=DEV)
    ops = dict(xc=rec.xc, yc=rec.yc, iCC=iCC, iCC_mask=iCC_mask, iU=iU)
    xs, ys = PP.compute_spike_positions(np.zeros((0, 6)), torch.zeros((0, 10, 6)), ops, device=DEV)
    assert xs.shape == (0,) and ys.shape == (0,)
