"""The peel with many templates and wide Gram unions, against the oracle's peel
on the same scores.

The column update reduces group keys (32 templates each) in passes of 8
groups; these cases reach the less common shapes:

* 600 templates on 48 channels: two covering spikes whose Gram ranges lie far
  apart need more than 16 groups (three passes per column);
* 2200 templates (69 groups, more than one per lane of a wave): the keys of
  the groups a column leaves untouched come from several loads per lane.

Both sides start from the same scores B (the oracle's float32 projection),
Gram and norms, and must agree exactly on (time, template) and to rounding on
amplitudes and the residual.
"""
import sys

import numpy as np
import pytest
import torch

from kilosort_amd import _lib, synth
from kilosort_amd.batches import ArrayBatches
from oracle import tm_oracle as O
from tests.gpu_peel import gpu_peel

pytestmark = pytest.mark.gpu

CPU = torch.device("cpu")
NT = 3000


def say(msg):
    print(f"[gpu-wide] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _lib.load()
    yield


@pytest.mark.parametrize("C,K,seed", [(48, 600, 71), (64, 2200, 72)])
def test_wide_peel_matches_oracle(C, K, seed):
    say(f"C={C} K={K} NT={NT} seed={seed}")
    rec = synth.make_recording(seed, C, K, NT + 2 * 61, rate_range=(20.0, 60.0),
                               amp_range=(10.0, 30.0))
    X = ArrayBatches(rec.data, NT=NT, nt=61, device=CPU).padded_batch_to_torch(0)
    U = torch.from_numpy(rec.U)
    ops = dict(nt=61, Th_learned=8, max_peels=100, wPCA=torch.from_numpy(rec.wPCA))
    B = O.project(ops, X, U).float().contiguous()
    got, ctc, nm = gpu_peel(ops, X, U, B.clone())
    ref = O.peel(ops, B.clone(), X, U, ctc, nm)
    del ctc
    st, amps, th, Xres = got
    say(f"GPU {len(st)} spikes, oracle {len(ref[0])}")
    assert len(ref[0]) > 50
    np.testing.assert_array_equal(st.numpy(), ref[0].numpy())
    np.testing.assert_allclose(amps.numpy(), ref[1].numpy(), rtol=1e-6)
    np.testing.assert_allclose(th.numpy(), ref[2].numpy(), rtol=1e-6)
    scale = float(np.abs(X.numpy()).max())
    assert np.abs(Xres.numpy() - ref[3].numpy()).max() <= 1e-5 * scale
