"""Exact score ties in the peel, against the oracle's peel on the same scores.

Exact ties make the reference accept spikes closer than nt + 1 columns; where
the windows of one subtraction phase then overlap, the reference's indexed
subtraction keeps the last spike's write (torch's CPU index_put).  These cases
plant ties in B directly, so both sides start from identical scores, Gram and
norms and must agree exactly on (time, template) and to rounding on
amplitudes and the residual:

* a tied pair (x, x+1) of one template: two spikes one column apart, their
  residual windows overlapping;
* a tied pair plus a spike at x+63: columns x+2 .. x+61 covered by three;
* two tied pairs: columns covered by four;
* a three-way tie: three overlapping residual windows.
"""
import numpy as np
import pytest
import torch

from kilosort_amd import _lib, synth
from kilosort_amd.batches import ArrayBatches
from oracle import tm_oracle as O
from tests.gpu_peel import gpu_peel

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda")
CPU = torch.device("cpu")
NT = 3000


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _lib.load()
    yield


def _case(plants, seed=61):
    rec = synth.make_recording(seed, 32, 12, NT + 2 * 61, rate_range=(0.01, 0.02))
    X = ArrayBatches(rec.data, NT=NT, nt=61, device=CPU).padded_batch_to_torch(0)
    U = torch.from_numpy(rec.U)
    ops = dict(nt=61, Th_learned=8, max_peels=20, wPCA=torch.from_numpy(rec.wPCA))
    K, P, C = U.shape
    T = X.shape[1]
    # small background scores, then planted peaks
    g = torch.Generator().manual_seed(seed)
    B = (torch.rand((K, T), generator=g) - 0.5).float()
    for (j, x, v) in plants:
        B[j, x] = v
    return ops, X, U, B


CASES = {
    "pair": [(3, 1500, 40.0), (3, 1501, 40.0)],
    "pair_plus_third": [(3, 1500, 40.0), (3, 1501, 40.0), (7, 1564, 30.0)],
    "two_pairs": [(3, 1500, 40.0), (3, 1501, 40.0), (7, 1564, 30.0), (7, 1565, 30.0)],
    "three_way": [(5, 2000, 35.0), (5, 2001, 35.0), (5, 2002, 35.0)],
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_exact_ties_follow_the_reference(name):
    print(f"[gpu-ties] {name}", flush=True)
    ops, X, U, B = _case(CASES[name])
    got, ctc_ref, nm = gpu_peel(ops, X, U, B, capacity=10000)
    ref = O.peel(ops, B.clone(), X, U, ctc_ref, nm, capacity=10000)
    st, amps, th, Xres = got
    assert len(ref[0]) >= len(CASES[name])
    np.testing.assert_array_equal(st.numpy(), ref[0].numpy())
    np.testing.assert_allclose(amps.numpy(), ref[1].numpy(), rtol=1e-6)
    np.testing.assert_allclose(th.numpy(), ref[2].numpy(), rtol=1e-6)
    scale = float(np.abs(X.numpy()).max())
    assert np.abs(Xres.numpy() - ref[3].numpy()).max() <= 1e-5 * scale
