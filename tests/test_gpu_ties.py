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
* a three-way tie: three overlapping residual windows;
* a plateau of 300 tied columns: more accepted spikes than a 256-column chunk
  would hold with a per-chunk cap, 123 windows covering a column and 61
  overlapping residual windows;
* the replicated-edge plateau of an unfiltered partial last batch
  (io.py:816-823: the columns after the data repeat its last sample), scored
  above threshold, through the peel and through run_matching end to end.
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
    "plateau": [(4, x, 30.0) for x in range(1000, 1300)],
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


def _partial_last_batch(seed=71, C=32, K=12):
    """The last, partial batch of a recording whose final sample is large on
    template 0's channels: its edge replication is a plateau of tied scores
    far above threshold."""
    rec = synth.make_recording(seed, C, K, NT + 1000, rate_range=(2.0, 6.0),
                               amp_range=(10.0, 20.0))
    data = rec.data.copy()
    T0 = np.einsum("pc,pk->ck", rec.U[0], rec.wPCA)          # template 0 [C, nt]
    s0 = T0.sum(1)
    A = 30.0 / np.sqrt((s0 ** 2).sum())                       # plateau score ~ 30 sqrt(nm)
    data[-1] = (A * s0).astype(np.float32)
    bf = ArrayBatches(data, NT=NT, nt=61, device=CPU)
    assert bf.n_batches == 2
    X = bf.padded_batch_to_torch(1)
    ops = dict(nt=61, Th_learned=8, max_peels=20, wPCA=torch.from_numpy(rec.wPCA))
    return ops, X, torch.from_numpy(rec.U)


def test_replicated_edge_plateau_peel_matches_oracle():
    print("[gpu-ties] partial last batch, replicated-edge plateau (same scores)", flush=True)
    ops, X, U = _partial_last_batch()
    B = O.project(ops, X, U)
    got, ctc_ref, nm = gpu_peel(ops, X, U, B, capacity=100000)
    ref = O.peel(ops, B.clone(), X, U, ctc_ref, nm, capacity=100000)
    st, amps, th, Xres = got
    print(f"[gpu-ties] plateau: {len(ref[0])} spikes (oracle), {len(st)} (device)", flush=True)
    assert len(ref[0]) > 1000                      # the plateau is accepted column by column
    np.testing.assert_array_equal(st.numpy(), ref[0].numpy())
    np.testing.assert_allclose(amps.numpy(), ref[1].numpy(), rtol=1e-6)
    np.testing.assert_allclose(th.numpy(), ref[2].numpy(), rtol=1e-6)
    scale = float(np.abs(X.numpy()).max())
    assert np.abs(Xres.numpy() - ref[3].numpy()).max() <= 1e-5 * scale


def test_replicated_edge_plateau_run_matching():
    """End to end on the device's own scores: no error, and the same spikes
    as the oracle's run_matching (the plateau's scores are tied on both
    sides, each side's own float32 values)."""
    from kilosort_amd import template_matching as tm
    print("[gpu-ties] partial last batch, run_matching end to end", flush=True)
    ops, X, U = _partial_last_batch()
    ref = O.run_matching(ops, X, U, O.prepare_matching(ops, U))
    o = dict(ops, wPCA=ops["wPCA"].to(DEV))
    Ud = U.to(DEV)
    st, amps, th, Xres = tm.run_matching(o, X.to(DEV), Ud, tm.prepare_matching(o, Ud), device=DEV)
    st = st.cpu().numpy()
    print(f"[gpu-ties] run_matching: {len(st)} spikes, oracle {len(ref[0])}", flush=True)
    assert len(ref[0]) > 1000
    np.testing.assert_array_equal(st, ref[0].numpy())
    np.testing.assert_allclose(amps.cpu().numpy(), ref[1].numpy(), rtol=1e-4)
