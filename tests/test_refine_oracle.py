"""The refinement-loss oracle, pinned without the GPU: finite differences, a
closed form, and the matching path's own subtraction."""
import numpy as np
import torch

from kilosort_amd import synth
from kilosort_amd.batches import ArrayBatches
from oracle import refine_oracle as RO
from oracle import tm_oracle as O


def tiny(seed=3, C=4, K=3, T=50, nt=11, P=3):
    g = torch.Generator().manual_seed(seed)
    X = torch.randn((C, T), generator=g, dtype=torch.float64)
    U = torch.randn((K, P, C), generator=g, dtype=torch.float64)
    W = torch.randn((P, nt), generator=g, dtype=torch.float64)
    # overlapping spikes, one clipped at each edge
    st = torch.tensor([[2, 0], [20, 1], [24, 2], [27, 1], [47, 0]], dtype=torch.int64)
    amps = torch.rand(5, generator=g, dtype=torch.float64) + 0.5
    return X, U, W, amps, st, nt


def test_oracle_gradients_match_finite_differences():
    X, U, W, amps, st, nt = tiny()
    f = lambda X, U, W, amps: RO.residual(X, U, W, amps, st, nt)[1]
    args = tuple(t.clone().requires_grad_(True) for t in (X, U, W, amps))
    assert torch.autograd.gradcheck(f, args, eps=1e-6, atol=1e-7, rtol=1e-6)


def test_single_spike_closed_form():
    """X = 0, one interior spike, orthonormal W: loss = a^2 |U_i|^2,
    dL/da = 2 a |U_i|^2, dL/dU_i = 2 a^2 U_i."""
    P, nt, C = 6, 61, 8
    W = torch.from_numpy(synth.make_wpca(nt, P, 20).astype(np.float64))
    W = torch.linalg.qr(W.T).Q.T.contiguous()          # orthonormal to float64 rounding
    assert torch.allclose(W @ W.T, torch.eye(P, dtype=torch.float64), atol=1e-12)
    U = torch.randn((2, P, C), dtype=torch.float64, generator=torch.Generator().manual_seed(4))
    U.requires_grad_(True)
    a = torch.tensor([1.7], dtype=torch.float64, requires_grad=True)
    st = torch.tensor([[100, 1]])
    X = torch.zeros((C, 300), dtype=torch.float64)
    _, loss = RO.residual(X, U, W, a, st, nt)
    loss.backward()
    n2 = float((U[1] ** 2).sum())
    assert abs(float(loss) - 1.7 ** 2 * n2) < 1e-10 * n2
    assert abs(float(a.grad[0]) - 2 * 1.7 * n2) < 1e-10 * n2
    assert torch.allclose(U.grad[1], 2 * 1.7 ** 2 * U[1].detach(), atol=1e-10)
    assert torch.all(U.grad[0] == 0)


def test_model_is_the_peel_subtraction():
    """With the peel's own spikes and amplitudes, the model residual equals the
    peel's residual (windows of isolated spikes do not overlap)."""
    rec = synth.make_recording(5, 16, 8, 20000 + 122, rate_range=(2.0, 6.0))
    X = ArrayBatches(rec.data, NT=20000, nt=61, device=torch.device("cpu")).padded_batch_to_torch(0)
    U = torch.from_numpy(rec.U)
    ops = dict(nt=61, Th_learned=8, max_peels=100, wPCA=torch.from_numpy(rec.wPCA))
    st, amps, _, Xres = O.run_matching(ops, X, U, O.prepare_matching(ops, U))
    assert len(st) > 10
    xs = st[:, 0].numpy()
    order = np.argsort(xs)
    assert np.diff(xs[order]).min() >= 61, "fixture needs isolated spikes"
    R, _ = RO.residual(X.double(), U.double(), ops["wPCA"].double(), amps.double()[:, 0], st, 61)
    scale = float(X.abs().max())
    assert float((R - Xres.double()).abs().max()) < 1e-5 * scale
