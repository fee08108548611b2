"""Template refinement kernels (residual and its gradients) against the
float64 autograd oracle, finite differences, and the averaging update.

Bars: float64 kernels within 1e-10 of the oracle (relative to the largest
entry of each output); float32 kernels within 2e-5; the float64 backward
passes torch.autograd.gradcheck; both passes are bit-reproducible.
"""
import sys

import numpy as np
import pytest
import torch

from kilosort_amd import _lib, refine, synth
from kilosort_amd.batches import ArrayBatches
from oracle import refine_oracle as RO

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda")
CPU = torch.device("cpu")


def say(msg):
    print(f"[gpu-refine] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _lib.load()
    yield


def instance(seed, C, K, NT, rate=(20.0, 60.0)):
    """A padded batch with its true spikes (times in padded-batch columns)."""
    rec = synth.make_recording(seed, C, K, NT + 122, rate_range=rate, overlap_frac=0.2)
    X = ArrayBatches(rec.data, NT=NT, nt=61, device=CPU).padded_batch_to_torch(0)
    x = np.asarray(rec.times) + 61 - rec.nt0min + 61 // 2   # trough-aligned -> column
    keep = (x >= 0) & (x < NT + 122)
    st = torch.from_numpy(np.stack([x[keep], np.asarray(rec.ids)[keep]], 1).astype(np.int64))
    amps = torch.from_numpy(np.asarray(rec.amps, np.float64)[keep])
    return X.double(), torch.from_numpy(rec.U).double(), torch.from_numpy(rec.wPCA).double(), \
        amps, st


def rel_err(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 2e-5)])
def test_residual_and_gradients_match_oracle(dtype, tol):
    X, U, W, amps, st = instance(91, 32, 16, 3000)
    say(f"C=32 K=16 NT=3000 {len(st)} spikes, {dtype}")
    assert len(st) > 30
    # oracle (float64 autograd on the CPU)
    args = [t.clone().requires_grad_(True) for t in (X, U, W, amps)]
    R64, L64 = RO.residual(*args, st, 61)
    L64.backward()
    g64 = [a.grad for a in args]
    # device
    dargs = [t.to(DEV, dtype).requires_grad_(True) for t in (X, U, W, amps)]
    R, L = refine.residual(*[a.detach() for a in dargs], st.to(DEV), 61)
    assert rel_err(R.cpu(), R64.detach()) <= tol
    assert abs(float(L) - float(L64)) <= tol * float(L64)
    loss = refine.residual_loss(*dargs, st.to(DEV), 61)
    loss.backward()
    for name, a, g in zip(("dX", "dU", "dW", "damp"), dargs, g64):
        e = rel_err(a.grad.cpu(), g)
        say(f"{name}: rel err {e:.2e}")
        assert e <= tol, name


def test_gradcheck_float64():
    """Finite differences of the device loss (float64 kernels)."""
    X, U, W, amps, st = instance(92, 32, 16, 3000)
    say(f"gradcheck C=32 K=16 NT=3000 {len(st)} spikes")
    Xd = X.to(DEV)
    std = st.to(DEV)
    f = lambda U, W, amps: refine.residual_loss(Xd, U, W, amps, std, 61)
    args = tuple(t.to(DEV).requires_grad_(True) for t in (U, W, amps))
    # the loss is quadratic in each argument alone, so central differences
    # are exact up to rounding: a wide step keeps that rounding small
    assert torch.autograd.gradcheck(f, args, eps=1e-3, atol=1e-5, rtol=1e-5, fast_mode=True)
    # full Jacobian on a smaller instance
    X, U, W, amps, st = instance(93, 8, 4, 400, rate=(200.0, 400.0))
    assert len(st) > 5
    Xd, std = X.to(DEV), st.to(DEV)
    f = lambda U, W, amps: refine.residual_loss(Xd, U, W, amps, std, 61)
    args = tuple(t.to(DEV).requires_grad_(True) for t in (U, W, amps))
    assert torch.autograd.gradcheck(f, args, eps=1e-3, atol=1e-5, rtol=1e-5)


def test_backward_is_reproducible():
    X, U, W, amps, st = instance(94, 48, 24, 6000)
    dargs = [t.to(DEV, torch.float32) for t in (X, U, W, amps)]
    out = []
    for _ in range(2):
        a = [t.clone().requires_grad_(True) for t in dargs]
        refine.residual_loss(*a, st.to(DEV), 61).backward()
        out.append([t.grad.cpu().numpy() for t in a])
    for g0, g1 in zip(*out):
        np.testing.assert_array_equal(g0, g1)


def averaging_cpu(batches, U, W, nt):
    """The averaging update written out on the CPU in float64: each template
    becomes sum_s amp_s F_s / sum_s amp_s^2 with F_s the projection on W of
    the spike's residual window plus its own contribution."""
    U = U.double().cpu()
    W = W.double().cpu()
    h = nt // 2
    num = torch.zeros_like(U)
    a2 = torch.zeros(U.shape[0], dtype=torch.float64)
    for X, st, am in batches:
        X, st, am = X.double().cpu(), st.cpu(), am.double().cpu()
        R, _ = RO.residual(X, U, W, am, st, nt)
        T = X.shape[1]
        for (x, i), a in zip(st.tolist(), am.tolist()):
            lo, hi = max(0, x - h), min(T, x + h + 1)
            k0 = lo - (x - h)
            win = R[:, lo:hi] + a * (U[i].T @ W)[:, k0:k0 + hi - lo]
            F = win @ W[:, k0:k0 + hi - lo].T                    # [C, P]
            num[i] += a * F.T
            a2[i] += a * a
    out = U.clone()
    keep = a2 > 0
    out[keep] = num[keep] / a2[keep][:, None, None]
    return out


def test_gradient_refinement_vs_averaging_update():
    """From perturbed templates: the averaging update (device) matches its
    explicit CPU form, and 20 preconditioned gradient steps (backtracking
    where a step would raise the loss) reach a residual no larger than it
    (the first step is that update wherever no window is clipped by the batch
    edges), the loss never rising from step to step."""
    say("refinement: 2 batches C=64 K=32 NT=20000, 20 steps")
    rec = synth.make_recording(95, 64, 32, 2 * 20000 + 122, rate_range=(10.0, 30.0),
                               overlap_frac=0.2)
    bf = ArrayBatches(rec.data, NT=20000, nt=61, device=DEV)
    x = np.asarray(rec.times) + 61 - rec.nt0min + 61 // 2
    ids = np.asarray(rec.ids)
    a = np.asarray(rec.amps, np.float32)
    batches = []
    for b in range(2):
        lo = b * 20000
        keep = (x - lo >= 0) & (x - lo < 20000 + 122)
        st = torch.from_numpy(np.stack([x[keep] - lo, ids[keep]], 1).astype(np.int64)).to(DEV)
        batches.append((bf.padded_batch_to_torch(b), st, torch.from_numpy(a[keep]).to(DEV)))
    W = torch.from_numpy(rec.wPCA).to(DEV)
    g = torch.Generator().manual_seed(3)
    U0 = torch.from_numpy(rec.U)
    U0 = (U0 + 0.3 * U0.abs().max() * torch.randn(U0.shape, generator=g) *
          (U0 != 0)).to(DEV)
    Uavg = refine.averaging_update(batches, U0, W, 61)
    assert rel_err(Uavg.cpu(), averaging_cpu(batches, U0, W, 61)) <= 1e-4
    _, loss0 = refine.refine_step(batches, U0, W, 61)
    loss = lambda U: sum(float(refine.residual_loss(X, U, W, am, st, 61))
                         for X, st, am in batches)
    l_avg = loss(Uavg)
    U, hist = refine.refine(batches, U0, W, 61, steps=20)
    assert all(b <= a for a, b in zip(hist, hist[1:]))
    l_gd = loss(U)
    say(f"loss: initial {loss0:.6g}, averaging {l_avg:.6g}, 20 gradient steps {l_gd:.6g}")
    assert l_avg < loss0
    assert l_gd <= l_avg * (1 + 1e-6)


def test_full_size_gradients_and_refinement():
    """The differentiation workload at its full size: 8 batches of NT=60000
    at 383 channels, 500 templates from the 64-channel workload's generator,
    spike times and ids fixed by a forward FP32 matching pass.  Gradients of
    the float32 kernels match the float64 kernels over every batch (2e-5 of
    each output's largest entry); the float64 kernels match the float64
    autograd oracle on a 6000-column slice of batch 0 (1e-10); and from
    perturbed templates, 20 gradient steps reach a residual no larger than
    the averaging update's.  At this density spikes of different templates
    overlap often, so the unscaled per-template step overshoots (the loss
    diverges without backtracking); refine halves it where it would raise
    the loss."""
    from kilosort_amd import template_matching as tm
    NT, C, K, NB = 60000, 383, 500, 8
    say(f"differentiation workload: {NB} batches x NT={NT}, C={C}, K={K}")
    rec = synth.make_recording(96, C, K, NB * NT + 122, rate_range=(2.0, 20.0),
                               amp_range=(10.0, 30.0), with_data=False)
    data = synth.synthesize_torch(rec.n_samples, C, rec.times, rec.ids, rec.amps, rec.U,
                                  rec.wPCA, rec.nt0min, DEV, seed=96)
    bf = ArrayBatches(data, NT=NT, nt=61, device=DEV)
    W = torch.from_numpy(rec.wPCA).to(DEV)
    Ut = torch.from_numpy(rec.U).to(DEV)
    ops = dict(nt=61, Th_learned=8, max_peels=100, wPCA=W)
    ctc = tm.prepare_matching(ops, Ut)
    batches = []
    for b in range(NB):
        X = bf.padded_batch_to_torch(b).contiguous()
        st, amps, _, _ = tm.run_matching(ops, X, Ut, ctc, device=DEV)
        batches.append((X, st, amps[:, 0].contiguous()))
    nsp = sum(len(s) for _, s, _ in batches)
    say(f"forward FP32 pass: {nsp} spikes")
    assert nsp > 50000
    # float32 kernels against float64 kernels, every batch
    g = torch.Generator().manual_seed(4)
    U0 = torch.from_numpy(rec.U)
    U0 = (U0 + 0.3 * U0.abs().max() * torch.randn(U0.shape, generator=g) * (U0 != 0)).to(DEV)
    worst = 0.0
    for X, st, am in batches:
        grads = []
        for dt in (torch.float32, torch.float64):
            a = [t.detach().to(dt).clone().requires_grad_(True) for t in (U0, W, am)]
            refine.residual_loss(X.to(dt), *a, st, 61).backward()
            grads.append([t.grad.cpu() for t in a])
        for g32, g64 in zip(*grads):
            worst = max(worst, rel_err(g32, g64))
    say(f"float32 vs float64 kernels, worst relative error over batches: {worst:.2e}")
    assert worst <= 2e-5
    # float64 kernels against the oracle on a slice of batch 0
    X, st, am = batches[0]
    lo, n = 20000, 6000
    keep = ((st[:, 0] >= lo) & (st[:, 0] < lo + n)).cpu()
    sts = st.cpu()[keep] - torch.tensor([lo, 0])
    Xs = X[:, lo:lo + n].double().cpu()
    args = [t.clone().requires_grad_(True) for t in (Xs, U0.double().cpu(), W.double().cpu(),
                                                      am.cpu()[keep].double())]
    _, L64 = RO.residual(*args, sts, 61)
    L64.backward()
    dargs = [t.to(DEV).requires_grad_(True) for t in (Xs, U0.double().cpu(), W.double().cpu(),
                                                       am.cpu()[keep].double())]
    refine.residual_loss(*dargs, sts.to(DEV), 61).backward()
    say(f"oracle slice: {int(keep.sum())} spikes")
    for name, a, o in zip(("dX", "dU", "dW", "damp"), dargs, args):
        e = rel_err(a.grad.cpu(), o.grad)
        say(f"  {name}: rel err {e:.2e}")
        assert e <= 1e-10, name
    # 20 gradient steps against the averaging update
    loss = lambda U: sum(float(refine.residual_loss(X, U, W, am, st, 61))
                         for X, st, am in batches)
    Uavg = refine.averaging_update(batches, U0, W, 61)
    l0, l_avg = loss(U0), loss(Uavg)
    U, hist = refine.refine(batches, U0, W, 61, steps=20)
    say(f"losses of the 20 steps: {hist[1]:.6g} .. {hist[-1]:.6g}")
    assert all(b <= a for a, b in zip(hist, hist[1:]))
    l_gd = loss(U)
    say(f"residual: initial {l0:.6g}, averaging {l_avg:.6g}, 20 gradient steps {l_gd:.6g}")
    assert l_avg < l0
    assert l_gd <= l_avg * (1 + 1e-6)
