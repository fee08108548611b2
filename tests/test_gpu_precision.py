"""Reduced-precision projection forms (BF16, split BF16 and FP8 on MFMA).

These forms are not bit-exact with the reference by construction; what they
must satisfy:
* the projection obeys the rounding-error bound of its operand format,
  checked elementwise against a float64 projection and the float64 sum of
  |U|.|W (*) X| (the bound scales with that sum, not with |B|);
* matching agrees with the FP32 oracle on (time, template) for nearly every
  spike, and the amplitudes of agreeing spikes are FP32 values (they are
  re-scored from the residual in FP32).
"""
import sys

import numpy as np
import pytest
import torch
import torch.nn.functional as F

from kilosort_amd import _lib, synth
from kilosort_amd import template_matching as tm
from kilosort_amd.batches import ArrayBatches
from oracle import tm_oracle as O

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda")
CPU = torch.device("cpu")
NT = 60000


def say(msg):
    print(f"[gpu-precision] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _lib.load()
    yield


@pytest.fixture(scope="module")
def batch1():
    rec = synth.config1_recording(seed=41, n_batches=1)
    X = ArrayBatches(rec.data, NT=NT, nt=61, device=CPU).padded_batch_to_torch(0)
    ops = dict(nt=rec.nt, nt0min=rec.nt0min, Th_learned=8, max_peels=100, batch_size=NT,
               wPCA=torch.from_numpy(rec.wPCA).to(DEV))
    return rec, ops, X


def _project(X, U, W, precision):
    lib = _lib.load()
    C, T = X.shape
    K, P, _ = U.shape
    ldbt = (K + 63) // 64 * 64
    Bt = torch.full(((T + 127) // 128 * 128, ldbt), float("nan"), device=DEV)
    nb = lib.ksm_project_workspace(C, T, K, P, 61, precision)
    ws = torch.empty(nb, dtype=torch.uint8, device=DEV)
    _lib.check(lib.ksm_project(X.data_ptr(), T, W.data_ptr(), U.data_ptr(), None, C, T, K, P,
                               61, Bt.data_ptr(), ldbt, precision, ws.data_ptr(), nb,
                               _lib.stream_of(DEV)), "ksm_project")
    torch.cuda.synchronize()
    return Bt[:T, :K].T.cpu().double().numpy()


@pytest.mark.parametrize("name,prec,unit", [("bf16", _lib.PREC_BF16, 2.0 ** -8),
                                            ("bf16x3", _lib.PREC_BF16X3, 2.0 ** -16)])
def test_projection_error_bound(batch1, name, prec, unit):
    """|B - B64| <= unit * sum|U||Bc| + FP32 slack, elementwise.

    unit: BF16 rounds each operand to 8 significant bits (relative error
    2^-9 each, 2^-8 for a product); the split form keeps hi + lo of each
    operand (2^-18) and drops lo*lo (2^-18): 3 * 2^-18 < 2^-16.
    """
    rec, ops, X = batch1
    say(f"projection {name} C=64 K=64 T=60122 vs float64")
    U = torch.from_numpy(rec.U)
    got = _project(X.to(DEV), U.to(DEV), ops["wPCA"], prec)
    Xd, Ud, Wd = X.double(), U.double(), ops["wPCA"].cpu().double()
    Bc = F.conv1d(Xd.unsqueeze(1), Wd.unsqueeze(1), padding=30)      # [C, P, T]
    ref = torch.einsum("kpc,cpt->kt", Ud, Bc).numpy()
    bound = torch.einsum("kpc,cpt->kt", Ud.abs(), Bc.abs()).numpy()
    err = np.abs(got - ref)
    slack = 2e-6 * np.abs(ref).max()
    ratio = err / (unit * bound + slack)
    say(f"{name}: max err {err.max():.3e}, max |B| {np.abs(ref).max():.3e}, "
        f"max err/bound {ratio.max():.3f}")
    assert np.isfinite(got).all()
    assert ratio.max() <= 1.0
    if prec == _lib.PREC_BF16:
        # a BF16 product is not exact: the error is visible
        assert err.max() > 1e-4 * np.abs(ref).max()


def test_fp8_projection_error_bound(batch1):
    """FP8 form: each operand is scaled into e4m3's range and rounded to 4
    significant bits (relative error <= 2^-4 for normal values, absolute
    error <= 2^-10 in the subnormal range), so per term of r = (c, p)
        |err| <= (2^-3 + 2^-8) |U||Bc| + 2^-10 (1 + 2^-4) (u |Bc| / s + s |U|)
                 + 2^-20 u,
    with s = max_t |X[c]| * sum_k |W[p, k]| / 448 the correlation row's scale
    and u = max_r |U[m, r] s_r| / 448 the template's; the scales are taken in
    float64 here, the absolute terms get 25 % slack for their float32 values."""
    rec, ops, X = batch1
    say("projection fp8 C=64 K=64 T=60122 vs float64")
    U = torch.from_numpy(rec.U)
    got = _project(X.to(DEV), U.to(DEV), ops["wPCA"], _lib.PREC_FP8)
    Xd, Ud, Wd = X.double(), U.double(), ops["wPCA"].cpu().double()
    Bc = F.conv1d(Xd.unsqueeze(1), Wd.unsqueeze(1), padding=30)      # [C, P, T]
    ref = torch.einsum("kpc,cpt->kt", Ud, Bc).numpy()
    s = Xd.abs().amax(1)[:, None] * Wd.abs().sum(1)[None, :] / 448.0  # [C, P]
    u = (Ud.abs() * s.T[None]).amax((1, 2)) / 448.0                  # [K]
    rel = torch.einsum("kpc,cpt->kt", Ud.abs(), Bc.abs()).numpy()
    a1 = torch.einsum("cp,cpt->t", 1.0 / s, Bc.abs()).numpy()[None, :] * u.numpy()[:, None]
    a2 = torch.einsum("kpc,cp->k", Ud.abs(), s).numpy()[:, None]
    nterm = U.shape[1] * U.shape[2]
    bound = (2.0 ** -3 + 2.0 ** -8) * rel + 1.25 * (
        2.0 ** -10 * (1 + 2.0 ** -4) * (a1 + a2) + 2.0 ** -20 * nterm * u.numpy()[:, None])
    err = np.abs(got - ref)
    slack = 2e-6 * np.abs(ref).max()
    ratio = err / (bound + slack)
    say(f"fp8: max err {err.max():.3e}, max |B| {np.abs(ref).max():.3e}, "
        f"median rel err {np.median(err / np.maximum(np.abs(ref), 1e-30)):.3e}, "
        f"max err/bound {ratio.max():.3f}")
    assert np.isfinite(got).all()
    assert ratio.max() <= 1.0
    # the e4m3 rounding is visible
    assert err.max() > 1e-3 * np.abs(ref).max()


def _pairs(st):
    return set(map(tuple, np.asarray(st).tolist()))


def _near(a, b):
    """Spikes of a with a spike of the same template in b within one sample."""
    return sum(1 for t, i in a if (t, i) in b or (t - 1, i) in b or (t + 1, i) in b)


# exact (time, template) agreement; agreement within one sample
@pytest.mark.parametrize("name,min_agree,min_near", [("bf16x3", 0.999, 0.999),
                                                     ("bf16", 0.99, 0.99),
                                                     ("fp8", 0.85, 0.9)])
def test_matching_agreement_with_fp32_oracle(batch1, name, min_agree, min_near):
    rec, ops, X = batch1
    say(f"run_matching {name} C=64 K=64 vs FP32 oracle")
    U = torch.from_numpy(rec.U)
    ops_c = dict(ops)
    ops_c["wPCA"] = ops["wPCA"].cpu()
    ref_st, ref_amps, ref_th, _ = O.run_matching(ops_c, X, U, O.prepare_matching(ops_c, U))
    o = dict(ops, matching_precision=name)
    Ud = U.to(DEV)
    st, amps, th, Xres = tm.run_matching(o, X.to(DEV), Ud, tm.prepare_matching(o, Ud),
                                         device=DEV)
    st = st.cpu().numpy()
    ref_st = ref_st.numpy()
    a, b = _pairs(st), _pairs(ref_st)
    agree = len(a & b) / max(len(a), len(b))
    near = _near(a, b) / max(len(a), len(b))
    say(f"{name}: {len(st)} spikes, oracle {len(ref_st)}, agreement {100 * agree:.3f} %, "
        f"within one sample {100 * near:.3f} %")
    assert len(ref_st) > 1000
    assert agree >= min_agree
    assert near >= min_near
    # amplitudes of the spikes found by both: FP32 re-scored values
    ref_idx = {p: i for i, p in enumerate(map(tuple, ref_st.tolist()))}
    both = [(i, ref_idx[p]) for i, p in enumerate(map(tuple, st.tolist())) if p in ref_idx]
    gi = np.array([i for i, _ in both])
    ri = np.array([j for _, j in both])
    ga = amps.cpu().numpy()[gi, 0].astype(np.float64)
    ra = ref_amps.numpy()[ri, 0].astype(np.float64)
    rel = np.abs(ga - ra) / np.abs(ra)
    say(f"{name}: amplitude rel err median {np.median(rel):.2e} max {rel.max():.2e}")
    assert np.median(rel) <= 1e-5
    gt = th.cpu().numpy()[gi, 0].astype(np.float64)
    rt = ref_th.numpy()[ri, 0].astype(np.float64)
    assert np.median(np.abs(gt - rt) / np.abs(rt)) <= 1e-5
    assert torch.isfinite(Xres).all()


def test_unknown_precision_is_rejected(batch1):
    rec, ops, X = batch1
    U = torch.from_numpy(rec.U).to(DEV)
    with pytest.raises(ValueError, match="matching_precision"):
        tm.run_matching(dict(ops, matching_precision="fp4"), X.to(DEV), U,
                        tm.prepare_matching(ops, U), device=DEV)
