"""Reduced-precision projection forms (BF16, split BF16 and FP8 on MFMA).

These forms are not bit-exact with the reference by construction; what they
must satisfy:
* the projection obeys the rounding-error bound of its operand format,
  checked elementwise against a float64 projection and the float64 sum of
  |U|.|W (*) X| (the bound scales with that sum, not with |B|);
* the bound ksm_project_bounded returns holds for every template and column
  (per unit template norm);
* matching takes the FP32 decisions: every decision the bound leaves open is
  taken on FP32 scores re-scored from the residual, so (time, template)
  agree with the FP32 oracle and the FP64 oracle to >= 99.9 % (the FP32
  form's own agreement with FP64), and the amplitudes of agreeing spikes are
  FP32 values.
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


def _bound_sections(C, T, K, nt):
    """Byte offsets of the bound buffer's sections (include/ksm.h)."""
    nsb = (C + 31) // 32
    out, o = {}, 0
    for name, n in (("hdr", 256), ("colD", 4 * T), ("colA", 4 * T), ("S", 4 * nsb * T),
                    ("Wt", 4 * nsb * K), ("srng", 8 * K), ("T", 4 * K * C * nt)):
        out[name] = o
        o = (o + n + 255) // 256 * 256
    return out, nsb, o


@pytest.mark.parametrize("name,prec", [("bf16", _lib.PREC_BF16), ("bf16x3", _lib.PREC_BF16X3),
                                       ("fp8", _lib.PREC_FP8)])
def test_bound_holds_per_template(batch1, name, prec):
    """|B~[r,t] - B64[r,t]| / sqrt(nm_r) <= D_r(t) + colA[t] <= colD[t] for every
    template and column (BF16 forms: a proven bound, so never violated; FP8: a
    6-sigma margin of the rounding model, which a batch should not exceed
    either), and the re-scoring table holds each template's waveforms on its
    significant channels."""
    rec, ops, X = batch1
    lib = _lib.load()
    C, T = X.shape
    U = torch.from_numpy(rec.U)
    K, P, _ = U.shape
    Xd, Ud, Wd = X.to(DEV), U.to(DEV), ops["wPCA"]
    ldbt = (K + 63) // 64 * 64
    rows = (T + 127) // 128 * 128
    Bt = torch.full((rows, ldbt), float("nan"), device=DEV)
    nbb = lib.ksm_bound_bytes(C, T, K, P, 61)
    sec, nsb, total = _bound_sections(C, T, K, 61)
    assert nbb == total
    bb = torch.full((nbb,), 0xFF, dtype=torch.uint8, device=DEV)
    nb = lib.ksm_project_workspace(C, T, K, P, 61, prec)
    ws = torch.empty(nb, dtype=torch.uint8, device=DEV)
    _lib.check(lib.ksm_project_bounded(Xd.data_ptr(), T, Wd.data_ptr(), Ud.data_ptr(), None, C, T,
                                       K, P, 61, Bt.data_ptr(), ldbt, prec, 0, bb.data_ptr(), nbb,
                                       ws.data_ptr(), nb, _lib.stream_of(DEV)),
               "ksm_project_bounded")
    raw = bb.cpu().numpy()

    def f32(name, n):
        return raw[sec[name]:sec[name] + 4 * n].view(np.float32).astype(np.float64)

    mode = int(raw[:4].view(np.int32)[0])
    assert mode == (1 if prec == _lib.PREC_FP8 else 0)
    colD, colA = f32("colD", T), f32("colA", T)
    S = f32("S", nsb * T).reshape(nsb, T)
    Wt = f32("Wt", nsb * K).reshape(nsb, K)
    D = Wt.T @ S
    if mode == 1:
        D = np.sqrt(D)
    D = D + colA[None, :]
    got = Bt[:T, :K].T.cpu().double().numpy()
    Xq, Uq, Wq = X.double(), U.double(), ops["wPCA"].cpu().double()
    Bc = F.conv1d(Xq.unsqueeze(1), Wq.unsqueeze(1), padding=30)
    ref = torch.einsum("kpc,cpt->kt", Uq, Bc).numpy()
    nrm = np.sqrt((U.double() ** 2).sum((1, 2)).numpy())[:, None]
    err = np.abs(got - ref) / nrm
    # columns whose correlation is exactly zero (the zero tail of a padded
    # batch shorter than NT + 2 nt) have a zero bound and no error
    ratio = np.where(err == 0, 0.0, err / np.where(D > 0, D, 1e-300))
    say(f"{name}: max error / bound {ratio.max():.3f}, median bound {np.median(D):.3e} "
        f"(per column {np.median(colD):.3e}), {int((colD == 0).sum())} exactly zero columns")
    assert np.isfinite(D).all() and (D >= 0).all()
    assert ratio.max() <= 1.0
    assert (D <= colD[None, :] * (1 + 1e-5)).all()
    # re-scoring table: significant channels and waveforms T_r = U_r^T W
    srng = raw[sec["srng"]:sec["srng"] + 8 * K].view(np.int32).reshape(K, 2)
    Tt = raw[sec["T"]:sec["T"] + 4 * K * C * 61].view(np.float32).reshape(K, C, 61)
    for r in range(0, K, 7):
        lo, hi = srng[r]
        e = (U[r].double() ** 2).sum(0).numpy()
        assert lo < hi and e[:lo].sum() + e[hi:].sum() <= 1e-18 * e.sum()
        want = torch.einsum("pc,pk->ck", U[r].double(), Wq).numpy()[lo:hi]
        np.testing.assert_allclose(Tt[r, :hi - lo], want, rtol=1e-5, atol=1e-6 * np.abs(want).max())
    with pytest.raises(ValueError):
        _lib.check(lib.ksm_project_bounded(Xd.data_ptr(), T, Wd.data_ptr(), Ud.data_ptr(), None,
                                           C, T, K, P, 61, Bt.data_ptr(), ldbt, _lib.PREC_FP32, 0,
                                           bb.data_ptr(), nbb, ws.data_ptr(), nb,
                                           _lib.stream_of(DEV)), "ksm_project_bounded")


# exact (time, template) agreement; agreement within one sample
@pytest.mark.parametrize("name,min_agree,min_near", [("bf16x3", 0.999, 0.999),
                                                     ("bf16", 0.999, 0.999),
                                                     ("fp8", 0.999, 0.999)])
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
        f"within one sample {100 * near:.3f} %, {tm._iters.rescored} FP32 re-scores")
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


def _pairs_agree(st, ref_st):
    a, b = _pairs(st), _pairs(ref_st)
    return len(a & b) / max(len(a), len(b), 1)


def test_forms_agree_with_fp64_oracle_at_383_channels():
    """One 383-channel x 700-template batch (Neuropixels-sized workload): every
    form against the FP64 oracle, >= 99.9 % exact (time, template) agreement
    and <= 1e-3 median relative amplitude error."""
    rec = synth.config2_recording(seed=5, n_batches=1)
    X = ArrayBatches(rec.data, NT=NT, nt=61, device=CPU).padded_batch_to_torch(0)
    U = torch.from_numpy(rec.U)
    W = torch.from_numpy(rec.wPCA)
    say("FP64 oracle run_matching C=383 K=700 (CPU)")
    torch.set_num_threads(min(16, torch.get_num_threads()))
    ops64 = dict(nt=61, Th_learned=8, max_peels=100, wPCA=W.double())
    st64, amps64, _, _ = O.run_matching(ops64, X.double(), U.double(),
                                        O.prepare_matching(ops64, U.double()))
    st64 = st64.numpy()
    ref_idx = {p: i for i, p in enumerate(map(tuple, st64.tolist()))}
    Ud, Wd, Xd = U.to(DEV), W.to(DEV), X.to(DEV)
    for name in ("fp32", "bf16x3", "bf16", "fp8"):
        o = dict(nt=61, Th_learned=8, max_peels=100, wPCA=Wd, matching_precision=name)
        st, amps, th, _ = tm.run_matching(o, Xd, Ud, tm.prepare_matching(o, Ud), device=DEV)
        st = st.cpu().numpy()
        agree = _pairs_agree(st, st64)
        both = [(i, ref_idx[p]) for i, p in enumerate(map(tuple, st.tolist())) if p in ref_idx]
        gi = np.array([i for i, _ in both])
        ri = np.array([j for _, j in both])
        rel = np.abs(amps.cpu().numpy()[gi, 0] - amps64.numpy()[ri, 0]) / np.abs(
            amps64.numpy()[ri, 0])
        say(f"{name}: {len(st)} spikes (FP64 {len(st64)}), exact agreement {100 * agree:.3f} %, "
            f"median amplitude error {np.median(rel):.2e}, "
            f"{getattr(tm._iters, 'rescored', 0)} FP32 re-scores")
        assert len(st64) > 10000
        assert agree >= 0.999, name
        assert np.median(rel) <= 1e-3, name
