"""ksm_peel_static (the peel entry point that allocates nothing) against
ksm_peel (the same peel with a caller's workspace), through the C-ABI.

Two calls in a row on one GPU, the second with a smaller batch, so that the
static scratch keeps the first call's contents when the second starts; every
output (spikes, amplitudes, th_amps, status, iteration count) and the
in-place scores and residual must be bit-identical.
"""
import sys

import numpy as np
import pytest
import torch

from kilosort_amd import _lib, synth
from kilosort_amd import template_matching as tm
from kilosort_amd.batches import ArrayBatches

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda")


def say(msg):
    print(f"[gpu-no-alloc] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _lib.load()
    yield


def _inputs(seed, C, K, NT, precision=_lib.PREC_FP32):
    """Projected scores and everything else ksm_peel takes, for one batch."""
    rec = synth.make_recording(seed, C, K, NT, rate_range=(10.0, 40.0), amp_range=(10.0, 30.0))
    X = ArrayBatches(rec.data, NT=NT, nt=61, device=DEV).padded_batch_to_torch(0)
    W = torch.from_numpy(rec.wPCA).to(DEV)
    U = torch.from_numpy(rec.U).to(DEV)
    ops = dict(nt=61, Th_learned=8, max_peels=100, wPCA=W)
    lib = _lib.load()
    stream = _lib.stream_of(DEV)
    T = X.shape[1]
    ctc = tm.prepare_matching(ops, U)
    g = tm._grams.get(ctc, U, 61, DEV)
    nm = torch.empty((K,), dtype=torch.float32, device=DEV)
    _lib.check(lib.ksm_template_norms(U.data_ptr(), K, 6, C, nm.data_ptr(), stream), "norms")
    ldbt = (K + 63) // 64 * 64
    Bt = torch.zeros(((T + 127) // 128 * 128, ldbt), dtype=torch.float32, device=DEV)
    nb = lib.ksm_project_workspace(C, T, K, 6, 61, precision)
    pws = torch.empty((nb,), dtype=torch.uint8, device=DEV)
    bound = None
    if precision == _lib.PREC_FP32:
        _lib.check(lib.ksm_project(X.data_ptr(), T, W.data_ptr(), U.data_ptr(), g.perm.data_ptr(),
                                   C, T, K, 6, 61, Bt.data_ptr(), ldbt, precision, pws.data_ptr(),
                                   nb, stream), "ksm_project")
    else:
        nbb = lib.ksm_bound_bytes(C, T, K, 6, 61)
        bound = torch.empty((nbb,), dtype=torch.uint8, device=DEV)
        _lib.check(lib.ksm_project_bounded(X.data_ptr(), T, W.data_ptr(), U.data_ptr(),
                                           g.perm.data_ptr(), C, T, K, 6, 61, Bt.data_ptr(), ldbt,
                                           precision, 0, bound.data_ptr(), nbb, pws.data_ptr(), nb,
                                           stream), "ksm_project_bounded")
    return dict(X=X, W=W, U=U, g=g, nm=nm, Bt=Bt, ldbt=ldbt, C=C, K=K, T=T, ctc=ctc,
                bound=bound)


def _peel(inp, static, flags=0, capacity=40000):
    lib = _lib.load()
    stream = _lib.stream_of(DEV)
    Bt = inp["Bt"].clone()
    Xres = inp["X"].clone()
    g = inp["g"]
    st = torch.full((capacity, 2), -7, dtype=torch.int64, device=DEV)
    amps = torch.full((capacity, 1), -7.0, dtype=torch.float32, device=DEV)
    th = torch.full((capacity, 1), -7.0, dtype=torch.float32, device=DEV)
    info = torch.full((4,), -7, dtype=torch.int32, device=DEV)
    args = [Bt.data_ptr(), inp["ldbt"], Xres.data_ptr(), inp["T"], inp["U"].data_ptr(),
            inp["W"].data_ptr(), inp["nm"].data_ptr(), g.perm.data_ptr(), g.rank.data_ptr(),
            g.cc.data_ptr(), g.lohi.data_ptr(), g.off.data_ptr(), inp["K"], 6, inp["C"],
            inp["T"], 61, 8.0, 100, flags,
            None if inp["bound"] is None else inp["bound"].data_ptr(),
            st.data_ptr(), amps.data_ptr(), th.data_ptr(), capacity, info.data_ptr()]
    if static:
        _lib.check(lib.ksm_peel_static(*args, stream), "ksm_peel_static")
    else:
        nb = lib.ksm_peel_workspace(inp["K"], inp["T"], 100)
        ws = torch.full((nb,), 0xA5, dtype=torch.uint8, device=DEV)
        _lib.check(lib.ksm_peel(*args, ws.data_ptr(), nb, stream), "ksm_peel")
    torch.cuda.synchronize()
    k, status, iters, _ = (int(v) for v in info.cpu())
    _lib.check(status, "peel (device)")
    return dict(k=k, iters=iters, st=st[:k].cpu(), amps=amps[:k].cpu(), th=th[:k].cpu(),
                Bt=Bt.cpu(), Xres=Xres.cpu())


def _same(a, b, what):
    assert a["k"] == b["k"] and a["iters"] == b["iters"], (what, a["k"], b["k"])
    for key in ("st", "amps", "th", "Bt", "Xres"):
        assert torch.equal(a[key], b[key]), (what, key)


@pytest.mark.parametrize("flags,precision", [(0, _lib.PREC_FP32),
                                             (_lib.PEEL_RESCORE, _lib.PREC_BF16)])
def test_static_peel_matches_workspace_peel_two_calls(flags, precision):
    shapes = [(11, 64, 64, 60000), (12, 32, 16, 6000)]   # second call smaller
    results = []
    for seed, C, K, NT in shapes:
        say(f"peel C={C} K={K} NT={NT} flags={flags}: workspace, then static")
        inp = _inputs(seed, C, K, NT, precision)
        ref = _peel(inp, static=False, flags=flags)
        got = _peel(inp, static=True, flags=flags)
        say(f"  {ref['k']} spikes in {ref['iters']} iterations; static {got['k']}")
        assert ref["k"] > 50
        _same(ref, got, (C, K, NT))
        results.append(got)
    # the first, larger call again after the smaller one: still identical
    seed, C, K, NT = shapes[0]
    inp = _inputs(seed, C, K, NT, precision)
    _same(_peel(inp, static=False, flags=flags), _peel(inp, static=True, flags=flags), "again")


def test_static_peel_rejects_sizes_beyond_its_scratch():
    lib = _lib.load()
    cap = lib.ksm_peel_static_capacity()
    assert cap == 32 << 20
    assert lib.ksm_peel_workspace(2048, 60122, 100) <= cap
    assert lib.ksm_peel_workspace(4096, 60122, 100) > cap
    dummy = torch.zeros(16, device=DEV)
    p = dummy.data_ptr()
    st = lib.ksm_peel_static(p, 4096, p, 60122, p, p, p, p, p, p, p, p, 4096, 6, 383, 60122, 61,
                             8.0, 100, 0, None, p, p, p, 10, p, _lib.stream_of(DEV))
    assert st == _lib.KSM_ERR_WORKSPACE
    assert b"ksm_peel_static" in lib.ksm_last_error()
    assert b"scratch" in lib.ksm_last_error()
