"""The batched form: many recordings of different channel counts, template
counts and lengths through one pipeline (``extract_many``), in BF16 with
certified decisions, with the autotuner's choices (GEMM tile, peel schedule,
streams) and graph-replayed batches, against ``extract`` of each recording
(eager, the same form) and against the FP32 oracle's extract.

Graph replays are checked replay by replay: a recording of 8 batches on two
streams replays each lane's graph three times, between eager batches of
other recordings and host synchronisations, and every batch must equal the
eager result.
"""
import sys

import numpy as np
import pytest
import torch

from kilosort_amd import _lib, autotune, synth
from kilosort_amd import template_matching as tm
from kilosort_amd.batches import ArrayBatches
from oracle import tm_oracle as O

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda")
CPU = torch.device("cpu")
NT = 6000

# (seed, channels, templates, batches): a reduced mix of the batched workload
MIX = [(91, 64, 50, 8), (92, 128, 200, 3), (93, 32, 20, 1), (94, 96, 120, 4), (95, 64, 80, 2)]


def say(msg):
    print(f"[gpu-many] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _lib.load()
    yield


def _job(seed, C, K, nb, device, precision):
    rec = synth.make_recording(seed, C, K, nb * NT, rate_range=(5.0, 30.0))
    ops = dict(nt=61, nt0min=rec.nt0min, Th_learned=8, max_peels=100, batch_size=NT,
               wPCA=torch.from_numpy(rec.wPCA).to(device), xc=rec.xc, yc=rec.yc,
               matching_precision=precision,
               settings=dict(nearest_chans=10, position_limit=100.0, n_pcs=6))
    return rec, ops, ArrayBatches(rec.data, NT=NT, nt=61, device=device), \
        torch.from_numpy(rec.U).to(device)


def _pairs(st):
    return set(map(tuple, np.round(st[:, :2]).astype(np.int64).tolist()))


def test_many_recordings_bf16_autotuned_graphs():
    jobs, recs = [], []
    for seed, C, K, nb in MIX:
        rec, ops, bf, U = _job(seed, C, K, nb, DEV, "bf16")
        jobs.append((ops, bf, U))
        recs.append(rec)
    say("autotune: tile per shape, peel schedule per recording, streams")
    streams, plans, times = autotune.tune_many(jobs, device=DEV, stream_candidates=(2, 3))
    say(f"streams {streams} ({times}); plans {plans}")
    assert streams in (2, 3)
    assert all(p["tile_m"] in (64, 128) for p in plans)
    assert all(2 <= p["graph_peels"] <= 100 for p in plans)
    # the 8-batch recording replays graphs on every lane
    got = tm.extract_many([(dict(o), b, U) for o, b, U in jobs], device=DEV, streams=2,
                          graphs=None)
    assert jobs[0][0]["matching_graphs"] or streams == 3
    for r, ((ops, bf, U), (st, tF, _)) in enumerate(zip(jobs, got)):
        seed, C, K, nb = MIX[r]
        say(f"recording {r}: C={C} K={K} batches={nb}: {len(st)} spikes")
        # the same recording through extract (eager batches, same form)
        ref_st, ref_tF, _ = tm.extract(dict(ops, matching_graphs=False), bf, U, device=DEV)
        np.testing.assert_array_equal(st, ref_st)
        np.testing.assert_array_equal(tF.numpy(), ref_tF.numpy())
        # and against the FP32 oracle on the CPU: the certified decisions
        ops_c = dict(ops, wPCA=ops["wPCA"].cpu())
        for key in ("matching_precision", "matching_tile_m", "matching_graph_peels",
                    "matching_graphs"):
            ops_c.pop(key, None)
        bf_c = ArrayBatches(recs[r].data, NT=NT, nt=61, device=CPU)
        o_st, o_tF, _ = O.extract(ops_c, bf_c, U.cpu())
        a, b = _pairs(st), _pairs(o_st)
        agree = len(a & b) / max(len(a), len(b), 1)
        say(f"recording {r}: agreement with the FP32 oracle {100 * agree:.3f} %")
        assert len(b) > 50
        assert agree >= 0.999


def test_graph_replays_equal_eager_batches():
    """One recording of 8 batches on two streams: each lane's first batch runs
    eagerly, the next three as replays; every batch equals run_matching +
    features of the same batch, and the replays stay equal when eager work of
    another recording and host synchronisations come between them."""
    rec, ops, bf, U = _job(96, 64, 64, 8, DEV, "bf16x3")
    other = _job(97, 32, 16, 2, DEV, "fp32")
    jobs = [(dict(ops), bf, U), (dict(other[1]), other[2], other[3]), (dict(ops), bf, U)]
    say("8 batches graph-replayed, an eager recording between, the same 8 again")
    got = tm.extract_many(jobs, device=DEV, streams=2, graphs=True)
    torch.cuda.synchronize()
    ref = tm.extract(dict(ops), bf, U, device=DEV)
    for idx in (0, 2):
        np.testing.assert_array_equal(got[idx][0], ref[0])
        np.testing.assert_array_equal(got[idx][1].numpy(), ref[1].numpy())
    ref_o = tm.extract(dict(other[1]), other[2], other[3], device=DEV)
    np.testing.assert_array_equal(got[1][0], ref_o[0])


def test_recorder_rejects_misuse():
    """The graph recorder's error paths: ending without a recording, a second
    recording on one thread, an empty recording, and ksm_peel (whose host
    follows the device) on a recorded stream."""
    import ctypes
    lib = _lib.load()
    s = _lib.stream_of(DEV)
    h = ctypes.c_void_p()
    assert lib.ksm_graph_end(s, ctypes.byref(h)) == _lib.KSM_ERR_INVALID and not h.value
    assert lib.ksm_graph_begin(s) == _lib.KSM_OK
    try:
        assert lib.ksm_graph_begin(s) == _lib.KSM_ERR_INVALID
        p = torch.zeros(16, device=DEV).data_ptr()
        st = lib.ksm_peel(p, 64, p, 6000, p, p, p, p, p, p, p, p, 16, 6, 32, 6000, 61, 8.0,
                          10, 0, None, p, p, p, 10, p, p, 1 << 30, s)
        assert st == _lib.KSM_ERR_INVALID
        assert b"recorded" in lib.ksm_last_error()
    finally:
        assert lib.ksm_graph_end(s, ctypes.byref(h)) == _lib.KSM_ERR_INVALID   # nothing recorded
    assert not h.value
    torch.cuda.synchronize()
