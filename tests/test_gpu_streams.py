"""Batches matched concurrently from two host threads, each on its own stream,
give the same results as one after the other (no scratch memory is shared
between streams)."""
import sys
import threading

import numpy as np
import pytest
import torch

from kilosort_amd import _lib, synth
from kilosort_amd import template_matching as tm
from kilosort_amd.batches import ArrayBatches

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda")
NT = 20000


def say(msg):
    print(f"[gpu-streams] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _lib.load()
    yield


def test_concurrent_streams_match_sequential():
    say("4 batches C=64 K=64 NT=20000: sequential, then 2 threads x 2 streams")
    rec = synth.make_recording(81, 64, 64, 4 * NT, rate_range=(5.0, 30.0))
    bf = ArrayBatches(rec.data, NT=NT, nt=61, device=DEV)
    batches = [bf.padded_batch_to_torch(i).contiguous() for i in range(4)]
    ops = dict(nt=61, Th_learned=8, max_peels=100, wPCA=torch.from_numpy(rec.wPCA).to(DEV))
    U = torch.from_numpy(rec.U).to(DEV)
    ctc = tm.prepare_matching(ops, U)

    def run(X):
        st, amps, th, Xres = tm.run_matching(ops, X, U, ctc, device=DEV)
        return st.cpu().numpy(), amps.cpu().numpy(), Xres.cpu().numpy()

    seq = [run(X) for X in batches]
    torch.cuda.synchronize()
    got = {}
    errors = []

    def worker(first):
        try:
            s = torch.cuda.Stream(device=DEV)
            with torch.cuda.stream(s):
                for i in (first, first + 2):
                    got[i] = run(batches[i])
        except Exception as e:          # reported by the main thread
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(j,)) for j in (0, 1)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    torch.cuda.synchronize()
    assert not errors, errors
    for i in range(4):
        say(f"batch {i}: {len(seq[i][0])} spikes")
        assert len(seq[i][0]) > 10
        np.testing.assert_array_equal(got[i][0], seq[i][0])
        np.testing.assert_array_equal(got[i][1], seq[i][1])
        np.testing.assert_array_equal(got[i][2], seq[i][2])


@pytest.mark.parametrize("streams", [1, 2, 4])
def test_extract_streams_identical(streams):
    """extract over 5 batches gives the same st and tF for any number of
    device streams (batches spread over lanes, results back in order)."""
    say(f"extract C=64 K=64 5x20000 with matching_streams={streams} vs 3")
    rec = synth.make_recording(82, 64, 64, 5 * NT, rate_range=(5.0, 30.0))
    ops = dict(nt=61, nt0min=rec.nt0min, Th_learned=8, max_peels=100, batch_size=NT,
               wPCA=torch.from_numpy(rec.wPCA).to(DEV), xc=rec.xc, yc=rec.yc,
               settings=dict(nearest_chans=10, position_limit=100.0, n_pcs=6))
    U = torch.from_numpy(rec.U).to(DEV)
    bf = ArrayBatches(rec.data, NT=NT, nt=61, device=DEV)
    st3, tF3, _ = tm.extract(dict(ops, matching_streams=3), bf, U, device=DEV)
    st, tF, _ = tm.extract(dict(ops, matching_streams=streams), bf, U, device=DEV)
    assert len(st3) > 100
    np.testing.assert_array_equal(st, st3)
    np.testing.assert_array_equal(tF.numpy(), tF3.numpy())


def test_extract_many_matches_extract():
    """Recordings of different channel and template counts and lengths through
    one pipeline give exactly what extract gives for each."""
    say("extract_many: 3 recordings (C, K, batches) = (64, 40, 3), (32, 12, 1), (96, 70, 2)")
    jobs, refs = [], []
    for seed, C, K, nb in ((83, 64, 40, 3), (84, 32, 12, 1), (85, 96, 70, 2)):
        # the second recording ends within a batch: its last batch is padded
        # by repeating its last sample (tied scores where they pass threshold)
        n = nb * NT + (1700 if nb == 1 else 0)
        rec = synth.make_recording(seed, C, K, n, rate_range=(5.0, 30.0))
        ops = dict(nt=61, nt0min=rec.nt0min, Th_learned=8, max_peels=100, batch_size=NT,
                   wPCA=torch.from_numpy(rec.wPCA).to(DEV), xc=rec.xc, yc=rec.yc,
                   settings=dict(nearest_chans=10, position_limit=100.0, n_pcs=6))
        U = torch.from_numpy(rec.U).to(DEV)
        bf = ArrayBatches(rec.data, NT=NT, nt=61, device=DEV)
        refs.append(tm.extract(dict(ops), bf, U, device=DEV))
        jobs.append((dict(ops), bf, U))
    got = tm.extract_many(jobs, device=DEV)
    for (st, tF, o), (st_r, tF_r, o_r) in zip(got, refs):
        assert len(st_r) > 20
        np.testing.assert_array_equal(st, st_r)
        np.testing.assert_array_equal(tF.numpy(), tF_r.numpy())
        np.testing.assert_array_equal(o["iCC"].cpu().numpy(), o_r["iCC"].cpu().numpy())
