"""The matching path at other temporal widths and PC counts than the
defaults (nt = 61, 6 PCs): nt = 41 with 3 PCs and nt = 81 with 8 PCs, the
largest PC count the kernels take.  At these shapes the temporal-PC
correlation runs its generic kernel, the peel works on windows of 2 nt + 1
columns and the features on nC x P = 30 and 80 values per spike.  The FP32
form must equal the oracle's extract as it does at the defaults (spike times
and ids exactly, amplitudes and features within 1e-5).  The reduced-precision
projections are built for the default shape only and refuse others with an
error, never a result.
"""
import sys

import numpy as np
import pytest
import torch

from kilosort_amd import _lib, synth
from kilosort_amd import template_matching as tm
from kilosort_amd.batches import ArrayBatches
from oracle import tm_oracle as O

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda")
CPU = torch.device("cpu")
NT = 20000


def say(msg):
    print(f"[gpu-shapes] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _lib.load()
    yield


@pytest.mark.parametrize("seed,nt,P,nt0min", [(61, 41, 3, 13), (62, 81, 8, 27)])
def test_extract_other_widths_and_pc_counts(seed, nt, P, nt0min):
    C, K, nb = 48, 40, 2
    say(f"C={C} K={K} {nb}x{NT}, nt={nt}, P={P}")
    rec = synth.make_recording(seed, C, K, nb * NT, nt=nt, nt0min=nt0min, P=P,
                               rate_range=(5.0, 30.0))
    assert rec.wPCA.shape == (P, nt)
    ops = dict(nt=nt, nt0min=nt0min, Th_learned=8, max_peels=100, batch_size=NT,
               wPCA=torch.from_numpy(rec.wPCA), xc=rec.xc, yc=rec.yc,
               settings=dict(nearest_chans=10, position_limit=100.0, n_pcs=P))
    U = torch.from_numpy(rec.U)
    st_ref, tF_ref, _ = O.extract(dict(ops), ArrayBatches(rec.data, NT=NT, nt=nt, device=CPU), U)
    say(f"oracle: {len(st_ref)} spikes")
    assert len(st_ref) > 500
    assert tF_ref.shape[1:] == (10, P)
    ops_d = dict(ops, wPCA=ops["wPCA"].to(DEV))
    bf = ArrayBatches(rec.data, NT=NT, nt=nt, device=DEV)
    st, tF, _ = tm.extract(dict(ops_d), bf, U.to(DEV), device=DEV)
    say(f"fp32: {len(st)} spikes")
    np.testing.assert_array_equal(st[:, :2], st_ref[:, :2])
    err = np.abs(st[:, 2] - st_ref[:, 2]) / np.maximum(np.abs(st_ref[:, 2]), 1e-30)
    assert err.max() <= 1e-5
    tf, tfr = tF.numpy().astype(np.float64), tF_ref.numpy().astype(np.float64)
    scale = np.abs(tfr).reshape(len(tfr), -1).max(1)
    ferr = np.abs(tf - tfr).reshape(len(tfr), -1).max(1)
    assert np.all(ferr <= 1e-5 * np.maximum(scale, 1e-30)), (ferr / scale).max()
    for form in ("bf16", "fp8"):
        with pytest.raises(ValueError, match="P=6, nt=61"):
            tm.extract(dict(ops_d, matching_precision=form), bf, U.to(DEV), device=DEV)
