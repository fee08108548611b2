"""Write golden vectors for the matching path from the reference itself.

Run in a container that has the reference checkout (it is never read on the
GPU host):

    python tests/golden/make_golden.py [/path/to/reference]

The reference package cannot be imported as a whole offline
(``kilosort/__init__.py`` pulls in faiss and numba, which are absent), so this
script loads ``kilosort/template_matching.py`` and ``kilosort/utils.py`` as
single modules.  ``template_matching.py`` imports ``kilosort.CCG`` only for
``merging_function`` (not on the matching path); a placeholder module object is
registered under that name so the import statement succeeds.  The four matching
functions then run unmodified on torch CPU tensors.

Outputs (``tests/golden/*.npz``): inputs and expected outputs of
``prepare_extract``, ``prepare_matching``, ``run_matching`` and ``extract`` on
small seeded synthetic recordings.
"""
from __future__ import annotations

import importlib.util
import os
import sys
import types

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from kilosort_amd import synth  # noqa: E402
from kilosort_amd.batches import ArrayBatches  # noqa: E402


def load_reference(root):
    kdir = os.path.join(root, "kilosort")
    pkg = types.ModuleType("kilosort")
    pkg.__path__ = [kdir]
    sys.modules["kilosort"] = pkg
    ccg = types.ModuleType("kilosort.CCG")
    sys.modules["kilosort.CCG"] = ccg
    pkg.CCG = ccg
    mods = {}
    for name in ("utils", "template_matching"):
        spec = importlib.util.spec_from_file_location(
            f"kilosort.{name}", os.path.join(kdir, f"{name}.py"))
        mod = importlib.util.module_from_spec(spec)
        sys.modules[f"kilosort.{name}"] = mod
        spec.loader.exec_module(mod)
        setattr(pkg, name, mod)
        mods[name] = mod
    return mods["template_matching"]


def base_ops(rec, NT, nC=10, position_limit=100.0, Th=8, max_peels=100):
    settings = dict(nearest_chans=nC, position_limit=position_limit, n_pcs=6)
    return dict(nt=rec.nt, nt0min=rec.nt0min, Th_learned=Th, max_peels=max_peels,
                batch_size=NT, wPCA=torch.from_numpy(rec.wPCA),
                xc=rec.xc, yc=rec.yc, settings=settings)


def main(root):
    torch.set_num_threads(1)
    tm = load_reference(root)
    cpu = torch.device("cpu")

    # 1. run_matching on one padded batch: C=16, K=8, NT=3000
    NT = 3000
    rec = synth.make_recording(101, 16, 8, NT, rate_range=(20.0, 60.0),
                               amp_range=(10.0, 30.0))
    ops = base_ops(rec, NT)
    U = torch.from_numpy(rec.U)
    bf = ArrayBatches(rec.data, NT=NT, nt=rec.nt, device=cpu)
    X = bf.padded_batch_to_torch(0, ops)
    ctc = tm.prepare_matching(ops, U)
    st, amps, th, Xres = tm.run_matching(ops, X.clone(), U, ctc, device=cpu)
    np.savez_compressed(os.path.join(HERE, "run_matching_c16_k8.npz"),
                        X=X.numpy(), U=rec.U, wPCA=rec.wPCA, nt=rec.nt, Th=8,
                        max_peels=100, ctc=ctc.numpy(), st=st.numpy(),
                        amps=amps.numpy(), th_amps=th.numpy(), Xres=Xres.numpy())

    # 2. run_matching with max_peels=1 on the same batch
    ops1 = dict(ops, max_peels=1)
    st1, amps1, th1, _ = tm.run_matching(ops1, X.clone(), U, ctc, device=cpu)
    np.savez_compressed(os.path.join(HERE, "run_matching_c16_k8_peel1.npz"),
                        st=st1.numpy(), amps=amps1.numpy(), th_amps=th1.numpy())

    # 3. extract over 3 batches (last one partial): C=24, K=12, NT=2000
    NT = 2000
    n_samples = 2 * NT + 900
    rec = synth.make_recording(202, 24, 12, n_samples, rate_range=(20.0, 60.0),
                               amp_range=(10.0, 30.0))
    ops = base_ops(rec, NT, nC=6, position_limit=50.0)
    U = torch.from_numpy(rec.U)
    bf = ArrayBatches(rec.data, NT=NT, nt=rec.nt, device=cpu)
    stx, tFx, opsx = tm.extract(ops, bf, U, device=cpu)
    np.savez_compressed(os.path.join(HERE, "extract_c24_k12.npz"),
                        data=rec.data, U=rec.U, wPCA=rec.wPCA, xc=rec.xc,
                        yc=rec.yc, NT=NT, nt=rec.nt, nt0min=rec.nt0min, nC=6,
                        position_limit=50.0, st=stx, tF=tFx.numpy(),
                        iCC=opsx["iCC"].numpy(), iCC_mask=opsx["iCC_mask"].numpy(),
                        iU=opsx["iU"].numpy(), n_batches=int(bf.n_batches))

    # 4. prepare_extract on the benchmark probes (pins the argsort tie order)
    for C, K in ((64, 64), (383, 700)):
        xc, yc = synth.probe_2col(C)
        rng = np.random.default_rng(303 + C)
        Uk, _ = synth.make_templates(rng, K, xc, yc)
        iCC, iCC_mask, iU, Ucc = tm.prepare_extract(
            xc, yc, torch.from_numpy(Uk), 10, 100.0, device=cpu)
        np.savez_compressed(os.path.join(HERE, f"prepare_extract_c{C}.npz"),
                            seed=303 + C, K=K, iCC=iCC.numpy(),
                            iCC_mask=iCC_mask.numpy(), iU=iU.numpy(),
                            Ucc=Ucc.numpy())
    print("golden vectors written to", HERE)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
