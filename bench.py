"""Benchmark of the learned-template matching path on one MI355X.

Workload: 383 channels (2-column staggered probe) x 700 templates, padded
batches of NT=60000 (+2*61) samples, Th=8, max_peels=100, synthetic whitened
data (noise + Poisson spikes with 20 % near-coincident overlaps), resident in
HBM before the timed region.

A step = one batch through the hot path exactly as ``extract`` runs it:
run_matching (temporal-PC correlation, projection, peel) + PC features + the
device-to-host copies of st/th_amps/tF.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--precision fp32]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "matched channel·samples/s, FP32 vs BF16-MFMA, + spike agreement % at 383ch"
C, K, NT, NTAP, P = 383, 700, 60000, 61, 6
PEAK_FP32_MFMA = 157.3e12
PEAK_HBM = 8.0e12


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_workload(n_batches, seed, device):
    from kilosort_amd import synth
    rec = synth.make_recording(seed, C, K, n_batches * NT + 2 * NTAP,
                               rate_range=(1.0, 20.0), amp_range=(8.0, 40.0),
                               overlap_frac=0.2, with_data=False)
    data = synth.synthesize_torch(rec.n_samples, C, rec.times, rec.ids, rec.amps, rec.U,
                                  rec.wPCA, rec.nt0min, device, seed=seed)
    from kilosort_amd.batches import ArrayBatches
    bf = ArrayBatches(data, NT=NT, nt=NTAP, device=device)
    batches = [bf.padded_batch_to_torch(i).contiguous() for i in range(n_batches)]
    del data
    ops = dict(nt=NTAP, nt0min=rec.nt0min, Th_learned=8, max_peels=100, batch_size=NT,
               wPCA=torch.from_numpy(rec.wPCA).to(device), xc=rec.xc, yc=rec.yc,
               settings=dict(nearest_chans=10, position_limit=100.0, n_pcs=P))
    return rec, ops, batches


def cpu_baseline(ops, X, U, spikes_gpu):
    """Oracle (torch-CPU restatement) on one batch: run_matching + features."""
    from oracle import tm_oracle as O
    cores = os.cpu_count() or 1
    allowed = os.environ.get("OMP_NUM_THREADS")
    if allowed and allowed.isdigit():
        cores = min(cores, int(allowed))
    torch.set_num_threads(cores)
    ops_c = dict(ops)
    ops_c["wPCA"] = ops["wPCA"].cpu()
    Xc = X.cpu()
    Uc = U.cpu()
    iCC, _, iU, Ucc = O.prepare_extract(ops["xc"], ops["yc"], Uc, 10, 100.0)
    ctc = O.prepare_matching(ops_c, Uc)
    t0 = time.perf_counter()
    st, amps, th, Xres = O.run_matching(ops_c, Xc, Uc, ctc)
    O.features(ops_c, Xres, st, amps, iCC, iU, Ucc)
    dt = time.perf_counter() - t0
    ref = set(map(tuple, st.numpy().tolist()))
    got = set(map(tuple, spikes_gpu.tolist()))
    agree = len(ref & got) / max(len(ref), len(got), 1)
    return dict(value=C * NT / dt, unit="channel·samples/s", cores=cores, kind="port",
                sample="1 padded batch (383 ch x 60122 samples, 700 templates): oracle "
                       "run_matching + features, torch CPU float32",
                seconds=dt), dict(fp32_gpu_vs_cpu_pct=100.0 * agree, n_cpu=len(ref),
                                  n_gpu=len(got))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="fp32")
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-peel", action="store_true",
                    help="also time the peel kernels with events (perturbs the timing)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the matching path is sharded by batch "
              "and no multi-GPU run is part of this build", file=sys.stderr)
        sys.exit(2)

    from kilosort_amd import _lib
    from kilosort_amd import template_matching as tm

    device = torch.device("cuda:0")
    torch.cuda.set_device(device)
    lib = _lib.load()
    n_distinct = min(args.steps + args.warmup, 12)
    log(f"generating {n_distinct} synthetic batches")
    rec, ops, batches = make_workload(n_distinct, args.seed, device)
    ops["matching_precision"] = args.precision
    U = torch.from_numpy(rec.U).to(device)
    iCC, _, iU, Ucc = tm.prepare_extract(ops["xc"], ops["yc"], U, 10, 100.0, device=device)
    ctc = tm.prepare_matching(ops, U)
    torch.cuda.synchronize()

    log(f"warmup {args.warmup}")
    for i in range(args.warmup):
        tm._extract_batch(ops, batches[i % n_distinct], U, ctc, iCC, iU, Ucc, device)
    torch.cuda.synchronize()

    lib.ksm_profile_reset()
    # projection kernels only by default: no events inside the peel loop
    lib.ksm_profile_enable(3 if args.profile_peel else 1)
    nsp = 0
    iters = 0
    log(f"timing {args.steps} steps")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        st, th, tf = tm._extract_batch(ops, batches[(args.warmup + i) % n_distinct], U, ctc,
                                       iCC, iU, Ucc, device)
        nsp += len(st)
        iters += tm.last_peel_iterations
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    lib.ksm_profile_enable(0)

    kern = {}
    names = ["conv_pc", "gemm_f32"]
    if args.profile_peel:
        names += ["colmax_all", "peel_select", "peel_update"]
    for name in names:
        ms, n = _lib.profile_query(name)
        kern[name] = (ms, n)
    lib.ksm_profile_reset()
    T = NT + 2 * NTAP
    gemm_ms, gemm_n = kern["gemm_f32"]
    gemm_avg = gemm_ms / max(gemm_n, 1) / 1e3
    flops_gemm = 2.0 * K * P * C * T
    achieved = flops_gemm / gemm_avg if gemm_avg > 0 else 0.0
    ms_step = elapsed / args.steps * 1e3
    value = C * NT * args.steps / elapsed
    log("kernel totals over the timed steps (ms, launches): " +
        ", ".join(f"{k}={v[0]:.2f}/{v[1]}" for k, v in kern.items()))

    cpu = None
    agreement = None
    if not args.no_cpu_baseline:
        Xs = batches[0]
        stg = tm.run_matching(ops, Xs, U, ctc, device=device)[0].cpu().numpy()
        log("cpu baseline (oracle on one batch)")
        cpu, agreement = cpu_baseline(ops, Xs, U, stg)
        log(f"cpu baseline {cpu['seconds']:.1f}s on {cpu['cores']} threads")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "channel·samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else args.precision,
        "data": "synthetic",
        "config": {"workload": "383 ch x 700 templates, NT=60000 (+2x61 pad) batches, "
                               "Th=8, max_peels=100, Poisson 1-20 Hz, 20% overlaps",
                   "channels": C, "templates": K, "batch_samples": NT,
                   "precision": args.precision, "distinct_batches": n_distinct},
        "roofline": {"kernel": "gemm_f32_mfma (projection)", "bound": "mfma",
                     "achieved": achieved / 1e12, "peak": PEAK_FP32_MFMA / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_MFMA,
                     "traffic": None,
                     "flops_per_launch": flops_gemm, "avg_launch_ms": gemm_avg * 1e3},
        "kernel_ms_per_step": {k: v[0] / args.steps for k, v in kern.items()},
        "spikes_per_batch": nsp / args.steps,
        "peel_iterations_per_batch": iters / args.steps,
        "cpu_baseline": cpu,
        "agreement": agreement,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
