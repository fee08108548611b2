"""Benchmark of the learned-template matching path on one MI355X.

Workload: 383 channels (2-column staggered probe) x 700 templates, padded
batches of NT=60000 (+2*61) samples, Th=8, max_peels=100, synthetic whitened
data (noise + Poisson spikes with 20 % near-coincident overlaps), resident in
HBM before the timed region.

A step = one batch through the hot path as ``extract`` runs it once the batch
is in HBM: run_matching (temporal-PC correlation, projection, peel) + PC
features + the device-to-host copies of st/th_amps; the features stay on the
device (``extract`` copies them to the host when it assembles ``tF``; the
``extract_wall`` legs time whole ``extract`` calls, that copy included).
Prints one JSON line.

Every form (fp32: the parity form; bf16x3, bf16: projection on BF16 MFMA;
fp8: projection on FP8 e4m3 MFMA) is timed the same way; `value` is the
--precision form.  A separate pass with events around each kernel gives
per-kernel times for the roofline objects, and batch 0 is matched by the CPU oracle in FP32 (the timed CPU baseline) and
FP64 (spike agreement and amplitude error of every form).

    python bench.py [--steps K] [--warmup W] [--precision fp32|bf16x3|bf16|fp8]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

# HIP hardware queues of this process, read when HIP starts (before torch
# touches the GPU): at least eight, so that the library's four matching
# streams and the default stream each have a queue of their own (with HIP's
# default of four the library runs three streams).  A runtime setting of the
# caller's process, as extract() users would set it; it changes no result.
try:
    _hwq = int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))
except ValueError:
    _hwq = 4
if _hwq < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "matched channel·samples/s, FP32 vs BF16-MFMA, + spike agreement % at 383ch"
C, K, NT, NTAP, P = 383, 700, 60000, 61, 6
FORMS = ("fp32", "bf16x3", "bf16", "fp8")
# timed by default: at 383 channels FP8's error margin leaves most decisions
# open, so its peel re-scores nearly every candidate column in FP32 (seconds
# per batch); time it on request (--forms fp32,fp8) with few steps
DEFAULT_FORMS = ("fp32", "bf16x3", "bf16")
T_PAD = NT + 2 * 61          # padded batch columns
PEAK_MFMA = {"fp32": 157.3e12, "bf16x3": 2.5e15, "bf16": 2.5e15,    # dense MFMA peaks
             "fp8": 5.0e15}
GEMM_KERNEL = {"fp32": "gemm_f32", "bf16x3": "gemm_bf16", "bf16": "gemm_bf16",
               "fp8": "gemm_fp8"}
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


def oracle_runs(ops, X, U, need64=True):
    """Oracle (torch-CPU restatement) on one batch: FP32 run_matching + features
    (timed: the CPU baseline) and an FP64 run_matching (the accuracy yardstick
    of the precision forms)."""
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
    cpu = dict(value=C * NT / dt, unit="channel·samples/s", cores=cores, kind="port",
               sample="1 padded batch (383 ch x 60122 samples, 700 templates): oracle "
                      "run_matching + features, torch CPU float32",
               seconds=dt)
    ref32 = (st.numpy(), amps.numpy()[:, 0])
    ref64 = None
    if need64:
        ops64 = dict(ops_c)
        ops64["wPCA"] = ops_c["wPCA"].double()
        t0 = time.perf_counter()
        st64, amps64, _, _ = O.run_matching(ops64, Xc.double(), Uc.double(),
                                            O.prepare_matching(ops64, Uc.double()))
        log(f"fp64 oracle {time.perf_counter() - t0:.1f}s")
        ref64 = (st64.numpy(), amps64.numpy()[:, 0])
    return cpu, ref32, ref64


def agreement(st, amps, ref):
    """Exact (time, template) agreement over max(n, n_ref); +-1-sample agreement
    (same template); median relative amplitude error over exact matches."""
    rst, ramps = ref
    a = {tuple(p): i for i, p in enumerate(np.asarray(st).tolist())}
    b = {tuple(p): i for i, p in enumerate(np.asarray(rst).tolist())}
    n = max(len(a), len(b), 1)
    both = [(a[p], b[p]) for p in a.keys() & b.keys()]
    near = 0
    for (t, i) in a:
        if (t, i) in b or (t - 1, i) in b or (t + 1, i) in b:
            near += 1
    if both:
        gi, ri = np.array(both).T
        rel = np.abs(amps[gi].astype(np.float64) - ramps[ri]) / np.abs(ramps[ri])
        med = float(np.median(rel))
    else:
        med = None
    return dict(exact_pct=100.0 * len(both) / n, within1_pct=100.0 * near / n,
                n=len(a), n_ref=len(b), median_rel_amp_err=med)


def time_form(tm, lib, ops, batches, U, ctc, iCC, iU, Ucc, device, precision, steps,
              warmup, lanes=None):
    """Untimed warmup, then `steps` batches end to end -- matching, features
    (left on the device) and the host copies of st / th_amps -- through the
    library's batch pipeline (batches spread over its device streams, one
    host thread each), with no events in the loop."""
    o = dict(ops, matching_precision=precision)
    nb = len(batches)
    pipe = tm._BatchPipeline(o, U, ctc, iCC, iU, Ucc, device, lanes=lanes)
    n_lanes = len(pipe.lanes)
    try:
        for i in range(warmup):
            pipe.submit(batches[i % nb])
        pipe.flush()
        torch.cuda.synchronize()
        it0 = pipe.peel_iterations
        nsp = 0
        t0 = time.perf_counter()
        for i in range(steps):
            for res in pipe.submit(batches[(warmup + i) % nb]):
                nsp += len(res[1])
        for res in pipe.flush():
            nsp += len(res[1])
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        iters = pipe.peel_iterations - it0
    finally:
        pipe.close()
    return dict(value=C * NT * steps / elapsed, ms_per_step=elapsed / steps * 1e3,
                spikes_per_batch=nsp / steps, peel_iterations_per_batch=iters / steps,
                streams=n_lanes)


def time_preprocess(device, seed, nbatch=3, reps=5, cpu_baseline=True):
    """Batch preprocessing at 383 channels (int16 raw, NT=60000): channel means,
    CAR, the default 300 Hz high-pass filter and a whitening matrix.
    * device rate: raw batch resident in HBM, ksm_filter_batch alone;
    * end-to-end rate: FilteredBatches.padded_batch_to_torch from a host
      int16 array (slice, pinned staging, host-to-device copy, filter);
    * per-kernel times (library events) and the high-pass product's MFMA
      roofline; the oracle (torch CPU, FFT filter) on one batch as its CPU
      baseline."""
    from kilosort_amd import _lib, preprocess as PP
    lib = _lib.load()
    n = nbatch * NT + 1000
    rng = np.random.default_rng(seed)
    data = np.clip(np.round(rng.standard_normal((n, C), dtype=np.float32) * 40.0
                            + rng.uniform(-200, 200, C).astype(np.float32)),
                   -32768, 32767).astype(np.int16)
    A = np.eye(C, dtype=np.float32) + 0.02 * rng.standard_normal((C, C), dtype=np.float32)
    W = (A @ A.T / np.sqrt(C)).astype(np.float32)
    hp = PP.get_highpass_filter(device=device)
    fb = PP.FilteredBatches(n_chan_bin=C, NT=NT, file_object=data, device=device,
                            hp_filter=hp, whiten_mat=W)
    T = NT + 2 * NTAP
    fb.padded_batch_to_torch(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        for ib in range(1, nbatch):
            fb.padded_batch_to_torch(ib)
    torch.cuda.synchronize()
    e2e_ms = (time.perf_counter() - t0) * 1e3 / (reps * (nbatch - 1))
    # raw resident: one middle batch's window
    raw = torch.from_numpy(np.ascontiguousarray(data[NT - NTAP:NT - NTAP + T])).to(device)
    taps, J = fb._taps_for(T)
    Wd = fb._whiten
    X = torch.empty((C, T), device=device)
    flags = _lib.FILTER_CAR | _lib.FILTER_WHITEN
    nb = lib.ksm_filter_workspace(C, T, J, flags)
    ws = torch.empty(nb, dtype=torch.uint8, device=device)
    s = _lib.stream_of(device)

    def one():
        _lib.check(lib.ksm_filter_batch(raw.data_ptr(), _lib.RAW_I16, C, T, 0, None, C, T,
                                        flags, 0.0, 0.0, taps.data_ptr(), J, float("inf"),
                                        Wd.data_ptr(), X.data_ptr(), T, ws.data_ptr(), nb, s),
                   "ksm_filter_batch")
    for _ in range(3):
        one()
    torch.cuda.synchronize()
    k = 20
    t0 = time.perf_counter()
    for _ in range(k):
        one()
    torch.cuda.synchronize()
    dev_ms = (time.perf_counter() - t0) * 1e3 / k
    lib.ksm_profile_reset()
    lib.ksm_profile_enable(3)
    for _ in range(5):
        one()
    torch.cuda.synchronize()
    lib.ksm_profile_enable(0)
    kern = {}
    for name in ("filter_car", "filter_highpass", "filter_whiten"):
        ms, cnt = _lib.profile_query(name)
        if cnt:
            kern[name] = ms / cnt
    lib.ksm_profile_reset()
    flops_hp = 2.0 * C * T * (2 * J + 1)
    # tiles of 128 channels x 64 samples, depth steps of 32 over 2J + 64
    nks = -(-(2 * J + 64) // 32)
    flops_hp_exec = 2.0 * (-(-C // 128) * 128) * (-(-T // 64) * 64) * 32 * nks
    flops_wh = 2.0 * C * C * T
    out = dict(value=C * T / (dev_ms / 1e3), unit="channel·samples/s", ms_per_batch=dev_ms,
               end_to_end_value=C * T / (e2e_ms / 1e3), end_to_end_ms_per_batch=e2e_ms,
               taps=2 * J + 1, kernel_ms=kern,
               config="383 ch int16, NT=60000 (+2x61), CAR, 300 Hz high-pass "
                      f"({2 * J + 1} taps), whitening")
    if "filter_highpass" in kern:
        t_hp = kern["filter_highpass"] / 1e3
        out["roofline_highpass"] = dict(
            kernel="mm_mfma_kernel<Toeplitz> (high-pass filter)", bound="mfma",
            achieved=flops_hp / t_hp / 1e12, peak=157.3, unit="TFLOP/s",
            frac=flops_hp / t_hp / 157.3e12, flops_per_launch=flops_hp,
            executed_flops_per_launch=flops_hp_exec, avg_launch_ms=kern["filter_highpass"])
    if "filter_whiten" in kern:
        t_w = kern["filter_whiten"] / 1e3
        out["roofline_whiten"] = dict(achieved=flops_wh / t_w / 1e12, peak=157.3,
                                      unit="TFLOP/s", frac=flops_wh / t_w / 157.3e12)
    if cpu_baseline:
        from oracle import filter_oracle as FO
        Xc, _ = FO.padded_batch(data, 1, NT, NTAP, int(fb.n_batches), imax=int(fb.imax))
        t0 = time.perf_counter()
        FO.filter_batch(Xc, hp_filter=hp.cpu(), whiten_mat=torch.from_numpy(W))
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = dict(value=C * T / dt, unit="channel·samples/s",
                                   cores=torch.get_num_threads(), kind="port",
                                   sample="1 padded batch: oracle filter (torch CPU FFT)",
                                   seconds=dt)
    return out


def time_templates(ops, U, device, reps=5, cpu_baseline=True):
    """Template Gram consumers on the bench's 700 templates x 383 channels:
    the all-pairs similarity over lags (CCG.similarity) and align_U, with the
    oracle's similarity as its CPU baseline."""
    from kilosort_amd import templates as TM
    Wall = U.transpose(1, 2).contiguous()
    W = ops["wPCA"]
    out = {}
    for name, fn in (("similarity", lambda: TM.similarity_matrix(Wall, W, NTAP, device)),
                     ("align_U", lambda: TM.align_U(Wall, dict(nt=NTAP, wPCA=W, wTEMP=W),
                                                    device=device))):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        out[name + "_ms"] = (time.perf_counter() - t0) * 1e3 / reps
    out["templates"] = int(U.shape[0])
    out["note"] = "wall time per call including the host copy of the [K, K] result"
    if cpu_baseline:
        from oracle import templates_oracle as TO
        Wc = Wall.cpu()
        t0 = time.perf_counter()
        TO.similarity(Wc, W.cpu(), NTAP)
        out["cpu_baseline"] = dict(similarity_s=time.perf_counter() - t0,
                                   cores=torch.get_num_threads(), kind="port",
                                   sample="oracle similarity of the 700 templates (torch CPU)")
    return out


def time_extract(tm, n_batches, seed, device, U, ops, runs=5):
    """The whole ``extract`` call as a caller sees it: the recording in pinned
    host memory (float32 [samples, 383]), batches copied to the device one by
    one, prepare_extract + prepare_matching inside, st / tF returned to the
    host.  One untimed call, then the median of `runs` calls."""
    from kilosort_amd import synth
    from kilosort_amd.batches import ArrayBatches
    # exactly n_batches batches: a short last batch would be padded with its
    # last sample (io.py:816-818), a constant run that only filtering makes
    # harmless, and this recording is not filtered
    rec = synth.make_recording(seed, C, K, n_batches * NT,
                               rate_range=(1.0, 20.0), amp_range=(8.0, 40.0),
                               overlap_frac=0.2, with_data=False)
    data = synth.synthesize_torch(rec.n_samples, C, rec.times, rec.ids, rec.amps, rec.U,
                                  rec.wPCA, rec.nt0min, device, seed=seed)
    host = torch.empty(data.shape, dtype=data.dtype, pin_memory=True)
    host.copy_(data)
    del data
    bf = ArrayBatches(host, NT=NT, nt=NTAP, device=device)
    nb = int(bf.n_batches)
    tm.extract(dict(ops), bf, U, device=device)
    torch.cuda.synchronize()
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        st, tF, _ = tm.extract(dict(ops), bf, U, device=device)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    return dict(value=C * NT * nb / med, unit="channel·samples/s", batches=nb,
                ms_per_batch=med * 1e3 / nb, median_s=med, runs=runs,
                spikes=int(len(st)),
                note="wall time of extract(): prepare_extract + prepare_matching, per-batch "
                     "host-to-device copies of float32 batches from pinned memory, matching, "
                     "features, st / tF on the host; median of the runs after one untimed call")


def time_extract_int16(tm, n_batches, seed, device, U, ops, runs=5):
    """``extract`` over a raw int16 recording in host memory through the device
    preprocessing (FilteredBatches: int16 host-to-device copy, channel means,
    CAR, the 300 Hz high-pass filter, whitening), as a sorting run streams
    it.  The recording is the bench's synthetic whitened data scaled by 20 and
    rounded to int16, with whitening matrix I / 20.  One untimed call, then the
    median of `runs` calls."""
    from kilosort_amd import synth, preprocess as PPr
    log(f"extract int16: synthesising {n_batches} batches")
    rec = synth.make_recording(seed, C, K, n_batches * NT,
                               rate_range=(1.0, 20.0), amp_range=(8.0, 40.0),
                               overlap_frac=0.2, with_data=False)
    data = synth.synthesize_torch(rec.n_samples, C, rec.times, rec.ids, rec.amps, rec.U,
                                  rec.wPCA, rec.nt0min, device, seed=seed)
    raw = (data * 20.0).round_().clamp_(-32768, 32767).to(torch.int16).cpu().numpy()
    del data
    fb = PPr.FilteredBatches(n_chan_bin=C, NT=NT, file_object=raw, device=device,
                             hp_filter=PPr.get_highpass_filter(device=device),
                             whiten_mat=np.eye(C, dtype=np.float32) / 20.0)
    nb = int(fb.n_batches)
    log(f"extract int16: {nb} batches, one untimed call")
    tm.extract(dict(ops), fb, U, device=device)
    torch.cuda.synchronize()
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        st, tF, _ = tm.extract(dict(ops), fb, U, device=device)
        times.append(time.perf_counter() - t0)
        log(f"extract int16: {times[-1]:.2f} s")
    med = float(np.median(times))
    return dict(value=C * NT * nb / med, unit="channel·samples/s", batches=nb,
                ms_per_batch=med * 1e3 / nb, median_s=med, runs=runs, spikes=int(len(st)),
                note="wall time of extract() over FilteredBatches of a host int16 recording: "
                     "int16 host-to-device copies, CAR, high-pass, whitening, matching, "
                     "features, st / tF on the host; median of the runs after one untimed call")


def phy_workload(ops, iCC, iU, iCC_mask, n, seed, n_clu=500, hours=1.0):
    """A sorted recording's export inputs: n spikes over `hours` from the
    bench's 700 detection templates at 383 channels, n_clu clusters (each
    template owned by one cluster, 10 % of spikes reassigned at random), 3 %
    artefact duplicates 4 samples apart, features [n, 10, 6]."""
    rng = np.random.default_rng(seed)
    K = int(iU.shape[0])
    t = np.sort(rng.integers(0, int(hours * 3600 * 30000), n)).astype(np.float64)
    dup = rng.uniform(size=n) < 0.03
    t[1:][dup[1:]] = t[:-1][dup[1:]] + 4
    t = np.sort(t)
    tmpl = rng.integers(0, K, n)
    owner = rng.permutation(np.arange(K) % n_clu)
    clu = owner[tmpl]
    split = rng.uniform(size=n) < 0.1
    clu[split] = rng.integers(0, n_clu, split.sum())
    st = np.stack([t, tmpl, rng.uniform(8, 30, n)], 1)
    tF = torch.from_numpy(rng.standard_normal((n, 10, 6)).astype(np.float32))
    pops = dict(iU=iU, iCC=iCC, iCC_mask=iCC_mask, xc=ops["xc"].astype(np.float32),
                yc=ops["yc"].astype(np.float32), nearest_chans=10, dmin=20.0, dminx=32.0,
                fs=30000.0, duplicate_spike_bins=7)
    return st, clu.astype(np.int32), tF, pops


def time_phy(ops, iCC, iU, iCC_mask, device, n=2_000_000, n_sample=200_000, seed=9,
             cpu_baseline=True):
    """The Phy export's per-spike work on n spikes (positions and amplitudes,
    duplicate removal, refractoriness, PC features): wall time per function
    with features resident on the device; the oracle on the first n_sample
    spikes as the CPU baseline."""
    from kilosort_amd import ccg as CG, postprocessing as PP
    st, clu, tF, pops = phy_workload(ops, iCC, iU, iCC_mask, n, seed)
    tFd = tF.to(device)
    t_s = st[:, 0].astype(np.int64)
    steps = {
        "positions_amplitudes": lambda: PP.positions_and_amplitudes(st, tFd, pops, device),
        "remove_duplicates": lambda: PP.remove_duplicates(t_s, clu, 7, device=device),
        "refract": lambda: CG.refract(clu, t_s / 30000.0, device=device),
        "make_pc_features": lambda: PP.make_pc_features(pops, st[:, 1].astype(np.int32), clu,
                                                        tFd.clone(), device=device),
    }
    out = {"spikes": n, "clusters": int(clu.max()) + 1, "templates": int(iU.shape[0]),
           "ms": {}}
    for name, fn in steps.items():
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        out["ms"][name] = (time.perf_counter() - t0) * 1e3
    total = sum(out["ms"].values())
    out["value"] = n / (total / 1e3)
    out["unit"] = "spikes/s"
    out["note"] = ("wall time of one call each, features resident on the device, results "
                   "copied to host numpy as the reference returns them")
    if cpu_baseline:
        from oracle import phy_oracle as PH, positions_oracle as PO
        m = n_sample
        sts, clus, tFs = st[:m], clu[:m], tF[:m].clone()
        cops = dict(pops, iU=iU.cpu(), iCC=iCC.cpu(), iCC_mask=iCC_mask.cpu())
        t0 = time.perf_counter()
        PO.compute_spike_positions(sts, tFs, cops)
        torch.norm(tFs, dim=[-2, -1])
        PH.remove_duplicates(sts[:, 0].astype(np.int64), clus, dt=7)
        PH.refract(clus, sts[:, 0].astype(np.int64) / 30000.0)
        PH.make_pc_features(cops, sts[:, 1].astype(np.int32), clus, tFs)
        sec = time.perf_counter() - t0
        out["cpu_baseline"] = dict(value=m / sec, unit="spikes/s", seconds=sec,
                                   cores=torch.get_num_threads(), kind="port",
                                   sample=f"first {m} spikes of the same workload: oracle "
                                          "positions, amplitudes, remove_duplicates, refract, "
                                          "make_pc_features")
    return out


def time_detect(ops, batches, device, reps=10, cpu_baseline=True):
    """Universal-template detection (spikedetect.template_match + the
    per-batch features of run) on the bench's 383-channel batches: device time
    per batch (the spike count is read back each batch, as run does),
    per-kernel times and the oracle on one batch as its CPU baseline."""
    from kilosort_amd import _lib, spikedetect as SD
    lib = _lib.load()
    C_ = len(ops["xc"])
    dops = dict(xc=ops["xc"], yc=ops["yc"], kcoords=np.zeros(C_), nt=NTAP, Th_universal=9.0,
                max_channel_distance=32.0, fs=30000.0, batch_size=NT,
                wTEMP=ops["wPCA"], wPCA=ops["wPCA"],
                settings=dict(dmin=None, dminx=32.0, nearest_chans=10, nearest_templates=100,
                              min_template_size=10.0, template_sizes=5, nt0min=20,
                              n_templates=6, n_pcs=6))
    iC, iC2, weigh = SD.prepare(dict(dops), device=device)
    det = SD._Detector(dops, iC, iC2, weigh, device)
    for i in range(2):
        det.run(batches[i % len(batches)])
    torch.cuda.synchronize()
    nsp = 0
    t0 = time.perf_counter()
    for r in range(reps):
        nsp += len(det.run(batches[r % len(batches)])[0])
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    lib.ksm_profile_reset()
    lib.ksm_profile_enable(3)
    for r in range(3):
        det.run(batches[r % len(batches)])
    torch.cuda.synchronize()
    lib.ksm_profile_enable(0)
    kern = {}
    for name in ("conv_pc", "detect_score", "detect_mask", "detect_emit", "detect_features"):
        t_ms, cnt = _lib.profile_query(name)
        if cnt:
            kern[name] = t_ms / 3
    lib.ksm_profile_reset()
    T = NT + 2 * NTAP
    nf = int(iC.shape[1])
    out = dict(value=C_ * T / (ms / 1e3), unit="channel·samples/s", ms_per_batch=ms,
               spikes_per_batch=nsp / reps, templates=nf, kernel_ms=kern,
               config=f"{C_} ch, {nf} universal-template positions x 5 sizes x 6 waveforms, "
                      "nC=10, nC2=100, Th=9")
    if "detect_score" in kern:
        # algorithmic work: 2 * nsizes * nk * nC flops per (template, column)
        flops = 2.0 * 5 * 6 * 10 * nf * T
        out["detect_score_tflops"] = flops / (kern["detect_score"] / 1e3) / 1e12
    if cpu_baseline:
        from oracle import detect_oracle as DO
        Xc = batches[0].cpu()
        cops = dict(dops, wTEMP=ops["wPCA"].cpu(), wPCA=ops["wPCA"].cpu())
        ciC, ciC2, cweigh = DO.setup(dict(cops))
        t0 = time.perf_counter()
        DO.detect_batch(Xc, cops, ciC, ciC2, cweigh)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = dict(value=C_ * T / dt, unit="channel·samples/s",
                                   cores=torch.get_num_threads(), kind="port",
                                   sample="1 padded batch: oracle template_match + features "
                                          "(torch CPU)", seconds=dt)
    return out


def time_refine(tm, device, nbatch=8, reps=3, steps=20, seed=96):
    """The differentiation workload: `nbatch` batches of NT=60000 at 383
    channels, 500 templates from the 64-channel workload's generator (Poisson
    2-20 Hz, amplitudes U(10, 30)), spike times, ids and amplitudes fixed by a
    forward FP32 matching pass.  Times the residual and the loss with its
    gradients for U, wPCA and the amplitudes (float32 kernels, events), and
    `steps` gradient-refinement steps on U from perturbed templates (wall
    time, line-search evaluations included), with the residual loss they
    reach beside the averaging update's from the same templates."""
    from kilosort_amd import refine, synth
    from kilosort_amd.batches import ArrayBatches
    Kr = 500
    rec = synth.make_recording(seed, C, Kr, nbatch * NT + 2 * NTAP, rate_range=(2.0, 20.0),
                               amp_range=(10.0, 30.0), with_data=False)
    data = synth.synthesize_torch(rec.n_samples, C, rec.times, rec.ids, rec.amps, rec.U,
                                  rec.wPCA, rec.nt0min, device, seed=seed)
    bf = ArrayBatches(data, NT=NT, nt=NTAP, device=device)
    W = torch.from_numpy(rec.wPCA).to(device)
    Ut = torch.from_numpy(rec.U).to(device)
    ops = dict(nt=NTAP, Th_learned=8, max_peels=100, wPCA=W, matching_precision="fp32")
    ctc = tm.prepare_matching(ops, Ut)
    data_b = []
    for i in range(nbatch):
        X = bf.padded_batch_to_torch(i).contiguous()
        st, amps, _, _ = tm.run_matching(ops, X, Ut, ctc, device=device)
        data_b.append((X, st, amps[:, 0].contiguous()))
    del data
    nsp = sum(len(d[1]) for d in data_b)
    g = torch.Generator().manual_seed(4)
    U0 = torch.from_numpy(rec.U)
    U0 = (U0 + 0.3 * U0.abs().max() * torch.randn(U0.shape, generator=g) * (U0 != 0)).to(device)

    def one_pass():
        Ur = U0.detach().requires_grad_(True)
        Wr = W.detach().requires_grad_(True)
        for X, st, a in data_b:
            ar = a.detach().requires_grad_(True)
            refine.residual_loss(X, Ur, Wr, ar, st, NTAP).backward()

    one_pass()
    torch.cuda.synchronize()
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    e0.record()
    for _ in range(reps):
        for X, st, a in data_b:
            refine.residual(X, U0, W, a, st, NTAP)
    e1.record()
    for _ in range(reps):
        one_pass()
    e2.record()
    torch.cuda.synchronize()
    fwd = e0.elapsed_time(e1) / (reps * nbatch)
    both = e1.elapsed_time(e2) / (reps * nbatch)

    def loss(Uc):
        return sum(float(refine.residual_loss(X, Uc, W, a, st, NTAP)) for X, st, a in data_b)

    loss0 = loss(U0)
    l_avg = loss(refine.averaging_update(data_b, U0, W, NTAP))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    U_gd, hist = refine.refine(data_b, U0, W, NTAP, steps=steps)
    torch.cuda.synchronize()
    t_steps = time.perf_counter() - t0
    l_gd = loss(U_gd)
    return {"batches": nbatch, "channels": C, "templates": Kr,
            "spikes_per_batch": nsp / nbatch,
            "residual_ms_per_batch": fwd, "loss_and_gradients_ms_per_batch": both,
            "refinement_steps": len(hist) - 1, "refinement_s": t_steps,
            "loss_initial": loss0, "loss_averaging_update": l_avg, "loss_after_steps": l_gd,
            "note": "float32; loss = sum over the batches of the squared residual of a padded "
                    "batch, gradients for U, wPCA and amplitudes with spike times fixed "
                    "(backward recomputes the residual); refinement from templates perturbed "
                    "by 0.3 max|U| Gaussian noise on their footprint, each step a "
                    "preconditioned gradient step with backtracking"}


def executed_gemm_flops(U, perm, precision):
    """MFMA flops the projection GEMM executes and the fraction of reduction
    steps it keeps.  Each form skips, per tile of templates (peel order), the
    reduction steps where the tile's U is zero:
      fp32:         64-template tiles, 16-row steps of r = p*C + c,
                    time columns padded to 256;
      bf16, bf16x3: 128-template tiles, 64-column steps of r = c*P + p
                    (bf16x3 runs three products per step), columns padded
                    to 128;
      fp8:          128-template tiles, 128-column steps of r = c*P + p."""
    Un = U.detach().cpu().numpy()[perm.cpu().numpy()]
    Kt, Pt, Ct = Un.shape
    rup = lambda a, b: -(-a // b) * b
    if precision == "fp32":
        tm_, step, cols, nseg = 64, 16, rup(T_PAD, 256), 1
        A = Un.reshape(Kt, Pt * Ct)                              # r = p*C + c
        R = rup(Pt * Ct, 16)
    else:
        tm_, cols = 128, rup(T_PAD, 128)
        step, nseg = {"bf16": (64, 1), "bf16x3": (64, 3), "fp8": (128, 1)}[precision]
        A = Un.transpose(0, 2, 1).reshape(Kt, Ct * Pt)          # r = c*P + p
        R = rup(Pt * Ct, 192) if step == 64 else rup(Pt * Ct, Pt * step)
    Kp = rup(Kt, tm_)
    Ap = np.zeros((Kp, R), bool)
    Ap[:Kt, :Pt * Ct] = A != 0
    steps = Ap.reshape(Kp // tm_, tm_, R // step, step).any(axis=(1, 3)).sum()
    return (2.0 * tm_ * step * float(steps) * cols * nseg,
            float(steps) / ((Kp // tm_) * (R // step)))


def ragged_workload(n_rec, max_batches, seed, device):
    """Recordings of uneven size: channels from {64, 128, 192, 256, 384},
    templates uniform in [50, 1000], 1 .. max_batches batches of NT, from the
    64-channel workload's generator (Poisson 2-20 Hz, amplitudes U(10, 30))
    with a seed per recording; the data is synthesised on the device."""
    from kilosort_amd import synth
    from kilosort_amd.batches import ArrayBatches
    rng = np.random.default_rng(seed)
    jobs = []
    for r in range(n_rec):
        if r % 8 == 0:
            log(f"ragged: synthesising recordings {r}..{min(r + 8, n_rec) - 1}")
        Cr = int(rng.choice([64, 128, 192, 256, 384]))
        Kr = int(rng.integers(50, 1001))
        nb = int(rng.integers(1, max_batches + 1))
        rec = synth.make_recording(seed * 1000 + r, Cr, Kr, nb * NT, rate_range=(2.0, 20.0),
                                   amp_range=(10.0, 30.0), with_data=False)
        data = synth.synthesize_torch(rec.n_samples, Cr, rec.times, rec.ids, rec.amps, rec.U,
                                      rec.wPCA, rec.nt0min, device, seed=seed * 1000 + r)
        bf = ArrayBatches(data, NT=NT, nt=NTAP, device=device)
        ops = dict(nt=NTAP, nt0min=rec.nt0min, Th_learned=8, max_peels=100, batch_size=NT,
                   wPCA=torch.from_numpy(rec.wPCA).to(device), xc=rec.xc, yc=rec.yc,
                   settings=dict(nearest_chans=10, position_limit=100.0, n_pcs=P))
        jobs.append((ops, bf, torch.from_numpy(rec.U).to(device)))
    return jobs


def _spike_pairs(st):
    return set(map(tuple, np.round(st[:, :2]).astype(np.int64).tolist()))


def time_ragged(tm, jobs, device, precision="bf16", runs=2):
    """Many uneven recordings.  Sequential: extract() of each recording in
    turn in FP32 (the library's default streams within a recording).
    Batched: all of them through one extract_many in `precision` with
    certified decisions, after the autotuner has picked per recording the
    projection tile, the peel schedule of the recorded graphs and whether
    batches launch graphs, and the stream count; the same pipeline is also
    timed in FP32 (batched_fp32).  All include each
    recording's preparation (prepare_extract + prepare_matching) and return
    st/tF on the host; times are the best of `runs` after one untimed run.
    Spike (time, template) agreement of the batched form with the sequential
    FP32 form is reported per recording."""
    from kilosort_amd import autotune
    work = sum(int(U.shape[2]) * NT * int(bf.n_batches) for _, bf, U in jobs)

    seq_st = []

    def seq(keep=False):
        for ops, bf, U in jobs:
            st, _, _ = tm.extract(dict(ops), bf, U, device=device)
            if keep:
                seq_st.append(st)

    jobs_b = [(dict(o, matching_precision=precision), b, U) for o, b, U in jobs]
    got = []

    def many(keep=False):
        out = tm.extract_many([(dict(o), b, U) for o, b, U in jobs_b], device=device,
                              streams=best)
        if keep:
            got.extend(st for st, _, _ in out)

    def timed(f):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        f()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    seq(keep=True)
    t_seq = min(timed(seq) for _ in range(runs))
    log(f"ragged: sequential fp32 {t_seq:.2f} s; autotuning")
    t0 = time.perf_counter()
    best, plans, tried = autotune.tune_many(jobs_b, device=device)
    t_tune = time.perf_counter() - t0
    many(keep=True)
    t_many = min(timed(many) for _ in range(runs))
    # the same pipeline in FP32 (the autotuner's streams and graph schedules),
    # which separates the pipeline's effect from the form's cost per batch
    jobs_f = [(dict(o, matching_precision="fp32"), b, U) for o, b, U in jobs]
    for (o32, _, _), (ob, _, _) in zip(jobs_f, jobs_b):
        o32["matching_graph_peels"] = ob.get("matching_graph_peels")
        o32["matching_graphs"] = ob.get("matching_graphs")
    t_many32 = min(timed(lambda: tm.extract_many([(dict(o), b, U) for o, b, U in jobs_f],
                                                 device=device, streams=best))
                   for _ in range(runs))
    agree = []
    for a_st, b_st in zip(seq_st, got):
        a, b = _spike_pairs(a_st), _spike_pairs(b_st)
        agree.append(len(a & b) / max(len(a), len(b), 1))
    nb = [int(bf.n_batches) for _, bf, _ in jobs]
    return {"recordings": len(jobs), "batches": int(sum(nb)),
            "timing": f"best of {runs} runs after one untimed run",
            "shapes": [[int(U.shape[2]), int(U.shape[0]), n] for (_, _, U), n in zip(jobs, nb)],
            "sequential_value": work / t_seq, "sequential_s": t_seq,
            "sequential_precision": "fp32",
            "batched_value": work / t_many, "batched_s": t_many, "batched_precision": precision,
            "batched_streams": best, "speedup": t_seq / t_many,
            "batched_fp32_value": work / t_many32, "batched_fp32_s": t_many32,
            "speedup_fp32": t_seq / t_many32,
            "autotune_s": t_tune, "batched_value_with_autotune": work / (t_many + t_tune),
            "autotune_streams_s": {str(k): v for k, v in tried.items()},
            "plans": plans,
            "graphed_recordings": int(sum(bool(p["graphs"]) for p in plans)),
            "spikes": int(sum(len(s) for s in seq_st)),
            "agreement_vs_fp32": {"min": float(min(agree)), "mean": float(np.mean(agree))},
            "unit": "channel·samples/s"}


PROFILED = ["conv_pc", "gemm_f32", "gemm_bf16", "gemm_fp8", "colmax_all", "peel_select",
            "peel_update", "features"]


def profile_form(tm, lib, ops, batches, U, ctc, iCC, iU, Ucc, device, precision, nbatch):
    """A separate pass with library events around every kernel: per-kernel
    totals (ms) and launch counts per batch, and the spike ids found."""
    o = dict(ops, matching_precision=precision)
    torch.cuda.synchronize()
    lib.ksm_profile_reset()
    lib.ksm_profile_enable(3)
    ids = []
    for i in range(nbatch):
        st, th, tf = tm._extract_batch(o, batches[i % len(batches)], U, ctc, iCC, iU, Ucc,
                                       device)
        ids.append(st[:, 1].numpy())
    torch.cuda.synchronize()
    lib.ksm_profile_enable(0)
    from kilosort_amd import _lib
    kern = {}
    for name in PROFILED:
        ms, n = _lib.profile_query(name)
        if n:
            kern[name] = dict(ms_per_batch=ms / nbatch, launches_per_batch=n / nbatch,
                              avg_launch_ms=ms / n)
    lib.ksm_profile_reset()
    return kern, np.concatenate(ids) if ids else np.zeros(0, np.int64)


TRAFFIC = os.path.join(ROOT, "profiles", "traffic_383ch_700t.json")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    runs on this workload (FETCH_SIZE + WRITE_SIZE, calibrated with a plain
    copy); None when absent."""
    try:
        with open(TRAFFIC) as fh:
            k = json.load(fh)["kernels"][kernel]
        return k["read_bytes"] + k["write_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def trace_avg_ms(kernel):
    """Average duration (ms) of `kernel` in the committed rocprofv3 kernel
    trace of this workload; None when absent."""
    try:
        with open(TRAFFIC) as fh:
            v = json.load(fh)["kernels"][kernel]["trace_avg_us"]
        return None if v is None else v / 1e3
    except (OSError, KeyError, ValueError):
        return None


def peel_update_bytes(ids, nnzK, nnzC, nbatch):
    """Algorithmic HBM bytes of the subtraction per batch: for each accepted
    spike of template j, 12 B per (template i, lag) of its nonzero Gram rows
    (ctc read + score read-modify-write over 2nt+1 lags) and 8 B per
    (channel, lag) of its nonzero channels (residual read-modify-write)."""
    return float((12 * (2 * NTAP + 1) * nnzK[ids] + 8 * NTAP * nnzC[ids]).sum()) / nbatch


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=6)
    ap.add_argument("--precision", default="fp32", choices=FORMS,
                    help="form timed for `value`; every form is also timed under `forms`")
    ap.add_argument("--forms", default=",".join(DEFAULT_FORMS),
                    help="comma list of forms to time beside `value` ('' for none)")
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--ragged", type=int, default=64,
                    help="recordings of the ragged (many uneven recordings) run; 0 skips it")
    ap.add_argument("--ragged-batches", type=int, default=40,
                    help="most batches of one recording in the ragged run")
    ap.add_argument("--streams", type=int, default=None,
                    help="device streams of the batch pipeline (default: the library's)")
    ap.add_argument("--no-cpu-baseline", action="store_true",
                    help="skip the CPU oracle runs (cpu_baseline and FP64 agreement)")
    ap.add_argument("--preprocess-only", action="store_true",
                    help="time only the batch preprocessing and print its JSON object")
    ap.add_argument("--detect-only", action="store_true",
                    help="time only the universal-template detection and print its JSON object")
    ap.add_argument("--phy-only", action="store_true",
                    help="time only the Phy export's per-spike work and print its JSON object")
    ap.add_argument("--matching-only", action="store_true",
                    help="only the matching forms (timing, per-kernel times, rooflines): the "
                         "run a kernel profile of the matching path uses")
    ap.add_argument("--templates-only", action="store_true",
                    help="time only template similarity and alignment and print their JSON object")
    ap.add_argument("--extract-batches", type=int, default=300,
                    help="batches of the int16 recording extract() streams (10 min at 30 kHz)")
    ap.add_argument("--extract-only", action="store_true",
                    help="run only the int16 extract() leg and print its JSON object")
    ap.add_argument("--ragged-only", action="store_true",
                    help="run only the many-recordings comparison and print its JSON object")
    ap.add_argument("--refine-only", action="store_true",
                    help="time only the template refinement and print its JSON object")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the matching path is sharded by batch "
              "and no multi-GPU run is part of this build", file=sys.stderr)
        sys.exit(2)

    from kilosort_amd import _lib
    from kilosort_amd import template_matching as tm

    if args.preprocess_only:
        device = torch.device("cuda:0")
        _lib.load()
        print(json.dumps(time_preprocess(device, args.seed + 200,
                                         cpu_baseline=not args.no_cpu_baseline)), flush=True)
        return

    device = torch.device("cuda:0")
    torch.cuda.set_device(device)
    lib = _lib.load()
    if args.extract_only:
        rec, ops, _ = make_workload(1, args.seed, device)
        U = torch.from_numpy(rec.U).to(device)
        print(json.dumps(time_extract_int16(tm, args.extract_batches, args.seed, device, U, ops,
                                            runs=2)), flush=True)
        return
    if args.ragged_only:
        jobs = ragged_workload(args.ragged, args.ragged_batches, args.seed + 100, device)
        print(json.dumps(time_ragged(tm, jobs, device)), flush=True)
        return
    n_distinct = min(args.steps + args.warmup, 12)
    log(f"generating {n_distinct} synthetic batches")
    rec, ops, batches = make_workload(n_distinct, args.seed, device)
    if args.detect_only:
        print(json.dumps(time_detect(ops, batches, device,
                                     cpu_baseline=not args.no_cpu_baseline)), flush=True)
        return
    U = torch.from_numpy(rec.U).to(device)
    iCC, iCC_mask, iU, Ucc = tm.prepare_extract(ops["xc"], ops["yc"], U, 10, 100.0,
                                                device=device)
    if args.phy_only:
        print(json.dumps(time_phy(ops, iCC, iU, iCC_mask, device,
                                  cpu_baseline=not args.no_cpu_baseline)), flush=True)
        return
    ctc = tm.prepare_matching(ops, U)
    if args.templates_only:
        print(json.dumps(time_templates(ops, U, device, cpu_baseline=not args.no_cpu_baseline)),
              flush=True)
        return
    if args.refine_only:
        print(json.dumps(time_refine(tm, device)), flush=True)
        return
    nnzK = (ctc != 0).any(dim=2).sum(dim=0).cpu().numpy().astype(np.int64)
    nnzC = (U != 0).any(dim=1).sum(dim=1).cpu().numpy().astype(np.int64)
    torch.cuda.synchronize()

    forms = [f for f in args.forms.split(",") if f]
    if args.precision not in forms:
        forms.insert(0, args.precision)
    results = {}
    for f in forms:
        log(f"timing {f}: warmup {args.warmup}, {args.steps} steps")
        results[f] = time_form(tm, lib, ops, batches, U, ctc, iCC, iU, Ucc, device, f,
                               args.steps, args.warmup, lanes=args.streams)
        log(f"{f}: {results[f]['ms_per_step']:.3f} ms/step on "
            f"{results[f]['streams']} streams")
    main_res = results[args.precision]

    nprof = min(n_distinct, 4)
    prof = {}
    for f in forms:
        prof[f] = profile_form(tm, lib, ops, batches, U, ctc, iCC, iU, Ucc, device, f, nprof)
        log(f"{f} kernels per batch (ms, launches): " + ", ".join(
            f"{k}={v['ms_per_batch']:.3f}/{v['launches_per_batch']:.0f}"
            for k, v in prof[f][0].items()))
    kern, ids = prof[args.precision]
    T = NT + 2 * NTAP
    upd = kern["peel_update"]
    upd_bytes = peel_update_bytes(ids, nnzK, nnzC, nprof)
    upd_achieved = upd_bytes / (upd["ms_per_batch"] / 1e3)
    roofline = {"kernel": "peel_update_kernel (subtraction + column maxima, per batch)",
                "bound": "hbm", "achieved": upd_achieved / 1e9, "peak": PEAK_HBM / 1e9,
                "unit": "GB/s", "frac": upd_achieved / PEAK_HBM,
                "traffic": measured_traffic("peel_update_kernel"),
                "traffic_source": os.path.relpath(TRAFFIC, ROOT),
                "bytes_per_batch": upd_bytes,
                "bytes_per_launch": upd_bytes / upd["launches_per_batch"],
                "avg_launch_ms": upd["avg_launch_ms"],
                "launches_per_batch": upd["launches_per_batch"]}
    # the committed kernel trace of the same workload: events stamped through
    # hipExtLaunchKernel read a few microseconds longer per launch than the
    # trace's dispatch timestamps (see the profile summary)
    tr = trace_avg_ms("peel_update_kernel")
    if tr:
        roofline["rocprof_avg_launch_ms"] = tr
        roofline["frac_at_rocprof_duration"] = (upd_bytes / upd["launches_per_batch"]) / (
            tr / 1e3) / PEAK_HBM
    gname = GEMM_KERNEL[args.precision]
    g = kern[gname]
    flops_gemm = 2.0 * K * P * C * T
    flops_run, kept = executed_gemm_flops(U, tm._grams.get(ctc, U, NTAP, device).perm,
                                          args.precision)
    g_ach = flops_run / (g["avg_launch_ms"] / 1e3)
    peak = PEAK_MFMA[args.precision]
    roofline_projection = {"kernel": f"{gname} (projection GEMM)", "bound": "mfma",
                           "achieved": g_ach / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                           "frac": g_ach / peak, "flops_per_launch": flops_run,
                           "dense_flops": flops_gemm, "steps_kept": kept,
                           "dense_equivalent_tflops": flops_gemm / (g["avg_launch_ms"] / 1e3) / 1e12,
                           "traffic": measured_traffic(f"{gname}_mfma"),
                           "avg_launch_ms": g["avg_launch_ms"]}

    # per form: the projection GEMM against its MFMA peak and the peel update
    # against HBM (the update's bytes do not depend on the form)
    form_rooflines = {}
    perm = tm._grams.get(ctc, U, NTAP, device).perm
    for f in forms:
        kf = prof[f][0]
        gk = kf.get(GEMM_KERNEL[f])
        fl, kept_f = executed_gemm_flops(U, perm, f)
        r = {}
        if gk:
            ach = fl / (gk["avg_launch_ms"] / 1e3)
            # the same launch against HBM: its algorithmic bytes are the
            # correlation operand read once in the form's storage plus the
            # FP32 scores written
            Kf, Pf, Cf = U.shape
            Tb = int(batches[0].shape[1])
            op_bytes = {"fp32": 4, "bf16": 2, "bf16x3": 4, "fp8": 1}[f] * Pf * Cf * Tb
            hb = (op_bytes + 4 * Kf * Tb) / (gk["avg_launch_ms"] / 1e3)
            r["projection"] = {"kernel": f"{GEMM_KERNEL[f]}_mfma", "bound": "mfma",
                               "achieved": ach / 1e12, "peak": PEAK_MFMA[f] / 1e12,
                               "unit": "TFLOP/s", "frac": ach / PEAK_MFMA[f],
                               "flops_per_launch": fl, "steps_kept": kept_f,
                               "avg_launch_ms": gk["avg_launch_ms"],
                               "hbm": {"bytes_per_launch": op_bytes + 4 * Kf * Tb,
                                       "achieved": hb / 1e9, "peak": PEAK_HBM / 1e9,
                                       "unit": "GB/s", "frac": hb / PEAK_HBM}}
        uk = kf.get("peel_update")
        if uk:
            ub = peel_update_bytes(prof[f][1], nnzK, nnzC, nprof) / uk["ms_per_batch"] * 1e3
            r["peel_update"] = {"bound": "hbm", "achieved": ub / 1e9, "peak": PEAK_HBM / 1e9,
                                "unit": "GB/s", "frac": ub / PEAK_HBM}
        form_rooflines[f] = r

    if args.matching_only:
        print(json.dumps({"metric": METRIC, "forms": {f: dict(r, roofline=form_rooflines[f])
                                                      for f, r in results.items()},
                          "roofline": roofline, "roofline_projection": roofline_projection,
                          "kernels_per_batch": kern}), flush=True)
        return

    log("template refinement (residual loss and gradients)")
    refine_res = time_refine(tm, device)
    log(f"refinement ({refine_res['batches']} x {refine_res['channels']} ch x "
        f"{refine_res['templates']} templates): residual {refine_res['residual_ms_per_batch']:.3f} ms, "
        f"loss+gradients {refine_res['loss_and_gradients_ms_per_batch']:.3f} ms per batch, "
        f"{refine_res['refinement_steps']} steps {refine_res['refinement_s']:.2f} s, loss "
        f"{refine_res['loss_initial']:.4g} -> {refine_res['loss_after_steps']:.4g} "
        f"(averaging update {refine_res['loss_averaging_update']:.4g})")

    log("template similarity and alignment (700 templates)")
    tmpl = time_templates(ops, U, device, cpu_baseline=not args.no_cpu_baseline)
    log(f"templates: similarity {tmpl['similarity_ms']:.2f} ms, align_U {tmpl['align_U_ms']:.2f} ms")

    log("extract() wall time, recording in host memory (12 batches)")
    ext = time_extract(tm, 12, args.seed, device, U, ops)
    log(f"extract: {ext['value']:.3e} channel·samples/s, {ext['ms_per_batch']:.3f} ms/batch "
        f"({ext['batches']} batches, median of {ext['runs']})")
    log(f"extract() over a raw int16 recording of {args.extract_batches} batches through the "
        "device preprocessing")
    ext16 = time_extract_int16(tm, args.extract_batches, args.seed, device, U, ops, runs=2)
    ext["int16_preprocessed"] = ext16
    log(f"extract int16: {ext16['value']:.3e} channel·samples/s, {ext16['ms_per_batch']:.3f} "
        f"ms/batch ({ext16['spikes']} spikes)")

    log("Phy export per-spike work (2 M spikes)")
    phy = time_phy(ops, iCC, iU, iCC_mask, device, cpu_baseline=not args.no_cpu_baseline)
    log("phy: " + ", ".join(f"{k}={v:.1f} ms" for k, v in phy["ms"].items()))

    log("universal-template detection at 383 channels")
    det = time_detect(ops, batches, device, cpu_baseline=not args.no_cpu_baseline)
    log(f"detection: {det['ms_per_batch']:.3f} ms/batch, {det['spikes_per_batch']:.0f} spikes; "
        "kernels " + ", ".join(f"{k}={v:.3f}" for k, v in det["kernel_ms"].items()))

    log("batch preprocessing (CAR, high-pass, whitening) at 383 channels")
    pre = time_preprocess(device, args.seed + 200, cpu_baseline=not args.no_cpu_baseline)
    log(f"preprocessing: {pre['ms_per_batch']:.3f} ms/batch resident, "
        f"{pre['end_to_end_ms_per_batch']:.3f} ms/batch from host int16; kernels "
        + ", ".join(f"{k}={v:.3f}" for k, v in pre["kernel_ms"].items()))

    ragged = None
    if args.ragged > 0:
        log(f"ragged: {args.ragged} recordings of 1-{args.ragged_batches} batches, sequential "
            "fp32 extract vs autotuned bf16 extract_many")
        jobs = ragged_workload(args.ragged, args.ragged_batches, args.seed + 100, device)
        ragged = time_ragged(tm, jobs, device)
        del jobs
        torch.cuda.empty_cache()
        log(f"ragged: sequential {ragged['sequential_value']:.3e}, batched "
            f"{ragged['batched_value']:.3e} channel·samples/s on {ragged['batched_streams']} "
            f"streams ({ragged['speedup']:.2f}x; batched fp32 {ragged['batched_fp32_value']:.3e}, "
            f"{ragged['speedup_fp32']:.2f}x), agreement min "
            f"{100 * ragged['agreement_vs_fp32']['min']:.3f} %")

    cpu = None
    agree = None
    if not args.no_cpu_baseline:
        Xs = batches[0]
        log("cpu baseline (oracle on one batch) and FP64 oracle")
        cpu, ref32, ref64 = oracle_runs(ops, Xs, U)
        log(f"cpu baseline {cpu['seconds']:.1f}s on {cpu['cores']} threads")
        agree = {}
        for f in forms:
            o = dict(ops, matching_precision=f)
            st, amps, _, _ = tm.run_matching(o, Xs, U, ctc, device=device)
            st = st.cpu().numpy()
            amps = amps.cpu().numpy()[:, 0]
            agree[f] = {"vs_fp64_oracle": agreement(st, amps, ref64),
                        "vs_fp32_oracle": agreement(st, amps, ref32)}
            log(f"{f}: vs fp64 {agree[f]['vs_fp64_oracle']['exact_pct']:.3f} %, "
                f"vs fp32 {agree[f]['vs_fp32_oracle']['exact_pct']:.3f} %")

    out = {
        "metric": METRIC,
        "value": main_res["value"],
        "unit": "channel·samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else args.precision,
        "data": "synthetic",
        "config": {"workload": "383 ch x 700 templates, NT=60000 (+2x61 pad) batches, "
                               "Th=8, max_peels=100, Poisson 1-20 Hz, 20% overlaps",
                   "channels": C, "templates": K, "batch_samples": NT,
                   "precision": args.precision, "distinct_batches": n_distinct,
                   "streams": main_res["streams"],
                   "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))},
        "roofline": roofline,
        "roofline_projection": roofline_projection,
        "kernels_per_batch": kern,
        "spikes_per_batch": main_res["spikes_per_batch"],
        "peel_iterations_per_batch": main_res["peel_iterations_per_batch"],
        "forms": {f: dict(value=results[f]["value"], ms_per_step=results[f]["ms_per_step"],
                          kernel_ms_per_batch={k: v["ms_per_batch"]
                                               for k, v in prof[f][0].items()},
                          roofline=form_rooflines[f])
                  for f in forms},
        "agreement_batch0": agree,
        "refinement": refine_res,
        "preprocess": pre,
        "detect": det,
        "templates": tmpl,
        "extract_wall": ext,
        "phy": phy,
        "ragged": ragged,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
