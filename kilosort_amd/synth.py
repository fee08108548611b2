"""Deterministic synthetic recordings for the learned-template matching path.

The reference has no offline generator for this path (its hybrid simulator in
``kilosort/simulation.py`` needs external waveform files), so this module
builds one from the description of the benchmark workloads:

* a 2-column staggered probe (x in {16, 48} um, 20 um row pitch);
* an orthonormal temporal basis ``wPCA`` [P, nt] made from Gaussian-derivative
  waveforms whose trough sits at ``nt0min`` (as ``spikedetect.py:94-98`` loads
  a [6, 61] basis with orthonormal rows);
* spatial components ``U`` [K, P, C] (the layout ``template_matching.py:170``
  takes) from a Gaussian footprint (sigma 30 um) at a uniform position times
  random PC coefficients, normalised to unit Frobenius norm;
* Poisson spike trains with a refractory period, optional forced overlaps;
* whitened data = i.i.d. N(0, 1) noise + sum of amp * (U_j^T wPCA) windows.

Ground-truth times follow the reference's output convention: a template whose
61-sample window starts at absolute sample ``w`` is reported by ``extract`` at
``w + nt0min`` (``template_matching.py:102-104`` with the padded-batch offset of
``io.py:771-783``).
"""
from __future__ import annotations

import dataclasses

import numpy as np


def probe_2col(C: int, pitch: float = 20.0, xcols=(16.0, 48.0)):
    """Channel coordinates of a 2-column staggered probe (float32, microns)."""
    c = np.arange(C)
    xc = np.where(c % 2 == 0, xcols[0], xcols[1]).astype(np.float32)
    yc = (pitch * (c // 2) + 0.5 * pitch * (c % 2)).astype(np.float32)
    return xc, yc


def make_wpca(nt: int = 61, P: int = 6, nt0min: int = 20,
              sigmas=(4.0, 6.0, 8.0)) -> np.ndarray:
    """Orthonormal temporal basis [P, nt] of Gaussian derivatives.

    Waveforms g, g', g'' ... of Gaussians centred on ``nt0min`` are stacked,
    orthonormalised with a QR decomposition in float64, and rounded to float32.
    """
    t = np.arange(nt, dtype=np.float64) - nt0min
    waves = []
    order = 0
    while len(waves) < P:
        for s in sigmas:
            g = np.exp(-t ** 2 / (2 * s * s))
            if order == 0:
                w = -g
            elif order == 1:
                w = t / (s * s) * g
            else:
                w = (1 - t ** 2 / (s * s)) * g
            waves.append(w)
            if len(waves) == P:
                break
        order = (order + 1) % 3
    A = np.stack(waves, 1)                     # [nt, P]
    Q, R = np.linalg.qr(A)
    Q = Q * np.sign(np.diag(R))[None, :]
    return np.ascontiguousarray(Q.T).astype(np.float32)


def make_templates(rng: np.random.Generator, K: int, xc, yc, P: int = 6,
                   sigma_um: float = 30.0):
    """Spatial components U [K, P, C] (float32) and template centres [K, 2]."""
    xc = np.asarray(xc, np.float64)
    yc = np.asarray(yc, np.float64)
    pos = np.stack([rng.uniform(xc.min(), xc.max(), K),
                    rng.uniform(yc.min(), yc.max(), K)], 1)
    d2 = (xc[None, :] - pos[:, :1]) ** 2 + (yc[None, :] - pos[:, 1:]) ** 2
    foot = np.exp(-d2 / (2 * sigma_um ** 2))                   # [K, C]
    coef = rng.standard_normal((K, P))
    coef[:, 0] += 2.0 * np.sign(coef[:, 0] + 1e-12)            # dominant first PC
    U = coef[:, :, None] * foot[:, None, :]
    U /= np.sqrt((U ** 2).sum((1, 2), keepdims=True))
    return U.astype(np.float32), pos


def poisson_train(rng, rate_hz: float, n_samples: int, fs: float,
                  refractory_samples: int) -> np.ndarray:
    """Sorted spike samples of a refractory Poisson process with mean rate."""
    mean_isi = fs / rate_hz
    extra = max(mean_isi - refractory_samples, 1.0)
    n_draw = int(n_samples / mean_isi * 1.5) + 16
    out = []
    t = rng.uniform(0, mean_isi)
    while True:
        isi = refractory_samples + rng.exponential(extra, n_draw)
        ts = t + np.cumsum(isi)
        out.append(ts)
        t = ts[-1]
        if t >= n_samples:
            break
    ts = np.concatenate(out)
    return np.floor(ts[ts < n_samples]).astype(np.int64)


@dataclasses.dataclass
class Recording:
    """A synthetic whitened recording plus its ground truth."""
    data: np.ndarray | None      # [n_samples, C] float32 (None if not synthesised)
    xc: np.ndarray               # [C] float32
    yc: np.ndarray               # [C] float32
    wPCA: np.ndarray             # [P, nt] float32
    U: np.ndarray                # [K, P, C] float32
    times: np.ndarray            # [S] int64, reference convention (window start + nt0min)
    ids: np.ndarray              # [S] int64
    amps: np.ndarray             # [S] float32
    n_samples: int
    nt: int
    nt0min: int


def make_spikes(rng, K, n_samples, fs, rate_range, amp_range, nt, nt0min,
                refractory_ms=2.0, overlap_frac=0.0, centres=None,
                overlap_um=60.0, overlap_ms=1.0):
    """Spike (time, id, amp) lists, optionally forcing near-coincident pairs."""
    refr = int(round(refractory_ms * fs / 1000.0))
    rates = rng.uniform(rate_range[0], rate_range[1], K)
    times, ids = [], []
    for j in range(K):
        tj = poisson_train(rng, rates[j], n_samples, fs, refr)
        times.append(tj)
        ids.append(np.full(len(tj), j, np.int64))
    times = np.concatenate(times) if times else np.zeros(0, np.int64)
    ids = np.concatenate(ids) if ids else np.zeros(0, np.int64)
    if overlap_frac > 0 and len(times) > 1 and centres is not None:
        # move a fraction of the spikes next to a spike of a nearby template
        d = np.sqrt(((centres[:, None, :] - centres[None, :, :]) ** 2).sum(-1))
        lag = int(round(overlap_ms * fs / 1000.0))
        pick = np.nonzero(rng.random(len(times)) < overlap_frac)[0]
        by_id = [times[ids == j].copy() for j in range(K)]
        for s in pick:
            j = ids[s]
            nb = np.nonzero((d[j] < overlap_um) & (np.arange(K) != j))[0]
            if len(nb) == 0:
                continue
            partner = nb[rng.integers(len(nb))]
            cand = by_id[partner]
            if len(cand) == 0:
                continue
            shift = 0
            while shift == 0:
                shift = int(rng.integers(-lag, lag + 1))
            times[s] = cand[rng.integers(len(cand))] + shift
        # restore the refractory period within each train
        keep = np.ones(len(times), bool)
        order = np.lexsort((times, ids))
        t_o, i_o = times[order], ids[order]
        bad = (np.diff(t_o) < refr) & (np.diff(i_o) == 0)
        keep[order[1:][bad]] = False
        times, ids = times[keep], ids[keep]
    # keep full windows inside the recording
    lo = nt0min
    hi = n_samples - (nt - nt0min)
    ok = (times >= lo) & (times < hi)
    times, ids = times[ok], ids[ok]
    order = np.argsort(times, kind="stable")
    times, ids = times[order], ids[order]
    amps = rng.uniform(amp_range[0], amp_range[1], len(times)).astype(np.float32)
    return times, ids, amps


def template_waveforms(U: np.ndarray, wPCA: np.ndarray) -> np.ndarray:
    """Full templates T[j] = U[j]^T wPCA, shape [K, C, nt], float64."""
    return np.einsum("kpc,pt->kct", U.astype(np.float64), wPCA.astype(np.float64))


def synthesize(rng, n_samples, C, times, ids, amps, U, wPCA, nt0min,
               noise_std=1.0) -> np.ndarray:
    """Whitened data [n_samples, C] float32: N(0, noise_std^2) + templates."""
    nt = wPCA.shape[1]
    X = (noise_std * rng.standard_normal((C, n_samples))).astype(np.float64)
    T = template_waveforms(U, wPCA)
    starts = times - nt0min
    for s0, j, a in zip(starts, ids, amps):
        X[:, s0:s0 + nt] += a * T[j]
    return np.ascontiguousarray(X.T).astype(np.float32)


def make_recording(seed: int, C: int, K: int, n_samples: int, *, fs=30000.0,
                   nt=61, nt0min=20, P=6, rate_range=(2.0, 20.0),
                   amp_range=(10.0, 30.0), overlap_frac=0.0,
                   refractory_ms=2.0, with_data=True, noise_std=1.0):
    """One seeded recording: probe, basis, templates, ground truth, data."""
    rng = np.random.default_rng(seed)
    xc, yc = probe_2col(C)
    wPCA = make_wpca(nt, P, nt0min)
    U, centres = make_templates(rng, K, xc, yc, P)
    times, ids, amps = make_spikes(rng, K, n_samples, fs, rate_range, amp_range,
                                   nt, nt0min, refractory_ms, overlap_frac,
                                   centres)
    data = None
    if with_data:
        data = synthesize(rng, n_samples, C, times, ids, amps, U, wPCA, nt0min,
                          noise_std)
    return Recording(data, xc, yc, wPCA, U, times, ids, amps, n_samples, nt,
                     nt0min)


# Named workloads.  Sizes are those of the benchmark configurations.
def config1_recording(seed: int = 1, n_batches: int = 2, NT: int = 60000):
    """64 channels, 64 templates, Poisson 2-20 Hz, amplitudes U(10, 30)."""
    return make_recording(seed, 64, 64, n_batches * NT, rate_range=(2.0, 20.0),
                          amp_range=(10.0, 30.0))


def config2_recording(seed: int = 2, n_batches: int = 1, NT: int = 60000,
                      with_data=True):
    """383 channels, 700 templates, Poisson 1-20 Hz, amplitudes U(8, 40),
    20 % of spikes moved within +-1 ms of a spike of a nearby template."""
    return make_recording(seed, 383, 700, n_batches * NT, rate_range=(1.0, 20.0),
                          amp_range=(8.0, 40.0), overlap_frac=0.2,
                          with_data=with_data)


def synthesize_torch(n_samples, C, times, ids, amps, U, wPCA, nt0min, device,
                     seed=0, noise_std=1.0, chunk=4096):
    """Device-side version of ``synthesize`` for large workloads: returns a
    [n_samples, C] float32 tensor on ``device`` (noise from torch's generator,
    so values differ from the numpy version for the same seed)."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    nt = wPCA.shape[1]
    X = noise_std * torch.randn((C, n_samples), generator=g, device=device)
    T = torch.from_numpy(template_waveforms(U, wPCA).astype(np.float32)).to(device)
    starts = torch.from_numpy(np.asarray(times) - nt0min).to(device)
    idt = torch.from_numpy(np.asarray(ids)).to(device)
    amt = torch.from_numpy(np.asarray(amps, np.float32)).to(device)
    lag = torch.arange(nt, device=device)
    for a in range(0, len(starts), chunk):
        s0 = starts[a:a + chunk]
        vals = amt[a:a + chunk, None, None] * T[idt[a:a + chunk]]        # [S, C, nt]
        idx = (s0[:, None] + lag[None, :]).reshape(-1)                   # [S*nt]
        X.index_add_(1, idx, vals.permute(1, 0, 2).reshape(C, -1))
    return X.T.contiguous()
