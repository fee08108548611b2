"""Phy export of the sorting results, with the per-spike work on the GPU.

``save_to_phy`` mirrors ``kilosort/io.py:248-469`` (same arguments, same
files, same return values).  The spike-level and template-level arithmetic
runs in ``libksm.so``:

* spike positions and amplitudes -- ``ksm_spike_positions`` (one pass over
  tF);
* duplicate removal -- ``ksm_group`` + ``ksm_remove_duplicates``;
* template similarity -- ``ksm_similarity``; template waveforms
  ``Wall @ wPCA`` -- ``ksm_gemm_f32``;
* PC features and their channel indices -- ``ksm_template_sums`` +
  ``ksm_gather_features``;
* contamination (refractoriness) -- ``ksm_acg`` autocorrelogram counts.

The host writes the ``.npy`` / ``.tsv`` / ``params.py`` files exactly as the
reference names and formats them.
"""
from __future__ import annotations

import logging
import shutil
from pathlib import Path

import numpy as np
import torch

from . import _lib
from . import ccg as CG
from . import postprocessing as PP
from . import templates as TM

logger = logging.getLogger(__name__)


def template_waveforms(Wall, wPCA, device):
    """(Wall [K, C, P] @ wPCA [P, nt]) as [K, nt, C] float32 numpy: the Phy
    ``templates.npy`` (``io.py:352-353``)."""
    lib = _lib.load()
    W = torch.as_tensor(Wall).to(device, torch.float32).contiguous()
    B = torch.as_tensor(wPCA).to(device, torch.float32).contiguous()
    K, C, P = W.shape
    nt = B.shape[1]
    out = torch.empty((K * C, nt), dtype=torch.float32, device=device)
    if K * C:
        _lib.check(lib.ksm_gemm_f32(W.data_ptr(), P, B.data_ptr(), nt, out.data_ptr(), nt,
                                    K * C, nt, P, _lib.stream_of(device)), "ksm_gemm_f32")
    return out.reshape(K, C, nt).cpu().numpy().transpose(0, 2, 1)


def save_to_phy(st, clu, tF, Wall, probe, ops, imin, results_dir=None, data_dtype=None,
                save_extra_vars=False, save_preprocessed_copy=False, skip_dat_path=False,
                device=None):
    """``io.save_to_phy``: writes the Phy files of the sorting into
    ``results_dir`` and returns (results_dir, similar_templates, is_ref,
    est_contam_rate, kept_spikes)."""
    device = PP._device_of(ops, device)
    if results_dir is None:
        results_dir = ops['data_dir'].joinpath('kilosort4')
    results_dir = Path(results_dir)
    results_dir.mkdir(exist_ok=True)

    # probe properties
    chan_map = probe['chanMap']
    channel_positions = np.stack((probe['xc'], probe['yc']), axis=-1)
    np.save(results_dir / 'channel_map.npy', chan_map)
    np.save(results_dir / 'channel_positions.npy', channel_positions)
    np.save(results_dir / 'channel_shanks.npy', probe['kcoords'])

    # whitening matrix
    whitening_mat = ops['Wrot']
    np.save(results_dir / 'whitening_mat_dat.npy', whitening_mat.cpu())
    whitening_mat_inv = torch.inverse(
        whitening_mat + 1e-5 * torch.eye(whitening_mat.shape[0]).to(whitening_mat.device))
    np.save(results_dir / 'whitening_mat.npy', whitening_mat.cpu())
    np.save(results_dir / 'whitening_mat_inv.npy', whitening_mat_inv.cpu())

    # spike properties
    spike_times = st[:, 0].astype('int64') + imin
    spike_templates = st[:, 1].astype('int32')
    spike_clusters = clu
    xs, ys, amps = PP.positions_and_amplitudes(st, tF, ops, device)
    pos_dtype = np.result_type(np.asarray(ops['xc']).dtype, np.float32)
    spike_positions = np.vstack([xs.cpu().numpy(), ys.cpu().numpy()]).T.astype(pos_dtype)
    amplitudes = amps.cpu().numpy()

    # remove duplicate (artifact) spikes
    spike_times, spike_clusters, kept_spikes = PP.remove_duplicates(
        spike_times, spike_clusters, dt=ops['duplicate_spike_bins'], device=device)
    amp = amplitudes[kept_spikes]
    spike_templates = spike_templates[kept_spikes]
    spike_positions = spike_positions[kept_spikes]
    np.save(results_dir / 'spike_times.npy', spike_times)
    np.save(results_dir / 'spike_templates.npy', spike_clusters)
    np.save(results_dir / 'spike_clusters.npy', spike_clusters)
    np.save(results_dir / 'spike_positions.npy', spike_positions)
    np.save(results_dir / 'spike_detection_templates.npy', spike_templates)
    np.save(results_dir / 'amplitudes.npy', amp)
    np.save(results_dir / 'kept_spikes.npy', kept_spikes)

    # template properties
    similar_templates = TM.similarity_matrix(Wall, ops['wPCA'], ops['nt'], device)[0]
    template_amplitudes = torch.norm(torch.as_tensor(Wall).cpu(), dim=[-2, -1]).numpy()
    templates = template_waveforms(Wall, ops['wPCA'], device)
    templates_ind = np.tile(np.arange(Wall.shape[1])[np.newaxis, :], (templates.shape[0], 1))
    np.save(results_dir / 'similar_templates.npy', similar_templates)
    np.save(results_dir / 'templates.npy', templates)
    np.save(results_dir / 'templates_ind.npy', templates_ind)

    # pc features
    if save_extra_vars:
        np.save(results_dir / 'tF.npy', tF.cpu().numpy())
    tF = tF[torch.from_numpy(kept_spikes).to(tF.device)]
    pc_features, pc_feature_ind = PP.make_pc_features(ops, spike_templates, spike_clusters, tF,
                                                      device=device)
    np.save(results_dir / 'pc_features.npy', pc_features.cpu().numpy())
    np.save(results_dir / 'pc_feature_ind.npy', pc_feature_ind)

    # contamination ratio
    acg_threshold = ops['settings']['acg_threshold']
    ccg_threshold = ops['settings']['ccg_threshold']
    is_ref, est_contam_rate = CG.refract(spike_clusters, spike_times / ops['fs'],
                                         acg_threshold=acg_threshold,
                                         ccg_threshold=ccg_threshold, device=device)

    # write properties to *.tsv
    stypes = ['ContamPct', 'Amplitude', 'KSLabel']
    ks_labels = [['mua', 'good'][int(r)] for r in is_ref]
    props = [est_contam_rate * 100, template_amplitudes, ks_labels]
    for stype, prop in zip(stypes, props):
        with open(results_dir / f'cluster_{stype}.tsv', 'w') as f:
            f.write(f'cluster_id\t{stype}\n')
            for i, p in enumerate(prop):
                if stype != 'KSLabel':
                    f.write(f'{i}\t{p:.1f}\n')
                else:
                    f.write(f'{i}\t{p}\n')
        if stype == 'KSLabel':
            shutil.copyfile(results_dir / f'cluster_{stype}.tsv',
                            results_dir / 'cluster_group.tsv')

    # params.py
    dtype = "'int16'" if data_dtype is None else f"'{data_dtype}'"
    params = {'n_channels_dat': ops['settings']['n_chan_bin'], 'offset': 0,
              'sample_rate': ops['settings']['fs']}
    if save_preprocessed_copy:
        dat_path = results_dir / 'temp_wh.dat'
        params['dtype'] = "'int16'"
        params['hp_filtered'] = True
        params['dat_path'] = f"'{dat_path.resolve().as_posix()}'"
    else:
        f = ops['settings']['filename']
        if not isinstance(f, list):
            f = [f]
        params['dtype'] = dtype
        params['hp_filtered'] = False
        params['dat_path'] = [Path(p).resolve().as_posix() for p in f]
    with open(results_dir / 'params.py', 'w') as f:
        for key in params.keys():
            if (key == 'dat_path') and skip_dat_path:
                f.write('dat_path = "no_path.bin"\n')
                f.write('# ')
            f.write(f'{key} = {params[key]}\n')

    if save_extra_vars:
        np.save(results_dir / 'Wall.npy', torch.as_tensor(Wall).cpu().numpy())
        np.save(results_dir / 'full_st.npy', st)
        np.save(results_dir / 'full_clu.npy', clu)
        np.save(results_dir / 'full_amp.npy', amplitudes)

    phy_cache_path = Path(results_dir / '.phy')
    if phy_cache_path.is_dir():
        shutil.rmtree(phy_cache_path)

    return results_dir, similar_templates, is_ref, est_contam_rate, kept_spikes
