"""Phy export of the sorting results, with the per-spike work on the GPU.

``save_to_phy`` has the signature and return values of
``kilosort/io.py:248-469`` and produces the file set that
``docs/export_files.rst`` specifies (names, shapes, dtypes and text formats),
so that Phy and the reference's loaders read it unchanged.

Design: the export is computed first, as an :class:`PhyExport` -- a manifest
of named arrays plus the per-cluster table and the ``params.py`` settings --
and then written by one generic writer.  The computation keeps every large
array on the device until the manifest is written:

* spike positions and amplitudes (``||tF||``) -- one pass over ``tF``
  (``ksm_spike_positions``);
* duplicate removal -- ``ksm_group`` + ``ksm_remove_duplicates``;
* template similarity -- ``ksm_similarity``; template waveforms
  ``Wall @ wPCA`` -- ``ksm_gemm_f32``;
* PC features and their channel indices -- ``ksm_template_sums`` +
  ``ksm_gather_features``;
* refractoriness / contamination -- ``ksm_acg`` autocorrelogram counts.
"""
from __future__ import annotations

import dataclasses
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

# Per-cluster text tables (export_files.rst: cluster_*.tsv): file column name
# and how one value is printed.  cluster_group.tsv repeats the labels.
CLUSTER_COLUMNS = (
    ("ContamPct", lambda v: f"{v:.1f}"),
    ("Amplitude", lambda v: f"{v:.1f}"),
    ("KSLabel", str),
)
LABEL_COPY = "cluster_group.tsv"


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


def cluster_table_text(column, values, fmt):
    """Text of one ``cluster_<column>.tsv``: a header row, then one
    ``<cluster id>\\t<value>`` row per cluster."""
    rows = [f"cluster_id\t{column}"]
    rows += [f"{i}\t{fmt(v)}" for i, v in enumerate(values)]
    return "".join(r + "\n" for r in rows)


def params_text(n_chan_bin, fs, filename, results_dir, data_dtype=None,
                save_preprocessed_copy=False, skip_dat_path=False):
    """Text of Phy's ``params.py`` (``io.py:432-457``): one ``key = value``
    line per setting in the order Phy's template-gui reads them.  With
    ``skip_dat_path`` the real paths are kept only as a comment after a
    placeholder ``dat_path``."""
    if save_preprocessed_copy:
        dat = f"'{(Path(results_dir) / 'temp_wh.dat').resolve().as_posix()}'"
        dtype, filtered = "'int16'", True
    else:
        files = filename if isinstance(filename, list) else [filename]
        dat = [Path(p).resolve().as_posix() for p in files]
        dtype = "'int16'" if data_dtype is None else f"'{data_dtype}'"
        filtered = False
    settings = [("n_channels_dat", n_chan_bin), ("offset", 0), ("sample_rate", fs),
                ("dtype", dtype), ("hp_filtered", filtered), ("dat_path", dat)]
    lines = []
    for key, value in settings:
        line = f"{key} = {value}"
        if key == "dat_path" and skip_dat_path:
            lines.append('dat_path = "no_path.bin"')
            line = "# " + line
        lines.append(line)
    return "".join(l + "\n" for l in lines)


@dataclasses.dataclass
class PhyExport:
    """Everything ``save_to_phy`` writes: ``arrays`` maps file names to
    arrays (``.npy``), ``clusters`` the per-cluster columns, ``params`` the
    ``params.py`` text."""
    arrays: dict
    clusters: dict
    params: str
    is_ref: np.ndarray
    contam: np.ndarray
    kept: np.ndarray
    similar: np.ndarray

    def write(self, results_dir):
        results_dir = Path(results_dir)
        for name, arr in self.arrays.items():
            np.save(results_dir / name, arr)
        for column, fmt in CLUSTER_COLUMNS:
            path = results_dir / f"cluster_{column}.tsv"
            path.write_text(cluster_table_text(column, self.clusters[column], fmt))
        shutil.copyfile(results_dir / "cluster_KSLabel.tsv", results_dir / LABEL_COPY)
        (results_dir / "params.py").write_text(self.params)
        # Phy's cache of an earlier export in the same folder is stale now
        cache = results_dir / ".phy"
        if cache.is_dir():
            shutil.rmtree(cache)


def _host(t):
    return t.cpu().numpy() if torch.is_tensor(t) else np.asarray(t)


def build_export(st, clu, tF, Wall, probe, ops, imin, results_dir, data_dtype=None,
                 save_extra_vars=False, save_preprocessed_copy=False, skip_dat_path=False,
                 device=None):
    """Computes the Phy export of one sorting (nothing is written)."""
    arrays = {}
    # probe and whitening (export_files.rst: channel_*, whitening_mat*)
    Wrot = torch.as_tensor(ops['Wrot'])
    eye = torch.eye(Wrot.shape[0], dtype=Wrot.dtype, device=Wrot.device)
    arrays['channel_map.npy'] = probe['chanMap']
    arrays['channel_positions.npy'] = np.stack((probe['xc'], probe['yc']), axis=-1)
    arrays['channel_shanks.npy'] = probe['kcoords']
    arrays['whitening_mat_dat.npy'] = _host(Wrot)
    arrays['whitening_mat.npy'] = _host(Wrot)
    arrays['whitening_mat_inv.npy'] = _host(torch.linalg.inv(Wrot + 1e-5 * eye))

    # per spike: sample index of the peak, detection template, position, ||tF||
    times_all = st[:, 0].astype(np.int64) + imin
    detect_all = st[:, 1].astype(np.int32)
    xs, ys, norms = PP.positions_and_amplitudes(st, tF, ops, device)
    pos_dtype = np.result_type(np.asarray(ops['xc']).dtype, np.float32)
    pos_all = np.stack([_host(xs), _host(ys)], 1).astype(pos_dtype)
    amp_all = _host(norms)

    # artefact duplicates out; every per-spike file holds the kept spikes
    times, labels, kept = PP.remove_duplicates(times_all, clu, dt=ops['duplicate_spike_bins'],
                                               device=device)
    detect = detect_all[kept]
    arrays['spike_times.npy'] = times
    arrays['spike_templates.npy'] = labels
    arrays['spike_clusters.npy'] = labels
    arrays['spike_positions.npy'] = pos_all[kept]
    arrays['spike_detection_templates.npy'] = detect
    arrays['amplitudes.npy'] = amp_all[kept]
    arrays['kept_spikes.npy'] = kept

    # per cluster ("template" in Phy's words): similarity, waveforms, norms
    similar = TM.similarity_matrix(Wall, ops['wPCA'], ops['nt'], device)[0]
    waves = template_waveforms(Wall, ops['wPCA'], device)
    K, nt_, C = waves.shape
    arrays['similar_templates.npy'] = similar
    arrays['templates.npy'] = waves
    arrays['templates_ind.npy'] = np.broadcast_to(np.arange(C), (K, C)).copy()
    cluster_norm = torch.linalg.vector_norm(torch.as_tensor(Wall).cpu(), dim=(-2, -1)).numpy()

    if save_extra_vars:
        arrays['tF.npy'] = _host(tF)            # before the feature rewrite below
    tF_kept = tF[torch.from_numpy(kept).to(tF.device)]
    pcf, pcf_ind = PP.make_pc_features(ops, detect, labels, tF_kept, device=device)
    arrays['pc_features.npy'] = _host(pcf)
    arrays['pc_feature_ind.npy'] = pcf_ind

    is_ref, contam = CG.refract(labels, times / ops['fs'],
                                acg_threshold=ops['settings']['acg_threshold'],
                                ccg_threshold=ops['settings']['ccg_threshold'], device=device)
    clusters = {"ContamPct": contam * 100, "Amplitude": cluster_norm,
                "KSLabel": [("mua", "good")[int(r)] for r in is_ref]}

    if save_extra_vars:
        arrays['Wall.npy'] = _host(torch.as_tensor(Wall))
        arrays['full_st.npy'] = st
        arrays['full_clu.npy'] = clu
        arrays['full_amp.npy'] = amp_all

    settings = ops['settings']
    params = params_text(settings['n_chan_bin'], settings['fs'], settings['filename'],
                         results_dir, data_dtype, save_preprocessed_copy, skip_dat_path)
    return PhyExport(arrays, clusters, params, is_ref, contam, kept, similar)


def save_to_phy(st, clu, tF, Wall, probe, ops, imin, results_dir=None, data_dtype=None,
                save_extra_vars=False, save_preprocessed_copy=False, skip_dat_path=False,
                device=None):
    """``io.save_to_phy``: writes the Phy files of the sorting into
    ``results_dir`` and returns (results_dir, similar_templates, is_ref,
    est_contam_rate, kept_spikes)."""
    device = PP._device_of(ops, device)
    results_dir = Path(ops['data_dir'].joinpath('kilosort4') if results_dir is None
                       else results_dir)
    results_dir.mkdir(exist_ok=True)
    export = build_export(st, clu, tF, Wall, probe, ops, imin, results_dir, data_dtype,
                          save_extra_vars, save_preprocessed_copy, skip_dat_path, device)
    export.write(results_dir)
    return results_dir, export.similar, export.is_ref, export.contam, export.kept
