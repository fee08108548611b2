"""Host-side text formats of the Phy export (no GPU): the per-cluster tables
and params.py, checked against the layouts io.save_to_phy writes
(io.py:414-457) and docs/export_files.rst describes."""
from pathlib import Path

import numpy as np

from kilosort_amd import phy


def test_cluster_tables():
    txt = phy.cluster_table_text("ContamPct", np.array([0.0, 12.345, 100.0]), lambda v: f"{v:.1f}")
    assert txt == "cluster_id\tContamPct\n0\t0.0\n1\t12.3\n2\t100.0\n"
    txt = phy.cluster_table_text("KSLabel", ["mua", "good"], str)
    assert txt == "cluster_id\tKSLabel\n0\tmua\n1\tgood\n"
    assert phy.cluster_table_text("Amplitude", [], str) == "cluster_id\tAmplitude\n"
    assert [c for c, _ in phy.CLUSTER_COLUMNS] == ["ContamPct", "Amplitude", "KSLabel"]


def test_params_from_the_recording(tmp_path):
    txt = phy.params_text(385, 30000.0, "/data/a.bin", tmp_path)
    assert txt.splitlines() == [
        "n_channels_dat = 385", "offset = 0", "sample_rate = 30000.0", "dtype = 'int16'",
        "hp_filtered = False", "dat_path = ['/data/a.bin']"]
    txt = phy.params_text(64, 25000, ["/d/x.bin", "/d/y.bin"], tmp_path, data_dtype="uint16")
    assert "dtype = 'uint16'" in txt.splitlines()
    assert txt.splitlines()[-1] == "dat_path = ['/d/x.bin', '/d/y.bin']"


def test_params_of_a_preprocessed_copy_and_skipped_path(tmp_path):
    txt = phy.params_text(64, 30000.0, "/data/a.bin", tmp_path, data_dtype="float32",
                          save_preprocessed_copy=True)
    lines = txt.splitlines()
    assert lines[3] == "dtype = 'int16'" and lines[4] == "hp_filtered = True"
    assert lines[5] == f"dat_path = '{(Path(tmp_path) / 'temp_wh.dat').resolve().as_posix()}'"
    txt = phy.params_text(64, 30000.0, "/data/a.bin", tmp_path, skip_dat_path=True)
    lines = txt.splitlines()
    assert lines[-2] == 'dat_path = "no_path.bin"'
    assert lines[-1] == "# dat_path = ['/data/a.bin']"
    # the file is valid Python that Phy can exec
    ns = {}
    exec(txt, ns)
    assert ns["dat_path"] == "no_path.bin" and ns["n_channels_dat"] == 64
