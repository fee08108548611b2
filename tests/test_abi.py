"""The C-ABI library loads and exports exactly what include/ksm.h declares.

CPU only: no compute call is made (no GPU here).
"""
import ctypes
import os
import re

import pytest

from kilosort_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ksm.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in re.finditer(r"\b(?:int|size_t|const char \*)\s*(ksm_\w+)\s*\(([^)]*)\)\s*;", text):
        args = m.group(2).strip()
        n = 0 if args in ("", "void") else len(args.split(","))
        out[m.group(1)] = n
    return out


def test_header_parses():
    fns = header_functions()
    assert "ksm_peel" in fns and "ksm_project" in fns and "ksm_features" in fns
    assert len(fns) >= 15


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libksm.so not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in header_functions():
        assert hasattr(lib, name), name


def test_ctypes_signatures_match_header():
    fns = header_functions()
    assert set(fns) == set(_lib.SIGNATURES), set(fns) ^ set(_lib.SIGNATURES)
    for name, n in fns.items():
        assert len(_lib.SIGNATURES[name][1]) == n, name


def test_status_strings_and_workspace_queries():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libksm.so not built")
    lib = _lib.load()
    assert lib.ksm_version() >= 100
    assert lib.ksm_status_string(0) == b"ok"
    assert lib.ksm_status_string(2) == b"spike capacity exceeded"
    assert lib.ksm_project_workspace(383, 60122, 700, 6, 61, 0) > 383 * 6 * 60122 * 4
    assert lib.ksm_peel_workspace(700, 60122, 100) > 60122 * 11 * 8
    assert lib.ksm_project_workspace(0, 10, 10, 6, 61, 0) == 0


def test_invalid_arguments_are_rejected_without_gpu_work():
    """Argument validation happens on the host, before any launch."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libksm.so not built")
    lib = _lib.load()
    st = lib.ksm_template_norms(None, 0, 6, 10, None, None)
    assert st == _lib.KSM_ERR_INVALID
    assert b"ksm_template_norms" in lib.ksm_last_error()
    st = lib.ksm_peel(None, 0, None, 0, None, None, None, None, None, None, None, None, 1, 6, 4,
                      10, 61, 8.0, 100, 0, None, None, None, None, 10, None, None, 0, None)
    assert st == _lib.KSM_ERR_INVALID
    with pytest.raises(ValueError):
        _lib.check(st, "ksm_peel")


def test_cpu_device_is_refused():
    import torch
    from kilosort_amd import template_matching as tm
    with pytest.raises(RuntimeError):
        tm.run_matching({"Th_learned": 8, "nt": 61, "max_peels": 1,
                         "wPCA": torch.zeros(6, 61)}, torch.zeros(4, 200),
                        torch.zeros(2, 6, 4), None, device=torch.device("cpu"))
