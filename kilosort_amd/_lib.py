"""ctypes binding of libksm.so (the C ABI declared in include/ksm.h).

The library is built in-tree by ``python -m kilosort_amd.build``.  There is no
fallback: if the library is missing or no ROCm GPU is present, every call
raises.
"""
from __future__ import annotations

import ctypes
import os

import torch

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libksm.so")

KSM_OK = 0
KSM_ERR_INVALID = 1
KSM_ERR_CAPACITY = 2
KSM_ERR_HIP = 3
KSM_ERR_WORKSPACE = 4
KSM_ERR_TIES = 5

PREC_FP32 = 0
PREC_BF16 = 1
PREC_BF16X3 = 2
PREC_FP8 = 3

PEEL_RESCORE = 1

COPY_D2D = 0
COPY_D2H = 1

DTYPE_F32 = 0
DTYPE_F64 = 1

RAW_I16 = 0
RAW_U16 = 1
RAW_F32 = 2
RAW_F32_CM = 3

FILTER_INVERT = 1
FILTER_CAR = 2
FILTER_SKIP = 4
FILTER_SCALE = 8
FILTER_SHIFT = 16
FILTER_EDGE_FIRST = 32
FILTER_EDGE_LAST = 64
FILTER_WHITEN = 128

_p = ctypes.c_void_p
_i = ctypes.c_int
_f = ctypes.c_float
_sz = ctypes.c_size_t
_d = ctypes.c_double
_i64 = ctypes.c_int64

# name -> (restype, argtypes); must match include/ksm.h
SIGNATURES = {
    "ksm_version": (_i, []),
    "ksm_status_string": (ctypes.c_char_p, [_i]),
    "ksm_last_error": (ctypes.c_char_p, []),
    "ksm_prepare_extract": (_i, [_p, _i, _i, _i, _p, _i, _p, _p, _p]),
    "ksm_template_norms": (_i, [_p, _i, _i, _i, _p, _p]),
    "ksm_prepare_matching_workspace": (_sz, [_i, _i, _i, _i]),
    "ksm_prepare_matching": (_i, [_p, _p, _i, _i, _i, _i, _p, _i, _p, _sz, _p]),
    "ksm_ctc_rows": (_i, [_p, _i, _i, _i, _p, _p, _p]),
    "ksm_template_order": (_i, [_p, _i, _i, _i, _p, _p, _p]),
    "ksm_gram_ranges": (_i, [_p, _p, _p, _i, _i, _p, _p, _p]),
    "ksm_gram_compact": (_i, [_p, _i, _p, _p, _p, _i, _i, _p, _p]),
    "ksm_project_workspace": (_sz, [_i, _i, _i, _i, _i, _i]),
    "ksm_project": (_i, [_p, _i, _p, _p, _p, _i, _i, _i, _i, _i, _p, _i, _i, _p, _sz, _p]),
    "ksm_bound_bytes": (_sz, [_i, _i, _i, _i, _i]),
    "ksm_project_bounded": (_i, [_p, _i, _p, _p, _p, _i, _i, _i, _i, _i, _p, _i, _i, _i, _p,
                                 _sz, _p, _sz, _p]),
    "ksm_peel_workspace": (_sz, [_i, _i, _i]),
    "ksm_peel": (_i, [_p, _i, _p, _i, _p, _p, _p, _p, _p, _p, _p, _p, _i, _i, _i, _i, _i, _f,
                      _i, _i, _p, _p, _p, _p, _i, _p, _p, _sz, _p]),
    "ksm_peel_static": (_i, [_p, _i, _p, _i, _p, _p, _p, _p, _p, _p, _p, _p, _i, _i, _i, _i, _i,
                             _f, _i, _i, _p, _p, _p, _p, _i, _p, _p]),
    "ksm_peel_static_capacity": (_sz, []),
    "ksm_peel_range": (_i, [_p, _i, _p, _i, _p, _p, _p, _p, _p, _p, _p, _p, _i, _i, _i, _i, _i,
                            _f, _i, _i, _p, _p, _p, _p, _i, _p, _p, _sz, _i, _i, _p]),
    "ksm_features_counted": (_i, [_p, _i, _p, _p, _p, _p, _i, _p, _p, _p, _i, _i, _i, _i, _i, _p,
                                  _p]),
    "ksm_features": (_i, [_p, _i, _p, _p, _p, _i, _p, _p, _p, _i, _i, _i, _i, _i, _p, _p]),
    "ksm_graph_begin": (_i, [_p]),
    "ksm_graph_end": (_i, [_p, ctypes.POINTER(ctypes.c_void_p)]),
    "ksm_graph_launch": (_i, [_p, _p]),
    "ksm_graph_destroy": (_i, [_p]),
    "ksm_copy_async": (_i, [_p, _p, _sz, _i, _p]),
    "ksm_residual_workspace": (_sz, [_i, _i, _i, _i, _i, _i, _i]),
    "ksm_residual": (_i, [_p, _i, _p, _p, _p, _p, _p, _i, _i, _i, _i, _i, _i, _p, _i, _p, _i,
                          _p, _sz, _p]),
    "ksm_residual_backward": (_i, [_p, _i, _p, _p, _p, _p, _p, _p, _i, _i, _i, _i, _i, _i, _d,
                                   _p, _p, _p, _i, _p, _sz, _p]),
    "ksm_filter_workspace": (_sz, [_i, _i, _i, _i]),
    "ksm_filter_batch": (_i, [_p, _i, _i64, _i, _i, _p, _i, _i, _i, _f, _f, _p, _i, _f, _p, _p,
                              _i, _p, _sz, _p]),
    "ksm_gemm_f32": (_i, [_p, _i, _p, _i, _p, _i, _i, _i, _i, _p]),
    "ksm_similarity_workspace": (_sz, [_i, _i, _i, _i]),
    "ksm_similarity": (_i, [_p, _p, _i, _i, _i, _i, _p, _p, _p, _sz, _p]),
    "ksm_align_workspace": (_sz, [_i, _i, _i, _i, _i]),
    "ksm_align_templates": (_i, [_p, _p, _p, _i, _i, _i, _i, _i, _p, _p, _p, _sz, _p]),
    "ksm_snippets_workspace": (_sz, [_i, _i]),
    "ksm_snippets": (_i, [_p, _i, _i, _i, _i, _i, _f, _i, _i, _i, _i, _i, _p, _p, _p, _p, _sz,
                          _p]),
    "ksm_spike_positions": (_i, [_p, _p, ctypes.c_longlong, _i, _i, _p, _p, _p, _i, _p, _p, _p,
                                 _p, _p, _p]),
    "ksm_group_workspace": (_sz, [ctypes.c_longlong, _i]),
    "ksm_group": (_i, [_p, ctypes.c_longlong, _i, _p, _p, _p, _sz, _p]),
    "ksm_remove_duplicates": (_i, [_p, ctypes.c_longlong, _i, _p, _p, ctypes.c_longlong, _p,
                                   _p]),
    "ksm_acg": (_i, [_p, ctypes.c_longlong, _i, _p, _p, _p, _i, ctypes.c_double, _p, _p]),
    "ksm_ccg": (_i, [_p, ctypes.c_longlong, _p, ctypes.c_longlong, _i, ctypes.c_double, _p,
                     _p]),
    "ksm_pc_map": (_i, [_p, _p, ctypes.c_longlong, _i, _i, _p, _p]),
    "ksm_template_sums": (_i, [_p, _p, _p, ctypes.c_longlong, _i, _i, _i, _i, _p, _p, _p, _p,
                               _sz, _p]),
    "ksm_template_sums_workspace": (_sz, [ctypes.c_longlong, _i]),
    "ksm_pc_channel_rank": (_i, [_p, _p, _p, _p, _i, _i, _i, _i, _i, _p, _p, _p, _p]),
    "ksm_gather_features": (_i, [_p, _p, ctypes.c_longlong, _i, _i, _p, _p, _i, _p]),
    "ksm_detect_workspace": (_sz, [_i, _i, _i, _i]),
    "ksm_detect": (_i, [_p, _i, _i, _i, _p, _i, _i, _p, _i, _p, _i, _p, _i, _i, _i, _f, _p, _p,
                        _i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _sz, _p]),
    "ksm_profile_enable": (_i, [_i]),
    "ksm_profile_reset": (_i, []),
    "ksm_profile_query": (_i, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                               ctypes.POINTER(ctypes.c_int)]),
    "ksm_profile_launch": (_i, [_i, ctypes.c_char_p, _i, ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_double)]),
}

_lib = None


class KsmError(RuntimeError):
    """A libksm call returned a non-zero status."""

    def __init__(self, status, message):
        super().__init__(message)
        self.status = status


def load():
    """Load libksm.so (once) and declare its signatures."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m kilosort_amd.build` "
                "(there is no CPU fallback)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(status: int, what: str):
    if status == KSM_OK:
        return
    lib = load()
    detail = lib.ksm_last_error().decode(errors="replace")
    text = lib.ksm_status_string(status).decode(errors="replace")
    msg = f"{what}: {text}" + (f" ({detail})" if detail else "")
    if status == KSM_ERR_INVALID:
        raise ValueError(msg)
    raise KsmError(status, msg)


def require_gpu(device) -> torch.device:
    """The matching path runs on a ROCm GPU only."""
    device = torch.device(device) if device is not None else torch.device("cuda")
    if device.type != "cuda":
        raise RuntimeError(f"kilosort_amd runs on a ROCm GPU only (got device {device})")
    if not torch.cuda.is_available():
        raise RuntimeError("kilosort_amd needs a ROCm GPU; torch.cuda.is_available() is False")
    return device


def stream_of(device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t) -> int:
    return 0 if t is None else t.data_ptr()


def profile_query(name: str):
    """(total_ms, launches) of one kernel recorded while profiling was on."""
    lib = load()
    ms = ctypes.c_double(0.0)
    n = ctypes.c_int(0)
    check(lib.ksm_profile_query(name.encode(), ctypes.byref(ms), ctypes.byref(n)),
          "ksm_profile_query")
    return ms.value, n.value


def profile_timeline():
    """Every recorded launch in launch order: (kernel name, start ms relative
    to the first launch, duration ms)."""
    lib = load()
    out = []
    name = ctypes.create_string_buffer(64)
    t0 = ctypes.c_double(0.0)
    d = ctypes.c_double(0.0)
    i = 0
    while True:
        rc = lib.ksm_profile_launch(i, name, 64, ctypes.byref(t0), ctypes.byref(d))
        if rc != 0:
            break
        out.append((name.value.decode(), t0.value, d.value))
        i += 1
    return out
