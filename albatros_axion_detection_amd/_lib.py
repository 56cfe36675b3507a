"""ctypes binding of libalbatros_hip.so (the C ABI declared in include/albatros_hip.h).

The library is loaded from the package directory.  There is no fallback: if
the library is missing every entry point raises, so a GPU run can never
silently compute somewhere else.
"""
import ctypes
import os

import numpy as np

_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_DIR, "libalbatros_hip.so")

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int
_f32 = ctypes.c_float
_f64 = ctypes.c_double

# name -> (restype, argtypes); mirrors include/albatros_hip.h
SIGNATURES = {
    "alb_last_error": (ctypes.c_char_p, []),
    "alb_version": (_i32, []),
    "alb_stream_synchronize": (_i32, [_vp]),
    "alb_fill_zero": (_i32, [_vp, _i64, _vp]),
    "alb_prepare_transform": (_i32, [_i64]),
    "alb_prepare_repfb": (_i32, [_i64]),
    "alb_unpack_4bit_float": (_i32, [_vp, _vp, _vp, _i64, _i64, _vp, _i64, _i64, _vp]),
    "alb_unpack_1bit_float": (_i32, [_vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64, _vp]),
    "alb_sortpols": (_i32, [_vp, _vp, _vp, _i64, _i64, _i64, _i64, _i32, _i64, _i64, _vp]),
    "alb_hist": (_i32, [_vp, _vp, _i64, _i64, _i64, _i32, _i32, _vp]),
    "alb_make_continuous": (_i32, [_vp, _i64, _vp, _i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp]),
    "alb_dropped_packets": (_i32, [_vp, _vp, _i64, _i64, _i64, _i32, _vp, _vp]),
    "alb_xcorr_4bit_pair": (_i32, [_vp, _vp, _i64, _i64, _vp, _i32, _vp, _vp]),
    "alb_match_rows": (_i32, [_vp, _vp, _i64, _i64, _i64, _i64, _vp, _vp, _vp]),
    "alb_xcorr_1bit_pair": (_i32, [_vp, _vp, _i64, _i64, _i64, _vp, _i32, _vp, _vp]),
    "alb_xcorr_4bit_allpairs": (_i32, [_vp, _i32, _i64, _i64, _i64, _vp, _vp]),
    "alb_xcorr_1bit_allpairs": (_i32, [_vp, _i32, _i64, _i64, _i64, _i64, _vp, _vp]),
    "alb_xcorr_all_ant": (_i32, [_vp, _i64, _i64, _i64, _vp, _f32, _vp]),
    "alb_xcorr_all_ant_prec": (_i32, [_vp, _i64, _i64, _i64, _vp, _f32, _i32, _vp]),
    "alb_xcorr_all_ant_strided": (_i32, [_vp, _i64, _i64, _i64, _i64, _i64, _i64, _i64, _vp, _f32, _i32, _vp]),
    "alb_xcorr_all_ant_f64acc": (_i32, [_vp, _i64, _i64, _i64, _i64, _i64, _i64, _vp, _i64, _i64, _i64, _vp]),
    "alb_cgemm_strided_batched": (_i32, [_vp, _vp, _vp, _i64, _i64, _i64, _i64, _vp]),
    "alb_matft_filter": (_i32, [_vp, _i32, _i64, _i64, _f64, _vp, _vp, _vp]),
    "alb_filter_from_matft": (_i32, [_vp, _i64, _f64, _vp, _vp]),
    "alb_irfft_rows": (_i32, [_vp, _i64, _vp, _i64, _i64, _i64, _vp]),
    "alb_rfft_rows": (_i32, [_vp, _i64, _vp, _i64, _i64, _i64, _i64, _i64, _vp]),
    "alb_fft_rows": (_i32, [_vp, _i64, _vp, _i64, _i64, _i64, _i32, _vp]),
    "alb_fft_large": (_i32, [_vp, _vp, _vp, _i64, _i64, _i32, _vp]),
    "alb_lag_xcorr": (_i32, [_vp, _vp, _i64, _i64, _i64, _i64, _vp, _vp, _vp]),
    "alb_apply_delay": (_i32, [_vp, _vp, _vp, _i64, _i64, _vp, _vp]),
    "alb_pfb_small": (_i32, [_vp, _i64, _i64, _vp, _i32, _i64, _i64, _vp, _i64, _vp]),
    "alb_ipfb_columns": (_i32, [_vp, _i64, _i64, _i64, _i64, _i64, _vp, _i64, _i32, _vp, _i64, _i64, _vp]),
    "alb_ipfb_columns_any": (_i32, [_vp, _i64, _i64, _i64, _i64, _i64, _vp, _i64, _i32, _vp, _i64, _i64, _vp, _vp,
                                     _i64, _vp]),
    "alb_ipfb_rows_packed4": (_i32, [_vp, _vp, _vp, _i64, _i64, _i32, _i64, _i64, _vp, _vp]),
    "alb_ipfb_rows_packed4_mfma": (_i32, [_vp, _vp, _vp, _i64, _i64, _i32, _i64, _i64, _vp, _vp, _vp]),
    "alb_ipfb_rows_packed4_blocked": (_i32, [_vp, _vp, _vp, _i64, _i64, _i32, _i64, _i64, _vp, _vp]),
    "alb_ipfb_rows_coeff_bytes": (_i64, [_i64]),
    "alb_ipfb_rows_coeff": (_i32, [_vp, _i64, _vp, _vp]),
    "alb_repfb": (_i32, [_vp, _vp, _i32, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _i64, _i64, _i64, _vp]),
    "alb_repfb_workspace_bytes": (_i64, [_i64, _i64]),
    "alb_repfb_noalloc": (_i32, [_vp, _vp, _i32, _i64, _i64, _i64, _i64, _vp, _i64, _i64, _i64, _i64, _vp]),
    "alb_repfb_noalloc_max_bytes": (_i64, []),
    "alb_repfb_stage1": (_i32, [_vp, _vp, _i32, _i64, _i64, _vp, _vp]),
    "alb_repfb_stage2": (_i32, [_vp, _i64, _i64, _i64, _i64, _vp, _i64, _i64, _i64, _i64, _vp]),
    "alb_repfb_stage2_inputs": (_i32, [_vp, _i64, _i64, _i64, _i64, _i64, _i64, _vp, _i64, _i64, _i64, _i64, _i64,
                                       _vp]),
    "alb_ipfb_columns_repfb": (_i32, [_vp, _i64, _i64, _i64, _i64, _vp, _i64, _i32, _vp, _i64, _i64, _vp, _i32, _i64,
                                      _i64, _vp, _vp]),
    "alb_ipfb_chan_fft_packed4": (_i32, [_vp, _vp, _vp, _i64, _i64, _i32, _i64, _i64, _vp, _vp]),
    "alb_ipfb_rows_spectral": (_i32, [_vp, _vp, _i64, _i64, _i64, _vp, _vp]),
    "alb_ipfb_columns_repfb_split": (_i32, [_vp, _i64, _i64, _i64, _i64, _vp, _i64, _i32, _i64, _i64, _vp, _i32, _i64,
                                            _i64, _vp, _vp, _i64, _i64, _vp]),
    "alb_ipfb_columns_repfb_spectral_split": (_i32, [_vp, _i64, _i64, _vp, _i64, _i32, _i64, _i64, _vp, _i32, _i64,
                                                     _i64, _vp, _vp, _i64, _i64, _vp]),
    "alb_repfb_stage2_split": (_i32, [_vp, _i64, _vp, _i64, _i64, _i64, _i64, _i64, _i64, _vp, _i64, _i64, _i64, _i64,
                                      _i64, _i64, _i64, _i64, _i32, _vp]),
    "alb_repfb_split_slot_bytes": (_i64, [_i64]),
    "alb_repfb_split_blocks": (_i32, [_i64, _i64, _i64, _i32, _i64, _vp, _vp, _vp]),
    "alb_repfb_split_slot": (_i64, [_i64, _i64, _i64, _i64, _i32, _i64]),
    "alb_ipfb_columns_repfb_spectral": (_i32, [_vp, _i64, _i64, _vp, _i64, _i32, _vp, _i64, _i64, _vp, _i32, _i64,
                                               _i64, _vp, _vp]),
    "alb_repfb_interior": (_i32, [_i64, _i64, _i64, _i32, _i64, _vp, _vp]),
    "alb_repfb_stage1_list": (_i32, [_vp, _vp, _i32, _vp, _i64, _i64, _vp, _vp]),
    "alb_tune_set": (_i32, [ctypes.c_char_p, _i64]),
    "alb_tune_get": (_i64, [ctypes.c_char_p]),
    "alb_build_rowmap": (_i32, [_vp, _i64, _i64, _i64, _i64, _vp, _vp]),
}

_lib = None


def load():
    """Load the library (once).  Raises ImportError if it has not been built.

    torch carries its own HIP runtime; the library uses the system one.  With
    a GPU present, torch's device state is initialised before the library is
    loaded: loaded the other way round, the library's launches later found no
    device."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                "libalbatros_hip.so is not built; run `python -m albatros_axion_detection_amd.build` "
                "(there is no CPU fallback)")
        t = torch()
        if t.cuda.is_available():
            t.cuda.init()
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class AlbError(RuntimeError):
    """HIP runtime failure inside libalbatros_hip."""


def call(name, *args):
    """Call an entry point and turn a non-zero status into an exception
    (1 -> ValueError, 3 -> NotImplementedError, 2 -> AlbError)."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.alb_last_error().decode(errors="replace")
        if rc == 1:
            raise ValueError(msg)
        if rc == 3:
            raise NotImplementedError(msg)
        raise AlbError("%s failed: %s" % (name, msg))


# ----------------------------------------------------------------- torch glue
def torch():
    import torch as _t
    return _t


def device():
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("albatros_axion_detection_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return t.device("cuda", t.cuda.current_device())


def stream():
    """hipStream_t of torch's current stream."""
    return _vp(torch().cuda.current_stream().cuda_stream)


def ptr(t):
    return _vp(t.data_ptr()) if t is not None else _vp(0)


def to_device(a, dtype=None):
    """numpy array / torch tensor -> contiguous device tensor (copied if needed)."""
    t = torch()
    if isinstance(a, t.Tensor):
        x = a
    else:
        x = t.from_numpy(np.ascontiguousarray(a))
    if dtype is not None and x.dtype != dtype:
        x = x.to(dtype)
    dev = device()
    if x.device != dev:
        x = x.to(dev, non_blocking=False)
    return x.contiguous()


def is_numpy(a):
    return isinstance(a, np.ndarray)


def host(x):
    return x.detach().cpu().numpy()
