"""ctypes binding of ``libpsx.so`` (declarations: include/psx.h).

The library is built in-tree by ``pyspectral_amd.build``; nothing here falls back to another
implementation: a missing or unloadable library raises.
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_double, c_int, c_longlong, c_void_p
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libpsx.so"


class FlowParams(ctypes.Structure):
    _fields_ = [("n", c_int), ("dim", c_int), ("form", c_int), ("fp64", c_int),
                ("nu", c_double), ("alpha", c_double), ("ka2", c_double), ("klo2", c_double),
                ("khi2", c_double), ("dealias", c_int)]


class Layout(ctypes.Structure):
    _fields_ = [("state_elems", c_longlong), ("work_elems", c_longlong), ("nfields", c_int),
                ("ldw", c_int), ("jm", c_int), ("four_step", c_int),
                ("e_line", c_int), ("e_spec", c_int), ("e_outer", c_int), ("e_mid", c_int),
                ("t_line", c_int), ("t_spec", c_int), ("t_outer", c_int), ("t_mid", c_int),
                ("state_pitch", c_int)]


class RunParams(ctypes.Structure):
    _fields_ = [("t0", c_double), ("t_end", c_double), ("dt_fixed", c_double), ("cfl", c_double),
                ("dt_max", c_double), ("diag_dt", c_double), ("power", c_double),
                ("amp0", c_double), ("kmax", c_longlong), ("cfl_every", c_int),
                ("cap_rec", c_int), ("cap_step", c_int), ("record_start", c_int),
                ("next_diag", c_double), ("last_rec_t", c_double)]


class RunState(ctypes.Structure):
    _fields_ = [("t", c_double), ("amp", c_double), ("cur_dt", c_double),
                ("next_diag", c_double), ("k", c_longlong), ("nrec", c_longlong),
                ("done", c_int)]


class LesArgs(ctypes.Structure):
    _fields_ = [("inp", c_void_p), ("in2", c_void_p), ("out", c_void_p), ("n", c_longlong),
                ("L", c_int), ("L2", c_int), ("pitch", c_int), ("scale", c_double),
                ("scale2", c_double), ("scale3", c_double)]


class ConvArgs(ctypes.Structure):
    _fields_ = [("x", c_void_p), ("w", c_void_p), ("bias", c_void_p), ("mask", c_void_p),
                ("y", c_void_p), ("yf", c_void_p), ("n", c_int), ("nb", c_int), ("relu", c_int)]


class AdamArgs(ctypes.Structure):
    _fields_ = [("w", c_void_p), ("m", c_void_p), ("v", c_void_p), ("g", c_void_p),
                ("wb", c_void_p), ("wt", c_void_p), ("n", c_longlong), ("lr", c_double),
                ("b1", c_double), ("b2", c_double), ("eps", c_double), ("c1", c_double),
                ("c2", c_double), ("is_w", c_int)]


class SgsDesc(ctypes.Structure):
    _fields_ = [("model", c_int), ("src", c_void_p), ("full", c_void_p), ("grad", c_void_p),
                ("tau", c_void_p), ("x", c_void_p), ("w", c_void_p * 6), ("b", c_void_p * 6),
                ("act", c_void_p * 5), ("out", c_void_p), ("g_scale", c_double),
                ("tau_scale", c_double), ("smag", c_double), ("backscatter", c_double),
                ("shell_gain", c_void_p)]


SGS_MODELS = {"smagorinsky": 1, "net": 2}  # PSX_SGS_*

LES_OPS = {"expand": 0, "gradpack": 1, "products": 2, "truncate": 3, "pairup": 4,
           "features": 5, "smagorinsky": 6, "tau_from_net": 7, "target": 8, "source": 9,
           "sumsq": 10}  # PSX_LES_*


# name -> (restype, argtypes); the exported C ABI
SIGNATURES = {
    "psx_query_layout": (c_int, [POINTER(FlowParams), POINTER(Layout)]),
    "psx_create": (c_int, [POINTER(FlowParams), c_void_p, c_void_p, c_void_p,
                           POINTER(c_void_p)]),
    "psx_destroy": (c_int, [c_void_p]),
    "psx_prepare": (c_int, [c_void_p, c_void_p]),
    "psx_set_step": (c_int, [c_void_p, c_double, c_double, c_void_p]),
    "psx_set_phase_shift": (c_int, [c_void_p, POINTER(c_double), POINTER(c_double), c_void_p]),
    "psx_step": (c_int, [c_void_p, c_void_p]),
    "psx_step_graph": (c_int, [c_void_p, c_void_p]),
    "psx_velocity_max": (c_int, [c_void_p, POINTER(c_double), c_void_p]),
    "psx_reduce": (c_int, [c_void_p, c_void_p, c_int, c_double, c_double, POINTER(c_double),
                           c_void_p]),
    "psx_spectrum": (c_int, [c_void_p, c_void_p, c_int, POINTER(c_double), c_int, c_void_p]),
    "psx_timing_begin": (c_int, [c_void_p]),
    "psx_timing_end": (c_int, [c_void_p, c_void_p, POINTER(c_double), POINTER(c_int)]),
    "psx_run_begin": (c_int, [c_void_p, POINTER(RunParams), c_void_p]),
    "psx_run_chunk": (c_int, [c_void_p, c_int, c_int, c_void_p]),
    "psx_run_peek": (c_int, [c_void_p, c_int, POINTER(RunState)]),
    "psx_run_poll": (c_int, [c_void_p, POINTER(RunState), c_void_p, c_void_p, c_void_p]),
    "psx_run_end": (c_int, [c_void_p, c_void_p]),
    "psx_ens_variants": (c_int, [c_int, POINTER(c_int), c_int]),
    "psx_ens_query_layout": (c_int, [POINTER(FlowParams), c_int, POINTER(Layout)]),
    "psx_ens_create": (c_int, [c_int, POINTER(FlowParams), POINTER(c_int), POINTER(c_void_p),
                               POINTER(c_void_p), POINTER(c_void_p), POINTER(c_void_p)]),
    "psx_ens_destroy": (c_int, [c_void_p]),
    "psx_ens_set_grid": (c_int, [c_void_p, c_int]),
    "psx_ens_prepare": (c_int, [c_void_p, c_void_p]),
    "psx_ens_set_step": (c_int, [c_void_p, POINTER(c_double), POINTER(c_double), c_void_p]),
    "psx_ens_step": (c_int, [c_void_p, c_void_p]),
    "psx_ens_step_graph": (c_int, [c_void_p, c_void_p]),
    "psx_ens_timing_begin": (c_int, [c_void_p]),
    "psx_ens_timing_end": (c_int, [c_void_p, c_void_p, POINTER(c_double), POINTER(c_int)]),
    "psx_ens_run_begin": (c_int, [c_void_p, POINTER(RunParams), c_void_p]),
    "psx_ens_run_chunk": (c_int, [c_void_p, c_int, c_int, c_void_p]),
    "psx_ens_run_peek": (c_int, [c_void_p, c_int, POINTER(RunState)]),
    "psx_ens_run_poll": (c_int, [c_void_p, POINTER(RunState), c_void_p, c_void_p, c_void_p]),
    "psx_ens_run_end": (c_int, [c_void_p, c_void_p]),
    "psx_set_source": (c_int, [c_void_p, c_void_p]),
    "psx_set_sgs": (c_int, [c_void_p, POINTER(SgsDesc)]),
    "psx_sgs_update": (c_int, [c_void_p, c_void_p]),
    "psx_fft3": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p]),
    "psx_les_op": (c_int, [c_int, POINTER(LesArgs), c_void_p]),
    "psx_cnn_conv": (c_int, [POINTER(ConvArgs), c_void_p]),
    "psx_cnn_wgrad": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p]),
    "psx_cnn_wgrad_noalloc": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int,
                                      c_void_p]),
    "psx_cnn_wgrad_scratch_bytes": (ctypes.c_ulong, []),
    "psx_cnn_mse": (c_int, [c_void_p, c_void_p, c_int, c_int, c_longlong, c_void_p, c_void_p,
                            c_void_p]),
    "psx_cnn_adam": (c_int, [POINTER(AdamArgs), c_void_p]),
    "psx_fft_lines": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p]),
    "psx_fft_twiddles": (c_int, [c_int, c_int, c_void_p, c_int, c_void_p, c_void_p]),
    "psx_last_error": (c_char_p, []),
    "psx_device_info": (c_int, [c_char_p, c_int]),
}

KINDS = ("line", "outer_fwd", "spec", "outer_inv", "mid_fwd", "mid_inv", "ctrl", "sgs")  # PSX_KIND_*

_LIB: ctypes.CDLL | None = None


def load() -> ctypes.CDLL:
    """Load ``libpsx.so`` once and declare its signatures."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m pyspectral_amd.build`")
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
    return _LIB


def check(rc: int) -> None:
    if rc != 0:
        msg = load().psx_last_error()
        raise RuntimeError((msg or b"psx error").decode())


def stream_ptr() -> c_void_p:
    """The current torch (HIP) stream as a ``hipStream_t``."""
    import torch

    return c_void_p(torch.cuda.current_stream().cuda_stream)


def device_info() -> tuple[bool, str]:
    buf = ctypes.create_string_buffer(256)
    rc = load().psx_device_info(buf, 256)
    return rc == 0, buf.value.decode() if rc == 0 else (load().psx_last_error() or b"").decode()
