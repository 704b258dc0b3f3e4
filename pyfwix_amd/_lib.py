"""ctypes binding of libpyfwix_hip.so (the C ABI declared in include/fwix_hip.h).

The product path has no CPU fallback: importing this module when the shared
library is missing raises ImportError, and every failing call raises
RuntimeError with the library's own message.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_LIB_PATH = Path(os.environ.get("FWIX_LIB", Path(__file__).resolve().parent / "libpyfwix_hip.so"))

if not _LIB_PATH.exists():
    raise ImportError(
        f"{_LIB_PATH} is missing: build it with `python -m pyfwix_amd.build` "
        "(there is no CPU fallback for the depth-stepping operators)")

# torch ships its own HIP runtime; loading it first makes this library bind to
# that same runtime (one libamdhip64 per process).  Loading this library first
# would bring in a second copy, and torch then finds no GPU.
import torch  # noqa: E402,F401

lib = C.CDLL(str(_LIB_PATH))

FWIX_PSPI = 0
FWIX_NSPS = 1
FWIX_PSPI_PLAIN = 2
FWIX_PREC_F32 = 0
FWIX_PREC_F64_SCATTER = 1


class fwix_dims(C.Structure):
    _fields_ = [("NX", C.c_int), ("NY", C.c_int), ("NW", C.c_int), ("NS", C.c_int),
                ("nx", C.c_int), ("ny", C.c_int),
                ("ox", C.c_float), ("dx", C.c_float), ("oy", C.c_float), ("dy", C.c_float),
                ("ow", C.c_float), ("dw", C.c_float), ("oz", C.c_float), ("dz", C.c_float),
                ("nz", C.c_int)]


class fwix_params(C.Structure):
    _fields_ = [("nref", C.c_int), ("taperx", C.c_int), ("tapery", C.c_int),
                ("eps", C.c_float), ("eps_scat", C.c_float), ("taylor", C.c_int),
                ("planes_per_chunk", C.c_int)]


_P = C.c_void_p
_I = C.c_int
_F = C.c_float
_OPP = C.POINTER(C.c_void_p)
_D = C.POINTER(fwix_dims)
_PP = C.POINTER(fwix_params)

# name -> (argtypes); all return int status except where noted
_SIGS = {
    "fwix_op_destroy": [_P],
    "fwix_fft2d_create": [_D, _OPP],
    "fwix_fft2d_forward": [_P, _I, _P, _P, _P],
    "fwix_fft2d_adjoint": [_P, _I, _P, _P, _P],
    "fwix_phase_shift_create": [_D, _F, _F, _OPP],
    "fwix_phase_shift_set_slow": [_P, _P],
    "fwix_phase_shift_forward": [_P, _I, _P, _P, _P],
    "fwix_phase_shift_adjoint": [_P, _I, _P, _P, _P],
    "fwix_phase_shift_inverse": [_P, _I, _P, _P, _P],
    "fwix_selector_create": [_D, _OPP],
    "fwix_selector_set_labels": [_P, _P],
    "fwix_selector_set_value": [_P, _I],
    "fwix_selector_forward": [_P, _I, _P, _P, _P],
    "fwix_selector_adjoint": [_P, _I, _P, _P, _P],
    "fwix_taper_create": [_D, _I, _I, _OPP],
    "fwix_taper_forward": [_P, _I, _P, _P, _P],
    "fwix_taper_adjoint": [_P, _I, _P, _P, _P],
    "fwix_onestep_create": [_I, _D, _PP, _OPP],
    "fwix_onestep_set_refs": [_P, _P, _P],
    "fwix_onestep_set_depth": [_P, _I],
    "fwix_onestep_set_frequencies": [_P, _P],
    "fwix_onestep_set_shot_planes": [_P, _I, _P],
    "fwix_onestep_forward": [_P, _I, _P, _P, _P],
    "fwix_onestep_adjoint": [_P, _I, _P, _P, _P],
    "fwix_onestep_apply_noalloc": [_I, _I, _D, _PP, _P, _P, _I, _P, _P, _P],
    "fwix_injection_create": [_D, _I, _OPP],
    "fwix_injection_set_coords": [_P, _P, _P, _P, _P],
    "fwix_injection_set_depth": [_P, _I],
    "fwix_injection_forward": [_P, _I, _P, _P, _P],
    "fwix_injection_adjoint": [_P, _I, _P, _P, _P],
    "fwix_scatter_create": [_D, _PP, _OPP],
    "fwix_scatter_set_slow": [_P, _P],
    "fwix_scatter_set_depth": [_P, _I],
    "fwix_scatter_forward": [_P, _I, _P, _P, _P],
    "fwix_scatter_adjoint": [_P, _I, _P, _P, _P],
    "fwix_ic_create": [_D, _OPP],
    "fwix_ic_set_background": [_P, _P],
    "fwix_ic_forward": [_P, _I, _P, _P, _P],
    "fwix_ic_adjoint": [_P, _I, _P, _P, _P],
    "fwix_oneway_create": [_I, _I, _D, _PP, _OPP],
    "fwix_oneway_set_refs": [_P, _P, _P],
    "fwix_oneway_set_step_chaining": [_P, _I],
    "fwix_oneway_set_frequencies": [_P, _P],
    "fwix_oneway_set_shot_planes": [_P, _I, _P],
    "fwix_oneway_forward": [_P, _I, _P, _P, _P, _P],
    "fwix_oneway_adjoint": [_P, _I, _P, _P, _P],
    "fwix_born_create": [_D, _PP, _I, _OPP],
    "fwix_born_set_refs": [_P, _P, _P],
    "fwix_born_set_frequencies": [_P, _P],
    "fwix_born_set_shot_planes": [_P, _I, _P],
    "fwix_born_set_method": [_P, _I],
    "fwix_born_set_slow": [_P, _P],
    "fwix_born_set_sources": [_P, _I, _P, _P, _P, _P],
    "fwix_born_set_receivers": [_P, _I, _P, _P, _P, _P],
    "fwix_born_set_wavelet": [_P, _P],
    "fwix_born_model": [_P, _I, _P, _P],
    "fwix_born_forward": [_P, _I, _P, _P, _P],
    "fwix_born_adjoint": [_P, _I, _P, _P, _P],
    "fwix_born_migrate": [_P, _I, _P, _P, _P],
    "fwix_born_set_density": [_P, _P],
    "fwix_born_set_codec": [_P, _I],
    "fwix_born_set_precision": [_P, _I],
    "fwix_born_set_step_chaining": [_P, _I],
    "fwix_bfp16_pack": [_P, _P, _P, C.c_size_t, _P],
    "fwix_bfp16_unpack": [_P, _P, _P, C.c_size_t, _P],
    "fwix_born_model_up": [_P, _I, _P, _P],
    "fwix_born_extended_forward": [_P, _I, _P, _P, _P, _P],
    "fwix_born_extended_adjoint": [_P, _I, _P, _P, _P, _P],
    "fwix_born_reflect_forward": [_P, _I, _P, _P, _P],
    "fwix_born_reflect_adjoint": [_P, _I, _P, _P, _P],
    "fwix_born_reflect_migrate": [_P, _I, _P, _P, _P],
    "fwix_reflect_coef": [_D, _P, _P, _P, _P],
    "fwix_timing_enable": [_I],
    "fwix_timing_select": [C.c_char_p, _I],
    "fwix_tune_columns_set": [_I, _I, _I, _I],
    "fwix_tune_columns_get": [_I, _I, _I],
    "fwix_tune_columns_layout_set": [_I, _I, _I, _I],
    "fwix_tune_columns_layout_get": [_I, _I, _I],
    "fwix_tune_scratch_layout_set": [_I, _I],
    "fwix_tune_scratch_layout_get": [_I],
    "fwix_copy": [_P, _P, C.c_size_t, _P],
    "fwix_timing_overhead": [C.c_void_p, _I, C.POINTER(C.c_float)],
    "fwix_timing_collect": [],
    "fwix_timing_reset": [],
    "fwix_timing_report": [C.c_char_p, _I],
}

EXPORTS = ["fwix_version", "fwix_last_error", "fwix_params_default", "fwix_fft_size_supported",
           "fwix_born_steps", "fwix_born_trace_samples", *_SIGS.keys()]

for _name, _args in _SIGS.items():
    _fn = getattr(lib, _name, None)
    if _fn is not None:
        _fn.argtypes = _args
        _fn.restype = C.c_int

lib.fwix_version.restype = C.c_int
lib.fwix_last_error.restype = C.c_char_p
lib.fwix_params_default.argtypes = [_PP]
lib.fwix_params_default.restype = None
lib.fwix_fft_size_supported.argtypes = [_I]
lib.fwix_fft_size_supported.restype = C.c_int
lib.fwix_born_steps.argtypes = [_P]
lib.fwix_born_steps.restype = C.c_longlong
lib.fwix_born_trace_samples.argtypes = [_P, _I]
lib.fwix_born_trace_samples.restype = C.c_longlong


def check(rc: int) -> None:
    if rc != 0:
        raise RuntimeError(lib.fwix_last_error().decode(errors="replace"))


def call(name: str, *args) -> None:
    check(getattr(lib, name)(*args))


def create(name: str, *args) -> C.c_void_p:
    h = C.c_void_p()
    check(getattr(lib, name)(*args, C.byref(h)))
    return h


class Handle:
    """Owns one operator handle; destroyed with the Python object."""

    def __init__(self, h: C.c_void_p):
        self.h = h

    def __del__(self):
        h, self.h = getattr(self, "h", None), None
        if h and lib is not None:
            lib.fwix_op_destroy(h)


def default_params(**kw) -> fwix_params:
    p = fwix_params()
    lib.fwix_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def make_dims(NX, NY, NW, NS, nx=None, ny=None, ox=0.0, dx=1.0, oy=0.0, dy=1.0,
              ow=1.0, dw=1.0, oz=0.0, dz=1.0, nz=1) -> fwix_dims:
    return fwix_dims(int(NX), int(NY), int(NW), int(NS), int(nx or NX), int(ny or NY),
                     float(ox), float(dx), float(oy), float(dy), float(ow), float(dw),
                     float(oz), float(dz), int(nz))


def fft_size_supported(n: int) -> bool:
    return bool(lib.fwix_fft_size_supported(int(n)))


def timing_enable(reserve_pairs: int = 1 << 16) -> None:
    """Start recording per-pass HIP events (``reserve_pairs`` event pairs pre-created)."""
    call("fwix_timing_enable", int(reserve_pairs))


def timing_disable() -> None:
    call("fwix_timing_enable", 0)


def timing_collect() -> dict:
    """Synchronize recorded events; return {pass: {"ms", "launches", "bytes"}} accumulated so far."""
    import json
    call("fwix_timing_collect")
    buf = C.create_string_buffer(1 << 16)
    call("fwix_timing_report", buf, len(buf))
    return json.loads(buf.value.decode())


def timing_reset() -> None:
    call("fwix_timing_reset")


def timing_select(pass_name: str | None = None, every: int = 1) -> None:
    """Record only ``pass_name`` (None: every pass), every ``every``-th launch of it."""
    call("fwix_timing_select", None if pass_name is None else pass_name.encode(), int(every))


def timing_overhead(stream: int, pairs: int = 200) -> float:
    """Mean elapsed microseconds of an empty event pair on ``stream``."""
    us = C.c_float()
    call("fwix_timing_overhead", C.c_void_p(stream), int(pairs), C.byref(us))
    return float(us.value)
