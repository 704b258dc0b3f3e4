"""Low-level device operators over the C ABI (include/fwix_hip.h).

Arguments are torch complex64 tensors resident on the GPU (torch supplies the
device memory and the current stream; all arithmetic runs in the HIP
kernels of libpyfwix_hip.so).  Shapes follow the reference layout:
wavefield (NS, NW, NY, NX), labels (nz, NW, NY, NX) uint8, reference
slownesses (nz, nref, NW).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import FWIX_NSPS, FWIX_PSPI, Handle, call, create, default_params, make_dims

__all__ = ["FFT2d", "PhaseShift", "Selector", "Taper", "OneStep", "ptr", "stream",
           "make_dims", "default_params", "FWIX_PSPI", "FWIX_NSPS"]


def ptr(t: torch.Tensor | None) -> C.c_void_p:
    if t is None:
        return C.c_void_p(0)
    if not t.is_cuda:
        raise ValueError("device operators take GPU tensors")
    if not t.is_contiguous():
        raise ValueError("device operators take contiguous tensors")
    return C.c_void_p(t.data_ptr())


def stream() -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _wshape(d) -> tuple:
    return (d.NS, d.NW, d.NY, d.NX)


def _check(t: torch.Tensor, shape, dtype=torch.complex64, name="array"):
    if tuple(t.shape) != tuple(shape) or t.dtype != dtype:
        raise ValueError(f"{name}: expected {tuple(shape)} {dtype}, got {tuple(t.shape)} {t.dtype}")


class _Op:
    _kind = ""

    def __init__(self, dims):
        self.dims = dims

    def _fwd(self, add, m, d):
        s = _wshape(self.dims)
        _check(m, s, name="model")
        _check(d, s, name="data")
        call(f"fwix_{self._kind}_forward", self._h.h, int(bool(add)), ptr(m), ptr(d), stream())
        return d

    def _adj(self, add, m, d):
        s = _wshape(self.dims)
        _check(m, s, name="model")
        _check(d, s, name="data")
        call(f"fwix_{self._kind}_adjoint", self._h.h, int(bool(add)), ptr(m), ptr(d), stream())
        return m

    def forward(self, add: bool, model: torch.Tensor, data: torch.Tensor) -> torch.Tensor:
        return self._fwd(add, model, data)

    def adjoint(self, add: bool, model: torch.Tensor, data: torch.Tensor) -> torch.Tensor:
        return self._adj(add, model, data)


class FFT2d(_Op):
    """Unitary batched 2-D FFT (``operator/src/FFT.h:10-37``)."""
    _kind = "fft2d"

    def __init__(self, dims):
        super().__init__(dims)
        self._h = Handle(create("fwix_fft2d_create", C.byref(dims)))


class PhaseShift(_Op):
    """Phase shift exp(-i kz dz) for one reference slowness (``propagator/src/PhaseShift.h:11-72``)."""
    _kind = "phase_shift"

    def __init__(self, dims, dz: float, eps: float = 0.0):
        super().__init__(dims)
        self._h = Handle(create("fwix_phase_shift_create", C.byref(dims), C.c_float(dz), C.c_float(eps)))

    def set_slow(self, sref) -> None:
        a = np.ascontiguousarray(np.asarray(sref, dtype=np.complex64).reshape(self.dims.NW))
        call("fwix_phase_shift_set_slow", self._h.h, a.ctypes.data_as(C.c_void_p))

    def inverse(self, add: bool, model: torch.Tensor, data: torch.Tensor) -> torch.Tensor:
        s = _wshape(self.dims)
        _check(model, s, name="model")
        _check(data, s, name="data")
        call("fwix_phase_shift_inverse", self._h.h, int(bool(add)), ptr(model), ptr(data), stream())
        return model


class Selector(_Op):
    """Reference-label mask (``propagator/src/Selector.h:8-58``); self-adjoint."""
    _kind = "selector"

    def __init__(self, dims):
        super().__init__(dims)
        self._h = Handle(create("fwix_selector_create", C.byref(dims)))

    def set_labels(self, labels) -> None:
        a = np.ascontiguousarray(np.asarray(labels, dtype=np.int32))
        if a.size != self.dims.NW * self.dims.NY * self.dims.NX:
            raise ValueError("labels must be (NW, NY, NX)")
        call("fwix_selector_set_labels", self._h.h, a.ctypes.data_as(C.c_void_p))

    def set_value(self, value: int) -> None:
        call("fwix_selector_set_value", self._h.h, int(value))


class Taper(_Op):
    """Cosine edge taper (``propagator/src/Taper.h:8-42``); self-adjoint."""
    _kind = "taper"

    def __init__(self, dims, taperx: int = 0, tapery: int = 0):
        super().__init__(dims)
        self._h = Handle(create("fwix_taper_create", C.byref(dims), int(taperx), int(tapery)))


class OneStep(_Op):
    """One PSPI or NSPS depth step (``propagator/src/OneStep.h:11-111``).

    ``set_refs`` borrows device tensors sref (nz, nref, NW) complex64 and
    labels (nz, NW, NY, NX) uint8 (kept alive by this object).
    ``forward(add, m, d)`` with ``m is d`` runs the step in place.
    """
    _kind = "onestep"

    def __init__(self, method: int, dims, params=None, **kw):
        super().__init__(dims)
        self.params = params if params is not None else default_params(**kw)
        self._h = Handle(create("fwix_onestep_create", int(method), C.byref(dims), C.byref(self.params)))
        self._refs = None

    def set_refs(self, sref: torch.Tensor, labels: torch.Tensor) -> None:
        d, p = self.dims, self.params
        _check(sref, (d.nz, p.nref, d.NW), name="sref")
        _check(labels, (d.nz, d.NW, d.NY, d.NX), torch.uint8, name="labels")
        self._refs = (sref, labels)
        call("fwix_onestep_set_refs", self._h.h, ptr(sref), ptr(labels))

    def set_depth(self, iz: int) -> None:
        call("fwix_onestep_set_depth", self._h.h, int(iz))
