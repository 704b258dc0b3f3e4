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
from ._lib import (FWIX_NSPS, FWIX_PREC_F32, FWIX_PREC_F64_SCATTER, FWIX_PSPI, FWIX_PSPI_PLAIN, Handle, call, create,
                   default_params, make_dims)

__all__ = ["FFT2d", "PhaseShift", "Selector", "Taper", "OneStep", "Injection", "Scatter",
           "ImagingCondition", "OneWay", "Born", "reflect_coef", "bfp16_pack", "bfp16_unpack", "ptr", "stream",
           "make_dims", "default_params", "FWIX_PSPI", "FWIX_NSPS", "FWIX_PSPI_PLAIN", "FWIX_PREC_F32",
           "FWIX_PREC_F64_SCATTER"]


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
        self.nmodel = dims.NW

    def set_shot_planes(self, counts) -> None:
        """Shot batch (NS == 1): shot k owns the next counts[k] planes; references
        (nz, nref, nshot) and labels (nz, nshot, NY, NX) are then per shot."""
        c = _i32(counts, None, "plane counts")
        call("fwix_onestep_set_shot_planes", self._h.h, len(c), _vp(c))
        self.nmodel = len(c)
        self._refs = None

    def set_refs(self, sref: torch.Tensor, labels: torch.Tensor) -> None:
        d, p = self.dims, self.params
        _check(sref, (d.nz, p.nref, self.nmodel), name="sref")
        _check(labels, (d.nz, self.nmodel, d.NY, d.NX), torch.uint8, name="labels")
        self._refs = (sref, labels)
        call("fwix_onestep_set_refs", self._h.h, ptr(sref), ptr(labels))

    def set_depth(self, iz: int) -> None:
        call("fwix_onestep_set_depth", self._h.h, int(iz))

    def set_frequencies(self, f_hz) -> None:
        """Per-plane frequencies (NW values, Hz) in place of the regular axis."""
        call("fwix_onestep_set_frequencies", self._h.h, _vp(_f32(f_hz, self.dims.NW, "frequencies")))


def _f32(a, n=None, name="array") -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32).ravel())
    if n is not None and a.size != n:
        raise ValueError(f"{name}: expected {n} values, got {a.size}")
    return a


def _i32(a, n=None, name="array") -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(a, dtype=np.int32).ravel())
    if n is not None and a.size != n:
        raise ValueError(f"{name}: expected {n} values, got {a.size}")
    return a


def _vp(a: np.ndarray) -> C.c_void_p:
    return a.ctypes.data_as(C.c_void_p)


class Injection:
    """Trace injection / recording at one depth step (``propagator/src/Injection.h:10-55``).

    ``forward(add, traces, wfld)``: wfld (+)= Inj_iz(traces); ``adjoint(add, traces,
    wfld)``: traces (+)= Rec_iz^T wfld.  traces (ntrace, NW), wfld (NS, NW, NY, NX).
    """

    def __init__(self, dims, cx, cy, cz, ids):
        self.dims = dims
        self.ntrace = len(np.ravel(cx))
        self._h = Handle(create("fwix_injection_create", C.byref(dims), self.ntrace))
        self.set_coords(cx, cy, cz, ids)

    def set_coords(self, cx, cy, cz, ids) -> None:
        n = self.ntrace
        self._coords = (_f32(cx, n, "cx"), _f32(cy, n, "cy"), _f32(cz, n, "cz"), _i32(ids, n, "ids"))
        call("fwix_injection_set_coords", self._h.h, *map(_vp, self._coords))

    def set_depth(self, iz: int) -> None:
        call("fwix_injection_set_depth", self._h.h, int(iz))

    def forward(self, add, traces, wfld):
        _check(traces, (self.ntrace, self.dims.NW), name="traces")
        _check(wfld, _wshape(self.dims), name="wavefield")
        call("fwix_injection_forward", self._h.h, int(bool(add)), ptr(traces), ptr(wfld), stream())
        return wfld

    def adjoint(self, add, traces, wfld):
        _check(traces, (self.ntrace, self.dims.NW), name="traces")
        _check(wfld, _wshape(self.dims), name="wavefield")
        call("fwix_injection_adjoint", self._h.h, int(bool(add)), ptr(traces), ptr(wfld), stream())
        return traces


class Scatter(_Op):
    """Born scattering operator of one depth (``propagator/src/Scatter.h:12-104``).

    ``set_slow`` borrows the unpadded background slowness (nz, NW, ny, nx).
    """
    _kind = "scatter"

    def __init__(self, dims, params=None, **kw):
        super().__init__(dims)
        self.params = params if params is not None else default_params(**kw)
        self._h = Handle(create("fwix_scatter_create", C.byref(dims), C.byref(self.params)))
        self._slow = None

    def set_slow(self, slow: torch.Tensor) -> None:
        d = self.dims
        _check(slow, (d.nz, d.NW, d.ny, d.nx), name="slowness")
        self._slow = slow
        call("fwix_scatter_set_slow", self._h.h, ptr(slow))

    def set_depth(self, iz: int) -> None:
        call("fwix_scatter_set_depth", self._h.h, int(iz))


class ImagingCondition:
    """Imaging condition of one depth (``propagator/src/ImagingCondition.h:11-43``).

    model: extended image (NW, ny, nx); data: wavefield (NS, NW, NY, NX).
    """

    def __init__(self, dims):
        self.dims = dims
        self._h = Handle(create("fwix_ic_create", C.byref(dims)))
        self._bg = None

    def set_background(self, bg: torch.Tensor) -> None:
        _check(bg, _wshape(self.dims), name="background")
        self._bg = bg
        call("fwix_ic_set_background", self._h.h, ptr(bg))

    def forward(self, add, img, wfld):
        d = self.dims
        _check(img, (d.NW, d.ny, d.nx), name="image")
        _check(wfld, _wshape(d), name="wavefield")
        call("fwix_ic_forward", self._h.h, int(bool(add)), ptr(img), ptr(wfld), stream())
        return wfld

    def adjoint(self, add, img, wfld):
        d = self.dims
        _check(img, (d.NW, d.ny, d.nx), name="image")
        _check(wfld, _wshape(d), name="wavefield")
        call("fwix_ic_adjoint", self._h.h, int(bool(add)), ptr(img), ptr(wfld), stream())
        return img


class OneWay:
    """Downward (direction 0) / Upward (direction 1) recursion (``propagator/src/OneWay.h:14-245``)."""

    def __init__(self, direction: int, method: int, dims, params=None, **kw):
        self.dims = dims
        self.params = params if params is not None else default_params(**kw)
        self._h = Handle(create("fwix_oneway_create", int(direction), int(method), C.byref(dims),
                                C.byref(self.params)))
        self._refs = None
        self.nmodel = dims.NW

    def set_shot_planes(self, counts) -> None:
        """Shot batch (NS == 1): shot k owns the next counts[k] planes; references
        (nz, nref, nshot) and labels (nz, nshot, NY, NX) are then per shot."""
        c = _i32(counts, None, "plane counts")
        call("fwix_oneway_set_shot_planes", self._h.h, len(c), _vp(c))
        self.nmodel = len(c)
        self._refs = None

    def set_refs(self, sref: torch.Tensor, labels: torch.Tensor) -> None:
        d, p = self.dims, self.params
        _check(sref, (d.nz, p.nref, self.nmodel), name="sref")
        _check(labels, (d.nz, self.nmodel, d.NY, d.NX), torch.uint8, name="labels")
        self._refs = (sref, labels)
        call("fwix_oneway_set_refs", self._h.h, ptr(sref), ptr(labels))

    def set_frequencies(self, f_hz) -> None:
        call("fwix_oneway_set_frequencies", self._h.h, _vp(_f32(f_hz, self.dims.NW, "frequencies")))

    def set_step_chaining(self, on: bool) -> None:
        """Chain consecutive PSPI-forward / NSPS-adjoint steps through the row spectrum
        (default on); off runs every step whole (same results to float32 rounding)."""
        call("fwix_oneway_set_step_chaining", self._h.h, int(bool(on)))

    def forward(self, add, model, data, checkpoints: torch.Tensor | None = None):
        s = _wshape(self.dims)
        _check(model, s, name="model")
        _check(data, s, name="data")
        if checkpoints is not None:
            _check(checkpoints, (self.dims.nz, *s), name="checkpoints")
        call("fwix_oneway_forward", self._h.h, int(bool(add)), ptr(model), ptr(data), ptr(checkpoints),
             stream())
        return data

    def adjoint(self, add, model, data):
        s = _wshape(self.dims)
        _check(model, s, name="model")
        _check(data, s, name="data")
        call("fwix_oneway_adjoint", self._h.h, int(bool(add)), ptr(model), ptr(data), stream())
        return model


class Born:
    """Background modeling + linearized Born operator (down-scattering channel) and migration.

    ``propagator/src/Propagator.cpp:155-180`` and ``propagator/src/ExtendedBorn.cpp:152-208``.
    Model ds (nz, NW, ny, nx); data (nrec, NW); image (nz, ny, nx).
    ``ckpt_every`` > 1 keeps every ckpt_every-th background wavefield and
    recomputes the others in the adjoint.

    After ``set_shot_planes(counts)`` (NS == 1) the planes are a batch of shots:
    shot k owns the next counts[k] planes; slowness and model become
    (nz, nshot, ny, nx), references (nz, nref, nshot), labels
    (nz, nshot, NY, NX), and wavelet / data flat ragged traces (each trace
    one sample per plane of its shot, traces one after another).
    """

    def __init__(self, dims, params=None, ckpt_every: int = 1, **kw):
        self.dims = dims
        self.params = params if params is not None else default_params(**kw)
        self._h = Handle(create("fwix_born_create", C.byref(dims), C.byref(self.params), int(ckpt_every)))
        self._keep = {}
        self.nsrc = self.nrec = 0
        self.nmodel = dims.NW        # model planes: NW, or the shots of a batch
        self.shot_planes = None

    def model_shape(self):
        """Shape of slowness, model and gradient: (nz, NW, ny, nx), or (nz, nshot, ny, nx) for a shot batch."""
        d = self.dims
        return (d.nz, self.nmodel, d.ny, d.nx)

    def trace_shape(self, receivers: bool):
        """Shape of the receiver (or source) traces: (n, NW), or the flat ragged (samples,) of a shot batch."""
        n = self.nrec if receivers else self.nsrc
        if self.shot_planes is None:
            return (n, self.dims.NW)
        return (int(_lib.lib.fwix_born_trace_samples(self._h.h, int(receivers))),)

    _model_shape = model_shape
    _trace_shape = trace_shape

    def set_method(self, method: int) -> None:
        """FWIX_PSPI (fused steps) or FWIX_PSPI_PLAIN (separate operators, one shot per launch)."""
        call("fwix_born_set_method", self._h.h, int(method))

    def set_shot_planes(self, counts) -> None:
        """Lay a batch of shots along the plane axis (shot k: the next counts[k] planes)."""
        c = _i32(counts, None, "plane counts")
        call("fwix_born_set_shot_planes", self._h.h, len(c), _vp(c))
        self.nmodel = len(c)
        self.shot_planes = c.copy()
        self.nsrc = self.nrec = 0
        self._keep.clear()

    def set_refs(self, sref, labels):
        d, p = self.dims, self.params
        _check(sref, (d.nz, p.nref, self.nmodel), name="sref")
        _check(labels, (d.nz, self.nmodel, d.NY, d.NX), torch.uint8, name="labels")
        self._keep["refs"] = (sref, labels)
        call("fwix_born_set_refs", self._h.h, ptr(sref), ptr(labels))

    def set_slow(self, slow):
        _check(slow, self._model_shape(), name="slowness")
        self._keep["slow"] = slow
        call("fwix_born_set_slow", self._h.h, ptr(slow))

    def set_frequencies(self, f_hz) -> None:
        call("fwix_born_set_frequencies", self._h.h, _vp(_f32(f_hz, self.dims.NW, "frequencies")))

    def set_sources(self, x, y, z, ids):
        n = len(np.ravel(x))
        a = (_f32(x, n, "x"), _f32(y, n, "y"), _f32(z, n, "z"), _i32(ids, n, "ids"))
        call("fwix_born_set_sources", self._h.h, n, *map(_vp, a))
        self.nsrc = n

    def set_receivers(self, x, y, z, ids):
        n = len(np.ravel(x))
        a = (_f32(x, n, "x"), _f32(y, n, "y"), _f32(z, n, "z"), _i32(ids, n, "ids"))
        call("fwix_born_set_receivers", self._h.h, n, *map(_vp, a))
        self.nrec = n

    def set_wavelet(self, wavelet):
        _check(wavelet, self._trace_shape(False), name="wavelet")
        self._keep["wavelet"] = wavelet
        call("fwix_born_set_wavelet", self._h.h, ptr(wavelet))

    def model(self, data: torch.Tensor | None = None, add: bool = False):
        if data is not None:
            _check(data, self._trace_shape(True), name="data")
        call("fwix_born_model", self._h.h, int(bool(add)), ptr(data), stream())
        return data

    def forward(self, add, dslow, data):
        _check(dslow, self._model_shape(), name="model")
        _check(data, self._trace_shape(True), name="data")
        call("fwix_born_forward", self._h.h, int(bool(add)), ptr(dslow), ptr(data), stream())
        return data

    def adjoint(self, add, dslow, data):
        _check(dslow, self._model_shape(), name="model")
        _check(data, self._trace_shape(True), name="data")
        call("fwix_born_adjoint", self._h.h, int(bool(add)), ptr(dslow), ptr(data), stream())
        return dslow

    def migrate(self, add, image, data):
        d = self.dims
        _check(image, (d.nz, d.ny, d.nx), name="image")
        _check(data, self._trace_shape(True), name="data")
        call("fwix_born_migrate", self._h.h, int(bool(add)), ptr(image), ptr(data), stream())
        return image

    def set_codec(self, codec: int) -> None:
        """Checkpoint codec: 0 lossless, 1 16-bit block floating point (set before model())."""
        call("fwix_born_set_codec", self._h.h, int(codec))

    def set_precision(self, precision: int) -> None:
        """FWIX_PREC_F32 (default) or FWIX_PREC_F64_SCATTER: the Born steps' scattered-field
        transforms in float64 arithmetic (complex64 storage), for low frequencies on fine
        grids where the first Taylor term's k^2 / (w^2 s) weight amplifies float32 rounding."""
        call("fwix_born_set_precision", self._h.h, int(precision))

    def set_step_chaining(self, on: bool) -> None:
        """Chain the background recursion's PSPI steps through the row spectrum (default
        on): each step's select pass leaves the next step's row spectrum, so the next
        step skips its first row pass.  Off runs every step whole (same results to
        float32 rounding)."""
        call("fwix_born_set_step_chaining", self._h.h, int(bool(on)))

    def set_density(self, den):
        d = self.dims
        _check(den, (d.nz, d.NW, d.ny, d.nx), name="density")
        self._keep["den"] = den
        call("fwix_born_set_density", self._h.h, ptr(den))

    def model_up(self, data: torch.Tensor | None = None, add: bool = False):
        """Up pass of the nonlinear modeling (``Propagator.cpp:181-211``); keeps every V_z."""
        if data is not None:
            _check(data, (self.nrec, self.dims.NW), name="data")
        call("fwix_born_model_up", self._h.h, int(bool(add)), ptr(data), stream())
        return data

    def extended_forward(self, add, dslow, dden, data):
        """All four ExtendedBorn channels (``ExtendedBorn.cpp:80-118``)."""
        d = self.dims
        for t, nm in ((dslow, "dslow"), (dden, "dden")):
            _check(t, (d.nz, d.NW, d.ny, d.nx), name=nm)
        _check(data, (self.nrec, d.NW), name="data")
        call("fwix_born_extended_forward", self._h.h, int(bool(add)), ptr(dslow), ptr(dden), ptr(data), stream())
        return data

    def extended_adjoint(self, add, dslow, dden, data):
        d = self.dims
        for t, nm in ((dslow, "dslow"), (dden, "dden")):
            _check(t, (d.nz, d.NW, d.ny, d.nx), name=nm)
        _check(data, (self.nrec, d.NW), name="data")
        call("fwix_born_extended_adjoint", self._h.h, int(bool(add)), ptr(dslow), ptr(dden), ptr(data), stream())
        return dslow, dden

    def reflect_forward(self, add, refl, data):
        """data (+)= up-going recording of refl * U_z (``Propagator.cpp:185-208``)."""
        d = self.dims
        _check(refl, (d.nz, d.NW, d.ny, d.nx), name="reflectivity")
        _check(data, (self.nrec, d.NW), name="data")
        call("fwix_born_reflect_forward", self._h.h, int(bool(add)), ptr(refl), ptr(data), stream())
        return data

    def reflect_adjoint(self, add, refl, data):
        d = self.dims
        _check(refl, (d.nz, d.NW, d.ny, d.nx), name="reflectivity")
        _check(data, (self.nrec, d.NW), name="data")
        call("fwix_born_reflect_adjoint", self._h.h, int(bool(add)), ptr(refl), ptr(data), stream())
        return refl

    def reflect_migrate(self, add, image, data):
        """image_z (+)= sum_{w,s} conj(U_z) mu_z: migration of surface data."""
        d = self.dims
        _check(image, (d.nz, d.ny, d.nx), name="image")
        _check(data, (self.nrec, d.NW), name="data")
        call("fwix_born_reflect_migrate", self._h.h, int(bool(add)), ptr(image), ptr(data), stream())
        return image

    @property
    def steps(self) -> int:
        return int(_lib.lib.fwix_born_steps(self._h.h))


def bfp16_pack(x: torch.Tensor):
    """16-bit block floating point code of a complex64 device array: (q int16 pairs, scales)."""
    n = x.numel()
    q = torch.empty((n, 2), dtype=torch.int16, device=x.device)
    sc = torch.empty(((n + 63) // 64,), dtype=torch.float32, device=x.device)
    call("fwix_bfp16_pack", ptr(x.contiguous()), ptr(q), ptr(sc), C.c_size_t(n), stream())
    return q, sc


def bfp16_unpack(q: torch.Tensor, sc: torch.Tensor, shape) -> torch.Tensor:
    out = torch.empty(shape, dtype=torch.complex64, device=q.device)
    call("fwix_bfp16_unpack", ptr(q), ptr(sc), ptr(out), C.c_size_t(out.numel()), stream())
    return out


def reflect_coef(dims, slow: torch.Tensor, den: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
    """Reflection coefficients of a (slowness, density) model (``reflect.cu:42-45``), (nz, NW, ny, nx)."""
    shape = (dims.nz, dims.NW, dims.ny, dims.nx)
    _check(slow, shape, name="slowness")
    _check(den, shape, name="density")
    out = torch.empty_like(slow) if out is None else out
    _check(out, shape, name="reflectivity")
    call("fwix_reflect_coef", C.byref(dims), ptr(slow), ptr(den), ptr(out), stream())
    return out
