"""Operator classes with the names and signatures of the reference's Python
front end (``fwix/CudaWEM.py:7-157``), running on the HIP library.

Hypercubes follow the reference: a wavefield has axes (x NX, y NY, frequency
NW, shot NS); a slowness model has axes (x nx, y ny, frequency nw, depth nz),
``s = 1/v^2`` per frequency; traces have axes (frequency NW, trace).
Reference slownesses and labels for every depth are derived once from the
slowness model (deterministic quantile references, nearest-reference
labels, edge-replicated into the padding) and stay resident on the GPU.

Differences from the reference classes (kept deliberately):

* ``Downward.forward(data)`` / ``PSPI.forward(data)`` (one argument)
  propagate in place, ``u <- P u``; the reference's host form accumulates
  into the aliased buffer (``operator/src/CudaOperator.h:141-151``).
* ``ExtendedBorn``'s linearized reflection uses the exact derivative of the
  reflection coefficient (the reference kernel's slowness terms carry the
  fourth root of s1/s0 instead of the square root, ``lin_reflect.cu:69-79``).
* ``nbatches`` splits shots and frequency bands as the reference's
  streaming layer does; the batches run one after another on one GPU.
* Windows and padding are rounded up to the next supported FFT size.
"""
from __future__ import annotations

import numpy as np

from . import ops
from .CudaOperator import CudaOperator, Operator
from .refsampler import RefSampler as _RefSampler, select_refs
from .sep import complexVector, hypercube, paramObj

__all__ = ["Propagator", "ExtendedBorn", "PhaseShift", "RefSampler", "PSPI", "NSPS", "Injection",
           "Downward", "Upward"]


def _hyp(v) -> hypercube:
    return v.getHyper() if hasattr(v, "getHyper") else v


def _dims(wfld: hypercube, slow: hypercube | None):
    ax = [wfld.getAxis(i) for i in (1, 2, 3, 4)]
    if slow is None:
        return ops.make_dims(ax[0].n, ax[1].n, ax[2].n, ax[3].n, ax[0].n, ax[1].n, ax[0].o, ax[0].d,
                             ax[1].o, ax[1].d, ax[2].o, ax[2].d)
    sx, sy, sw, sz = (slow.getAxis(i) for i in (1, 2, 3, 4))
    if sw.n != ax[2].n:
        raise RuntimeError(f"slowness has {sw.n} frequencies, wavefield {ax[2].n}")
    if sx.n > ax[0].n or sy.n > ax[1].n:
        raise RuntimeError("wavefield grid must be the slowness grid plus padding")
    return ops.make_dims(ax[0].n, ax[1].n, ax[2].n, ax[3].n, sx.n, sy.n, ax[0].o, ax[0].d, ax[1].o, ax[1].d,
                         ax[2].o, ax[2].d, sz.o, sz.d, sz.n)


def _slow_array(slow) -> np.ndarray:
    a = slow.numpy() if hasattr(slow, "numpy") else np.asarray(slow)
    return np.ascontiguousarray(a, np.complex64)


def _refs_on_device(slow_arr, nref, padx, pady, device):
    import torch
    sref, labels = select_refs(slow_arr, nref, padx, pady)
    return torch.from_numpy(sref).to(device), torch.from_numpy(labels).to(device)


class PhaseShift(CudaOperator):
    """``fwix/CudaWEM.py:65-77``: PhaseShift(model, data, dz, eps=0) + set_slow."""

    def __init__(self, model, data, dz, eps=0, device=None):
        super().__init__(model, data, device)
        self.cppMode = ops.PhaseShift(_dims(_hyp(model), None), float(dz), float(eps))

    def set_slow(self, slow):
        self.cppMode.set_slow(np.asarray(slow))

    def cu_forward(self, add, m, d):
        self.cppMode.forward(add, m, d)

    def cu_adjoint(self, add, m, d):
        self.cppMode.adjoint(add, m, d)

    def cu_inverse(self, add, m, d):
        self.cppMode.inverse(add, m, d)


class RefSampler:
    """``fwix/CudaWEM.py:80-88``: reference slownesses and labels of a slowness model."""

    def __init__(self, slow, par):
        par = par if isinstance(par, paramObj) else paramObj(par)
        self.cppMode = _RefSampler(_slow_array(slow), par.getInt("nref"), par.getInt("padx", 0),
                                   par.getInt("pady", 0), par.getString("ref_method", "quantile"),
                                   par.getInt("ref_seed", 0))

    def get_ref_slow(self, iz, iref):
        return self.cppMode.get_ref_slow(iz, iref)

    def get_ref_labels(self, iz):
        return self.cppMode.get_ref_labels(iz)


class _OneStepOp(CudaOperator):
    method = ops.FWIX_PSPI

    def __init__(self, model, data, slow, par, device=None):
        super().__init__(model, data, device)
        par = par if isinstance(par, paramObj) else paramObj(par)
        h, sh = _hyp(model), _hyp(slow)
        dims = _dims(h, sh)
        nref = par.getInt("nref", 1)
        taper = self.method == ops.FWIX_PSPI
        self.cppMode = ops.OneStep(self.method, dims, nref=nref,
                                   taperx=par.getInt("taperx", 0) if taper else 0,
                                   tapery=par.getInt("tapery", 0) if taper else 0,
                                   eps=par.getFloat("eps", 0.0))
        self.cppMode.set_refs(*_refs_on_device(_slow_array(slow), nref, dims.NX - dims.nx, dims.NY - dims.ny,
                                               self.device))

    def set_depth(self, iz):
        self.cppMode.set_depth(iz)

    def cu_forward(self, add, m, d):
        self.cppMode.forward(add, m, d)

    def cu_adjoint(self, add, m, d):
        self.cppMode.adjoint(add, m, d)

    def cu_forward_inplace(self, m):
        self.cppMode.forward(False, m, m)

    def cu_adjoint_inplace(self, d):
        self.cppMode.adjoint(False, d, d)


class PSPI(_OneStepOp):
    """``fwix/CudaWEM.py:91-104``: PSPI(model, data, slow, par) + set_depth."""
    method = ops.FWIX_PSPI


class NSPS(_OneStepOp):
    """``fwix/CudaWEM.py:107-119``: NSPS(model, data, slow, par) + set_depth (true adjoint)."""
    method = ops.FWIX_NSPS


class Injection(CudaOperator):
    """``fwix/CudaWEM.py:121-138``: Injection(model=traces, data=wavefield, oz, dz, cx, cy, cz, ids)."""

    def __init__(self, model, data, oz, dz, cx, cy, cz, ids, device=None):
        super().__init__(model, data, device)
        h = _hyp(data)
        d = _dims(h, None)
        d.oz, d.dz = float(oz), float(dz)
        th = _hyp(model)
        if th.getAxis(1).n != d.NW:
            raise RuntimeError("trace hypercube axis 1 must be the frequency axis of the wavefield")
        self.cppMode = ops.Injection(d, cx, cy, cz, ids)

    def set_coords(self, cx, cy, cz, ids):
        self.cppMode.set_coords(cx, cy, cz, ids)

    def set_depth(self, iz):
        self.cppMode.set_depth(iz)

    def cu_forward(self, add, m, d):
        self.cppMode.forward(add, m, d)

    def cu_adjoint(self, add, m, d):
        self.cppMode.adjoint(add, m, d)


class _OneWayOp(CudaOperator):
    direction = 0

    def __init__(self, model, data, slow, par, device=None):
        super().__init__(model, data, device)
        par = par if isinstance(par, paramObj) else paramObj(par)
        dims = _dims(_hyp(model), _hyp(slow))
        nref = par.getInt("nref", 1)
        self.cppMode = ops.OneWay(self.direction, ops.FWIX_PSPI, dims, nref=nref, taperx=par.getInt("taperx", 0),
                                  tapery=par.getInt("tapery", 0), eps=par.getFloat("eps", 0.0))
        self.cppMode.set_refs(*_refs_on_device(_slow_array(slow), nref, dims.NX - dims.nx, dims.NY - dims.ny,
                                               self.device))

    def set_depth(self, iz):  # the recursion runs over every depth; kept for signature compatibility
        pass

    def cu_forward(self, add, m, d):
        self.cppMode.forward(add, m, d)

    def cu_adjoint(self, add, m, d):
        self.cppMode.adjoint(add, m, d)

    def cu_forward_inplace(self, m):
        self.cppMode.forward(False, m, m)

    def cu_adjoint_inplace(self, d):
        self.cppMode.adjoint(False, d, d)


class Downward(_OneWayOp):
    """``fwix/CudaWEM.py:141-157``: Downward(model, data, slow, par)."""
    direction = 0


class Upward(_OneWayOp):
    """``fwix/CudaWEM.py:160-171``: Upward(model, data, slow, par)."""
    direction = 1


def dense_shot_ids(s_ids, r_ids):
    """Map shot ids to 0..nshot-1 in sorted order; receivers must name a known shot
    (``propagator/src/Propagator.cpp:28-58``)."""
    s_ids = np.asarray(s_ids, np.int64).ravel()
    r_ids = np.asarray(r_ids, np.int64).ravel()
    uniq = np.unique(s_ids)
    pos = np.searchsorted(uniq, r_ids)
    pos_c = np.minimum(pos, len(uniq) - 1)
    if r_ids.size and not np.all(uniq[pos_c] == r_ids):
        raise RuntimeError("Error: Found a receiver associated with a Shot ID not present in the source list.")
    return np.searchsorted(uniq, s_ids).astype(np.int32), pos_c.astype(np.int32), len(uniq)


class Propagator(Operator):
    """``fwix/CudaWEM.py:7-27``: Propagator(model, data, wavelet, par, geometry, nbatches=(1, 1)).

    ``model`` is [slowness, density] (or a single slowness vector) with axes
    (nx, ny, nw, nz); ``data`` holds the recorded traces (nw, ntrace);
    ``wavelet`` the source traces (nw, nsrc).  ``par`` needs ``padx``/``pady``
    >= 0 (``Propagator.cpp:61-70``), ``nref``, and optionally
    ``taperx/tapery``, ``eps``, ``eps_scat``, ``taylor``, ``ginsu_x/ginsu_y``,
    ``ckpt_every``.  As the reference's front end, it runs through the
    streaming layer (``pyFWIX.StreamingPropagator``): source batches x
    frequency bands on ginsu windows of the model.
    """

    def __init__(self, model, data, wavelet, par, geometry, nbatches=(1, 1), device=None):
        from .streaming import StreamingPropagator
        slow = model[0] if isinstance(model, (list, tuple)) else model
        self.setDomainRange(slow, data)
        par = par if isinstance(par, paramObj) else paramObj(par)
        sh = _hyp(slow)
        axes = tuple((sh.getAxis(i).n, sh.getAxis(i).o, sh.getAxis(i).d) for i in (1, 2, 3, 4))
        keys = {"padx": -1, "pady": -1, "nref": 1, "taperx": 0, "tapery": 0, "taylor": 1, "ckpt_every": 1}
        params = {k: par.getInt(k, v) for k, v in keys.items()}
        params.update({k: par.getFloat(k, v) for k, v in (("eps", 0.0), ("eps_scat", 0.04), ("ginsu_x", 0.0),
                                                            ("ginsu_y", 0.0))})
        w = wavelet.numpy() if hasattr(wavelet, "numpy") else np.asarray(wavelet)
        self.cppMode = StreamingPropagator(axes, w, geometry, params, tuple(nbatches), device)
        self.nbatches = tuple(nbatches)

    @staticmethod
    def _parts(model):
        if isinstance(model, (list, tuple)):
            return _slow_array(model[0]), (_slow_array(model[1]) if len(model) > 1 else None)
        return _slow_array(model), None

    def forward(self, add, model, data):
        """Nonlinear modeling (``Propagator.cpp:121-211``): the down-going wavefield
        recorded at every depth, then the up-going wavefield reflected at the
        slowness/density contrasts (``reflect.cu:42-45``) recorded on its way up."""
        slow, den = self._parts(model)
        out = data.numpy().copy() if data.on_device else data.arr
        self.cppMode.forward(add, slow, den, out)
        if data.on_device:
            data.arr.copy_(__import__("torch").from_numpy(out))


class ExtendedBorn(Operator):
    """``fwix/CudaWEM.py:30-50``: ExtendedBorn(model, data, slow_den, propagator).

    Model [ds, drho] with the slowness axes (extended: one value per
    frequency); data the recorded traces.  All four channels of
    ``propagator/src/ExtendedBorn.cpp:80-150`` around the background of
    ``slow_den`` = [slowness, density], per batch of the propagator
    (``pyFWIX.StreamingExtendedBorn``).
    """

    def __init__(self, model, data, slow_den, propagator: Propagator):
        from .streaming import StreamingExtendedBorn
        parts = list(model) if isinstance(model, (list, tuple)) else [model]
        self.domain = [p.cloneSpace() for p in parts]
        self.range = data.cloneSpace()
        slow, den = Propagator._parts(slow_den)
        self.cppMode = StreamingExtendedBorn(propagator.cppMode, slow, den)

    def getDomain(self):
        return self.domain

    @staticmethod
    def _host(v, like):
        if v is None:
            return np.zeros(like.getHyper().shape, np.complex64)
        return v.numpy() if v.on_device else v.arr

    def forward(self, add, model, data):
        parts = list(model) if isinstance(model, (list, tuple)) else [model]
        ds = self._host(parts[0], self.domain[0])
        dr = self._host(parts[1] if len(parts) > 1 else None, self.domain[0])
        out = data.numpy().copy() if data.on_device else data.arr
        self.cppMode.forward(add, ds, dr, out)
        if data.on_device:
            data.arr.copy_(__import__("torch").from_numpy(out))

    def adjoint(self, add, model, data):
        parts = list(model) if isinstance(model, (list, tuple)) else [model]
        ds = parts[0].numpy().copy() if parts[0].on_device else parts[0].arr
        dr = (parts[1].numpy().copy() if parts[1].on_device else parts[1].arr) if len(parts) > 1 \
            else np.zeros_like(ds)
        self.cppMode.adjoint(add, ds, dr, self._host(data, self.range))
        for v, arr in zip(parts, (ds, dr)):
            if v.on_device:
                v.arr.copy_(__import__("torch").from_numpy(arr))

    def dotTest(self, verbose=False, rng=None):
        rng = rng if rng is not None else np.random.default_rng()
        m = [p.cloneSpace().random(rng) for p in self.domain]
        d = self.range.cloneSpace().random(rng)
        d1, m1 = self.range.cloneSpace(), [p.cloneSpace() for p in self.domain]
        errs = []
        for add in (False, True):
            self.forward(add, m, d1)
            self.adjoint(add, m1, d)
            rhs = sum(a.dot(b) for a, b in zip(m, m1))
            errs.append(abs((d.dot(d1).conjugate() / rhs).real - 1.0))
            if verbose:
                print(f"add={add}: error {errs[-1]:.3e}")
        return tuple(errs)
