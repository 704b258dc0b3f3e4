"""Minimal stand-ins for the SEPlib / genericIO types the reference's Python
front end passes around (``axis``, ``hypercube``, ``complexNDReg`` vectors,
``paramObj``), enough for the operator classes of ``CudaWEM`` and
``CudaOperator`` to keep the reference's signatures.

Axis 1 is the fastest (the reference's ``ND_TO_FLAT`` layout,
``operator/src/complex_vector.h:10-20``), so a hypercube with axes
(n1, n2, n3, n4) holds an array of numpy shape (n4, n3, n2, n1).

A vector lives on the host (numpy complex64) or on the GPU (torch complex64
tensor); operators accept either and copy host vectors through their own
device buffers, as ``CudaOperator<M,D>::forward`` does
(``operator/src/CudaOperator.h:74-100``).
"""
from __future__ import annotations

import copy

import numpy as np

__all__ = ["axis", "hypercube", "complexVector", "complex2DReg", "complex3DReg", "complex4DReg",
           "paramObj"]


class axis:
    def __init__(self, n: int = 1, o: float = 0.0, d: float = 1.0, label: str = "", unit: str = ""):
        self.n, self.o, self.d, self.label, self.unit = int(n), float(o), float(d), label, unit

    def __eq__(self, other):
        return isinstance(other, axis) and (self.n, self.o, self.d) == (other.n, other.o, other.d)

    def __repr__(self):
        return f"axis(n={self.n}, o={self.o}, d={self.d})"


class hypercube:
    """Axes 1..N, axis 1 fastest.  ``hypercube(axes=[...])`` or ``hypercube(ns=[...], os=[...], ds=[...])``."""

    def __init__(self, axes=None, ns=None, os=None, ds=None):
        if axes is None:
            ns = list(ns or [])
            os = list(os or [0.0] * len(ns))
            ds = list(ds or [1.0] * len(ns))
            axes = [axis(n, o, d) for n, o, d in zip(ns, os, ds)]
        self.axes = [copy.copy(a) for a in axes]

    def getAxis(self, i: int) -> axis:
        return self.axes[i - 1]

    @property
    def cppMode(self) -> "hypercube":
        """The object the compiled module (pyFWIX) receives, as SEP's ``.cppMode``."""
        return self

    def getAxes(self):
        return list(self.axes)

    def getNdim(self) -> int:
        return len(self.axes)

    def getNs(self):
        return [a.n for a in self.axes]

    def getN123(self) -> int:
        return int(np.prod(self.getNs(), dtype=np.int64))

    def clone(self) -> "hypercube":
        return hypercube(self.axes)

    @property
    def shape(self) -> tuple:
        """numpy shape (slowest axis first)."""
        return tuple(reversed(self.getNs()))

    def __eq__(self, other):
        return isinstance(other, hypercube) and self.axes == other.axes


class complexVector:
    """complex64 vector over a hypercube, on the host (numpy) or the GPU (torch)."""

    def __init__(self, hyper: hypercube, arr=None, device=None):
        self.hyper = hyper.clone()
        if arr is None:
            if device is None:
                arr = np.zeros(self.hyper.shape, np.complex64)
            else:
                import torch
                arr = torch.zeros(self.hyper.shape, dtype=torch.complex64, device=device)
        if tuple(arr.shape) != self.hyper.shape:
            raise ValueError(f"array shape {tuple(arr.shape)} does not match the hypercube {self.hyper.shape}")
        if isinstance(arr, np.ndarray):
            arr = np.ascontiguousarray(arr, np.complex64)
        self.arr = arr

    # --- SepVector-like accessors
    def getHyper(self) -> hypercube:
        return self.hyper

    @property
    def cppMode(self) -> "complexVector":
        """The object the compiled module (pyFWIX) receives, as SEP's ``.cppMode``."""
        return self

    def getNdArray(self):
        return self.arr

    @property
    def on_device(self) -> bool:
        return not isinstance(self.arr, np.ndarray)

    def zero(self) -> "complexVector":
        self.arr[...] = 0
        return self

    def set(self, value) -> "complexVector":
        self.arr[...] = value
        return self

    def clone(self) -> "complexVector":
        return complexVector(self.hyper, self.arr.copy() if not self.on_device else self.arr.clone())

    def cloneSpace(self) -> "complexVector":
        return complexVector(self.hyper, device=self.arr.device if self.on_device else None)

    def random(self, rng=None) -> "complexVector":
        rng = rng if rng is not None else np.random.default_rng()
        r = (rng.uniform(-1, 1, self.hyper.shape) + 1j * rng.uniform(-1, 1, self.hyper.shape)).astype(np.complex64)
        if self.on_device:
            import torch
            self.arr.copy_(torch.from_numpy(r))
        else:
            self.arr[...] = r
        return self

    def numpy(self) -> np.ndarray:
        return self.arr.cpu().numpy() if self.on_device else self.arr

    def dot(self, other: "complexVector") -> complex:
        """sum(conj(self) * other), accumulated in double precision."""
        a = self.numpy().astype(np.complex128).ravel()
        b = other.numpy().astype(np.complex128).ravel()
        return complex(np.vdot(a, b))

    def norm(self) -> float:
        return float(np.linalg.norm(self.numpy().astype(np.complex128).ravel()))

    def scale(self, s) -> "complexVector":
        self.arr *= s
        return self

    def scaleAdd(self, other: "complexVector", sc1=1.0, sc2=1.0) -> "complexVector":
        self.arr *= sc1
        self.arr += sc2 * other.arr
        return self

    def to(self, device) -> "complexVector":
        """A copy on ``device`` ("cpu" gives a host vector)."""
        import torch
        if device in (None, "cpu"):
            return complexVector(self.hyper, self.numpy().copy())
        t = torch.from_numpy(np.ascontiguousarray(self.numpy())).to(device) if not self.on_device \
            else self.arr.to(device, copy=True)
        return complexVector(self.hyper, t)


complex2DReg = complex3DReg = complex4DReg = complexVector


class paramObj:
    """Parameter lookup with defaults; a missing name without a default raises KeyError
    (as ``par->getInt("nref")`` does in ``propagator/src/RefSampler.cpp:16``)."""

    _missing = object()

    def __init__(self, params: dict | None = None, **kw):
        self.params = dict(params or {}, **kw)

    @property
    def cppMode(self) -> "paramObj":
        """The object the compiled module (pyFWIX) receives, as SEP's ``.cppMode``."""
        return self

    def _get(self, name, default, cast):
        if name in self.params:
            return cast(self.params[name])
        if default is self._missing:
            raise KeyError(f"parameter '{name}' is required")
        return default

    def getInt(self, name, default=_missing):
        return self._get(name, default, int)

    def getFloat(self, name, default=_missing):
        return self._get(name, default, float)

    def getString(self, name, default=_missing):
        return self._get(name, default, str)

    def getBool(self, name, default=_missing):
        return self._get(name, default, bool)
