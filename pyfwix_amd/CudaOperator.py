"""Operator base classes with the reference's host semantics, and ``cuFFT2d``.

``CudaOperator`` mirrors ``operator/src/CudaOperator.h:74-225`` (host
``forward/adjoint/inverse``, the one-argument in-place forms, ``dotTest``)
and the Python base of ``fwix/CudaOperator.py:8-11``:

* ``add=False`` overwrites the output, ``add=True`` accumulates into it;
* host vectors are copied into the operator's device buffers, the HIP
  entry point runs, and the result is copied back; GPU vectors are used in
  place with no copies;
* ``dotTest`` returns ``|Re(conj(<d, A m>) / <m, A^H d>) - 1|`` for
  ``add=False`` and for ``add=True``.

Subclasses implement ``cu_forward/cu_adjoint(add, m, d)`` on torch device
tensors (and ``cu_inverse`` / the one-argument forms where the reference
has them).
"""
from __future__ import annotations

import numpy as np

from .sep import complexVector, hypercube

__all__ = ["Operator", "CudaOperator", "cuFFT2d"]


class Operator:
    """Duck-typed stand-in of ``pyOperator.Operator``: domain/range bookkeeping."""

    def setDomainRange(self, model, data):
        self.domain = model.cloneSpace() if hasattr(model, "cloneSpace") else model
        self.range = data.cloneSpace() if hasattr(data, "cloneSpace") else data

    def getDomain(self):
        return self.domain

    def getRange(self):
        return self.range


def _hyper(v) -> hypercube:
    return v.getHyper() if hasattr(v, "getHyper") else v


class CudaOperator(Operator):
    def __init__(self, model, data, device=None):
        import torch
        self.setDomainRange(model, data)
        self.device = torch.device(device or "cuda:0")
        self._bufs = {}

    # -- device side: implemented by subclasses on torch tensors
    def cu_forward(self, add, m, d):
        raise NotImplementedError

    def cu_adjoint(self, add, m, d):
        raise NotImplementedError

    def cu_inverse(self, add, m, d):
        raise RuntimeError("cu_inverse not implemented in the derived class.")

    def cu_forward_inplace(self, m):
        raise RuntimeError("cu_forward not implemented in the derived class.")

    def cu_adjoint_inplace(self, d):
        raise RuntimeError("cu_adjoint not implemented in the derived class.")

    # -- host side
    def _buf(self, which: str, hyper: hypercube):
        import torch
        b = self._bufs.get(which)
        if b is None or tuple(b.shape) != hyper.shape:
            b = torch.empty(hyper.shape, dtype=torch.complex64, device=self.device)
            self._bufs[which] = b
        return b

    def _stage(self, vec: complexVector, which: str, load: bool):
        if vec.on_device:
            return vec.arr
        b = self._buf(which, vec.getHyper())
        if load:
            import torch
            b.copy_(torch.from_numpy(vec.arr))
        return b

    @staticmethod
    def _unstage(vec: complexVector, dev):
        if not vec.on_device:
            vec.arr[...] = dev.cpu().numpy()

    def _apply(self, fn, add, model, data, out_is_model):
        m = self._stage(model, "model", load=not out_is_model or add)
        d = self._stage(data, "data", load=out_is_model or add)
        fn(bool(add), m, d)
        self._unstage(model if out_is_model else data, m if out_is_model else d)

    def forward(self, *args):
        """``forward(add, model, data)``, or ``forward(data)``: propagate ``data`` in place."""
        if len(args) == 1:
            (v,) = args
            dv = self._stage(v, "data", load=True)
            self.cu_forward_inplace(dv)
            self._unstage(v, dv)
            return
        add, model, data = args
        self._apply(self.cu_forward, add, model, data, out_is_model=False)

    def adjoint(self, *args):
        """``adjoint(add, model, data)``, or ``adjoint(model)``: back-propagate ``model`` in place."""
        if len(args) == 1:
            (v,) = args
            dv = self._stage(v, "model", load=True)
            self.cu_adjoint_inplace(dv)
            self._unstage(v, dv)
            return
        add, model, data = args
        self._apply(self.cu_adjoint, add, model, data, out_is_model=True)

    def inverse(self, add, model, data):
        self._apply(self.cu_inverse, add, model, data, out_is_model=True)

    def dotTest(self, verbose: bool = False, rng=None):
        rng = rng if rng is not None else np.random.default_rng()
        m = self.domain.cloneSpace().random(rng)
        d = self.range.cloneSpace().random(rng)
        d1 = self.range.cloneSpace()
        m1 = self.domain.cloneSpace()
        errs = []
        for add in (False, True):
            self.forward(add, m, d1)
            self.adjoint(add, m1, d)
            lhs = d.dot(d1).conjugate()   # <A m, d>
            rhs = m.dot(m1)               # <m, A^H d>
            errs.append(abs((lhs / rhs).real - 1.0))
            if verbose:
                print(f"add={add}: <m,A'd> {rhs} <Am,d> {lhs} error {errs[-1]:.3e}")
        return tuple(errs)


class cuFFT2d(CudaOperator):
    """Unitary 2-D FFT over axes 1-2 of a 4-D wavefield (``operator/src/FFT.cpp:5-69``)."""

    def __init__(self, model, data=None, device=None):
        from . import ops
        data = data if data is not None else model
        super().__init__(model, data, device)
        h = _hyper(model)
        nx, ny, nw, ns = (h.getAxis(i).n for i in (1, 2, 3, 4))
        self.cppMode = ops.FFT2d(ops.make_dims(nx, ny, nw, ns))

    def cu_forward(self, add, m, d):
        self.cppMode.forward(add, m, d)

    def cu_adjoint(self, add, m, d):
        self.cppMode.adjoint(add, m, d)
