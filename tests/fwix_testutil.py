"""Shared helpers of the test suite: seeded inputs, grids, tolerances."""
from __future__ import annotations

import numpy as np

import oracle as O

# FP32 device path vs FP64 oracle (relative L2 error on wavefields / images)
TOL_WFLD = 1e-4
# dot-product tests of adjoint pairs (the reference's --tolerance default,
# operator/tests/operator_unit_test.cpp:9)
TOL_DOT = 1e-5


def crand(rng, shape, dtype=np.complex128):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype)


def rel(a, b) -> float:
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    nb = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (nb if nb > 0 else 1.0))


def dot_err(m, Am, d, Atd) -> float:
    """|Re(<Am, d> / <m, A^H d>) - 1| as the reference's dotTest (CudaOperator.h:190-225)."""
    lhs = O.dot(Am, d)
    rhs = O.dot(m, Atd)
    return abs((lhs / rhs).real - 1.0)


def grid(NX, NY, NW, NS, nx=None, ny=None, nz=4, dx=10.0, dy=10.0, dz=10.0, ow=5.0, dw=2.0):
    return O.Grid(NX=NX, NY=NY, NW=NW, NS=NS, nx=nx or NX, ny=ny or NY, ox=0.0, dx=dx, oy=0.0,
                  dy=dy, ow=ow, dw=dw, oz=0.0, dz=dz, nz=nz)


def smooth_slowness(rng, g, imag=0.0, vmin=1500.0, vmax=3000.0):
    """Smooth random s = 1/v^2, (nz, NW, ny, nx), frequency-invariant unless imag != 0."""
    nz, ny, nx = g.nz, g.ny, g.nx
    z = np.linspace(0, 1, nz)[:, None, None]
    y = np.linspace(0, 1, ny)[None, :, None]
    x = np.linspace(0, 1, nx)[None, None, :]
    ph = rng.uniform(0, 2 * np.pi, 3)
    v = vmin + (vmax - vmin) * (0.5 + 0.25 * np.sin(2 * np.pi * x + ph[0]) * np.cos(np.pi * y + ph[1])
                                + 0.25 * z * np.cos(3 * x + ph[2]))
    s = 1.0 / v ** 2
    s = np.broadcast_to(s[:, None], (nz, g.NW, ny, nx)).astype(np.complex128)
    if imag:
        s = s + 1j * imag * s.real
    return np.ascontiguousarray(s)


def dims_of(g):
    from pyfwix_amd._lib import make_dims
    return make_dims(g.NX, g.NY, g.NW, g.NS, g.nx, g.ny, g.ox, g.dx, g.oy, g.dy, g.ow, g.dw,
                     g.oz, g.dz, g.nz)
