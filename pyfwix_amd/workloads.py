"""Synthetic workloads of the benchmark configurations.

Each builder returns host-side numpy inputs (frequency-invariant slowness
s = 1/v^2 as ``fwix/utils.py:195-215`` prepares it, reference slownesses and
labels from :mod:`pyfwix_amd.refsampler`, geometry, wavelet, a model
perturbation) plus the grid description; nothing here touches the GPU.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .refsampler import select_refs_depth

__all__ = ["Workload", "born_workload", "ricker_spectrum", "CONFIGS"]


@dataclass
class Workload:
    name: str
    nx: int
    ny: int
    nz: int
    NX: int
    NY: int
    NW: int
    NS: int
    dx: float
    dy: float
    dz: float
    ow: float
    dw: float
    nref: int
    taper: int
    slow3: np.ndarray            # (nz, ny, nx) float64 slowness (frequency invariant)
    sref1: np.ndarray            # (nz, nref, 1) complex64
    labels1: np.ndarray          # (nz, 1, NY, NX) uint8
    src: tuple                   # (x, y, z, ids)
    rec: tuple
    wavelet: np.ndarray          # (nsrc, NW) complex64
    dslow3: np.ndarray           # (nz, ny, nx) float64 perturbation
    extra: dict = field(default_factory=dict)

    @property
    def points(self) -> int:     # wavefield points of one depth slice
        return self.NX * self.NY * self.NW * self.NS


def ricker_spectrum(f: np.ndarray, f0: float = 15.0, t0: float = 0.1) -> np.ndarray:
    """Ricker wavelet spectrum W(f) = (2/sqrt(pi)) f^2/f0^3 exp(-f^2/f0^2) exp(-i 2 pi f t0)."""
    return (2.0 / math.sqrt(math.pi)) * f ** 2 / f0 ** 3 * np.exp(-(f / f0) ** 2) * np.exp(-2j * np.pi * f * t0)


def _velocity(nx, ny, nz, dx, dy, dz):
    """v = 1500 + 0.8 z + 300 exp(-r^2 / (2 * 150^2)), r from the model centre (m/s)."""
    z = (np.arange(nz) * dz)[:, None, None]
    y = (np.arange(ny) * dy)[None, :, None]
    x = (np.arange(nx) * dx)[None, None, :]
    cx, cy, cz = (nx - 1) * dx / 2, (ny - 1) * dy / 2, (nz - 1) * dz / 2
    r2 = (x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2
    return 1500.0 + 0.8 * z + 300.0 * np.exp(-r2 / (2 * 150.0 ** 2))


def born_workload(name: str, nx: int, nz: int, nw: int, fmin: float, fmax: float, nref: int = 4,
                  taper: int = 20, d: float = 10.0, seed: int = 2, sigma: float = 3.0,
                  amp: float = 0.05, reflectors: bool = False) -> Workload:
    """Born forward + migration workload on an nx^2 x nz grid padded to 2nx.

    One Ricker 15 Hz source at the surface centre, receivers on every surface
    node.  The perturbation is a seeded Gaussian random field smoothed with
    ``sigma`` cells, ``amp`` * s of amplitude, or (``reflectors``) three flat
    +-5 % slowness reflectors at 1/4, 1/2, 3/4 of the depth.
    """
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    ny = nx
    NX = NY = 2 * nx                 # padx = nx, pady = ny (workers.py:51-52)
    v = _velocity(nx, ny, nz, d, d, d)
    s = 1.0 / v ** 2
    sref1 = np.empty((nz, nref, 1), np.complex64)
    labels1 = np.empty((nz, 1, NY, NX), np.uint8)
    for z in range(nz):
        sref1[z], labels1[z] = select_refs_depth(s[z][None].astype(np.complex64), nref, NX - nx, NY - ny)
    dw = (fmax - fmin) / (nw - 1) if nw > 1 else 1.0
    f = fmin + dw * np.arange(nw)
    wav = ricker_spectrum(f)[None].astype(np.complex64)
    src = (np.array([nx // 2 * d], np.float32), np.array([ny // 2 * d], np.float32),
           np.array([0.0], np.float32), np.array([0], np.int32))
    gx, gy = np.meshgrid(np.arange(nx) * d, np.arange(ny) * d)
    rec = (gx.ravel().astype(np.float32), gy.ravel().astype(np.float32),
           np.zeros(nx * ny, np.float32), np.zeros(nx * ny, np.int32))
    if reflectors:
        ds = np.zeros_like(s)
        for k, sign in zip((nz // 4, nz // 2, 3 * nz // 4), (1.0, -1.0, 1.0)):
            ds[k] = sign * 0.05 * s[k]
    else:
        g = gaussian_filter(rng.standard_normal((nz, ny, nx)), sigma)
        ds = amp * s * g / g.std()
    return Workload(name, nx, ny, nz, NX, NY, nw, 1, d, d, d, fmin, dw, nref, taper, s, sref1, labels1,
                    src, rec, wav, ds)


CONFIGS = {
    # plain PSPI shot migration (CPU-runnable size): 128^2 x 100, 32 freqs 5-30 Hz, nref 4
    "c1": dict(nx=128, nz=100, nw=32, fmin=5.0, fmax=30.0, nref=4, taper=20, seed=1, reflectors=True),
    # Born forward + adjoint: 256^2 x 200, 64 freqs 3-40 Hz, smoothed 5 % random perturbation
    "c2": dict(nx=256, nz=200, nw=64, fmin=3.0, fmax=40.0, nref=4, taper=20, seed=2),
}
