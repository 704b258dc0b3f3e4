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

__all__ = ["Workload", "born_workload", "ricker_spectrum", "CONFIGS", "SHOT_BATCH_CONFIGS", "uneven_shots",
           "shot_models", "FWI_CONFIGS", "fwi_models", "grid_shots"]


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
                  amp: float = 0.05, reflectors: bool = False, rec_depth: float = 0.0) -> Workload:
    """Born forward + migration workload on an nx^2 x nz grid padded to 2nx.

    One Ricker 15 Hz source at the surface centre, receivers on every node of
    the plane z = ``rec_depth`` (the surface by default).  The perturbation is
    a seeded Gaussian random field smoothed with ``sigma`` cells, ``amp`` * s
    of amplitude, or (``reflectors``) three flat +-5 % slowness reflectors at
    1/4, 1/2, 3/4 of the depth.

    The down-scattering Born channel records the scattered field where it
    arrives, below the scatterers: its data on the surface are zero, so a
    transmission workload puts the receivers on the deepest plane.
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
           np.full(nx * ny, rec_depth, np.float32), np.zeros(nx * ny, np.int32))
    extra = {}
    if reflectors:
        ds = np.zeros_like(s)
        refl = np.zeros_like(s)
        for k, sign in zip((nz // 4, nz // 2, 3 * nz // 4), (1.0, -1.0, 1.0)):
            ds[k] = sign * 0.05 * s[k]
            refl[k] = sign * 0.05          # reflectivity of the same three flat reflectors
        extra["refl3"] = refl
    else:
        g = gaussian_filter(rng.standard_normal((nz, ny, nx)), sigma)
        ds = amp * s * g / g.std()
    return Workload(name, nx, ny, nz, NX, NY, nw, 1, d, d, d, fmin, dw, nref, taper, s, sref1, labels1,
                    src, rec, wav, ds, extra)


def salt_velocity(nx, ny, nz, d, rng, layers: int = 8, vmin: float = 1500.0, vmax: float = 4500.0,
                  fluct: float = 0.03):
    """Layered vmin..vmax model + a vmax ellipsoid ("salt") + seeded fluctuation of ``fluct`` (m/s).

    Layers are equal-thickness, velocity rising linearly from vmin to vmax;
    the ellipsoid is centred laterally at 55 % of the depth with semi-axes
    (0.25 nx, 0.2 ny, 0.12 nz) cells; the fluctuation is a Gaussian random
    field smoothed over 2 cells, scaled to ``fluct`` relative standard
    deviation, and is not applied inside the salt.
    """
    from scipy.ndimage import gaussian_filter
    z = np.arange(nz)[:, None, None]
    y = np.arange(ny)[None, :, None]
    x = np.arange(nx)[None, None, :]
    layer = np.minimum(z * layers // nz, layers - 1)
    v = np.broadcast_to(vmin + (vmax - vmin) * layer / (layers - 1), (nz, ny, nx)).astype(np.float64)
    r = (((x - (nx - 1) / 2) / (0.25 * nx)) ** 2 + ((y - (ny - 1) / 2) / (0.2 * ny)) ** 2
         + ((z - 0.55 * (nz - 1)) / (0.12 * nz)) ** 2)
    salt = r <= 1.0
    g = gaussian_filter(rng.standard_normal((nz, ny, nx)), 2.0)
    v = v * (1.0 + fluct * g / g.std())
    return np.where(salt, vmax, v)


def nsps_workload(name: str, nx: int, nz: int, nw: int, fmin: float, fmax: float, nref: int = 8,
                  taper: int = 20, d: float = 10.0, seed: int = 3) -> Workload:
    """One-way propagation in a strongly heterogeneous model (salt_velocity) on an
    nx^2 x nz grid padded to 2nx; one Ricker 15 Hz source at the surface centre."""
    rng = np.random.default_rng(seed)
    ny = nx
    NX = NY = 2 * nx
    s = 1.0 / salt_velocity(nx, ny, nz, d, rng) ** 2
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
    return Workload(name, nx, ny, nz, NX, NY, nw, 1, d, d, d, fmin, dw, nref, taper, s, sref1, labels1,
                    src, rec, wav, np.zeros_like(s))


CONFIGS = {
    # plain PSPI shot migration (CPU-runnable size): 128^2 x 100, 32 freqs 5-30 Hz, nref 4
    "c1": dict(nx=128, nz=100, nw=32, fmin=5.0, fmax=30.0, nref=4, taper=20, seed=1, reflectors=True),
    # Born forward + adjoint: 256^2 x 200, 64 freqs 3-40 Hz, smoothed 5 % random perturbation,
    # receivers on the deepest plane (transmission)
    "c2": dict(nx=256, nz=200, nw=64, fmin=3.0, fmax=40.0, nref=4, taper=20, seed=2, rec_depth=1990.0),
}
# NSPS propagation in a layered + salt model: 384^2 x 300, 100 freqs 2-50 Hz, nref 8
NSPS_CONFIGS = {
    "c3": dict(nx=384, nz=300, nw=100, fmin=2.0, fmax=50.0, nref=8, taper=20, seed=3),
}

# Batched multi-shot modeling + migration: 64 shots, square apertures U{96..512}
# nodes padded to the next supported size >= 2n, nz 250, U{50..200} frequencies
# per shot over 3-40 Hz, one Ricker 15 Hz source at the aperture centre,
# receivers every 4th node on the deepest plane
SHOT_BATCH_CONFIGS = {
    "c4": dict(nshots=64, nmin=96, nmax=512, wmin=50, wmax=200, fmin=3.0, fmax=40.0, nz=250, d=10.0,
               rec_step=4, nref=4, taper=20, seed=4),
}


def uneven_shots(nshots, nmin, nmax, wmin, wmax, fmin, fmax, nz, d, rec_step, seed, **_):
    """Shots of uneven aperture and frequency count (window-local coordinates)."""
    from .batching import Shot
    rng = np.random.default_rng(seed)
    out = []
    for _k in range(nshots):
        n = int(rng.integers(nmin, nmax + 1))
        nw = int(rng.integers(wmin, wmax + 1))
        f = (fmin + (fmax - fmin) * np.arange(nw) / max(nw - 1, 1)).astype(np.float32)
        src = np.array([[n // 2 * d, n // 2 * d, 0.0]], np.float32)
        gx, gy = np.meshgrid(np.arange(0, n, rec_step) * d, np.arange(0, n, rec_step) * d)
        rec = np.stack([gx.ravel(), gy.ravel(), np.full(gx.size, (nz - 1) * d)], 1).astype(np.float32)
        wav = ricker_spectrum(f.astype(np.float64))[None].astype(np.complex64)
        out.append(Shot(n, n, f, src, wav, rec))
    return out


def shot_models(shot, nz, d):
    """Background slowness (the smooth family of the other workloads on the
    shot's window) and a model perturbation of three flat +-5 % reflectors at
    1/4, 1/2, 3/4 of the depth; both (nz, ny, nx) float32."""
    v = _velocity(shot.nx, shot.ny, nz, d, d, d)
    s = (1.0 / v ** 2).astype(np.float32)
    del v
    ds = np.zeros_like(s)
    for k, sign in zip((nz // 4, nz // 2, 3 * nz // 4), (1.0, -1.0, 1.0)):
        ds[k] = sign * 0.05 * s[k]
    return s, ds


# FWI-style slowness gradient: 512^2 x 400 at 10 m, 16 shots on a 4 x 4 surface
# grid, 128 frequencies over 2-40 Hz, checkpoints every 16 depths; true model
# = smooth background + seeded anomalies, start model = the true model
# Gaussian-smoothed with sigma = 20 cells (closed form for these Gaussians)
FWI_CONFIGS = {
    "c5": dict(nx=512, nz=400, nw=128, fmin=2.0, fmax=40.0, shots_per_side=4, ckpt=16, nref=4, taper=20,
               rec_step=4, d=10.0, smooth=20.0, anomalies=6, seed=5),
}


def _gauss_blobs(x, y, z, blobs, widen=0.0):
    """Sum of 3-D Gaussians a * exp(-r^2 / (2 s^2)); with ``widen`` (metres) each is
    convolved with a unit-mass Gaussian of that width: s^2 -> s^2 + widen^2 and
    a -> a * (s / s')^3 (exact on an unbounded grid)."""
    v = 0.0
    for (cx, cy, cz, a, sg) in blobs:
        s2 = sg * sg + widen * widen
        amp = a * (sg * sg / s2) ** 1.5
        v = v + amp * np.exp(-((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2) / (2 * s2))
    return v


def fwi_models(nx, nz, d, smooth, anomalies, seed, **_):
    """(true, start) slowness (nz, nx, nx) float32 of the FWI workload.

    Background: v = 1500 + 0.8 z + 300 exp(-r^2 / (2 * 150^2)) about the model
    centre; true = background + ``anomalies`` seeded Gaussians (+-100 m/s,
    100-300 m wide); start = true smoothed by a Gaussian of ``smooth`` cells
    (the linear trend is unchanged by smoothing, the Gaussians widen in closed
    form)."""
    rng = np.random.default_rng(seed)
    ny = nx
    z = (np.arange(nz) * d)[:, None, None]
    y = (np.arange(ny) * d)[None, :, None]
    x = (np.arange(nx) * d)[None, None, :]
    c = ((nx - 1) * d / 2, (ny - 1) * d / 2, (nz - 1) * d / 2)
    bg = [(c[0], c[1], c[2], 300.0, 150.0)]
    blobs = [(rng.uniform(0.15, 0.85) * (nx - 1) * d, rng.uniform(0.15, 0.85) * (ny - 1) * d,
              rng.uniform(0.2, 0.8) * (nz - 1) * d, rng.choice([-1.0, 1.0]) * 100.0, rng.uniform(100.0, 300.0))
             for _k in range(anomalies)]
    w = smooth * d
    v_true = 1500.0 + 0.8 * z + _gauss_blobs(x, y, z, bg + blobs)
    v_start = 1500.0 + 0.8 * z + _gauss_blobs(x, y, z, bg + blobs, widen=w)
    return (1.0 / v_true ** 2).astype(np.float32), (1.0 / v_start ** 2).astype(np.float32)


def grid_shots(nx, nz, nw, fmin, fmax, shots_per_side, rec_step, d, **_):
    """Shots on a regular surface grid over the whole nx^2 model (every shot's
    window is the model), receivers every ``rec_step`` nodes on the deepest plane."""
    from .batching import Shot
    f = (fmin + (fmax - fmin) * np.arange(nw) / max(nw - 1, 1)).astype(np.float32)
    wav = ricker_spectrum(f.astype(np.float64))[None].astype(np.complex64)
    gx, gy = np.meshgrid(np.arange(0, nx, rec_step) * d, np.arange(0, nx, rec_step) * d)
    rec = np.stack([gx.ravel(), gy.ravel(), np.full(gx.size, (nz - 1) * d)], 1).astype(np.float32)
    pos = (np.arange(shots_per_side) + 0.5) * nx / shots_per_side * d
    return [Shot(nx, nx, f, np.array([[sx, sy, 0.0]], np.float32), wav, rec)
            for sy in pos for sx in pos]


def workload_dims(w: Workload):
    """fwix_dims of a workload (origins at 0)."""
    from ._lib import make_dims
    return make_dims(w.NX, w.NY, w.NW, w.NS, w.nx, w.ny, 0.0, w.dx, 0.0, w.dy, w.ow, w.dw, 0.0, w.dz, w.nz)


def device_inputs(w: Workload, dev):
    """Upload a workload and expand its frequency-invariant arrays over NW on the GPU.

    Returns slowness and dslow (nz, NW, ny, nx) complex64, reference
    slownesses (nz, nref, NW), labels (nz, NW, NY, NX) uint8 and the
    wavelet (nsrc, NW).
    """
    import torch
    nz, NW = w.nz, w.NW
    s3 = torch.from_numpy(w.slow3.astype(np.float32)).to(dev)
    slow = torch.empty((nz, NW, w.ny, w.nx), dtype=torch.complex64, device=dev)
    slow.real.copy_(s3[:, None].expand(nz, NW, w.ny, w.nx))
    slow.imag.zero_()
    d3 = torch.from_numpy(w.dslow3.astype(np.float32)).to(dev)
    ds = torch.empty_like(slow)
    ds.real.copy_(d3[:, None].expand(nz, NW, w.ny, w.nx))
    ds.imag.zero_()
    sref = torch.from_numpy(w.sref1).to(dev).expand(nz, w.nref, NW).contiguous()
    labels = torch.from_numpy(w.labels1).to(dev).expand(nz, NW, w.NY, w.NX).contiguous()
    wav = torch.from_numpy(w.wavelet).to(dev)
    del s3, d3
    return slow, ds, sref, labels, wav
