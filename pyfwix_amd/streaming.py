"""Shot / frequency batching with ginsu windows (the reference's
``StreamingPropagator`` and ``StreamingExtendedBorn``).

* sources are split by unique id, sorted, as evenly as possible, the first
  batches taking the remainder; a batch holds every source and receiver trace
  of its ids (``propagator/src/StreamingPropagator.cpp:68-153``);
* frequencies are split the same way into contiguous bands (``:155-169``);
* each (source batch, frequency band) runs on a lateral window of the model:
  the batch's source and receiver extent widened by ``ginsu_x/ginsu_y``,
  floor/ceil to model nodes and clamped, the whole model when degenerate
  (``:327-401``), padded by ``padx/pady`` and then up to the next supported
  FFT size;
* data and gradients of the batches are accumulated into the full arrays
  (``:449-464``, ``StreamingExtendedBorn.cpp:162-185``).

On one GPU the batches run one after another on the current stream, each
with its own operator (background wavefields kept in HBM).
"""
from __future__ import annotations

import math

import numpy as np

from . import _lib, ops
from .refsampler import select_refs

__all__ = ["split_sources", "split_frequencies", "ginsu_window", "next_fft_size", "StreamingPropagator",
           "StreamingExtendedBorn"]


def split_sources(s_ids, r_ids, nbatches: int):
    """-> list of (source trace indices, receiver trace indices) per batch."""
    s_ids = np.asarray(s_ids).ravel()
    r_ids = np.asarray(r_ids).ravel()
    uniq = np.unique(s_ids)
    nb = max(1, min(int(nbatches), len(uniq)))
    per, rem = divmod(len(uniq), nb)
    out, k = [], 0
    for b in range(nb):
        ids = uniq[k:k + per + (1 if b < rem else 0)]
        k += len(ids)
        out.append((np.flatnonzero(np.isin(s_ids, ids)), np.flatnonzero(np.isin(r_ids, ids))))
    return out


def split_frequencies(nw: int, nbatches: int):
    """-> (starts, sizes) of contiguous frequency bands."""
    nb = max(1, min(int(nbatches), nw))
    per, rem = divmod(nw, nb)
    sizes = [per + (1 if b < rem else 0) for b in range(nb)]
    starts = list(np.cumsum([0] + sizes[:-1]))
    return [int(s) for s in starts], sizes


def ginsu_window(xs, ys, ginsu_x, ginsu_y, n_x, n_y, ox, dx, oy, dy):
    """Model-node window (min_ix, max_ix, min_iy, max_iy) covering the points ± ginsu."""
    xs, ys = np.asarray(xs, np.float32), np.asarray(ys, np.float32)
    min_x, max_x = float(xs.min()) - ginsu_x, float(xs.max()) + ginsu_x
    min_y, max_y = float(ys.min()) - ginsu_y, float(ys.max()) + ginsu_y
    min_ix = max(0, int(math.floor((min_x - ox) / dx)))
    max_ix = min(n_x - 1, int(math.ceil((max_x - ox) / dx)))
    min_iy = max(0, int(math.floor((min_y - oy) / dy)))
    max_iy = min(n_y - 1, int(math.ceil((max_y - oy) / dy)))
    if max_ix <= min_ix:
        min_ix, max_ix = 0, n_x - 1
    if max_iy <= min_iy:
        min_iy, max_iy = 0, n_y - 1
    return min_ix, max_ix, min_iy, max_iy


def next_fft_size(n: int) -> int:
    """Smallest supported lateral FFT size >= n."""
    for m in range(max(8, n), 1025):
        if _lib.fft_size_supported(m):
            return m
    raise ValueError(f"no supported FFT size >= {n} (largest is 1024)")


class _Batch:
    pass


class StreamingPropagator:
    """Nonlinear modeling over (source batch x frequency band) windows.

    ``slow_axes``: ((nx, ox, dx), (ny, oy, dy), (nw, ow, dw), (nz, oz, dz));
    ``geometry`` holds sx, sy, sz, s_ids, rx, ry, rz, r_ids; ``wavelet``
    (nsrc, nw); ``params``: nref, padx, pady (>= 0), taperx/tapery, eps,
    eps_scat, taylor, ginsu_x/ginsu_y, ckpt_every, and pad_to_window (padding
    of each batch = its window's size, as ``fwix/workers.py:50-52`` sets it).
    """

    def __init__(self, slow_axes, wavelet, geometry, params: dict, nbatches=(1, 1), device=None):
        import torch
        from .CudaWEM import dense_shot_ids
        self.device = torch.device(device or "cuda:0")
        (nx, ox, dx), (ny, oy, dy), (nw, ow, dw), (nz, oz, dz) = slow_axes
        self.axes = slow_axes
        padx, pady = int(params.get("padx", -1)), int(params.get("pady", -1))
        if padx < 0 or pady < 0:
            raise RuntimeError("Error: negative padx/pady values")
        self.params = params
        self.nref = int(params.get("nref", 1))
        g = {k: np.asarray(v) for k, v in geometry.items()}
        wavelet = np.asarray(wavelet)
        self.nrec = len(g["rx"])
        self.nw = nw
        starts, sizes = split_frequencies(nw, nbatches[1])
        self.batches = []
        for si, ri in split_sources(g["s_ids"], g["r_ids"], nbatches[0]):
            pts_x = np.concatenate([g["sx"][si], g["rx"][ri]])
            pts_y = np.concatenate([g["sy"][si], g["ry"][ri]])
            x0, x1, y0, y1 = ginsu_window(pts_x, pts_y, float(params.get("ginsu_x", 0.0)),
                                          float(params.get("ginsu_y", 0.0)), nx, ny, ox, dx, oy, dy)
            wx, wy = x1 - x0 + 1, y1 - y0 + 1
            s_dense, r_dense, nshot = dense_shot_ids(g["s_ids"][si], g["r_ids"][ri])
            for f0, nf in zip(starts, sizes):
                b = _Batch()
                b.src, b.rec, b.f0, b.nf = si, ri, f0, nf
                b.x0, b.y0, b.wx, b.wy = x0, y0, wx, wy
                # pad_to_window: padding = the window's own size (workers.py:50-52)
                px, py = (wx, wy) if params.get("pad_to_window") else (padx, pady)
                NX, NY = next_fft_size(wx + px), next_fft_size(wy + py)
                b.dims = ops.make_dims(NX, NY, nf, nshot, wx, wy, ox + x0 * dx, dx, oy + y0 * dy, dy,
                                       ow + f0 * dw, dw, oz, dz, nz)
                b.op = ops.Born(b.dims, nref=self.nref, taperx=int(params.get("taperx", 0)),
                                tapery=int(params.get("tapery", 0)), eps=float(params.get("eps", 0.0)),
                                eps_scat=float(params.get("eps_scat", 0.04)), taylor=int(params.get("taylor", 1)),
                                ckpt_every=int(params.get("ckpt_every", 1)))
                b.op.set_sources(g["sx"][si], g["sy"][si], g["sz"][si], s_dense)
                b.op.set_receivers(g["rx"][ri], g["ry"][ri], g["rz"][ri], r_dense)
                b.wav = torch.from_numpy(np.ascontiguousarray(wavelet[si, f0:f0 + nf], np.complex64)).to(self.device)
                b.op.set_wavelet(b.wav)
                self.batches.append(b)

    def _window(self, b, arr):
        return np.ascontiguousarray(arr[:, b.f0:b.f0 + b.nf, b.y0:b.y0 + b.wy, b.x0:b.x0 + b.wx], np.complex64)

    def set_background_model(self, slow, den=None):
        """Window the model for every batch and model its background (down and up passes)."""
        import torch
        slow = np.asarray(slow)
        den = np.ones_like(slow) if den is None else np.asarray(den)
        for b in self.batches:
            s = self._window(b, slow)
            sref, labels = select_refs(s, self.nref, b.dims.NX - b.dims.nx, b.dims.NY - b.dims.ny)
            b.slow = torch.from_numpy(s).to(self.device)
            b.den = torch.from_numpy(self._window(b, den)).to(self.device)
            b.refs = (torch.from_numpy(sref).to(self.device), torch.from_numpy(labels).to(self.device))
            b.op.set_refs(*b.refs)
            b.op.set_slow(b.slow)
            b.op.set_density(b.den)

    def forward(self, add, slow, den, data: np.ndarray) -> np.ndarray:
        """data (+)= direct + reflected recordings, data (nrec, nw) on the host."""
        import torch
        if not add:
            data[...] = 0
        self.set_background_model(slow, den)
        for b in self.batches:
            d = torch.zeros((len(b.rec), b.nf), dtype=torch.complex64, device=self.device)
            b.op.model(d)
            b.op.model_up(d, add=True)
            data[b.rec[:, None], b.f0 + np.arange(b.nf)[None, :]] += d.cpu().numpy()
        return data


class StreamingExtendedBorn:
    """ExtendedBorn (all four channels) over the batches of a StreamingPropagator;
    model (ds, drho) each (nz, nw, ny, nx) on the host, data (nrec, nw)."""

    def __init__(self, prop: StreamingPropagator, slow, den=None):
        self.prop = prop
        prop.set_background_model(slow, den)
        for b in prop.batches:
            b.op.model()
            b.op.model_up()

    def forward(self, add, ds: np.ndarray, dr: np.ndarray, data: np.ndarray) -> np.ndarray:
        import torch
        p = self.prop
        if not add:
            data[...] = 0
        for b in p.batches:
            d = torch.zeros((len(b.rec), b.nf), dtype=torch.complex64, device=p.device)
            b.op.extended_forward(False, torch.from_numpy(p._window(b, ds)).to(p.device),
                                  torch.from_numpy(p._window(b, dr)).to(p.device), d)
            data[b.rec[:, None], b.f0 + np.arange(b.nf)[None, :]] += d.cpu().numpy()
        return data

    def adjoint(self, add, ds: np.ndarray, dr: np.ndarray, data: np.ndarray):
        import torch
        p = self.prop
        if not add:
            ds[...] = 0
            dr[...] = 0
        for b in p.batches:
            d = np.ascontiguousarray(data[b.rec][:, b.f0:b.f0 + b.nf], np.complex64)
            shape = (b.dims.nz, b.nf, b.wy, b.wx)
            gs = torch.zeros(shape, dtype=torch.complex64, device=p.device)
            gr = torch.zeros_like(gs)
            b.op.extended_adjoint(False, gs, gr, torch.from_numpy(d).to(p.device))
            sl = (slice(None), slice(b.f0, b.f0 + b.nf), slice(b.y0, b.y0 + b.wy), slice(b.x0, b.x0 + b.wx))
            ds[sl] += gs.cpu().numpy()
            dr[sl] += gr.cpu().numpy()
        return ds, dr
