"""Shot batches: many shots of uneven aperture and frequency count in one launch.

The reference runs one shot geometry per operator and its streaming layer
takes batches one after another (``propagator/src/StreamingPropagator.cpp:68-153``),
so a small shot (a 192^2 padded grid, 50 frequencies) leaves most of a large
GPU idle and pays every launch on its own.  Here the shots of one padded
grid size are laid one after another along the plane axis of a single
:class:`~pyfwix_amd.ops.Born` operator (``set_shot_planes``), each plane with
its own frequency:

* shot k owns the planes ``[first_k, first_k + nw_k)``; every fused depth-step
  launch covers planes of many shots;
* slowness, references, labels and the model perturbation are stored once per
  shot (the kernels map a plane to its shot), not once per frequency;
* sources and receivers carry the shot index as their id; their traces are
  ragged (one sample per plane of the shot);
* the interior grid of the launch is the largest aperture of its shots; a
  smaller shot's slowness is edge-replicated up to it, which is exactly the
  padding it gets on its own (the scattering term clamps to the interior
  edge), its model perturbation is zero there, and its gradient is cut back to
  its aperture, so every shot's result equals the shot run alone.

:func:`plan_launches` groups shots by padded grid and packs each group into
launches that fit an HBM budget, never raising a shot's checkpoint interval
above what it needs on its own.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import ops

__all__ = ["Shot", "padded_size", "shot_refs", "launch_bytes", "min_ckpt", "plan_launches", "ShotBatch"]


@dataclass
class Shot:
    """One shot on its own lateral window (coordinates in metres, window-local)."""
    nx: int                      # aperture (interior nodes)
    ny: int
    freqs: np.ndarray            # (nw,) Hz, all nonzero
    src: np.ndarray              # (nsrc, 3) x, y, z
    wavelet: np.ndarray          # (nsrc, nw) complex64
    rec: np.ndarray              # (nrec, 3) x, y, z
    x0: int = 0                  # window origin in model nodes (placing results)
    y0: int = 0

    @property
    def nw(self) -> int:
        return len(self.freqs)


def padded_size(n: int) -> int:
    """Smallest supported lateral FFT size >= 2n (the padding rule padx = nx)."""
    from .streaming import next_fft_size
    return next_fft_size(2 * n)


def shot_refs(slow3: np.ndarray, nref: int, NX: int, NY: int):
    """Reference slownesses and labels of one shot's frequency-invariant slowness
    (nz, ny, nx): slowness quantiles (k + 1/2) / nref per depth and the nearest
    reference per node, labels edge-replicated to the padded grid -- the rule of
    :func:`pyfwix_amd.refsampler.select_refs_depth`, vectorised over depth.
    -> sref (nz, nref) complex64, labels (nz, NY, NX) uint8."""
    s = np.asarray(slow3)
    nz, ny, nx = s.shape
    flat = s.reshape(nz, ny * nx).astype(np.complex64)
    n = flat.shape[1]
    k = np.round((np.arange(nref) + 0.5) / nref * (n - 1)).astype(np.int64)
    idx = np.argpartition(flat.real, k, axis=1)[:, k]
    refs = np.take_along_axis(flat, idx, axis=1)                       # (nz, nref)
    best = np.full(flat.shape, np.inf, np.float32)
    lab = np.zeros(flat.shape, np.uint8)
    for r in range(nref):
        dist = np.abs(flat - refs[:, r:r + 1]).astype(np.float32)
        upd = dist < best
        best[upd] = dist[upd]
        lab[upd] = r
    lab = lab.reshape(nz, ny, nx)
    iy = np.minimum(np.arange(NY), ny - 1)
    ix = np.minimum(np.arange(NX), nx - 1)
    return np.ascontiguousarray(refs), np.ascontiguousarray(lab[:, iy][:, :, ix])


def launch_bytes(planes: int, nshot: int, NX: int, NY: int, nx: int, ny: int, nz: int, ckpt: int) -> int:
    """HBM held by one shot-batch launch (Born forward + adjoint): the background
    checkpoints (every ckpt-th depth) and the recomputed segment, three working
    wavefields, the chunk scratch (at most 8 GiB plus its spectra) and the
    chained-step spectrum of all planes, the per-plane gradient of one depth,
    and the per-shot slowness, model, gradient and labels."""
    slice_ = planes * NX * NY * 8
    nck = -(-nz // ckpt)
    seg = ckpt if ckpt > 1 else 0
    scratch = min(slice_ * 3, 9 << 30) + slice_
    model = nz * nshot * (3 * nx * ny * 8 + NX * NY + 4 * NY)
    return slice_ * (nck + seg + 3) + scratch + planes * nx * ny * 8 + model


def min_ckpt(planes, nshot, NX, NY, nx, ny, nz, budget) -> int | None:
    """Smallest checkpoint interval whose launch fits the budget (None: none does)."""
    for k in range(1, nz + 1):
        if launch_bytes(planes, nshot, NX, NY, nx, ny, nz, k) <= budget:
            return k
    return None


def plan_launches(shots, nz: int, budget: int):
    """-> list of (shot indices, NX, NY, ckpt_every).

    Shots are grouped by padded grid; inside a group, by decreasing frequency
    count, a shot joins the open launch when the launch still fits the budget
    at a checkpoint interval no larger than the largest the launch's shots
    (the new one included) need on their own; otherwise it opens a new launch.
    """
    groups = {}
    for i, s in enumerate(shots):
        groups.setdefault((padded_size(s.nx), padded_size(s.ny)), []).append(i)
    out = []
    for (NX, NY), idx in sorted(groups.items()):
        idx = sorted(idx, key=lambda i: -shots[i].nw)
        alone = {}
        for i in idx:
            k = min_ckpt(shots[i].nw, 1, NX, NY, shots[i].nx, shots[i].ny, nz, budget)
            if k is None:
                raise ValueError(f"shot {i} ({NX}x{NY}, {shots[i].nw} planes) does not fit the budget")
            alone[i] = k
        cur, kcur = [], 1
        for i in idx:
            if cur:
                trial = cur + [i]
                k = min_ckpt(sum(shots[j].nw for j in trial), len(trial), NX, NY,
                             max(shots[j].nx for j in trial), max(shots[j].ny for j in trial), nz, budget)
                if k is not None and k <= max(kcur, alone[i]):
                    cur, kcur = trial, k
                    continue
                out.append((cur, NX, NY, kcur))
            cur, kcur = [i], alone[i]
        if cur:
            out.append((cur, NX, NY, kcur))
    return out


class ShotBatch:
    """Born modeling (forward) and migration (adjoint) of a batch of shots that
    share one padded grid, in one operator.

    ``params``: nref, taper, eps, eps_scat, taylor.  ``nx``/``ny`` default to
    the largest aperture; the padded grid defaults to :func:`padded_size` of
    it.  Results come back per shot: data (nrec_k, nw_k), gradients
    (nz, ny_k, nx_k).
    """

    def __init__(self, shots, nz: int, dz: float, dx: float, dy: float, params: dict, ckpt_every: int = 1,
                 device=None, NX: int | None = None, NY: int | None = None, planes_per_chunk: int = 0):
        import torch
        self.device = torch.device(device or "cuda:0")
        self.shots = list(shots)
        if not self.shots:
            raise ValueError("a shot batch needs at least one shot")
        self.nz, self.dx, self.dy, self.dz = nz, dx, dy, dz
        self.nx = max(s.nx for s in self.shots)
        self.ny = max(s.ny for s in self.shots)
        self.NX = NX or padded_size(self.nx)
        self.NY = NY or padded_size(self.ny)
        if NX is None and any(padded_size(s.nx) != self.NX or padded_size(s.ny) != self.NY for s in self.shots):
            raise ValueError("shots of a batch must share one padded grid (or pass NX, NY)")
        self.nref = int(params.get("nref", 4))
        self.counts = np.array([s.nw for s in self.shots], np.int32)
        self.first = np.concatenate([[0], np.cumsum(self.counts)[:-1]]).astype(np.int64)
        self.planes = int(self.counts.sum())
        freqs = np.concatenate([np.asarray(s.freqs, np.float32) for s in self.shots])
        if np.any(freqs == 0):
            raise ValueError("Frequency is zero in the scattering operator!")
        # ow/dw are placeholders: the per-plane frequency table replaces them
        dims = ops.make_dims(self.NX, self.NY, self.planes, 1, self.nx, self.ny, 0.0, dx, 0.0, dy,
                             float(freqs[0]), 0.0, 0.0, dz, nz)
        taper = int(params.get("taper", 0))
        self.op = ops.Born(dims, nref=self.nref, taperx=taper, tapery=taper, eps=float(params.get("eps", 0.0)),
                           eps_scat=float(params.get("eps_scat", 0.04)), taylor=int(params.get("taylor", 1)),
                           ckpt_every=int(ckpt_every), planes_per_chunk=int(planes_per_chunk))
        self.op.set_shot_planes(self.counts)
        self.op.set_frequencies(freqs)
        self.freqs = freqs
        self._set_geometry()
        self.slow = self.dslow = self.grad = None

    def _set_geometry(self) -> None:
        import torch
        ids = lambda attr: np.concatenate([np.full(len(getattr(s, attr)), k, np.int32)  # noqa: E731
                                           for k, s in enumerate(self.shots)])
        src = np.concatenate([np.asarray(s.src, np.float32).reshape(-1, 3) for s in self.shots])
        rec = np.concatenate([np.asarray(s.rec, np.float32).reshape(-1, 3) for s in self.shots])
        self.op.set_sources(src[:, 0], src[:, 1], src[:, 2], ids("src"))
        self.op.set_receivers(rec[:, 0], rec[:, 1], rec[:, 2], ids("rec"))
        wav = np.concatenate([np.asarray(s.wavelet, np.complex64).reshape(-1) for s in self.shots])
        self.wavelet = torch.from_numpy(wav).to(self.device)
        self.op.set_wavelet(self.wavelet)
        # ragged trace layout: shot k's receiver block starts at rec_off[k]
        nrec = np.array([len(s.rec) for s in self.shots], np.int64)
        self.rec_off = np.concatenate([[0], np.cumsum(nrec * self.counts)[:-1]])
        self.nsamp = int((nrec * self.counts).sum())

    def set_shots(self, shots) -> None:
        """Rebind the batch to other shots of the same layout (aperture, frequencies
        and plane count per shot): new sources, receivers and wavelets; slowness,
        references and scratch are kept.  The background must be modeled again."""
        shots = list(shots)
        if len(shots) != len(self.shots):
            raise ValueError(f"expected {len(self.shots)} shots, got {len(shots)}")
        for a, b in zip(shots, self.shots):
            if (a.nx, a.ny) != (b.nx, b.ny) or a.nw != b.nw or not np.array_equal(
                    np.asarray(a.freqs, np.float32), np.asarray(b.freqs, np.float32)):
                raise ValueError("set_shots keeps the batch layout: same apertures and frequencies per shot")
        self.shots = shots
        self._set_geometry()

    # -------------------------------------------------------------- model inputs
    def _embed(self, arrs, pad: str):
        """Per-shot (nz, ny_k, nx_k) arrays -> (nz, nshot, ny, nx) complex64 on the host:
        edge replication (slowness) or zeros (model perturbation) beyond each aperture."""
        out = np.zeros((self.nz, len(self.shots), self.ny, self.nx), np.complex64)
        for k, (s, a) in enumerate(zip(self.shots, arrs)):
            a = np.asarray(a)
            if a.shape != (self.nz, s.ny, s.nx):
                raise ValueError(f"shot {k}: expected ({self.nz}, {s.ny}, {s.nx}), got {a.shape}")
            if pad == "edge":
                iy = np.minimum(np.arange(self.ny), s.ny - 1)
                ix = np.minimum(np.arange(self.nx), s.nx - 1)
                out[:, k] = a[:, iy][:, :, ix]
            else:
                out[:, k, :s.ny, :s.nx] = a
        return out

    def set_slowness(self, slows, refs=None) -> None:
        """Background slowness per shot, (nz, ny_k, nx_k) each (frequency invariant);
        ``refs``: the shots' :func:`shot_refs` results when already computed."""
        import torch
        sref = np.empty((self.nz, self.nref, len(self.shots)), np.complex64)
        labels = np.empty((self.nz, len(self.shots), self.NY, self.NX), np.uint8)
        for k, a in enumerate(slows):
            sref[:, :, k], labels[:, k] = refs[k] if refs is not None else shot_refs(a, self.nref, self.NX, self.NY)
        self.refs = (torch.from_numpy(sref).to(self.device), torch.from_numpy(labels).to(self.device))
        self.op.set_refs(*self.refs)
        self.slow = torch.from_numpy(self._embed(slows, "edge")).to(self.device)
        self.op.set_slow(self.slow)

    def set_perturbation(self, dslows) -> None:
        """Model perturbation per shot, (nz, ny_k, nx_k) each."""
        import torch
        self.dslow = torch.from_numpy(self._embed(dslows, "zero")).to(self.device)

    # -------------------------------------------------------------- operators
    def model(self) -> None:
        """Background down-going modeling of every shot (kept for forward / adjoint)."""
        self.op.model()

    def forward(self, data=None):
        """Born data of every shot into the ragged trace array (allocated when None)."""
        import torch
        if data is None:
            data = torch.zeros(self.nsamp, dtype=torch.complex64, device=self.device)
        self.op.forward(False, self.dslow, data)
        return data

    def adjoint(self, data, grad=None):
        """Migration: the model-space gradient (nz, nshot, ny, nx), summed over each shot's planes."""
        import torch
        if grad is None:
            grad = torch.zeros((self.nz, len(self.shots), self.ny, self.nx), dtype=torch.complex64,
                               device=self.device)
        self.op.adjoint(False, grad, data)
        return grad

    # -------------------------------------------------------------- per-shot views
    def shot_data(self, data, k: int):
        s = self.shots[k]
        o = int(self.rec_off[k])
        return data[o:o + len(s.rec) * s.nw].view(len(s.rec), s.nw)

    def shot_grad(self, grad, k: int):
        s = self.shots[k]
        return grad[:, k, :s.ny, :s.nx]

    def pack_data(self, per_shot):
        """Per-shot (nrec_k, nw_k) data -> the ragged trace array."""
        import torch
        return torch.cat([torch.as_tensor(d, device=self.device).reshape(-1) for d in per_shot]).to(torch.complex64)

    @property
    def points(self) -> int:
        """Wavefield points of one depth slice (all planes of all shots)."""
        return self.planes * self.NX * self.NY
