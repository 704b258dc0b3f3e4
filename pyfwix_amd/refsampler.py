"""Reference-slowness selection: nref reference slownesses per (depth, frequency)
and the nearest-reference label of every lateral node.

Replaces the reference's ``RefSampler`` (``propagator/src/RefSampler.cpp:15-96``),
which runs OpenCV k-means (randomised, thread-dependent) per (z, w).  The
default selector is deterministic: the references are the slowness samples
at the real-part quantiles (r + 1/2)/nref, r = 0..nref-1, and each node takes
the nearest reference in the complex plane (lowest index on ties).
``method="kmeans"`` follows the reference's clustering instead, with a seed:
k-means++ seeding, Lloyd iterations stopped after 100 or when no center moves
by more than 1e-3 (``RefSampler.cpp:60-62``), one attempt.  Labels are
replicated into the padding from the last row / column exactly as
``RefSampler.cpp:69-86`` does.  Refs and labels are inputs fed identically to
the GPU operators and to the CPU oracle.
"""
from __future__ import annotations

import numpy as np

__all__ = ["select_refs", "select_refs_depth", "RefSampler"]


def kmeans_1slice(x: np.ndarray, nref: int, rng, max_iter: int = 100, eps: float = 1e-3) -> np.ndarray:
    """k-means centers of complex samples ``x`` (n,) -> (nref,) complex64.

    k-means++ seeding (each next center drawn with probability proportional
    to the squared distance to the nearest chosen one), then Lloyd updates
    until ``max_iter`` or the largest center shift is <= ``eps``; an empty
    cluster takes the sample farthest from its center."""
    x = np.asarray(x, np.complex128)
    n = x.size
    c = np.empty(nref, np.complex128)
    c[0] = x[rng.integers(n)]
    d2 = np.abs(x - c[0]) ** 2
    for k in range(1, nref):
        tot = d2.sum()
        c[k] = x[rng.integers(n)] if tot <= 0 else x[min(np.searchsorted(np.cumsum(d2), rng.random() * tot), n - 1)]
        d2 = np.minimum(d2, np.abs(x - c[k]) ** 2)
    for _ in range(max_iter):
        lab = np.argmin(np.abs(x[:, None] - c[None, :]), axis=1)
        new = c.copy()
        for k in range(nref):
            m = lab == k
            if m.any():
                new[k] = x[m].mean()
            else:
                far = np.argmax(np.abs(x - c[lab]))
                new[k] = x[far]
                lab[far] = k
        shift = np.max(np.abs(new - c))
        c = new
        if shift <= eps:
            break
    return c.astype(np.complex64)


def select_refs_depth(s: np.ndarray, nref: int, padx: int = 0, pady: int = 0, method: str = "quantile",
                      seed: int = 0):
    """One depth slice ``s`` (NW, ny, nx) -> (sref (nref, NW) c64, labels (NW, ny+pady, nx+padx) u8)."""
    if not 1 <= nref <= 255:
        raise ValueError("nref must be in [1, 255]")
    if method not in ("quantile", "kmeans"):
        raise ValueError("method must be 'quantile' or 'kmeans'")
    s = np.asarray(s)
    nw, ny, nx = s.shape
    # frequency-invariant slowness (the usual s = 1/v^2 replicated over w,
    # fwix/utils.py:211-215): select once and broadcast
    if nw > 1 and np.array_equal(s, np.broadcast_to(s[:1], s.shape)):
        r1, l1 = select_refs_depth(s[:1], nref, padx, pady, method, seed)
        return np.repeat(r1, nw, axis=1), np.repeat(l1, nw, axis=0)
    flat = s.reshape(nw, ny * nx).astype(np.complex64)
    n = flat.shape[1]
    if method == "kmeans":
        refs = np.stack([kmeans_1slice(flat[w], nref, np.random.default_rng([seed, w])) for w in range(nw)])
    else:
        k = np.round((np.arange(nref) + 0.5) / nref * (n - 1)).astype(np.int64)
        part = np.argpartition(flat.real, k, axis=1)
        idx = part[:, k]
        refs = np.take_along_axis(flat, idx, axis=1)                  # (NW, nref)
    best = np.full(flat.shape, np.inf, np.float32)
    lab = np.zeros(flat.shape, np.uint8)
    for r in range(nref):
        dist = np.abs(flat - refs[:, r:r + 1]).astype(np.float32)
        upd = dist < best
        best[upd] = dist[upd]
        lab[upd] = r
    lab = lab.reshape(nw, ny, nx)
    iy = np.minimum(np.arange(ny + pady), ny - 1)
    ix = np.minimum(np.arange(nx + padx), nx - 1)
    lab = lab[:, iy][:, :, ix]
    return np.ascontiguousarray(refs.T.astype(np.complex64)), np.ascontiguousarray(lab)


def select_refs(slow: np.ndarray, nref: int, padx: int = 0, pady: int = 0, method: str = "quantile",
                seed: int = 0):
    """Whole model ``slow`` (nz, NW, ny, nx) -> sref (nz, nref, NW) c64, labels (nz, NW, NY, NX) u8."""
    slow = np.asarray(slow)
    nz, nw, ny, nx = slow.shape
    sref = np.empty((nz, nref, nw), np.complex64)
    labels = np.empty((nz, nw, ny + pady, nx + padx), np.uint8)
    for z in range(nz):
        sref[z], labels[z] = select_refs_depth(slow[z], nref, padx, pady, method, seed + 7919 * z)
    return sref, labels


class RefSampler:
    """Same accessors as the reference's RefSampler (``propagator/src/RefSampler.h:29-36``)."""

    def __init__(self, slow: np.ndarray, nref: int, padx: int = 0, pady: int = 0, method: str = "quantile",
                 seed: int = 0):
        self.sref, self.labels = select_refs(slow, nref, padx, pady, method, seed)
        self.nref = nref

    def get_ref_slow(self, iz: int, iref: int) -> np.ndarray:
        return self.sref[iz, iref]

    def get_ref_labels(self, iz: int) -> np.ndarray:
        return self.labels[iz].astype(np.int32)
