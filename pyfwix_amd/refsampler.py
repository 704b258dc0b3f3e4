"""Reference-slowness selection: nref reference slownesses per (depth, frequency)
and the nearest-reference label of every lateral node.

Replaces the reference's ``RefSampler`` (``propagator/src/RefSampler.cpp:15-96``),
which runs OpenCV k-means (randomised, thread-dependent) per (z, w).  This
selector is deterministic: the references are the slowness samples at the
real-part quantiles (r + 1/2)/nref, r = 0..nref-1, and each node takes the
nearest reference in the complex plane (lowest index on ties).  Labels are
replicated into the padding from the last row / column exactly as
``RefSampler.cpp:69-86`` does.  Refs and labels are inputs fed identically to
the GPU operators and to the CPU oracle.
"""
from __future__ import annotations

import numpy as np

__all__ = ["select_refs", "select_refs_depth", "RefSampler"]


def select_refs_depth(s: np.ndarray, nref: int, padx: int = 0, pady: int = 0):
    """One depth slice ``s`` (NW, ny, nx) -> (sref (nref, NW) c64, labels (NW, ny+pady, nx+padx) u8)."""
    if not 1 <= nref <= 255:
        raise ValueError("nref must be in [1, 255]")
    s = np.asarray(s)
    nw, ny, nx = s.shape
    # frequency-invariant slowness (the usual s = 1/v^2 replicated over w,
    # fwix/utils.py:211-215): select once and broadcast
    if nw > 1 and np.array_equal(s, np.broadcast_to(s[:1], s.shape)):
        r1, l1 = select_refs_depth(s[:1], nref, padx, pady)
        return np.repeat(r1, nw, axis=1), np.repeat(l1, nw, axis=0)
    flat = s.reshape(nw, ny * nx).astype(np.complex64)
    n = flat.shape[1]
    k = np.round((np.arange(nref) + 0.5) / nref * (n - 1)).astype(np.int64)
    part = np.argpartition(flat.real, k, axis=1)
    idx = part[:, k]
    refs = np.take_along_axis(flat, idx, axis=1)                      # (NW, nref)
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


def select_refs(slow: np.ndarray, nref: int, padx: int = 0, pady: int = 0):
    """Whole model ``slow`` (nz, NW, ny, nx) -> sref (nz, nref, NW) c64, labels (nz, NW, NY, NX) u8."""
    slow = np.asarray(slow)
    nz, nw, ny, nx = slow.shape
    sref = np.empty((nz, nref, nw), np.complex64)
    labels = np.empty((nz, nw, ny + pady, nx + padx), np.uint8)
    for z in range(nz):
        sref[z], labels[z] = select_refs_depth(slow[z], nref, padx, pady)
    return sref, labels


class RefSampler:
    """Same accessors as the reference's RefSampler (``propagator/src/RefSampler.h:29-36``)."""

    def __init__(self, slow: np.ndarray, nref: int, padx: int = 0, pady: int = 0):
        self.sref, self.labels = select_refs(slow, nref, padx, pady)
        self.nref = nref

    def get_ref_slow(self, iz: int, iref: int) -> np.ndarray:
        return self.sref[iz, iref]

    def get_ref_labels(self, iz: int) -> np.ndarray:
        return self.labels[iz].astype(np.int32)
