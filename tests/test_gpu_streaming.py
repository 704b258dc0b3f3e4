"""Shot / frequency batching with ginsu windows (StreamingPropagator.cpp:68-464,
StreamingExtendedBorn.cpp:162-185) on the GPU:

* batched == unbatched when every window is the whole model
  (stream_unit_test.cpp:393-437), for the modeling and the four-channel Born;
* with ginsu windows, each batch equals the FP64 oracle run on its window,
  and the accumulated Born pair passes the dot test.
"""
import numpy as np
import pytest

import oracle as O
from fwix_testutil import TOL_DOT, TOL_WFLD, crand, rel, smooth_slowness, grid

pytestmark = pytest.mark.gpu


def _case(seed=0, nw=4):
    rng = np.random.default_rng(seed)
    g = grid(40, 24, nw, 1, nx=40, ny=24, nz=4, ow=6.0, dw=3.0)
    slow = smooth_slowness(rng, g).astype(np.complex64)
    den = (1.0 + 0.3 * rng.random(slow.shape)).astype(np.complex64)
    f32 = np.float32
    geom = {"sx": np.array([60.0, 300.0], f32), "sy": np.array([80.0, 150.0], f32), "sz": np.array([2.0, 2.0], f32),
            "s_ids": np.array([3, 9]), "rx": np.array([30.0, 80.0, 110.0, 270.0, 330.0, 350.0], f32),
            "ry": np.array([60.0, 100.0, 120.0, 130.0, 170.0, 200.0], f32), "rz": np.zeros(6, f32),
            "r_ids": np.array([3, 3, 3, 9, 9, 9])}
    wav = crand(rng, (2, nw)).astype(np.complex64)
    axes = ((g.nx, 0.0, g.dx), (g.ny, 0.0, g.dy), (nw, g.ow, g.dw), (g.nz, 0.0, g.dz))
    return g, slow, den, geom, wav, axes, rng


def _params(ginsu):
    return {"padx": 8, "pady": 8, "nref": 2, "taperx": 3, "tapery": 2, "ginsu_x": ginsu, "ginsu_y": ginsu}


def test_batches_equal_unbatched_on_full_windows(dev):
    from pyfwix_amd.streaming import StreamingExtendedBorn, StreamingPropagator
    g, slow, den, geom, wav, axes, rng = _case(1)
    outs = []
    ds = crand(rng, slow.shape).astype(np.complex64) * 1e-8
    dr = crand(rng, slow.shape).astype(np.complex64) * 0.05
    dd = crand(rng, (6, g.NW)).astype(np.complex64)
    for nb in ((1, 1), (2, 2), (2, 3)):
        p = StreamingPropagator(axes, wav, geom, _params(1e4), nb)
        d = np.zeros((6, g.NW), np.complex64)
        p.forward(False, slow, den, d)
        eb = StreamingExtendedBorn(p, slow, den)
        db = np.zeros_like(d)
        eb.forward(False, ds, dr, db)
        ms, mr = np.zeros_like(ds), np.zeros_like(dr)
        eb.adjoint(False, ms, mr, dd)
        outs.append((d, db, ms, mr))
    for o in outs[1:]:
        for a, b in zip(o, outs[0]):
            assert rel(a, b) < 1e-6


def test_ginsu_windows_match_the_oracle_per_batch(dev):
    from pyfwix_amd.refsampler import select_refs
    from pyfwix_amd.streaming import StreamingExtendedBorn, StreamingPropagator
    g, slow, den, geom, wav, axes, rng = _case(2)
    p = StreamingPropagator(axes, wav, geom, _params(20.0), (2, 2))
    assert any(b.wx < g.nx or b.wy < g.ny for b in p.batches)      # real windows
    d = np.zeros((6, g.NW), np.complex64)
    p.forward(False, slow, den, d)
    ref = np.zeros((6, g.NW), complex)
    for b in p.batches:
        gb = O.Grid(NX=b.dims.NX, NY=b.dims.NY, NW=b.nf, NS=b.dims.NS, nx=b.wx, ny=b.wy, ox=b.dims.ox, dx=g.dx,
                    oy=b.dims.oy, dy=g.dy, ow=b.dims.ow, dw=g.dw, oz=0.0, dz=g.dz, nz=g.nz)
        s = slow[:, b.f0:b.f0 + b.nf, b.y0:b.y0 + b.wy, b.x0:b.x0 + b.wx]
        rr = den[:, b.f0:b.f0 + b.nf, b.y0:b.y0 + b.wy, b.x0:b.x0 + b.wx]
        sref, lab = select_refs(s, 2, gb.NX - gb.nx, gb.NY - gb.ny)
        sids = np.searchsorted(np.unique(geom["s_ids"][b.src]), geom["s_ids"][b.src])
        rids = np.searchsorted(np.unique(geom["s_ids"][b.src]), geom["r_ids"][b.rec])
        src = (geom["sx"][b.src], geom["sy"][b.src], geom["sz"][b.src], sids)
        rec = (geom["rx"][b.rec], geom["ry"][b.rec], geom["rz"][b.rec], rids)
        w = wav[b.src][:, b.f0:b.f0 + b.nf].astype(complex)
        d0, bg = O.model_background(w, src, rec, sref.astype(complex), lab.astype(np.int64), gb, gb.dz, 0.0, 3, 2)
        du, _ = O.model_up(bg, s.astype(complex), rr.astype(complex), rec, sref.astype(complex),
                           lab.astype(np.int64), gb, gb.dz, 0.0, 3, 2)
        ref[b.rec[:, None], b.f0 + np.arange(b.nf)[None, :]] += d0 + du
    assert rel(d, ref) < TOL_WFLD
    # the accumulated four-channel Born pair is adjoint
    eb = StreamingExtendedBorn(p, slow, den)
    ds = crand(rng, slow.shape).astype(np.complex64)
    dr = crand(rng, slow.shape).astype(np.complex64)
    dd = crand(rng, (6, g.NW)).astype(np.complex64)
    Am = np.zeros_like(dd)
    eb.forward(False, ds, dr, Am)
    ms, mr = np.zeros_like(ds), np.zeros_like(dr)
    eb.adjoint(False, ms, mr, dd)
    lhs = np.vdot(Am.astype(complex), dd)
    rhs = np.vdot(ds.astype(complex), ms) + np.vdot(dr.astype(complex), mr)
    assert abs((lhs / rhs).real - 1) < TOL_DOT
