"""GPU parity of injection, scattering, imaging condition, depth recursions and
the Born operator against the FP64 oracle.

Pins besides the oracle: dot-product tests at 1e-5 for every adjoint pair
(propagator/tests/born_unit_test.cpp:19-307, prop_unit_test.cpp:199-340),
checkpoint/recompute == store-all, and run-to-run bitwise determinism.
"""
import numpy as np
import pytest

import oracle as O
from fwix_testutil import TOL_DOT, TOL_WFLD, crand, dims_of, dot_err, grid, rel, smooth_slowness

pytestmark = pytest.mark.gpu


def _t(x, dev):
    import torch
    if x.dtype == np.uint8:
        return torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    return torch.from_numpy(np.ascontiguousarray(x.astype(np.complex64))).to(dev)


def _z(shape, dev):
    import torch
    return torch.zeros(shape, dtype=torch.complex64, device=dev)


def _geometry(g, rng, nrec=40, nsrc=None):
    nsrc = nsrc or g.NS
    sx = rng.uniform(g.ox + 2 * g.dx, g.ox + (g.nx - 3) * g.dx, nsrc)
    sy = rng.uniform(g.oy + 2 * g.dy, g.oy + (g.ny - 3) * g.dy, nsrc)
    sz = np.full(nsrc, g.oz + 0.3 * g.dz)
    sid = np.arange(nsrc) % g.NS
    rx = rng.uniform(g.ox, g.ox + (g.nx - 1) * g.dx, nrec)
    ry = rng.uniform(g.oy, g.oy + (g.ny - 1) * g.dy, nrec)
    rz = rng.uniform(g.oz, g.oz + 2.5 * g.dz, nrec)
    # a few receivers on shared nodes, one outside the grid
    rx[:4] = g.ox + 3 * g.dx
    ry[:4] = g.oy + 5 * g.dy
    rx[-1] = g.ox - 5 * g.dx
    rid = np.arange(nrec) % g.NS
    f32 = np.float32
    return (sx.astype(f32), sy.astype(f32), sz.astype(f32), sid), (rx.astype(f32), ry.astype(f32),
                                                                  rz.astype(f32), rid)


def test_injection_matches_oracle(dev):
    from pyfwix_amd.ops import Injection
    rng = np.random.default_rng(31)
    g = grid(48, 32, 5, 2, nx=40, ny=24, nz=4)
    _, (rx, ry, rz, rid) = _geometry(g, rng)
    op = Injection(dims_of(g), rx, ry, rz, rid)
    tr = crand(rng, (len(rx), g.NW))
    wf = crand(rng, (g.NS, g.NW, g.NY, g.NX))
    shape = (g.NS, g.NW, g.NY, g.NX)
    for iz in range(4):
        op.set_depth(iz)
        out = _t(wf, dev)
        op.forward(True, _t(tr, dev), out)
        ref = O.injection_forward(tr, rx, ry, rz, rid, g, iz, wfld=wf, add=True, shape=shape)
        assert rel(out.cpu().numpy(), ref) < 1e-6
        t2 = _z((len(rx), g.NW), dev)
        op.adjoint(False, t2, _t(wf, dev))
        ref = O.injection_adjoint(wf, rx, ry, rz, rid, g, iz)
        assert rel(t2.cpu().numpy(), ref) < 1e-6
        # dot test
        Ax = _z(shape, dev)
        op.forward(False, _t(tr, dev), Ax)
        assert dot_err(tr, Ax.cpu().numpy(), wf, t2.cpu().numpy()) < TOL_DOT


def test_injection_edges_and_empty(dev):
    """Edge cases of injection.cu:38-75: a trace on the last interpolating node
    (ix = NX-2, iy = NY-2), traces one node further (skipped), above the first
    level or below the last (only their in-range level is used); then traces
    that are all skipped, and no traces at all."""
    import torch
    from pyfwix_amd.ops import Injection
    rng = np.random.default_rng(43)
    g = grid(32, 24, 3, 2, nx=28, ny=20, nz=3)
    shape = (g.NS, g.NW, g.NY, g.NX)
    f32 = np.float32
    cx = np.array([g.ox + (g.NX - 1.5) * g.dx, g.ox + (g.NX - 0.9) * g.dx, g.ox + 3.2 * g.dx,
                   g.ox + 4.7 * g.dx, g.ox + 5.1 * g.dx, g.ox - 0.2 * g.dx], f32)
    cy = np.array([g.oy + (g.NY - 1.6) * g.dy, g.oy + 2.0 * g.dy, g.oy + (g.NY - 0.5) * g.dy,
                   g.oy + 6.3 * g.dy, g.oy + 7.9 * g.dy, g.oy + 1.0 * g.dy], f32)
    cz = np.array([g.oz + 0.4 * g.dz, g.oz + 1.2 * g.dz, g.oz, g.oz - 0.3 * g.dz,
                   g.oz + (g.nz - 0.6) * g.dz, g.oz + 0.5 * g.dz], f32)
    ids = np.array([0, 1, 0, 1, 0, 1], np.int32)
    cases = {"edges": (cx, cy, cz, ids),
             "all skipped": (cx[[1, 2, 5]], cy[[1, 2, 5]], cz[[1, 2, 5]], ids[[1, 2, 5]]),
             "no traces": (cx[:0], cy[:0], cz[:0], ids[:0])}
    for name, (x, y, z, i) in cases.items():
        op = Injection(dims_of(g), x, y, z, i)
        tr = crand(rng, (len(x), g.NW))
        wf = crand(rng, shape)
        for iz in range(g.nz):
            op.set_depth(iz)
            out = _t(wf, dev)
            op.forward(True, _t(tr, dev), out)
            ref = O.injection_forward(tr, x, y, z, i, g, iz, wfld=wf, add=True, shape=shape)
            assert rel(out.cpu().numpy(), ref) < 1e-6, (name, iz)
            if name != "edges":
                assert torch.equal(out, _t(wf, dev)), (name, iz)  # nothing injected, nothing touched
            out = _t(wf, dev)
            op.forward(False, _t(tr, dev), out)
            assert rel(out.cpu().numpy(), O.injection_forward(tr, x, y, z, i, g, iz, shape=shape)) < 1e-6, (name, iz)
            t2 = _t(crand(rng, (len(x), g.NW)), dev)
            op.adjoint(False, t2, _t(wf, dev))
            ref = O.injection_adjoint(wf, x, y, z, i, g, iz)
            if name == "edges":
                assert rel(t2.cpu().numpy(), ref) < 1e-6, (name, iz)
            else:
                assert not t2.cpu().numpy().any(), (name, iz)


@pytest.mark.parametrize("taylor", [0, 1, 2])
def test_scatter_matches_oracle(dev, taylor):
    from pyfwix_amd.ops import Scatter
    rng = np.random.default_rng(7 + taylor)
    g = grid(64, 48, 4, 2, nx=48, ny=40, nz=3)
    slow = smooth_slowness(rng, g, imag=0.01)
    op = Scatter(dims_of(g), taylor=taylor, eps_scat=0.04)
    op.set_slow(_t(slow, dev))
    op.set_depth(2)
    shape = (g.NS, g.NW, g.NY, g.NX)
    x, y = crand(rng, shape), crand(rng, shape)
    out = _z(shape, dev)
    op.forward(False, _t(x, dev), out)
    ref = O.scatter_forward(x, slow[2].astype(np.complex64).astype(np.complex128), g, g.dz, 0.04, taylor)
    assert rel(out.cpu().numpy(), ref) < TOL_WFLD
    m = _t(x, dev)
    op.adjoint(True, m, _t(y, dev))
    ref = x + O.scatter_adjoint(y, slow[2].astype(np.complex64).astype(np.complex128), g, g.dz, 0.04, taylor)
    assert rel(m.cpu().numpy(), ref) < TOL_WFLD
    Ax, Aty = _z(shape, dev), _z(shape, dev)
    op.forward(False, _t(x, dev), Ax)
    op.adjoint(False, Aty, _t(y, dev))
    assert dot_err(x, Ax.cpu().numpy(), y, Aty.cpu().numpy()) < TOL_DOT


def test_imaging_condition_matches_oracle(dev):
    from pyfwix_amd.ops import ImagingCondition
    rng = np.random.default_rng(17)
    g = grid(32, 24, 3, 3, nx=20, ny=16)
    shape = (g.NS, g.NW, g.NY, g.NX)
    U = crand(rng, shape)
    img = crand(rng, (g.NW, g.ny, g.nx))
    D = crand(rng, shape)
    op = ImagingCondition(dims_of(g))
    Ud = _t(U, dev)
    op.set_background(Ud)
    out = _t(D, dev)
    op.forward(True, _t(img, dev), out)
    assert rel(out.cpu().numpy(), O.ic_forward(img, U, d=D, add=True)) < 1e-6
    out0 = _t(D, dev)
    op.forward(False, _t(img, dev), out0)
    assert rel(out0.cpu().numpy(), O.ic_forward(img, U)) < 1e-6
    im = _z((g.NW, g.ny, g.nx), dev)
    op.adjoint(False, im, _t(D, dev))
    assert rel(im.cpu().numpy(), O.ic_adjoint(D, U, g.nx, g.ny)) < 1e-6
    assert dot_err(img, out0.cpu().numpy(), D, im.cpu().numpy()) < TOL_DOT


def _refs(g, rng, nref, imag=0.0):
    from pyfwix_amd.refsampler import select_refs
    slow = smooth_slowness(rng, g, imag=imag)
    sref, labels = select_refs(slow, nref, g.NX - g.nx, g.NY - g.ny)
    return slow, sref, labels


@pytest.mark.parametrize("direction", [0, 1])
@pytest.mark.parametrize("method", ["pspi", "nsps"])
def test_oneway_matches_oracle(dev, direction, method):
    from pyfwix_amd.ops import FWIX_NSPS, FWIX_PSPI, OneWay
    rng = np.random.default_rng(41 + direction)
    g = grid(48, 64, 3, 2, nx=36, ny=48, nz=5)
    slow, sref, labels = _refs(g, rng, 3)
    tap = 6 if method == "pspi" else 0
    op = OneWay(direction, FWIX_PSPI if method == "pspi" else FWIX_NSPS, dims_of(g), nref=3,
                taperx=tap, tapery=tap)
    op.set_refs(_t(sref, dev), _t(labels, dev))
    shape = (g.NS, g.NW, g.NY, g.NX)
    x, y = crand(rng, shape), crand(rng, shape)
    s64, lab = sref.astype(np.complex128), labels.astype(np.int64)
    out = _z(shape, dev)
    ck = _z((g.nz, *shape), dev) if direction == 0 else None
    op.forward(False, _t(x, dev), out, ck)
    f = O.downward_forward if direction == 0 else O.upward_forward
    b = O.downward_adjoint if direction == 0 else O.upward_adjoint
    cks = [] if direction == 0 else None
    if direction == 0:
        ref = f(x, s64, lab, g, g.dz, 0.0, tap, tap, method, checkpoints=cks)
    else:
        ref = f(x, s64, lab, g, g.dz, 0.0, tap, tap, method)
    assert rel(out.cpu().numpy(), ref) < TOL_WFLD
    if direction == 0:
        assert rel(ck.cpu().numpy(), np.stack(cks)) < TOL_WFLD
    Aty = _z(shape, dev)
    op.adjoint(False, Aty, _t(y, dev))
    assert rel(Aty.cpu().numpy(), b(y, s64, lab, g, g.dz, 0.0, tap, tap, method)) < TOL_WFLD
    assert dot_err(x, out.cpu().numpy(), y, Aty.cpu().numpy()) < TOL_DOT


def _born_setup(dev, ckpt=1, seed=5, NS=2, tap=4, nz=6, size=(48, 48, 4, 32, 36), ow=6.0, chunk=0):
    from pyfwix_amd.ops import Born
    rng = np.random.default_rng(seed)
    NX, NY, NW, nx, ny = size
    g = grid(NX, NY, NW, NS, nx=nx, ny=ny, nz=nz, ow=ow, dw=3.0)
    slow, sref, labels = _refs(g, rng, 3)
    src, rec = _geometry(g, rng, nrec=48)
    born = Born(dims_of(g), nref=3, taperx=tap, tapery=tap, taylor=1, eps_scat=0.04, ckpt_every=ckpt,
                planes_per_chunk=chunk)
    born.set_refs(_t(sref, dev), _t(labels, dev))
    born.set_slow(_t(slow, dev))
    born.set_sources(*src)
    born.set_receivers(*rec)
    wav = crand(rng, (len(src[0]), g.NW))
    born.set_wavelet(_t(wav, dev))
    return g, born, slow, sref, labels, src, rec, wav, rng


def test_born_background_and_forward_match_oracle(dev):
    g, born, slow, sref, labels, src, rec, wav, rng = _born_setup(dev)
    s64, lab, sl = sref.astype(np.complex128), labels.astype(np.int64), slow.astype(np.complex64).astype(np.complex128)
    d0 = _z((len(rec[0]), g.NW), dev)
    born.model(d0)
    ref_d0, bg = O.model_background(wav, src, rec, s64, lab, g, g.dz, 0.0, 4, 4)
    assert rel(d0.cpu().numpy(), ref_d0) < TOL_WFLD
    ds = crand(rng, (g.nz, g.NW, g.ny, g.nx)) * 1e-7
    d = _z((len(rec[0]), g.NW), dev)
    born.forward(False, _t(ds, dev), d)
    ref = O.born_forward(ds, bg, sl, rec, s64, lab, g, g.dz, 0.0, 4, 4, 0.04, 1)
    assert rel(d.cpu().numpy(), ref) < TOL_WFLD
    dres = crand(rng, (len(rec[0]), g.NW))
    m = _z((g.nz, g.NW, g.ny, g.nx), dev)
    born.adjoint(False, m, _t(dres, dev))
    ref_m = O.born_adjoint(dres, bg, sl, rec, s64, lab, g, g.dz, 0.0, 4, 4, 0.04, 1)
    assert rel(m.cpu().numpy(), ref_m) < TOL_WFLD
    assert dot_err(ds, d.cpu().numpy(), dres, m.cpu().numpy()) < TOL_DOT
    img = _z((g.nz, g.ny, g.nx), dev)
    born.migrate(False, img, _t(dres, dev))
    assert rel(img.cpu().numpy(), O.migrate(dres, bg, rec, s64, lab, g, g.dz, 0.0, 4, 4)) < TOL_WFLD


def _born_vs_oracle(dev, size, ow, fp32_floor=False, precision=0):
    g, born, slow, sref, labels, src, rec, wav, rng = _born_setup(dev, NS=1, tap=12, nz=3, size=size, seed=9, ow=ow)
    born.set_precision(precision)
    s64, lab, sl = sref.astype(np.complex128), labels.astype(np.int64), slow.astype(np.complex64).astype(np.complex128)
    d0 = _z((len(rec[0]), g.NW), dev)
    born.model(d0)
    ref_d0, bg = O.model_background(wav, src, rec, s64, lab, g, g.dz, 0.0, 12, 12)
    ds = crand(rng, (g.nz, g.NW, g.ny, g.nx)) * 1e-7
    d = _z((len(rec[0]), g.NW), dev)
    born.forward(False, _t(ds, dev), d)
    ref_d = O.born_forward(ds, bg, sl, rec, s64, lab, g, g.dz, 0.0, 12, 12, 0.04, 1)
    dres = crand(rng, (len(rec[0]), g.NW))
    m = _z((g.nz, g.NW, g.ny, g.nx), dev)
    born.adjoint(False, m, _t(dres, dev))
    ref_m = O.born_adjoint(dres, bg, sl, rec, s64, lab, g, g.dz, 0.0, 12, 12, 0.04, 1)
    errs = (rel(d0.cpu().numpy(), ref_d0), rel(d.cpu().numpy(), ref_d), rel(m.cpu().numpy(), ref_m))
    floor = None
    if fp32_floor:  # the same formulas evaluated in complex64 (scipy FFTs keep the precision)
        c = np.complex64
        bg32 = [b.astype(c) for b in bg]
        f_d = O.born_forward(ds.astype(c), bg32, sl.astype(c), rec, s64.astype(c), lab, g, g.dz, 0.0, 12, 12, 0.04, 1)
        f_m = O.born_adjoint(dres.astype(c), bg32, sl.astype(c), rec, s64.astype(c), lab, g, g.dz, 0.0, 12, 12,
                             0.04, 1)
        floor = (rel(f_d, ref_d), rel(f_m, ref_m))
    return errs, floor


@pytest.mark.parametrize("size", [(768, 768, 2, 500, 600), (1024, 1024, 2, 700, 512), (384, 1024, 1, 300, 640)],
                         ids=lambda c: f"{c[0]}x{c[1]}")
def test_born_large_grids_match_oracle(dev, size):
    """Background, Born forward and Born adjoint on the largest padded grids
    (3 * 2^a and 1024 columns, every register budget at its default) vs the
    oracle, at 20 and 23 Hz."""
    errs, _ = _born_vs_oracle(dev, size, 20.0)
    assert max(errs) < TOL_WFLD, errs


@pytest.mark.parametrize("ow", [2.0, 6.0], ids=["2-5Hz", "6-9Hz"])
@pytest.mark.parametrize("size", [(768, 768, 2, 500, 600), (1024, 1024, 2, 700, 512)],
                         ids=lambda c: f"{c[0]}x{c[1]}")
def test_born_low_frequency_large_grids_float64_scatter(dev, size, ow):
    """At 2-9 Hz on 768^2 and 1024^2 grids the first Taylor term of the
    scattering operator weighs wavenumber k by k^2 / (w^2 s) (up to ~10^3),
    so float32 rounding of the amplified terms (white in k) lands on the
    propagating band: the same formulas evaluated in complex64 miss FP64 by
    ~2.6e-4 at 6 Hz.  With the scattered-field passes in float64 arithmetic
    (FWIX_PREC_F64_SCATTER, complex64 storage between passes) background,
    Born forward and Born adjoint meet 1e-4 (a host emulation of the pass
    structure gives ~1e-5 at 6 Hz and ~3e-5 at 2 Hz)."""
    from pyfwix_amd.ops import FWIX_PREC_F64_SCATTER
    errs, _ = _born_vs_oracle(dev, size, ow, precision=FWIX_PREC_F64_SCATTER)
    assert max(errs) < TOL_WFLD, errs


def test_born_low_frequency_float32_is_at_the_complex64_floor(dev):
    """The float32 (default) arithmetic at 6 and 9 Hz on a 768^2 grid: the
    complex64 evaluation of the same formulas misses FP64 by ~2.6e-4
    (forward) and ~1.0e-4 (adjoint); the fused float32 passes stay within a
    few times that floor (measured 2.4e-4 forward, 3.0e-4 adjoint)."""
    errs, floor = _born_vs_oracle(dev, (768, 768, 2, 500, 600), 6.0, fp32_floor=True)
    assert errs[0] < TOL_WFLD
    assert errs[1] < 1.5 * floor[0] + 1e-5, (errs, floor)
    assert errs[2] < 4.0 * floor[1] + 1e-5, (errs, floor)


def test_born_float64_scatter_matches_float32_where_both_are_exact(dev):
    """On a well-conditioned grid (48^2, 6 Hz up) both arithmetics agree with
    each other to float32 rounding (the float32 passes carry ~1e-5 relative
    rounding error here: 1.0e-5 measured on the adjoint, so the bound is 2e-5;
    at low frequency on large grids they differ by 1e-4 and more); the dot
    test holds in float64 mode too."""
    import torch
    from pyfwix_amd.ops import FWIX_PREC_F32, FWIX_PREC_F64_SCATTER
    out = {}
    for prec in (FWIX_PREC_F32, FWIX_PREC_F64_SCATTER):
        g, born, slow, sref, labels, src, rec, wav, rng = _born_setup(dev, seed=21)
        born.set_precision(prec)
        born.model()
        ds = crand(rng, (g.nz, g.NW, g.ny, g.nx)) * 1e-7
        d = _z((len(rec[0]), g.NW), dev)
        born.forward(False, _t(ds, dev), d)
        dres = crand(rng, (len(rec[0]), g.NW))
        m = _z((g.nz, g.NW, g.ny, g.nx), dev)
        born.adjoint(False, m, _t(dres, dev))
        torch.cuda.synchronize()
        out[prec] = (d.cpu().numpy(), m.cpu().numpy())
        assert dot_err(ds, out[prec][0], dres, out[prec][1]) < TOL_DOT
    assert rel(out[FWIX_PREC_F64_SCATTER][0], out[FWIX_PREC_F32][0]) < 2e-5
    assert rel(out[FWIX_PREC_F64_SCATTER][1], out[FWIX_PREC_F32][1]) < 2e-5


def test_born_dot_no_taper_single_shot(dev):
    g, born, *_rest = _born_setup(dev, NS=1, tap=0, seed=9)
    rng = _rest[-1]
    nrec = born.nrec
    born.model()
    ds = crand(rng, (g.nz, g.NW, g.ny, g.nx))
    dd = crand(rng, (nrec, g.NW))
    d = _z((nrec, g.NW), dev)
    m = _z((g.nz, g.NW, g.ny, g.nx), dev)
    born.forward(False, _t(ds, dev), d)
    born.adjoint(False, m, _t(dd, dev))
    assert dot_err(ds, d.cpu().numpy(), dd, m.cpu().numpy()) < TOL_DOT
    # add = true
    born.forward(True, _t(ds, dev), d)
    born.adjoint(True, m, _t(dd, dev))
    assert dot_err(ds, d.cpu().numpy(), dd, m.cpu().numpy()) < TOL_DOT


def test_born_checkpointing_is_exact_and_deterministic(dev):
    import torch
    outs = []
    for ck in (1, 3, 4):
        g, born, *_rest = _born_setup(dev, ckpt=ck, seed=13, nz=7)
        rng = np.random.default_rng(99)
        born.model()
        dd = crand(rng, (born.nrec, g.NW))
        ds = crand(rng, (g.nz, g.NW, g.ny, g.nx))
        m = _z((g.nz, g.NW, g.ny, g.nx), dev)
        born.adjoint(False, m, _t(dd, dev))
        d = _z((born.nrec, g.NW), dev)
        born.forward(False, _t(ds, dev), d)
        outs.append((m, d))
    for m, d in outs[1:]:
        assert torch.equal(m, outs[0][0])
        assert torch.equal(d, outs[0][1])
    # run to run
    g, born, *_rest = _born_setup(dev, ckpt=1, seed=13, nz=7)
    rng = np.random.default_rng(99)
    born.model()
    dd = crand(rng, (born.nrec, g.NW))
    m = _z((g.nz, g.NW, g.ny, g.nx), dev)
    born.adjoint(False, m, _t(dd, dev))
    assert torch.equal(m, outs[0][0])


@pytest.mark.parametrize("ck", [1, 3])
def test_reflect_channel_matches_oracle(dev, ck):
    """Up-going reflection recording, its adjoint and the surface-data migration image."""
    import torch
    g, born, slow, sref, labels, src, rec, wav, rng = _born_setup(dev, ckpt=ck, seed=21)
    s64, lab = sref.astype(np.complex128), labels.astype(np.int64)
    born.model()
    _, bg = O.model_background(wav, src, rec, s64, lab, g, g.dz, 0.0, 4, 4)
    refl = crand(rng, (g.nz, g.NW, g.ny, g.nx)) * 0.05
    d = _z((born.nrec, g.NW), dev)
    born.reflect_forward(False, _t(refl, dev), d)
    ref_d = O.reflect_forward(refl, bg, rec, s64, lab, g, g.dz, 0.0, 4, 4)
    assert rel(d.cpu().numpy(), ref_d) < TOL_WFLD
    dd = crand(rng, (born.nrec, g.NW))
    m = _z(refl.shape, dev)
    born.reflect_adjoint(False, m, _t(dd, dev))
    ref_m = O.reflect_adjoint(dd, bg, rec, s64, lab, g, g.dz, 0.0, 4, 4)
    assert rel(m.cpu().numpy(), ref_m) < TOL_WFLD
    assert dot_err(refl, d.cpu().numpy(), dd, m.cpu().numpy()) < TOL_DOT
    img = _z((g.nz, g.ny, g.nx), dev)
    born.reflect_migrate(False, img, _t(dd, dev))
    assert rel(img.cpu().numpy(), ref_m.sum(axis=1)) < TOL_WFLD
    # add = true accumulates
    born.reflect_forward(True, _t(refl, dev), d)
    assert rel(d.cpu().numpy(), 2 * ref_d) < TOL_WFLD
    torch.cuda.synchronize()


def test_reflect_coef_matches_oracle(dev):
    from pyfwix_amd.ops import reflect_coef
    rng = np.random.default_rng(4)
    g = grid(48, 40, 3, 1, nx=40, ny=32, nz=5)
    v = rng.uniform(1500, 4500, (g.nz, g.NW, g.ny, g.nx))
    s = (1.0 / v ** 2) * (1 - 0.03j * rng.random(v.shape))
    rho = rng.uniform(1.0, 2.5, v.shape) * (1 + 0.01j * rng.standard_normal(v.shape))
    r = reflect_coef(dims_of(g), _t(s, dev), _t(rho, dev))
    assert rel(r.cpu().numpy(), O.reflect_coef(s.astype(np.complex64), rho.astype(np.complex64))) < 1e-5


@pytest.mark.parametrize("ck", [1, 3])
def test_extended_born_four_channels_match_oracle(dev, ck):
    """Nonlinear up pass and ExtendedBorn with all four channels (ExtendedBorn.cpp:80-150)."""
    import torch
    g, born, slow, sref, labels, src, rec, wav, rng = _born_setup(dev, ckpt=ck, seed=23)
    # surface receivers: the up-going channels are what they record
    rec = (rec[0], rec[1], np.zeros_like(rec[2]), rec[3])
    born.set_receivers(*rec)
    s64, lab, sl = sref.astype(np.complex128), labels.astype(np.int64), slow.astype(np.complex64).astype(np.complex128)
    den = (1.0 + 0.4 * rng.random(slow.shape)) * (1 + 0.01j * rng.standard_normal(slow.shape))
    den = den.astype(np.complex64)
    born.set_density(_t(den, dev))
    born.model()
    dup = _z((born.nrec, g.NW), dev)
    born.model_up(dup)
    _, bg = O.model_background(wav, src, rec, s64, lab, g, g.dz, 0.0, 4, 4)
    d64 = den.astype(np.complex128)
    ref_up, vbg = O.model_up(bg, sl, d64, rec, s64, lab, g, g.dz, 0.0, 4, 4)
    assert rel(dup.cpu().numpy(), ref_up) < TOL_WFLD
    ds = crand(rng, (g.nz, g.NW, g.ny, g.nx)) * 1e-8
    dr = crand(rng, ds.shape) * 0.05
    d = _z((born.nrec, g.NW), dev)
    born.extended_forward(False, _t(ds, dev), _t(dr, dev), d)
    ref_d = O.extended_born_forward(ds, dr, bg, vbg, sl, d64, rec, s64, lab, g, g.dz, 0.0, 4, 4, 0.04, 1)
    assert rel(d.cpu().numpy(), ref_d) < TOL_WFLD
    dd = crand(rng, (born.nrec, g.NW))
    ms, mr = _z(ds.shape, dev), _z(ds.shape, dev)
    born.extended_adjoint(False, ms, mr, _t(dd, dev))
    ref_s, ref_r = O.extended_born_adjoint(dd, bg, vbg, sl, d64, rec, s64, lab, g, g.dz, 0.0, 4, 4, 0.04, 1)
    assert rel(ms.cpu().numpy(), ref_s) < TOL_WFLD
    assert rel(mr.cpu().numpy(), ref_r) < TOL_WFLD
    lhs = np.vdot(d.cpu().numpy().astype(complex), dd)
    rhs = np.vdot(ds, ms.cpu().numpy()) + np.vdot(dr, mr.cpu().numpy())
    assert abs((lhs / rhs).real - 1) < TOL_DOT
    torch.cuda.synchronize()


def test_bfp16_codec_matches_oracle(dev):
    import torch
    from pyfwix_amd.ops import bfp16_pack, bfp16_unpack
    rng = np.random.default_rng(14)
    x = (crand(rng, 5000) * np.repeat(10.0 ** rng.uniform(-6, 1, 79), 64)[:5000]).astype(np.complex64)
    q, sc = bfp16_pack(_t(x, dev))
    rq, rsc = O.bfp16_pack(x)
    assert np.array_equal(sc.cpu().numpy(), rsc)
    assert np.abs(q.cpu().numpy().astype(int) - rq.astype(int)).max() <= 1   # round-half ties aside
    y = bfp16_unpack(q, sc, x.shape)
    assert rel(y.cpu().numpy(), O.bfp16_unpack(rq, rsc)) < 1e-6
    torch.cuda.synchronize()


@pytest.mark.parametrize("ck", [1, 3])
def test_born_with_checkpoint_codec(dev, ck):
    """16-bit checkpoints: Born forward/adjoint within 1e-4 of the lossless ones, still adjoint."""
    outs = []
    for codec in (0, 1):
        g, born, slow, sref, labels, src, rec, wav, rng = _born_setup(dev, ckpt=ck, seed=27)
        born.set_codec(codec)
        born.model()
        rng = np.random.default_rng(3)
        ds = crand(rng, (g.nz, g.NW, g.ny, g.nx)) * 1e-7
        dd = crand(rng, (born.nrec, g.NW))
        d = _z((born.nrec, g.NW), dev)
        born.forward(False, _t(ds, dev), d)
        m = _z(ds.shape, dev)
        born.adjoint(False, m, _t(dd, dev))
        outs.append((d.cpu().numpy(), m.cpu().numpy()))
        if codec:
            assert dot_err(ds, outs[-1][0], dd, outs[-1][1]) < TOL_DOT
    assert rel(outs[1][0], outs[0][0]) < 1e-4
    assert rel(outs[1][1], outs[0][1]) < 1e-4


@pytest.mark.parametrize("ck", [1, 3])
def test_plain_born_equals_fused(dev, ck):
    """The reference-granularity Born (imaging condition, scattering operator and
    PSPI step as separate operators, one shot per launch) gives the fused
    results: background data, Born data, Born adjoint and migration image."""
    from pyfwix_amd.ops import FWIX_PSPI, FWIX_PSPI_PLAIN
    g, born, slow, sref, labels, src, rec, wav, rng = _born_setup(dev, ckpt=ck)
    ds = _t(crand(rng, (g.nz, g.NW, g.ny, g.nx)) * 1e-7, dev)
    dres = _t(crand(rng, (len(rec[0]), g.NW)), dev)
    outs = []
    for method in (FWIX_PSPI, FWIX_PSPI_PLAIN, FWIX_PSPI):
        born.set_method(method)
        d0 = _z((len(rec[0]), g.NW), dev)
        born.model(d0)
        d = _z((len(rec[0]), g.NW), dev)
        born.forward(False, ds, d)
        m = _z((g.nz, g.NW, g.ny, g.nx), dev)
        born.adjoint(False, m, dres)
        img = _z((g.nz, g.ny, g.nx), dev)
        born.migrate(False, img, dres)
        outs.append([x.cpu().numpy() for x in (d0, d, m, img)])
    for a, b in zip(outs[1], outs[0]):
        assert rel(a, b) < 1e-5
    for a, b in zip(outs[2], outs[0]):          # back to fused: bit-identical
        assert np.array_equal(a, b)


@pytest.mark.parametrize("ckpt,chunk,size", [(1, 0, None), (3, 0, None), (3, 3, None),
                                             (3, 4, (1024, 1024, 3, 96, 80))])
def test_born_step_chaining_matches_whole_steps(dev, ckpt, chunk, size):
    """Background steps chained through the row spectrum (a select pass leaves
    the next step's F_x(taper * U)) give the data and the Born forward and
    adjoint of whole steps, with every depth stored and with recomputed
    segments of 3, and a source injecting at depths 2-3 (where the chain must
    break) besides the one at the top.  chunk = 3 runs the chained steps'
    8 planes as launches of 3, 3 and 2 planes (the spectrum of every chunk
    kept at its own offset), against whole steps in one launch; the 1024^2
    case (6 planes as 4 + 2) runs the split column kernels and the
    quad-block scratch layout."""
    import torch
    out = {}
    for chain in (True, False):
        kw = {"size": size, "nz": 6} if size else {"nz": 8}
        g, born, slow, sref, labels, src, rec, wav, rng = _born_setup(dev, ckpt=ckpt, seed=33,
                                                                      chunk=chunk if chain else 0, **kw)
        sz = np.array([g.oz + 0.3 * g.dz, g.oz + 2.6 * g.dz], np.float32)
        born.set_sources(src[0], src[1], sz, src[3])
        born.set_step_chaining(chain)
        d0 = _z((len(rec[0]), g.NW), dev)
        born.model(d0)
        ds = crand(rng, (g.nz, g.NW, g.ny, g.nx)) * 1e-7
        d = _z((len(rec[0]), g.NW), dev)
        born.forward(False, _t(ds, dev), d)
        dres = crand(rng, (len(rec[0]), g.NW))
        m = _z((g.nz, g.NW, g.ny, g.nx), dev)
        born.adjoint(False, m, _t(dres, dev))
        torch.cuda.synchronize()
        out[chain] = [x.cpu().numpy() for x in (d0, d, m)]
    for a, b in zip(out[True], out[False]):
        assert np.abs(b).max() > 0
        assert rel(a, b) < 1e-6


@pytest.mark.parametrize("chunk", [0, 4])
@pytest.mark.parametrize("direction", [0, 1])
@pytest.mark.parametrize("method", ["pspi", "nsps"])
def test_oneway_step_chaining_matches_whole_steps(dev, direction, method, chunk):
    """Downward / Upward recursions with chained steps (PSPI forward and NSPS
    adjoint: the select pass leaves the next step's row spectrum) give the
    whole steps' results, with and without per-depth checkpoints; chunk = 4
    runs the chained steps' 6 planes as launches of 4 and 2 planes, against
    whole steps in one launch."""
    import torch
    from pyfwix_amd.ops import FWIX_NSPS, FWIX_PSPI, OneWay
    rng = np.random.default_rng(61 + direction)
    g = grid(48, 64, 3, 2, nx=36, ny=48, nz=6)
    slow, sref, labels = _refs(g, rng, 3)
    tap = 6 if method == "pspi" else 0
    shape = (g.NS, g.NW, g.NY, g.NX)
    x = _t(crand(rng, shape), dev)
    out = {}
    for chain in (True, False):
        op = OneWay(direction, FWIX_PSPI if method == "pspi" else FWIX_NSPS, dims_of(g), nref=3,
                    taperx=tap, tapery=tap, planes_per_chunk=chunk if chain else 0)
        op.set_refs(_t(sref, dev), _t(labels, dev))
        op.set_step_chaining(chain)
        f, f2, a = _z(shape, dev), _z(shape, dev), _z(shape, dev)
        ck = _z((g.nz, *shape), dev)
        op.forward(False, x, f)
        op.forward(False, x, f2, ck)
        op.adjoint(False, a, x)
        torch.cuda.synchronize()
        out[chain] = [t.cpu().numpy() for t in (f, f2, ck, a)]
    for p, q in zip(out[True], out[False]):
        assert np.abs(q).max() > 0
        assert rel(p, q) < 1e-6
