"""The compiled pyFWIX module (``fwix/src/pyFWIX.cpp`` classes and argument
lists, ``propagator/src/python/pyCudaWEM.cpp:22-242``,
``operator/src/python/pyCudaOperator.cpp:15-30``) on the GPU.  Every class is
built the way the reference's ``fwix/CudaWEM.py`` builds it -- hypercubes,
vectors and parameters passed as ``.cppMode`` -- and checked against the FP64
oracle; host and GPU vectors give the same numbers."""
import numpy as np
import pytest

import oracle as O
from fwix_testutil import TOL_DOT, TOL_WFLD, dot_err, grid, rel, smooth_slowness

pytestmark = pytest.mark.gpu


def _setup(seed=3, NS=2):
    from pyfwix_amd.sep import axis, complexVector, hypercube, paramObj
    rng = np.random.default_rng(seed)
    g = grid(48, 32, 3, NS, nx=40, ny=24, nz=4)
    wf = hypercube([axis(g.NX, g.ox, g.dx), axis(g.NY, g.oy, g.dy), axis(g.NW, g.ow, g.dw), axis(g.NS)])
    sl = hypercube([axis(g.nx, g.ox, g.dx), axis(g.ny, g.oy, g.dy), axis(g.NW, g.ow, g.dw), axis(g.nz, g.oz, g.dz)])
    slow = complexVector(sl, smooth_slowness(rng, g).astype(np.complex64))
    par = paramObj(nref=3, taperx=4, tapery=3, padx=g.NX - g.nx, pady=g.NY - g.ny, ginsu_x=1e4, ginsu_y=1e4)
    return g, wf, sl, slow, par, rng


def _refs(slow, g):
    from pyfwix_amd.refsampler import select_refs
    sref, labels = select_refs(slow.arr, 3, g.NX - g.nx, g.NY - g.ny)
    return sref.astype(np.complex128), labels.astype(np.int64)


@pytest.mark.parametrize("name", ["PSPI", "NSPS"])
def test_pyfwix_onestep(dev, name):
    from pyfwix_amd import pyFWIX
    from pyfwix_amd.sep import complexVector
    g, wf, sl, slow, par, rng = _setup()
    m = complexVector(wf).random(rng)
    d = complexVector(wf).random(rng)
    op = getattr(pyFWIX, name)(m.getHyper().cppMode, slow.cppMode, par.cppMode)
    op.set_depth(2)
    sref, lab = _refs(slow, g)
    M = m.arr.astype(np.complex128)
    if name == "PSPI":
        ref = O.pspi_forward(M, sref[2], lab[2], g, g.dz, 0.0, 4, 3)
    else:
        ref = O.nsps_forward(M, sref[2], lab[2], g, g.dz, 0.0)
    d0 = d.arr.astype(np.complex128)
    op.forward(True, m.cppMode, d.cppMode)
    assert rel(d.arr, d0 + ref) < TOL_WFLD
    op.forward(False, m.cppMode, d.cppMode)
    assert rel(d.arr, ref) < TOL_WFLD
    md, dd = m.to(dev), complexVector(wf, device=dev)
    op.forward(False, md.cppMode, dd.cppMode)
    assert np.array_equal(dd.numpy(), d.arr)
    y = complexVector(wf).random(rng)
    x = complexVector(wf)
    op.adjoint(False, x.cppMode, y.cppMode)
    if name == "PSPI":
        refa = O.pspi_adjoint(y.arr.astype(np.complex128), sref[2], lab[2], g, g.dz, 0.0, 4, 3)
    else:
        refa = O.nsps_adjoint(y.arr.astype(np.complex128), sref[2], lab[2], g, g.dz, 0.0)
    assert rel(x.arr, refa) < TOL_WFLD
    assert dot_err(m.arr, d.arr, y.arr, x.arr) < TOL_DOT
    if name == "PSPI":  # the one-argument form propagates in place
        u = m.clone()
        op.forward(u.cppMode)
        assert rel(u.arr, d.arr) < 1e-6


def test_pyfwix_fft_injection_oneway(dev):
    from pyfwix_amd import pyFWIX
    from pyfwix_amd.sep import axis, complexVector, hypercube
    g, wf, sl, slow, par, rng = _setup(seed=8)
    m = complexVector(wf).random(rng)
    out = complexVector(wf)
    fft = pyFWIX.cuFFT2d(m.getHyper().cppMode)
    fft.forward(False, m.cppMode, out.cppMode)
    assert rel(out.arr, O.fft2(m.arr.astype(np.complex128))) < 1e-5
    back = complexVector(wf)
    fft.adjoint(False, back.cppMode, out.cppMode)
    assert rel(back.arr, m.arr) < 1e-5
    # Injection(traces hyper, wavefield hyper, oz, dz, cx, cy, cz, ids)
    n = 12
    th = hypercube([axis(g.NW, g.ow, g.dw), axis(n)])
    cx = list(rng.uniform(0, (g.nx - 1) * g.dx, n))
    cy = list(rng.uniform(0, (g.ny - 1) * g.dy, n))
    cz = list(rng.uniform(0, 2 * g.dz, n))
    ids = [int(i) for i in np.arange(n) % g.NS]
    tr = complexVector(th).random(rng)
    inj = pyFWIX.Injection(th.cppMode, wf.cppMode, g.oz, g.dz, cx, cy, cz, ids)
    inj.set_depth(1)
    w = complexVector(wf)
    inj.forward(False, tr.cppMode, w.cppMode)
    f32 = lambda a: np.asarray(a, np.float32)  # noqa: E731
    ref = O.injection_forward(tr.arr.astype(np.complex128), f32(cx), f32(cy), f32(cz), np.asarray(ids, np.int32), g, 1,
                              shape=(g.NS, g.NW, g.NY, g.NX))
    assert rel(w.arr, ref) < 1e-6
    tr2 = complexVector(th)
    inj.adjoint(False, tr2.cppMode, m.cppMode)
    ref = O.injection_adjoint(m.arr.astype(np.complex128), f32(cx), f32(cy), f32(cz), np.asarray(ids, np.int32), g, 1)
    assert rel(tr2.arr, ref) < 1e-6
    # Downward / Upward(domain, slow, par)
    sref, lab = _refs(slow, g)
    down = pyFWIX.Downward(m.getHyper().cppMode, slow.cppMode, par.cppMode)
    down.forward(False, m.cppMode, out.cppMode)
    assert rel(out.arr, O.downward_forward(m.arr.astype(np.complex128), sref, lab, g, g.dz, 0.0, 4, 3)) < TOL_WFLD
    u = m.clone()
    down.forward(u.cppMode)
    assert rel(u.arr, out.arr) < 1e-6
    y = complexVector(wf).random(rng)
    x = complexVector(wf)
    down.adjoint(False, x.cppMode, y.cppMode)
    assert dot_err(m.arr, out.arr, y.arr, x.arr) < TOL_DOT
    up = pyFWIX.Upward(m.getHyper().cppMode, slow.cppMode, par.cppMode)
    up.forward(False, m.cppMode, out.cppMode)
    assert rel(out.arr, O.upward_forward(m.arr.astype(np.complex128), sref, lab, g, g.dz, 0.0, 4, 3)) < TOL_WFLD


def test_pyfwix_propagators_and_extended_born(dev):
    from pyfwix_amd import pyFWIX
    from pyfwix_amd.sep import axis, complexVector, hypercube
    g, wf, sl, slow, par, rng = _setup(seed=12, NS=1)
    f32 = np.float32
    rx, ry = np.meshgrid(np.arange(2, 38, 4) * g.dx, np.arange(2, 22, 4) * g.dy)
    nrec = rx.size
    sx, sy, sz, s_ids = [200.0], [110.0], [3.0], [7]
    rxl, ryl, rzl, r_ids = list(rx.ravel()), list(ry.ravel()), [6.0] * nrec, [7] * nrec
    wav = complexVector(hypercube([axis(g.NW, g.ow, g.dw), axis(1)])).random(rng)
    data = complexVector(hypercube([axis(g.NW, g.ow, g.dw), axis(nrec)]))
    den = complexVector(sl).set(1.0)
    den.arr[2:] = 1.2                                  # a density contrast reflects
    args = (wav.getHyper().cppMode, data.getHyper().cppMode, slow.getHyper().cppMode, wav.cppMode,
            sx, sy, sz, s_ids, rxl, ryl, rzl, r_ids, par.cppMode)
    prop = pyFWIX.Propagator(*args)
    prop.forward(False, [slow.cppMode, den.cppMode], data.cppMode)
    assert prop.get_compression_ratio() == 1.0
    sref, lab = _refs(slow, g)
    src = (np.array(sx, f32), np.array(sy, f32), np.array(sz, f32), np.zeros(1, np.int32))
    rec = (rx.ravel().astype(f32), ry.ravel().astype(f32), np.full(nrec, 6.0, f32), np.zeros(nrec, np.int32))
    d0, bg = O.model_background(wav.arr.astype(np.complex128), src, rec, sref, lab, g, g.dz, 0.0, 4, 3)
    sl64, den64 = slow.arr.astype(np.complex128), den.arr.astype(np.complex128)
    refl = O.reflect_coef(sl64, den64)
    d0 = d0 + O.reflect_forward(refl, bg, rec, sref, lab, g, g.dz, 0.0, 4, 3)
    assert rel(data.arr, d0) < TOL_WFLD
    # streaming over two frequency bands (windows cover the model): the same data
    sprop = pyFWIX.StreamingPropagator(*args, [1, 2])
    sdata = data.cloneSpace()
    sprop.forward(False, [slow.cppMode, den.cppMode], sdata.cppMode)
    assert rel(sdata.arr, data.arr) < 1e-6
    # ExtendedBorn over the propagator: four channels vs the oracle, dot test
    ds = complexVector(sl)
    ds.arr[...] = 1e-8 * (rng.standard_normal(ds.arr.shape) + 1j * rng.standard_normal(ds.arr.shape))
    drho = complexVector(sl)
    drho.arr[...] = 0.05 * rng.standard_normal(drho.arr.shape)
    born = pyFWIX.ExtendedBorn(ds.getHyper().cppMode, data.getHyper().cppMode, [slow.cppMode, den.cppMode], prop)
    dd = data.cloneSpace()
    born.forward(False, [ds.cppMode, drho.cppMode], dd.cppMode)
    _, vbg = O.model_up(bg, sl64, den64, rec, sref, lab, g, g.dz, 0.0, 4, 3)
    ref = O.extended_born_forward(ds.arr.astype(np.complex128), drho.arr.astype(np.complex128), bg, vbg, sl64, den64,
                                  rec, sref, lab, g, g.dz, 0.0, 4, 3, 0.04, 1)
    assert rel(dd.arr, ref) < TOL_WFLD
    r = data.cloneSpace().random(rng)
    gs, gr = complexVector(sl), complexVector(sl)
    born.adjoint(False, [gs.cppMode, gr.cppMode], r.cppMode)
    lhs = O.dot(dd.arr, r.arr)
    rhs = O.dot(ds.arr, gs.arr) + O.dot(drho.arr, gr.arr)
    assert abs((lhs / rhs).real - 1.0) < TOL_DOT
    # the streaming form over the same background gives the same numbers
    sborn = pyFWIX.StreamingExtendedBorn(sprop, [slow.cppMode, den.cppMode])
    sd = data.cloneSpace()
    sborn.forward(False, [ds.cppMode, drho.cppMode], sd.cppMode)
    assert rel(sd.arr, dd.arr) < 1e-6
    sgs, sgr = complexVector(sl), complexVector(sl)
    sborn.adjoint(False, [sgs.cppMode, sgr.cppMode], r.cppMode)
    assert rel(sgs.arr, gs.arr) < 1e-6 and rel(sgr.arr, gr.arr) < 1e-6
