"""The Python front-end mirror (``fwix/CudaWEM.py``, ``fwix/CudaOperator.py``
class names and signatures) on the GPU:

* host-vector semantics of ``CudaOperator<M,D>`` (``operator/src/CudaOperator.h:74-225``):
  add=False overwrites, add=True accumulates, results match the FP64 oracle;
* ``dotTest`` (both add flags) at the reference's 1e-5 for every operator;
* host vectors and GPU vectors give identical results.
"""
import numpy as np
import pytest

import oracle as O
from fwix_testutil import TOL_DOT, TOL_WFLD, grid, rel, smooth_slowness

pytestmark = pytest.mark.gpu


def _hypers(g):
    from pyfwix_amd.sep import axis, hypercube
    wf = hypercube([axis(g.NX, g.ox, g.dx), axis(g.NY, g.oy, g.dy), axis(g.NW, g.ow, g.dw), axis(g.NS)])
    sl = hypercube([axis(g.nx, g.ox, g.dx), axis(g.ny, g.oy, g.dy), axis(g.NW, g.ow, g.dw),
                    axis(g.nz, g.oz, g.dz)])
    return wf, sl


def _setup(seed=3, NS=2):
    from pyfwix_amd.sep import complexVector, paramObj
    rng = np.random.default_rng(seed)
    g = grid(48, 32, 3, NS, nx=40, ny=24, nz=4)
    wf, sl = _hypers(g)
    slow = complexVector(sl, smooth_slowness(rng, g).astype(np.complex64))
    # ginsu margins wider than the model: every batch window is the whole model
    par = paramObj(nref=3, taperx=4, tapery=3, padx=g.NX - g.nx, pady=g.NY - g.ny, ginsu_x=1e4, ginsu_y=1e4)
    return g, wf, sl, slow, par, rng


@pytest.mark.parametrize("name", ["PSPI", "NSPS"])
def test_onestep_classes_match_oracle_and_dot(dev, name):
    from pyfwix_amd import CudaWEM
    from pyfwix_amd.refsampler import select_refs
    from pyfwix_amd.sep import complexVector
    g, wf, sl, slow, par, rng = _setup()
    m = complexVector(wf).random(rng)
    d = complexVector(wf).random(rng)
    op = getattr(CudaWEM, name)(m, d, slow, par)
    op.set_depth(2)
    d0 = d.clone()
    op.forward(True, m, d)
    sref, labels = select_refs(slow.arr, 3, g.NX - g.nx, g.NY - g.ny)
    s64, lab = sref[2].astype(np.complex128), labels[2].astype(np.int64)
    M = m.arr.astype(np.complex128)
    if name == "PSPI":
        ref = O.pspi_forward(M, s64, lab, g, g.dz, 0.0, 4, 3)
    else:
        ref = O.nsps_forward(M, s64, lab, g, g.dz, 0.0)
    assert rel(d.arr, d0.arr.astype(np.complex128) + ref) < TOL_WFLD
    op.forward(False, m, d)
    assert rel(d.arr, ref) < TOL_WFLD
    # GPU vectors: same numbers, no host copies
    md, dd = m.to(dev), complexVector(wf, device=dev)
    op.forward(False, md, dd)
    assert np.array_equal(dd.numpy(), d.arr)
    # in place (one-argument form): u <- P u
    u = m.clone()
    op.forward(u)
    assert rel(u.arr, d.arr) < 1e-6
    e1, e2 = op.dotTest(rng=rng)
    assert e1 < TOL_DOT and e2 < TOL_DOT


def test_fft_phase_shift_injection_oneway_dot(dev):
    import torch
    from pyfwix_amd import CudaWEM
    from pyfwix_amd.CudaOperator import cuFFT2d
    from pyfwix_amd.sep import axis, complexVector, hypercube
    g, wf, sl, slow, par, rng = _setup(seed=8)
    m = complexVector(wf).random(rng)
    fft = cuFFT2d(m, m.cloneSpace())
    out = m.cloneSpace()
    fft.forward(False, m, out)
    assert rel(out.arr, O.fft2(m.arr.astype(np.complex128))) < 1e-5
    assert max(fft.dotTest(rng=rng)) < TOL_DOT
    ps = CudaWEM.PhaseShift(m, m.cloneSpace(), g.dz, 0.0)
    ps.set_slow((1.0 / 2000.0 ** 2) * (1 - 0.01j) * np.ones(g.NW))  # Im s <= eps Re s: one branch
    assert max(ps.dotTest(rng=rng)) < TOL_DOT
    back = m.cloneSpace()
    ps.forward(False, m, out)
    ps.inverse(False, back, out)
    assert rel(back.arr, m.arr) < 1e-5
    n = 12
    th = hypercube([axis(g.NW, g.ow, g.dw), axis(n)])
    cx = rng.uniform(0, (g.nx - 1) * g.dx, n)
    cy = rng.uniform(0, (g.ny - 1) * g.dy, n)
    cz = rng.uniform(0, 2 * g.dz, n)
    inj = CudaWEM.Injection(complexVector(th), m.cloneSpace(), g.oz, g.dz, cx, cy, cz, np.arange(n) % g.NS)
    inj.set_depth(1)
    assert max(inj.dotTest(rng=rng)) < TOL_DOT
    down = CudaWEM.Downward(m, m.cloneSpace(), slow, par)
    up = CudaWEM.Upward(m, m.cloneSpace(), slow, par)
    assert max(down.dotTest(rng=rng)) < TOL_DOT
    assert max(up.dotTest(rng=rng)) < TOL_DOT
    from pyfwix_amd.refsampler import select_refs
    sref, labels = select_refs(slow.arr, 3, g.NX - g.nx, g.NY - g.ny)
    down.forward(False, m, out)
    ref = O.downward_forward(m.arr.astype(np.complex128), sref.astype(np.complex128), labels.astype(np.int64),
                             g, g.dz, 0.0, 4, 3)
    assert rel(out.arr, ref) < TOL_WFLD
    torch.cuda.synchronize()


def test_propagator_and_extended_born(dev):
    from pyfwix_amd import CudaWEM
    from pyfwix_amd.refsampler import select_refs
    from pyfwix_amd.sep import axis, complexVector, hypercube
    g, wf, sl, slow, par, rng = _setup(seed=12, NS=1)
    f32 = np.float32
    rx, ry = np.meshgrid(np.arange(2, 38, 4) * g.dx, np.arange(2, 22, 4) * g.dy)
    nrec = rx.size
    geom = {"sx": np.array([200.0], f32), "sy": np.array([110.0], f32), "sz": np.array([3.0], f32),
            "s_ids": np.array([7]), "rx": rx.ravel().astype(f32), "ry": ry.ravel().astype(f32),
            "rz": np.full(nrec, 6.0, f32), "r_ids": np.full(nrec, 7)}
    wav = complexVector(hypercube([axis(g.NW, g.ow, g.dw), axis(1)])).random(rng)
    data = complexVector(hypercube([axis(g.NW, g.ow, g.dw), axis(nrec)]))
    den = complexVector(sl).set(1.0)
    prop = CudaWEM.Propagator([slow, den], data, wav, par, geom)
    prop.forward(False, [slow, den], data)
    sref, labels = select_refs(slow.arr, 3, g.NX - g.nx, g.NY - g.ny)
    src = tuple(np.asarray(geom[k]) for k in ("sx", "sy", "sz")) + (np.zeros(1, np.int32),)
    rec = tuple(np.asarray(geom[k]) for k in ("rx", "ry", "rz")) + (np.zeros(nrec, np.int32),)
    d0, bg = O.model_background(wav.arr.astype(np.complex128), src, rec, sref.astype(np.complex128),
                                labels.astype(np.int64), g, g.dz, 0.0, 4, 3)
    refl = O.reflect_coef(slow.arr, den.arr)
    d0 = d0 + O.reflect_forward(refl, bg, rec, sref.astype(np.complex128), labels.astype(np.int64), g, g.dz, 0.0,
                                4, 3)
    assert rel(data.arr, d0) < TOL_WFLD
    ds = complexVector(sl)
    ds.arr[...] = (1e-8 * (rng.standard_normal(ds.arr.shape) + 1j * rng.standard_normal(ds.arr.shape)))
    drho = complexVector(sl)
    drho.arr[...] = 0.05 * rng.standard_normal(drho.arr.shape)
    born = CudaWEM.ExtendedBorn([ds, drho], data, [slow, den], prop)
    dd = data.cloneSpace()
    born.forward(False, [ds, drho], dd)
    s64, lab = sref.astype(np.complex128), labels.astype(np.int64)
    sl64, den64 = slow.arr.astype(np.complex128), den.arr.astype(np.complex128)
    _, vbg = O.model_up(bg, sl64, den64, rec, s64, lab, g, g.dz, 0.0, 4, 3)
    ref = O.extended_born_forward(ds.arr.astype(np.complex128), drho.arr.astype(np.complex128), bg, vbg, sl64,
                                  den64, rec, s64, lab, g, g.dz, 0.0, 4, 3, 0.04, 1)
    assert rel(dd.arr, ref) < TOL_WFLD
    e1, e2 = born.dotTest(rng=rng)
    assert e1 < TOL_DOT and e2 < TOL_DOT
