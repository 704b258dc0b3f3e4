"""Autograd wrappers of the Born operator and the depth recursion.

* the backward of ``BornFunction`` is the HIP Born adjoint, and matches the
  FP64 oracle's adjoint (``propagator/src/ExtendedBorn.cpp:180-208``);
* the gradient of 0.5*||B ds - d||^2 agrees with central finite differences
  along 20 seeded directions to 1e-3, at the full 256^2 x 200 / 64-frequency
  workload with background checkpoints every 16 depths (recomputed in the
  adjoint), and the Born pair passes the dot test at 1e-5 there;
* the backward of ``OneWayFunction`` is the back-propagation of the depth
  recursion (``propagator/src/OneWay.cpp:6-79``).
"""
import numpy as np
import pytest

import oracle as O
from fwix_testutil import TOL_DOT, TOL_WFLD, crand, dims_of, dot_err, grid, rel, smooth_slowness

pytestmark = pytest.mark.gpu

TOL_FD = 1e-3


def _t(x, dev):
    import torch
    if x.dtype == np.uint8:
        return torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    return torch.from_numpy(np.ascontiguousarray(x.astype(np.complex64))).to(dev)


def _small_born(dev, ckpt=1):
    from pyfwix_amd.ops import Born
    from pyfwix_amd.refsampler import select_refs
    rng = np.random.default_rng(41)
    g = grid(48, 48, 4, 1, nx=32, ny=36, nz=6, ow=6.0, dw=3.0)
    slow = smooth_slowness(rng, g).astype(np.complex64)
    sref, labels = select_refs(slow, 3, g.NX - g.nx, g.NY - g.ny)
    f32 = np.float32
    src = (np.array([150.0], f32), np.array([170.0], f32), np.array([3.0], f32), np.array([0], np.int32))
    rx, ry = np.meshgrid(np.arange(2, 30, 3) * g.dx, np.arange(2, 34, 4) * g.dy)
    rec = (rx.ravel().astype(f32), ry.ravel().astype(f32), np.full(rx.size, 4.0, f32),
           np.zeros(rx.size, np.int32))
    wav = crand(rng, (1, g.NW)).astype(np.complex64)
    born = Born(dims_of(g), nref=3, taperx=4, tapery=4, taylor=1, eps_scat=0.04, ckpt_every=ckpt)
    born.set_refs(_t(sref, dev), _t(labels, dev))
    born.set_slow(_t(slow, dev))
    born.set_sources(*src)
    born.set_receivers(*rec)
    born.set_wavelet(_t(wav, dev))
    born.model()
    return g, born, slow, sref, labels, src, rec, wav, rng


def test_born_function_backward_is_the_oracle_adjoint(dev):
    import torch
    from pyfwix_amd.autograd import born, born_misfit_gradient
    g, op, slow, sref, labels, src, rec, wav, rng = _small_born(dev)
    s64, lab, sl = sref.astype(np.complex128), labels.astype(np.int64), slow.astype(np.complex128)
    _, bg = O.model_background(wav.astype(np.complex128), src, rec, s64, lab, g, g.dz, 0.0, 4, 4)
    ds_true = crand(rng, (g.nz, g.NW, g.ny, g.nx)) * 1e-7
    dobs = O.born_forward(ds_true, bg, sl, rec, s64, lab, g, g.dz, 0.0, 4, 4, 0.04, 1)
    x0 = 0.5 * ds_true + 0.2e-7 * crand(rng, ds_true.shape)
    # forward through the Function == oracle forward
    x = _t(x0, dev).requires_grad_(True)
    d = born(op, x)
    ref_d = O.born_forward(x0, bg, sl, rec, s64, lab, g, g.dz, 0.0, 4, 4, 0.04, 1)
    assert rel(d.detach().cpu().numpy(), ref_d) < TOL_WFLD
    # gradient of the misfit == oracle adjoint of the oracle residual
    loss, grad = born_misfit_gradient(op, _t(x0, dev), _t(dobs, dev))
    ref_g = O.born_adjoint(ref_d - dobs, bg, sl, rec, s64, lab, g, g.dz, 0.0, 4, 4, 0.04, 1)
    assert rel(grad.cpu().numpy(), ref_g) < TOL_WFLD
    assert abs(loss.item() - 0.5 * np.sum(np.abs(ref_d - dobs) ** 2)) <= 1e-4 * loss.item()
    # the backward is exactly the operator's adjoint entry point
    gout = _t(crand(rng, (op.nrec, g.NW)), dev)
    x.grad = None
    born(op, x).backward(gout)
    m = torch.zeros_like(x.detach())
    op.adjoint(False, m, gout)
    assert torch.equal(x.grad, m)


def test_born_function_checkpoint_every_k_gives_the_same_gradient(dev):
    import torch
    from pyfwix_amd.autograd import born_misfit_gradient
    grads = []
    for ck in (1, 4):
        g, op, *_rest, rng = _small_born(dev, ckpt=ck)
        rng = np.random.default_rng(5)
        x0 = _t(crand(rng, (g.nz, g.NW, g.ny, g.nx)) * 1e-7, dev)
        dobs = _t(crand(rng, (op.nrec, g.NW)), dev)
        grads.append(born_misfit_gradient(op, x0, dobs)[1])
    assert torch.equal(grads[0], grads[1])


def test_oneway_function_backward_is_back_propagation(dev):
    import torch
    from pyfwix_amd.autograd import oneway
    from pyfwix_amd.ops import FWIX_PSPI, OneWay
    from pyfwix_amd.refsampler import select_refs
    rng = np.random.default_rng(17)
    g = grid(32, 48, 3, 2, nx=24, ny=40, nz=5)
    slow = smooth_slowness(rng, g).astype(np.complex64)
    sref, labels = select_refs(slow, 2, g.NX - g.nx, g.NY - g.ny)
    s64, lab = sref.astype(np.complex128), labels.astype(np.int64)
    op = OneWay(0, FWIX_PSPI, dims_of(g), nref=2, taperx=3, tapery=3)
    op.set_refs(_t(sref, dev), _t(labels, dev))
    u0 = crand(rng, (g.NS, g.NW, g.NY, g.NX))
    gout = crand(rng, u0.shape)
    x = _t(u0, dev).requires_grad_(True)
    y = oneway(op, x)
    y.backward(_t(gout, dev))
    ref_y = O.downward_forward(u0, s64, lab, g, g.dz, 0.0, 3, 3)
    ref_g = O.downward_adjoint(gout, s64, lab, g, g.dz, 0.0, 3, 3)
    assert rel(y.detach().cpu().numpy(), ref_y) < TOL_WFLD
    assert rel(x.grad.cpu().numpy(), ref_g) < TOL_WFLD
    assert dot_err(u0, y.detach().cpu().numpy(), gout, x.grad.cpu().numpy()) < TOL_DOT
    assert x.grad.dtype == torch.complex64


def _blocky_direction(rng, shape, dev):
    """Seeded real, frequency-invariant random field, constant on 4x8x8 cells (a direction like ds)."""
    import torch
    nz, NW, ny, nx = shape
    c = rng.standard_normal((nz // 4 + 1, ny // 8 + 1, nx // 8 + 1)).astype(np.float32)
    f = c.repeat(4, 0).repeat(8, 1).repeat(8, 2)[:nz, :ny, :nx]
    v = torch.zeros(shape, dtype=torch.complex64, device=dev)
    v.real.copy_(torch.from_numpy(np.ascontiguousarray(f)).to(dev)[:, None].expand(nz, NW, ny, nx))
    return v


def test_c2_born_gradient_matches_finite_differences_and_dot(dev):
    """256^2 x 200 (padded 512^2), 64 frequencies, checkpoints every 16 depths."""
    import torch
    from pyfwix_amd.autograd import born, born_misfit_gradient
    from pyfwix_amd.ops import Born
    from pyfwix_amd.workloads import CONFIGS, born_workload, device_inputs, workload_dims
    w = born_workload("c2", **CONFIGS["c2"])
    slow, ds, sref, labels, wav = device_inputs(w, dev)
    op = Born(workload_dims(w), nref=w.nref, taperx=w.taper, tapery=w.taper, taylor=1, eps_scat=0.04,
              ckpt_every=16)
    op.set_refs(sref, labels)
    op.set_slow(slow)
    op.set_sources(*w.src)
    op.set_receivers(*w.rec)
    op.set_wavelet(wav)
    op.model()
    del sref, labels, wav
    with torch.no_grad():
        dobs = born(op, ds)
    x0 = 0.5 * ds
    del ds
    loss, grad = born_misfit_gradient(op, x0, dobs)
    scale = float(torch.linalg.vector_norm(x0))

    def misfit(x):
        with torch.no_grad():
            r = born(op, x) - dobs
            return 0.5 * float(torch.sum(r.real.double() ** 2 + r.imag.double() ** 2))

    rng = np.random.default_rng(2)
    errs = []
    for i in range(20):
        v = _blocky_direction(rng, tuple(x0.shape), dev)
        v *= scale / float(torch.linalg.vector_norm(v))
        h = 0.25
        fd = (misfit(x0 + h * v) - misfit(x0 - h * v)) / (2 * h)
        an = float(torch.sum(grad.real.double() * v.real.double() + grad.imag.double() * v.imag.double()))
        errs.append(abs(fd - an) / max(abs(an), 1e-30))
        print(f"direction {i}: fd {fd:.6e} autograd {an:.6e} rel {errs[-1]:.2e}", flush=True)
        del v
    assert max(errs) < TOL_FD, errs
    # dot test of the Born pair at full size
    v = _blocky_direction(rng, tuple(x0.shape), dev)
    dd = _t(crand(rng, (op.nrec, w.NW)), dev)
    Bv = born(op, v)
    Btd = torch.zeros_like(v)
    op.adjoint(False, Btd, dd)
    lhs = torch.sum(Bv.conj().to(torch.complex128) * dd.to(torch.complex128))
    rhs = torch.sum(v.conj().to(torch.complex128) * Btd.to(torch.complex128))
    assert abs((lhs / rhs).real.item() - 1.0) < TOL_DOT
