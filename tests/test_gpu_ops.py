"""GPU parity of the FFT and the point-wise operators against the FP64 oracle.

Known answers follow the reference's FFT tests (operator/tests/operator_unit_test.cpp):
forward-inverse round trip (:36-48), constant signal -> DC spike sqrt(n1*n2)
(:50-65), plane wave -> spike at (k1, k2) (:67-99), dot test (:101-105);
and the phase-shift inverse test (propagator/tests/prop_unit_test.cpp:56-67).
"""
import numpy as np
import pytest

import oracle as O
from fwix_testutil import TOL_DOT, crand, dims_of, dot_err, grid, rel

pytestmark = pytest.mark.gpu

FFT_SHAPES = [  # (NS, NW, NY, NX)
    (2, 3, 8, 16), (1, 2, 12, 24), (1, 2, 48, 96), (1, 3, 64, 128), (1, 2, 256, 192),
    (1, 2, 512, 512), (1, 1, 768, 384), (1, 1, 1024, 1024), (1, 2, 96, 768), (1, 2, 32, 1024),
]


def _t(x, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x.astype(np.complex64))).to(dev)


@pytest.mark.parametrize("shape", FFT_SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_fft2d_matches_oracle(dev, shape):
    import torch
    from pyfwix_amd.ops import FFT2d
    rng = np.random.default_rng(11)
    NS, NW, NY, NX = shape
    g = grid(NX, NY, NW, NS)
    op = FFT2d(dims_of(g))
    x = crand(rng, shape)
    m = _t(x, dev)
    d = torch.zeros_like(m)
    op.forward(False, m, d)
    assert rel(d.cpu().numpy(), O.fft2(x)) < 3e-6
    # adjoint = inverse, add semantics
    y = crand(rng, shape)
    md = _t(y, dev)
    op.adjoint(True, md, d)
    assert rel(md.cpu().numpy(), y + O.ifft2(O.fft2(x))) < 3e-6


def test_fft2d_known_answers(dev):
    import torch
    from pyfwix_amd.ops import FFT2d
    n1, n2, n3, n4 = 96, 64, 4, 2
    g = grid(n1, n2, n3, n4)
    op = FFT2d(dims_of(g))
    ones = torch.ones((n4, n3, n2, n1), dtype=torch.complex64, device=dev)
    out = torch.zeros_like(ones)
    op.forward(False, ones, out)
    o = out.cpu().numpy()
    assert np.allclose(o[:, :, 0, 0], n1 * n2 / np.sqrt(n1 * n2), atol=1e-4)
    o[:, :, 0, 0] = 0
    assert np.abs(o).max() < 1e-4
    k1, k2 = 10, 20
    i1 = np.arange(n1)[None, :]
    i2 = np.arange(n2)[:, None]
    pw = np.exp(2j * np.pi * (k1 * i1 / n1 + k2 * i2 / n2))
    x = np.broadcast_to(pw, (n4, n3, n2, n1))
    m = _t(x, dev)
    op.forward(False, m, out)
    o = out.cpu().numpy()
    assert np.allclose(o[:, :, k2, k1], np.sqrt(n1 * n2), atol=1e-3)
    o[:, :, k2, k1] = 0
    assert np.abs(o).max() < 1e-3
    # in place
    op.forward(False, m, m)
    assert rel(m.cpu().numpy(), O.fft2(x)) < 3e-6


def test_fft2d_dot(dev):
    import torch
    from pyfwix_amd.ops import FFT2d
    rng = np.random.default_rng(3)
    shape = (2, 3, 48, 64)
    op = FFT2d(dims_of(grid(64, 48, 3, 2)))
    x, y = crand(rng, shape), crand(rng, shape)
    Ax = torch.zeros(shape, dtype=torch.complex64, device=dev)
    Aty = torch.zeros_like(Ax)
    op.forward(False, _t(x, dev), Ax)
    op.adjoint(False, Aty, _t(y, dev))
    assert dot_err(x, Ax.cpu().numpy(), y, Aty.cpu().numpy()) < TOL_DOT


@pytest.mark.parametrize("mode", ["fwd", "adj", "inv"])
@pytest.mark.parametrize("imag,eps", [(0.0, 0.0), (0.0, 0.04), (0.02, 0.0)])
def test_phase_shift_matches_oracle(dev, mode, imag, eps):
    import torch
    from pyfwix_amd.ops import PhaseShift
    rng = np.random.default_rng(5)
    g = grid(64, 48, 6, 2)
    dz = 10.0
    op = PhaseShift(dims_of(g), dz, eps)
    sref = (1.0 / rng.uniform(1500, 3000, g.NW) ** 2) * (1 + 1j * imag)
    op.set_slow(sref)
    shape = (g.NS, g.NW, g.NY, g.NX)
    x = crand(rng, shape)
    out = torch.zeros(shape, dtype=torch.complex64, device=dev)
    kx, ky, w2 = O.kgrid(g.NX, g.dx), O.kgrid(g.NY, g.dy), O.w2grid(g.NW, g.ow, g.dw)
    ref = O.phase_shift(x, sref.astype(np.complex64).astype(np.complex128), kx, ky, w2, dz, eps, mode)
    if mode == "fwd":
        op.forward(False, _t(x, dev), out)
    elif mode == "adj":
        op.adjoint(False, out, _t(x, dev))
    else:
        op.inverse(False, out, _t(x, dev))
    assert rel(out.cpu().numpy(), ref) < 2e-6


def test_phase_shift_inverse_roundtrip(dev):
    """PS^-1(PS(x)) = x (prop_unit_test.cpp:56-67), slowness 1, dz 0.1, eps 0."""
    import torch
    from pyfwix_amd.ops import PhaseShift
    g = grid(96, 96, 10, 2, dx=1.0, dy=1.0, ow=1.0, dw=1.0)
    op = PhaseShift(dims_of(g), 0.1, 0.0)
    op.set_slow(np.ones(g.NW))
    rng = np.random.default_rng(8)
    x = crand(rng, (g.NS, g.NW, g.NY, g.NX))
    xd = _t(x, dev)
    y = torch.zeros_like(xd)
    z = torch.zeros_like(xd)
    for _ in range(3):
        op.forward(False, xd, y)
        op.inverse(False, z, y)
        assert rel(z.cpu().numpy(), x) < 1e-6


def test_phase_shift_dot(dev):
    import torch
    from pyfwix_amd.ops import PhaseShift
    rng = np.random.default_rng(9)
    g = grid(48, 64, 5, 2)
    op = PhaseShift(dims_of(g), 10.0, 0.04)
    op.set_slow(1.0 / rng.uniform(1500, 3000, g.NW) ** 2)
    shape = (g.NS, g.NW, g.NY, g.NX)
    x, y = crand(rng, shape), crand(rng, shape)
    Ax = torch.zeros(shape, dtype=torch.complex64, device=dev)
    Aty = torch.zeros_like(Ax)
    op.forward(False, _t(x, dev), Ax)
    op.adjoint(False, Aty, _t(y, dev))
    assert dot_err(x, Ax.cpu().numpy(), y, Aty.cpu().numpy()) < TOL_DOT


def test_selector_and_taper(dev):
    import torch
    from pyfwix_amd.ops import Selector, Taper
    rng = np.random.default_rng(12)
    g = grid(40, 24, 3, 2)
    shape = (g.NS, g.NW, g.NY, g.NX)
    labels = rng.integers(0, 4, (g.NW, g.NY, g.NX)).astype(np.int32)
    sel = Selector(dims_of(g))
    sel.set_labels(labels)
    x, y = crand(rng, shape), crand(rng, shape)
    for v in range(4):
        sel.set_value(v)
        out = _t(y, dev)
        sel.forward(True, _t(x, dev), out)
        assert rel(out.cpu().numpy(), y + O.select(x, labels, v)) < 1e-6
    tp = Taper(dims_of(g), 7, 5)
    out = torch.zeros(shape, dtype=torch.complex64, device=dev)
    tp.forward(False, _t(x, dev), out)
    assert rel(out.cpu().numpy(), O.taper(x, 7, 5)) < 1e-6
    out2 = _t(y, dev)
    tp.adjoint(True, out2, _t(x, dev))
    assert rel(out2.cpu().numpy(), y + O.taper(x, 7, 5)) < 1e-6
