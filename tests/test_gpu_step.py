"""GPU parity of the fused PSPI / NSPS depth step against the FP64 oracle.

Pins (besides the oracle itself): dot-product tests of every adjoint pair at
the reference's 1e-5 (propagator/tests/prop_unit_test.cpp:145-151), in-place
== out-of-place (PSPI.cpp:66-103 vs :6-44), and the add flag.
"""
import numpy as np
import pytest

import oracle as O
from fwix_testutil import TOL_DOT, TOL_WFLD, crand, dims_of, dot_err, grid, rel, smooth_slowness

pytestmark = pytest.mark.gpu

CASES = [  # NX, NY, NW, NS, nref, taper
    (64, 48, 5, 2, 3, 6),
    (96, 128, 4, 1, 4, 10),
    (256, 256, 2, 1, 4, 20),
    (24, 16, 3, 3, 2, 0),
    (512, 384, 1, 1, 5, 0),
    (768, 128, 1, 1, 2, 8),
    (128, 1024, 1, 1, 3, 0),
    (8, 12, 3, 2, 2, 2),      # smallest row / column transforms
    (12, 8, 2, 1, 3, 0),
    (1024, 8, 1, 2, 2, 4),    # 8-point columns in two 512-column tiles
]


def _t(x, dev, dtype=None):
    import torch
    if dtype is not None:
        return torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    return torch.from_numpy(np.ascontiguousarray(x.astype(np.complex64))).to(dev)


def _setup(dev, case, method, imag=0.0, eps=0.0, seed=1):
    from pyfwix_amd.ops import OneStep
    from pyfwix_amd.refsampler import select_refs
    NX, NY, NW, NS, nref, tap = case
    g = grid(NX, NY, NW, NS, nx=NX - NX // 4, ny=NY - NY // 4, nz=2)
    rng = np.random.default_rng(seed)
    slow = smooth_slowness(rng, g, imag=imag)
    sref, labels = select_refs(slow, nref, NX - g.nx, NY - g.ny)
    op = OneStep(method, dims_of(g), nref=nref, taperx=tap, tapery=tap, eps=eps)
    op.set_refs(_t(sref, dev), _t(labels, dev, np.uint8))
    op.set_depth(1)
    sref64 = sref.astype(np.complex128)
    return g, op, sref64[1], labels[1].astype(np.int64), rng


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c)))
@pytest.mark.parametrize("method", ["pspi", "nsps"])
def test_step_matches_oracle(dev, case, method):
    import torch
    from pyfwix_amd.ops import FWIX_NSPS, FWIX_PSPI
    g, op, sref, lab, rng = _setup(dev, case, FWIX_PSPI if method == "pspi" else FWIX_NSPS)
    tap = case[5] if method == "pspi" else 0
    shape = (g.NS, g.NW, g.NY, g.NX)
    x, y = crand(rng, shape), crand(rng, shape)
    out = torch.zeros(shape, dtype=torch.complex64, device=dev)
    op.forward(False, _t(x, dev), out)
    if method == "pspi":
        ref = O.pspi_forward(x, sref, lab, g, g.dz, 0.0, tap, tap)
    else:
        ref = O.nsps_forward(x, sref, lab, g, g.dz, 0.0)
    assert rel(out.cpu().numpy(), ref) < TOL_WFLD
    # adjoint with add
    m = _t(x, dev)
    op.adjoint(True, m, _t(y, dev))
    if method == "pspi":
        ref = x + O.pspi_adjoint(y, sref, lab, g, g.dz, 0.0, tap, tap)
    else:
        ref = x + O.nsps_adjoint(y, sref, lab, g, g.dz, 0.0)
    assert rel(m.cpu().numpy(), ref) < TOL_WFLD


@pytest.mark.parametrize("method", ["pspi", "nsps"])
@pytest.mark.parametrize("case", CASES[:4], ids=lambda c: "x".join(map(str, c)))
def test_step_dot(dev, case, method):
    import torch
    from pyfwix_amd.ops import FWIX_NSPS, FWIX_PSPI
    g, op, sref, lab, rng = _setup(dev, case, FWIX_PSPI if method == "pspi" else FWIX_NSPS,
                                   eps=0.04, seed=4)
    shape = (g.NS, g.NW, g.NY, g.NX)
    x, y = crand(rng, shape), crand(rng, shape)
    Ax = torch.zeros(shape, dtype=torch.complex64, device=dev)
    Aty = torch.zeros_like(Ax)
    op.forward(False, _t(x, dev), Ax)
    op.adjoint(False, Aty, _t(y, dev))
    assert dot_err(x, Ax.cpu().numpy(), y, Aty.cpu().numpy()) < TOL_DOT
    # add = true variant of the reference's dotTest (CudaOperator.h:212-215)
    op.forward(True, _t(x, dev), Ax)
    op.adjoint(True, Aty, _t(y, dev))
    assert dot_err(x, Ax.cpu().numpy(), y, Aty.cpu().numpy()) < TOL_DOT


@pytest.mark.parametrize("method", ["pspi", "nsps"])
def test_step_in_place(dev, method):
    import torch
    from pyfwix_amd.ops import FWIX_NSPS, FWIX_PSPI
    g, op, sref, lab, rng = _setup(dev, CASES[0], FWIX_PSPI if method == "pspi" else FWIX_NSPS)
    shape = (g.NS, g.NW, g.NY, g.NX)
    x = crand(rng, shape)
    out = torch.zeros(shape, dtype=torch.complex64, device=dev)
    op.forward(False, _t(x, dev), out)
    w = _t(x, dev)
    op.forward(False, w, w)
    assert torch.equal(w, out)
    op.adjoint(False, out, _t(x, dev))
    w = _t(x, dev)
    op.adjoint(False, w, w)
    assert torch.equal(w, out)


def test_step_chunking_invariant(dev):
    """Results do not depend on how planes are grouped into fused launches."""
    import torch
    from pyfwix_amd.ops import FWIX_PSPI, OneStep
    from pyfwix_amd.refsampler import select_refs
    g = grid(64, 64, 7, 3, nz=1)
    rng = np.random.default_rng(21)
    slow = smooth_slowness(rng, g)
    sref, labels = select_refs(slow, 3)
    x = _t(crand(rng, (g.NS, g.NW, g.NY, g.NX)), dev)
    outs = []
    for ppc in (0, 1, 5, 21):
        op = OneStep(FWIX_PSPI, dims_of(g), nref=3, taperx=4, tapery=4, planes_per_chunk=ppc)
        op.set_refs(_t(sref, dev), _t(labels, dev, np.uint8))
        op.set_depth(0)
        out = torch.zeros_like(x)
        op.forward(False, x, out)
        outs.append(out)
    for o in outs[1:]:
        assert torch.equal(o, outs[0])


def test_pspi_equal_refs_is_gazdag(dev):
    """All references equal and no taper: PSPI = NSPS = F^-1 Phi F (textbook phase shift)."""
    import torch
    from pyfwix_amd.ops import FWIX_NSPS, FWIX_PSPI, OneStep
    g = grid(64, 96, 4, 2, nz=1)
    rng = np.random.default_rng(2)
    s0 = 1.0 / 2000.0 ** 2
    sref = np.full((1, 3, g.NW), s0, np.complex64)
    labels = rng.integers(0, 3, (1, g.NW, g.NY, g.NX)).astype(np.uint8)
    x = crand(rng, (g.NS, g.NW, g.NY, g.NX))
    phi = O.phase_shift_mult(np.full(g.NW, np.complex64(s0)).astype(np.complex128),
                             O.kgrid(g.NX, g.dx), O.kgrid(g.NY, g.dy), O.w2grid(g.NW, g.ow, g.dw),
                             g.dz, 0.0)
    ref = O.ifft2(phi[None] * O.fft2(x))
    for method in (FWIX_PSPI, FWIX_NSPS):
        op = OneStep(method, dims_of(g), nref=3)
        op.set_refs(_t(sref, dev), _t(labels, dev, np.uint8))
        op.set_depth(0)
        out = torch.zeros((g.NS, g.NW, g.NY, g.NX), dtype=torch.complex64, device=dev)
        op.forward(False, _t(x, dev), out)
        assert rel(out.cpu().numpy(), ref) < 2e-6


@pytest.mark.parametrize("NX,NY", [(64, 48), (512, 96), (24, 32)])
@pytest.mark.parametrize("method", ["pspi", "nsps"])
def test_step_sparse_references(dev, NX, NY, method):
    """References absent from whole planes or rows (the presence masks skip them)."""
    import torch
    from pyfwix_amd.ops import FWIX_NSPS, FWIX_PSPI, OneStep
    g = grid(NX, NY, 3, 2, nz=2)
    rng = np.random.default_rng(77)
    nref = 5
    sref = (1.0 / rng.uniform(1500, 3000, (2, nref, g.NW)) ** 2).astype(np.complex64)
    labels = np.zeros((2, g.NW, g.NY, g.NX), np.uint8)
    labels[1, 0] = 2                                        # plane with a single reference
    labels[1, 1, : NY // 2] = 4                             # rows with 4, rows with 0
    labels[1, 1, NY // 2:, : NX // 3] = 1                   # rows mixing 0 and 1
    labels[1, 2] = rng.choice([0, 3], (g.NY, g.NX))         # refs 1, 2, 4 absent
    op = OneStep(FWIX_PSPI if method == "pspi" else FWIX_NSPS, dims_of(g), nref=nref, taperx=3, tapery=2)
    op.set_refs(_t(sref, dev), _t(labels, dev, np.uint8))
    op.set_depth(1)
    shape = (g.NS, g.NW, g.NY, g.NX)
    x, y = crand(rng, shape), crand(rng, shape)
    s64, lab = sref[1].astype(np.complex128), labels[1].astype(np.int64)
    out = torch.zeros(shape, dtype=torch.complex64, device=dev)
    op.forward(False, _t(x, dev), out)
    ref = (O.pspi_forward(x, s64, lab, g, g.dz, 0.0, 3, 2) if method == "pspi"
           else O.nsps_forward(x, s64, lab, g, g.dz, 0.0))
    assert rel(out.cpu().numpy(), ref) < TOL_WFLD
    m = torch.zeros_like(out)
    op.adjoint(False, m, _t(y, dev))
    ref = (O.pspi_adjoint(y, s64, lab, g, g.dz, 0.0, 3, 2) if method == "pspi"
           else O.nsps_adjoint(y, s64, lab, g, g.dz, 0.0))
    assert rel(m.cpu().numpy(), ref) < TOL_WFLD
    assert dot_err(x, out.cpu().numpy(), y, m.cpu().numpy()) < TOL_DOT


@pytest.mark.parametrize("nref", [32, 40])
@pytest.mark.parametrize("method", ["pspi", "nsps"])
def test_step_many_references(dev, nref, method):
    """32 references (the widest presence masks) and more than 32 (no masks:
    every reference's tile is transformed); labels reach the last reference."""
    import torch
    from pyfwix_amd.ops import FWIX_NSPS, FWIX_PSPI, OneStep
    g = grid(64, 48, 3, 1, nz=2)
    rng = np.random.default_rng(91)
    sref = (1.0 / rng.uniform(1500, 4500, (2, nref, g.NW)) ** 2).astype(np.complex64)
    labels = np.zeros((2, g.NW, g.NY, g.NX), np.uint8)
    labels[1, 0] = rng.integers(0, nref, (g.NY, g.NX))      # every reference present
    labels[1, 1, : g.NY // 2] = nref - 1                    # the last reference alone in some rows
    labels[1, 1, g.NY // 2:] = rng.choice([0, 17], (g.NY - g.NY // 2, g.NX))
    labels[1, 2] = rng.choice([nref - 2, 30], (g.NY, g.NX))  # most references absent
    op = OneStep(FWIX_PSPI if method == "pspi" else FWIX_NSPS, dims_of(g), nref=nref, taperx=4, tapery=3)
    op.set_refs(_t(sref, dev), _t(labels, dev, np.uint8))
    op.set_depth(1)
    shape = (g.NS, g.NW, g.NY, g.NX)
    x, y = crand(rng, shape), crand(rng, shape)
    s64, lab = sref[1].astype(np.complex128), labels[1].astype(np.int64)
    out = torch.zeros(shape, dtype=torch.complex64, device=dev)
    op.forward(False, _t(x, dev), out)
    ref = (O.pspi_forward(x, s64, lab, g, g.dz, 0.0, 4, 3) if method == "pspi"
           else O.nsps_forward(x, s64, lab, g, g.dz, 0.0))
    assert rel(out.cpu().numpy(), ref) < TOL_WFLD
    m = torch.zeros_like(out)
    op.adjoint(False, m, _t(y, dev))
    ref = (O.pspi_adjoint(y, s64, lab, g, g.dz, 0.0, 4, 3) if method == "pspi"
           else O.nsps_adjoint(y, s64, lab, g, g.dz, 0.0))
    assert rel(m.cpu().numpy(), ref) < TOL_WFLD
    assert dot_err(x, out.cpu().numpy(), y, m.cpu().numpy()) < TOL_DOT


@pytest.mark.parametrize("case", [CASES[0], CASES[2], CASES[3]], ids=lambda c: "x".join(map(str, c)))
def test_plain_path_matches_fused_and_oracle(dev, case):
    """The per-operator comparator path computes the same PSPI step."""
    import torch
    from pyfwix_amd.ops import FWIX_PSPI, FWIX_PSPI_PLAIN, OneStep
    from pyfwix_amd.refsampler import select_refs
    NX, NY, NW, NS, nref, tap = case
    g = grid(NX, NY, NW, NS, nx=NX - NX // 4, ny=NY - NY // 4, nz=2)
    rng = np.random.default_rng(5)
    slow = smooth_slowness(rng, g)
    sref, labels = select_refs(slow, nref, NX - g.nx, NY - g.ny)
    shape = (g.NS, g.NW, g.NY, g.NX)
    x = crand(rng, shape)
    outs = {}
    for name, method in (("fused", FWIX_PSPI), ("plain", FWIX_PSPI_PLAIN)):
        op = OneStep(method, dims_of(g), nref=nref, taperx=tap, tapery=tap)
        op.set_refs(_t(sref, dev), _t(labels, dev, np.uint8))
        op.set_depth(1)
        f = torch.zeros(shape, dtype=torch.complex64, device=dev)
        op.forward(False, _t(x, dev), f)
        a = _t(x, dev)
        op.adjoint(True, a, _t(x, dev))
        outs[name] = (f.cpu().numpy(), a.cpu().numpy())
    s64, lab = sref[1].astype(np.complex128), labels[1].astype(np.int64)
    ref_f = O.pspi_forward(x, s64, lab, g, g.dz, 0.0, tap, tap)
    ref_a = x + O.pspi_adjoint(x, s64, lab, g, g.dz, 0.0, tap, tap)
    for name in outs:
        assert rel(outs[name][0], ref_f) < TOL_WFLD
        assert rel(outs[name][1], ref_a) < TOL_WFLD
    assert rel(outs["plain"][0], outs["fused"][0]) < 2e-6
