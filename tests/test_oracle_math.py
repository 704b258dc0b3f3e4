"""Pins of the oracle that the mathematics fixes (CPU, no GPU).

Each check would fail for a plausible slip in the restatement that the dot
tests and the equal-reference (Gazdag) case cannot see:

* reference locality: in PSPI every output node takes the phase shift of
  its own reference only (propagator/src/PSPI.cpp:6-24 selects after the
  inverse FFT), and in NSPS every input node is shifted with its own
  reference only (NSPS.cpp:27-45 selects before the FFT).  Swapping the two
  mask placements breaks these statements, the equal-reference case does not.
* the Born forward (ExtendedBorn.cpp:152-178, Scattering.cpp:43-52) is the
  derivative of the background recursion (Propagator.cpp:155-180) in the
  slowness: for a laterally homogeneous model and a plane wave, a central
  finite difference of the background equals the Born data up to the
  Taylor truncation of dkz/ds, q^(T+1) with q = k^2 / (w^2 s)
  (Scatter.h:73).  Dropping the propagator from the scattered term, or
  scattering with the wrong depth's slowness, changes the data at O(1).
* the migration image (imaging_condition.cu:57-108 summed over w) is the
  Born adjoint's imaging core: with taylor = 0 and a constant slowness the
  scattering adjoint is the constant +i w dz / (2 sqrt(s)) per frequency,
  so the image is the frequency sum of the Born adjoint divided by it.
* the frequency tables: a vertically travelling wave (k = 0) turns by
  2 pi dz / lambda in one depth step, lambda = v / f.
"""
import math

import numpy as np
import pytest

import oracle as O
from fwix_testutil import crand, grid, rel

TOL = 1e-12


def _refs_random(rng, nref, g):
    sref = (1.0 / rng.uniform(1500.0, 3500.0, (nref, g.NW)) ** 2).astype(complex)
    labels = rng.integers(0, nref, (g.NW, g.NY, g.NX))
    return sref, labels


@pytest.mark.parametrize("method", ["pspi", "nsps"])
def test_reference_locality_fixes_the_mask_placement(method):
    rng = np.random.default_rng(21)
    g = grid(16, 12, 2, 2, nz=1)
    nref = 3
    sref, labels = _refs_random(rng, nref, g)
    x = crand(rng, (g.NS, g.NW, g.NY, g.NX))
    tap = 3 if method == "pspi" else 0
    if method == "pspi":
        fwd = lambda u, s: O.pspi_forward(u, s, labels, g, g.dz, 0.0, tap, tap)   # noqa: E731
        adj = lambda u, s: O.pspi_adjoint(u, s, labels, g, g.dz, 0.0, tap, tap)   # noqa: E731
    else:
        fwd = lambda u, s: O.nsps_forward(u, s, labels, g, g.dz, 0.0)             # noqa: E731
        adj = lambda u, s: O.nsps_adjoint(u, s, labels, g, g.dz, 0.0)             # noqa: E731
    for r in range(nref):
        others = sref.copy()
        others[[k for k in range(nref) if k != r]] *= 1.3     # change every other reference
        on_r = np.broadcast_to(labels == r, x.shape)
        xr = O.select(x, labels, r)                            # input living on label-r nodes only
        if method == "pspi":
            # output on label-r nodes: F^-1[Phi_r F(T x)] -- reference r only
            a, b = fwd(x, sref), fwd(x, others)
            assert np.max(np.abs(a[on_r] - b[on_r])) < TOL * np.max(np.abs(a))
            assert np.max(np.abs(a[~on_r] - b[~on_r])) > 1e-3 * np.max(np.abs(a))
            # adjoint of an input on label-r nodes: T F^-1[Phi_r^adj F(M_r y)] -- reference r only
            a, b = adj(xr, sref), adj(xr, others)
            assert np.max(np.abs(a - b)) < TOL * np.max(np.abs(a))
        else:
            # input on label-r nodes: F^-1[Phi_r F(M_r x)] -- reference r only
            a, b = fwd(xr, sref), fwd(xr, others)
            assert np.max(np.abs(a - b)) < TOL * np.max(np.abs(a))
            a, b = fwd(x, sref), fwd(x, others)
            assert np.max(np.abs(a[on_r] - b[on_r])) > 1e-3 * np.max(np.abs(a))
            # adjoint output on label-r nodes: M_r F^-1[Phi_r^adj F y] -- reference r only
            a, b = adj(x, sref), adj(x, others)
            assert np.max(np.abs(a[on_r] - b[on_r])) < TOL * np.max(np.abs(a))


def _homogeneous_plane_wave_case(nz=4, ow=30.0, dw=6.0):
    g = grid(64, 48, 2, 1, nz=nz, ow=ow, dw=dw)
    rng = np.random.default_rng(22)
    s_z = 1.0 / rng.uniform(1800.0, 2600.0, nz) ** 2                 # per depth, laterally constant
    jx, jy = 1, 1
    X, Y = np.arange(g.NX)[None, :], np.arange(g.NY)[:, None]
    pw = np.exp(2j * np.pi * (jx * X / g.NX + jy * Y / g.NY))
    U0 = np.broadcast_to(pw, (1, g.NW, g.NY, g.NX)).astype(complex) * np.array([1.0, 0.7 - 0.2j])[None, :, None, None]
    labels = np.zeros((nz, g.NW, g.NY, g.NX), int)
    # receivers on exact nodes of the deepest recorded level
    ix = rng.integers(0, g.NX - 1, 24)
    iy = rng.integers(0, g.NY - 1, 24)
    rec = ((ix * g.dx).astype(np.float32), (iy * g.dy).astype(np.float32),
           np.full(24, (nz - 1) * g.dz, np.float32), np.zeros(24, int))
    k2 = O.kgrid(g.NX, g.dx)[jx] ** 2 + O.kgrid(g.NY, g.dy)[jy] ** 2
    q = k2 / (O.wgrid(g.NW, g.ow, g.dw) ** 2 * s_z.min())
    return g, s_z, U0, labels, rec, float(q.max())


def _background(U0, s_z, labels, g):
    sref = np.repeat(s_z[:, None, None], g.NW, axis=2).astype(complex)   # nref = 1, the slowness itself
    ck = []
    O.downward_forward(U0, sref, labels, g, g.dz, 0.0, 0, 0, checkpoints=ck)
    return ck, sref


@pytest.mark.parametrize("taylor", [1, 3])
def test_born_forward_is_the_derivative_of_the_background(taylor):
    g, s_z, U0, labels, rec, q = _homogeneous_plane_wave_case()
    rng = np.random.default_rng(23)
    ds_z = s_z * rng.uniform(-0.05, 0.05, len(s_z))
    bg, sref = _background(U0, s_z, labels, g)
    slow = np.broadcast_to(s_z[:, None, None, None], (g.nz, g.NW, g.ny, g.nx)).astype(complex)
    dslow = np.broadcast_to(ds_z[:, None, None, None], slow.shape).astype(complex)
    born = O.born_forward(dslow, bg, slow, rec, sref, labels, g, g.dz, 0.0, 0, 0, eps_s=0.0, ntaylor=taylor)
    h = 1e-3

    def recorded(s):
        ck, _ = _background(U0, s, labels, g)
        return O.injection_adjoint(ck[-1], *rec, g, g.nz - 1)

    fd = (recorded(s_z + h * ds_z) - recorded(s_z - h * ds_z)) / (2 * h)
    err = rel(born, fd)
    assert np.linalg.norm(born) > 0
    assert err <= 2.0 * q ** (taylor + 1) + 1e-6, (err, q)


def test_born_forward_needs_the_propagator_on_the_scattered_term():
    """The same finite difference rejects dU_{z+1} = P_z dU_z + S_z(U_z ds_z)
    (the scattered term not propagated through the step it scatters in)."""
    g, s_z, U0, labels, rec, q = _homogeneous_plane_wave_case()
    bg, sref = _background(U0, s_z, labels, g)
    ds_z = 0.03 * s_z
    slow = np.broadcast_to(s_z[:, None, None, None], (g.nz, g.NW, g.ny, g.nx)).astype(complex)
    dslow = np.broadcast_to(ds_z[:, None, None, None], slow.shape).astype(complex)
    dU = np.zeros_like(bg[0])
    for z in range(g.nz - 1):
        sc = O.scatter_forward(O.ic_forward(dslow[z], bg[z]), slow[z], g, g.dz, 0.0, 3)
        dU = O.pspi_forward(dU, sref[z], labels[z], g, g.dz) + sc
    wrong = O.injection_adjoint(dU, *rec, g, g.nz - 1)
    right = O.born_forward(dslow, bg, slow, rec, sref, labels, g, g.dz, 0.0, 0, 0, eps_s=0.0, ntaylor=3)
    assert rel(wrong, right) > 0.1


@pytest.mark.parametrize("nref", [1, 3])
def test_migrate_is_the_frequency_sum_of_the_born_adjoint_core(nref):
    rng = np.random.default_rng(24)
    g = grid(16, 12, 3, 2, nx=12, ny=9, nz=4, ow=6.0, dw=3.0)
    s0 = 1.0 / 2100.0 ** 2
    slow = np.full((g.nz, g.NW, g.ny, g.nx), s0, complex)
    sref = (1.0 / rng.uniform(1800.0, 2600.0, (g.nz, nref, g.NW)) ** 2).astype(complex)
    labels = rng.integers(0, nref, (g.nz, g.NW, g.NY, g.NX))
    src = (np.array([55.0, 35.0], np.float32), np.array([45.0, 25.0], np.float32), np.array([2.0, 4.0], np.float32),
           np.array([0, 1]))
    rx, ry = np.meshgrid(np.arange(1, 11) * 10.0, np.arange(1, 8) * 10.0)
    rec = (rx.ravel().astype(np.float32), ry.ravel().astype(np.float32), np.full(rx.size, 25.0, np.float32),
           np.arange(rx.size) % 2)
    wav = crand(rng, (2, g.NW))
    _, bg = O.model_background(wav, src, rec, sref, labels, g, g.dz, 0.0, 2, 2)
    dd = crand(rng, (rx.size, g.NW))
    img = O.migrate(dd, bg, rec, sref, labels, g, g.dz, 0.0, 2, 2)
    m = O.born_adjoint(dd, bg, slow, rec, sref, labels, g, g.dz, 0.0, 2, 2, eps_s=0.0, ntaylor=0)
    f = 0.5j * O.wgrid(g.NW, g.ow, g.dw) * g.dz / math.sqrt(s0)        # the order-0 scattering adjoint
    assert np.linalg.norm(img) > 0
    assert rel(img, (m / f[None, :, None, None]).sum(axis=1)) < TOL


@pytest.mark.parametrize("frac, turn", [(0.1, -0.2 * math.pi), (0.25, -0.5 * math.pi), (0.5, math.pi)])
def test_vertical_wave_turns_by_its_travel_in_wavelengths(frac, turn):
    """v = 1500 m/s, f = 15 Hz: lambda = 100 m; a step of frac * lambda turns
    the k = 0 mode by -2 pi frac (e^{-i kz dz}), whatever the grid."""
    v, f = 1500.0, 15.0
    dz = frac * v / f
    kx, ky = O.kgrid(8, 10.0), O.kgrid(6, 10.0)
    phi = O.phase_shift_mult(np.array([1.0 / v ** 2]), kx, ky, O.w2grid(1, f, 1.0), dz, 0.0)[0, 0, 0]
    assert abs(phi - np.exp(1j * turn)) < 1e-12
    # and the scattering multiplier at k = 0 is -i dz (2 pi f) / (2 sqrt(s)) = -i pi f dz v
    g = grid(8, 6, 1, 1, nz=1, ow=f, dw=1.0, dz=dz)
    one = np.ones((1, 1, 6, 8), complex)
    sc = O.scatter_forward(one, np.full((1, 6, 8), 1.0 / v ** 2, complex), g, dz, eps_s=0.0, ntaylor=2)
    assert abs(sc[0, 0, 0, 0] - (-1j * math.pi * f * dz * v)) < 1e-9 * math.pi * f * dz * v
