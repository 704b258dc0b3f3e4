"""The reference workloads as parity cases on the GPU.

* c1 — PSPI shot migration, 128^2 x 100 padded to 256^2, 32 frequencies
  5-30 Hz, surface receivers, 3 flat reflectors: on 4 of the frequencies
  (every 10th) at full grid and depth, the direct data, the up-going
  reflection data and the Sum_w conj(S) R image match the FP64 oracle to
  1e-4; with all 32 frequencies the image focuses at the reflector depths.
* c2 — Born forward + adjoint, 256^2 x 200 padded to 512^2: on 2 of the 64
  frequencies and the first 20 depths at full lateral size (receiver plane
  moved into the window), background, Born data and Born adjoint match the
  FP64 oracle to 1e-4.
* c3 — NSPS in a layered + salt model, 384^2 x 300 padded to 768^2, 8
  references, 100 frequencies: on 4 frequencies and 2 depths at full lateral
  size, NSPS and PSPI forward/adjoint steps match the oracle; at full size
  the NSPS recursion passes the dot test and is compared with PSPI.
"""
import numpy as np
import pytest

import oracle as O
from fwix_testutil import TOL_DOT, TOL_WFLD, rel

pytestmark = pytest.mark.gpu


def _sub(w, idx, nz=None):
    """Grid, refs, labels, slowness and wavelet of a workload on frequency subset ``idx`` (uniform)."""
    idx = np.asarray(idx)
    nz = nz or w.nz
    ow = w.ow + w.dw * idx[0]
    dw = w.dw * (idx[1] - idx[0]) if len(idx) > 1 else w.dw
    NW = len(idx)
    g = O.Grid(NX=w.NX, NY=w.NY, NW=NW, NS=w.NS, nx=w.nx, ny=w.ny, ox=0.0, dx=w.dx, oy=0.0, dy=w.dy,
               ow=ow, dw=dw, oz=0.0, dz=w.dz, nz=nz)
    sref = np.ascontiguousarray(np.broadcast_to(w.sref1[:nz], (nz, w.nref, NW)))
    labels = np.ascontiguousarray(np.broadcast_to(w.labels1[:nz], (nz, NW, w.NY, w.NX)))
    slow = np.ascontiguousarray(np.broadcast_to(w.slow3[:nz, None], (nz, NW, w.ny, w.nx))).astype(np.complex64)
    ds = np.ascontiguousarray(np.broadcast_to(w.dslow3[:nz, None], (nz, NW, w.ny, w.nx))).astype(np.complex64)
    wav = np.ascontiguousarray(w.wavelet[:, idx])
    return g, sref, labels, slow, ds, wav


def _born_op(g, w, sref, labels, slow, wav, dev):
    import torch
    from pyfwix_amd._lib import make_dims
    from pyfwix_amd.ops import Born
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    op = Born(make_dims(g.NX, g.NY, g.NW, g.NS, g.nx, g.ny, g.ox, g.dx, g.oy, g.dy, g.ow, g.dw, g.oz, g.dz,
                        g.nz), nref=w.nref, taperx=w.taper, tapery=w.taper, taylor=1, eps_scat=0.04)
    op.set_refs(t(sref), t(labels))
    op.set_slow(t(slow))
    op.set_sources(*w.src)
    op.set_receivers(*w.rec)
    op.set_wavelet(t(wav))
    return op


def _oracle_case(g, w, sref, labels, slow, ds, wav):
    s64, lab = sref.astype(np.complex128), labels.astype(np.int64)
    d0, bg = O.model_background(wav.astype(np.complex128), w.src, w.rec, s64, lab, g, g.dz, 0.0, w.taper,
                                w.taper)
    d = O.born_forward(ds.astype(np.complex128), bg, slow.astype(np.complex128), w.rec, s64, lab, g, g.dz, 0.0,
                       w.taper, w.taper, 0.04, 1)
    return s64, lab, d0, bg, d


def test_c1_migration_window_matches_oracle(dev):
    """Reflection data of the 3 reflectors and the Sum_w conj(S) R image on every 10th frequency."""
    import torch
    from pyfwix_amd.workloads import CONFIGS, born_workload
    O.set_fft_workers(16)
    w = born_workload("c1", **CONFIGS["c1"])
    g, sref, labels, slow, _, wav = _sub(w, [0, 10, 20, 30])
    refl = np.ascontiguousarray(np.broadcast_to(w.extra["refl3"][:, None], (g.nz, g.NW, g.ny, g.nx)))
    op = _born_op(g, w, sref, labels, slow, wav, dev)
    d0 = torch.zeros((op.nrec, g.NW), dtype=torch.complex64, device=dev)
    op.model(d0)
    d = torch.zeros_like(d0)
    op.reflect_forward(False, torch.from_numpy(refl.astype(np.complex64)).to(dev), d)
    img = torch.zeros((g.nz, g.ny, g.nx), dtype=torch.complex64, device=dev)
    op.reflect_migrate(False, img, d)
    s64, lab = sref.astype(np.complex128), labels.astype(np.int64)
    rd0, bg = O.model_background(wav.astype(np.complex128), w.src, w.rec, s64, lab, g, g.dz, 0.0, w.taper, w.taper)
    rd = O.reflect_forward(refl.astype(np.complex128), bg, w.rec, s64, lab, g, g.dz, 0.0, w.taper, w.taper)
    rimg = O.reflect_migrate(d.cpu().numpy().astype(np.complex128), bg, w.rec, s64, lab, g, g.dz, 0.0, w.taper,
                             w.taper)
    e = (rel(d0.cpu().numpy(), rd0), rel(d.cpu().numpy(), rd), rel(img.cpu().numpy(), rimg))
    print("c1 window rel errors (direct data, reflection data, image):", e)
    assert max(e) < TOL_WFLD and np.linalg.norm(rd) > 0


def test_c1_full_image_focuses_at_the_reflectors(dev):
    import torch
    from pyfwix_amd.ops import Born
    from pyfwix_amd.workloads import CONFIGS, born_workload, device_inputs, workload_dims
    w = born_workload("c1", **CONFIGS["c1"])
    slow, _, sref, labels, wav = device_inputs(w, dev)
    op = Born(workload_dims(w), nref=w.nref, taperx=w.taper, tapery=w.taper, taylor=1, eps_scat=0.04)
    op.set_refs(sref, labels)
    op.set_slow(slow)
    op.set_sources(*w.src)
    op.set_receivers(*w.rec)
    op.set_wavelet(wav)
    op.model()
    refl = torch.zeros_like(slow)
    refl.real.copy_(torch.from_numpy(w.extra["refl3"].astype(np.float32)).to(dev)[:, None].expand_as(slow.real))
    d = torch.zeros((op.nrec, w.NW), dtype=torch.complex64, device=dev)
    op.reflect_forward(False, refl, d)
    img = torch.zeros((w.nz, w.ny, w.nx), dtype=torch.complex64, device=dev)
    op.reflect_migrate(False, img, d)
    # linearity of the migration in the data
    img2 = torch.zeros_like(img)
    op.reflect_migrate(False, img2, 2.0 * d)
    assert rel(img2.cpu().numpy(), 2.0 * img.cpu().numpy()) < 1e-6
    # depth profile of the image energy around the source: peaks at the reflectors
    c = w.nx // 2
    prof = (img.real[:, c - 24:c + 24, c - 24:c + 24] ** 2).sum(dim=(1, 2)).cpu().numpy()
    refl_z = [w.nz // 4, w.nz // 2, 3 * w.nz // 4]
    floor = float(np.median(prof[5:]))
    peaks = [float(prof[r - 3:r + 4].max()) for r in refl_z]
    print("c1 image energy near the reflectors", peaks, "median over depth", floor)
    assert all(p > 10.0 * floor for p in peaks)


def test_c2_window_matches_oracle(dev):
    import torch
    from pyfwix_amd.workloads import CONFIGS, born_workload
    O.set_fft_workers(16)
    w = born_workload("c2", **CONFIGS["c2"])
    g, sref, labels, slow, ds, wav = _sub(w, [20, 40], nz=20)
    # the receiver plane moved into the 20-depth window (z = 15 dz)
    w.rec = (w.rec[0], w.rec[1], np.full_like(w.rec[2], 15 * w.dz), w.rec[3])
    op = _born_op(g, w, sref, labels, slow, wav, dev)
    d0 = torch.zeros((op.nrec, g.NW), dtype=torch.complex64, device=dev)
    op.model(d0)
    d = torch.zeros_like(d0)
    op.forward(False, torch.from_numpy(ds).to(dev), d)
    rng = np.random.default_rng(22)
    res = (rng.standard_normal(d.shape) + 1j * rng.standard_normal(d.shape)).astype(np.complex64)
    m = torch.zeros((g.nz, g.NW, g.ny, g.nx), dtype=torch.complex64, device=dev)
    op.adjoint(False, m, torch.from_numpy(res).to(dev))
    s64, lab, rd0, bg, rd = _oracle_case(g, w, sref, labels, slow, ds, wav)
    rm = O.born_adjoint(res.astype(np.complex128), bg, slow.astype(np.complex128), w.rec, s64, lab, g, g.dz, 0.0,
                        w.taper, w.taper, 0.04, 1)
    e = (rel(d0.cpu().numpy(), rd0), rel(d.cpu().numpy(), rd), rel(m.cpu().numpy(), rm))
    print("c2 window rel errors (direct data, Born data, Born adjoint):", e)
    assert max(e) < TOL_WFLD


@pytest.mark.parametrize("method", ["nsps", "pspi"])
def test_c3_window_steps_match_oracle(dev, method):
    import torch
    from pyfwix_amd._lib import make_dims
    from pyfwix_amd.ops import FWIX_NSPS, FWIX_PSPI, OneStep
    from pyfwix_amd.workloads import NSPS_CONFIGS, nsps_workload
    O.set_fft_workers(16)
    w = nsps_workload("c3", **NSPS_CONFIGS["c3"])
    g, sref, labels, *_ = _sub(w, [0, 33, 66, 99], nz=2)
    tap = w.taper if method == "pspi" else 0
    op = OneStep(FWIX_NSPS if method == "nsps" else FWIX_PSPI,
                 make_dims(g.NX, g.NY, g.NW, g.NS, g.nx, g.ny, g.ox, g.dx, g.oy, g.dy, g.ow, g.dw, g.oz, g.dz, g.nz),
                 nref=w.nref, taperx=tap, tapery=tap)
    op.set_refs(torch.from_numpy(sref).to(dev), torch.from_numpy(labels).to(dev))
    rng = np.random.default_rng(3)
    shape = (g.NS, g.NW, g.NY, g.NX)
    u = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    fwd, adj = (O.nsps_forward, O.nsps_adjoint) if method == "nsps" else (O.pspi_forward, O.pspi_adjoint)
    extra = () if method == "nsps" else (tap, tap)
    for iz in range(2):
        op.set_depth(iz)
        out = torch.zeros(shape, dtype=torch.complex64, device=dev)
        op.forward(False, torch.from_numpy(u).to(dev), out)
        ref = fwd(u.astype(np.complex128), sref[iz].astype(np.complex128), labels[iz].astype(np.int64), g, g.dz,
                  0.0, *extra)
        assert rel(out.cpu().numpy(), ref) < TOL_WFLD
        op.adjoint(False, out, torch.from_numpy(u).to(dev))
        ref = adj(u.astype(np.complex128), sref[iz].astype(np.complex128), labels[iz].astype(np.int64), g, g.dz,
                  0.0, *extra)
        assert rel(out.cpu().numpy(), ref) < TOL_WFLD


def test_c3_full_nsps_dot_and_pspi_comparison(dev):
    import torch
    from pyfwix_amd.ops import FWIX_NSPS, FWIX_PSPI, Injection, OneWay
    from pyfwix_amd.workloads import NSPS_CONFIGS, device_inputs, nsps_workload, workload_dims
    w = nsps_workload("c3", **NSPS_CONFIGS["c3"])
    _, _, sref, labels, wav = device_inputs(w, dev)
    dims = workload_dims(w)
    shape = (w.NS, w.NW, w.NY, w.NX)
    # source wavefield at the surface
    u0 = torch.zeros(shape, dtype=torch.complex64, device=dev)
    inj = Injection(dims, *w.src)
    inj.set_depth(0)
    inj.forward(False, wav, u0)
    outs = {}
    for name, method in (("nsps", FWIX_NSPS), ("pspi", FWIX_PSPI)):
        tap = w.taper if method == FWIX_PSPI else 0
        op = OneWay(0, method, dims, nref=w.nref, taperx=tap, tapery=tap)
        op.set_refs(sref, labels)
        out = torch.zeros_like(u0)
        op.forward(False, u0, out)
        outs[name] = out
        if name == "nsps":
            g = torch.Generator(device="cpu").manual_seed(33)
            dd = torch.randn(shape, generator=g, dtype=torch.complex64).to(dev)
            m = torch.randn(shape, generator=g, dtype=torch.complex64).to(dev)
            Am = torch.zeros_like(m)
            op.forward(False, m, Am)
            Atd = torch.zeros_like(m)
            op.adjoint(False, Atd, dd)
            lhs = torch.sum(Am.conj().to(torch.complex128) * dd.to(torch.complex128))
            rhs = torch.sum(m.conj().to(torch.complex128) * Atd.to(torch.complex128))
            err = abs((lhs / rhs).real.item() - 1.0)
            print(f"c3 NSPS recursion dot error {err:.2e}")
            assert err < TOL_DOT
            del dd, m, Am, Atd
    a, b = outs["nsps"], outs["pspi"]
    assert torch.isfinite(a).all() and torch.isfinite(b).all()
    # both propagate the same source energy; their difference measures the
    # method's error in the salt model (reported, bounded loosely)
    na, nb = float(torch.linalg.vector_norm(a)), float(torch.linalg.vector_norm(b))
    diff = float(torch.linalg.vector_norm(a - b)) / nb
    print(f"c3 full propagation: |NSPS| {na:.4e} |PSPI| {nb:.4e} rel difference {diff:.3f}")
    assert 0.2 < na / nb < 5.0
