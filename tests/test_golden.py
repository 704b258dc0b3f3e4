"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the oracle still reproduces its committed vectors (regression pin).
GPU: the HIP operators reproduce them within the FP32 tolerance.
"""
import os

import numpy as np
import pytest

import oracle as O
from fwix_testutil import TOL_WFLD, rel

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["fft2", "pspi", "nsps", "scatter", "injection", "born"]


def load(name):
    with np.load(os.path.join(GOLD, f"{name}.npz"), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def grid_of(z):
    return O.Grid(**{k: (int(z[k]) if k in ("NX", "NY", "NW", "NS", "nx", "ny", "nz") else float(z[k]))
                     for k in ("NX", "NY", "NW", "NS", "nx", "ny", "ox", "dx", "oy", "dy", "ow", "dw", "oz",
                               "dz", "nz")})


def test_oracle_reproduces_golden():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLD, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    cases = mg.make_cases()
    for name in NAMES:
        z = load(name)
        for k, v in cases[name].items():
            assert np.allclose(np.asarray(v), z[k], rtol=1e-12, atol=0), (name, k)


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _dims(g):
    from pyfwix_amd._lib import make_dims
    return make_dims(g.NX, g.NY, g.NW, g.NS, g.nx, g.ny, g.ox, g.dx, g.oy, g.dy, g.ow, g.dw, g.oz, g.dz, g.nz)


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_matches_golden(dev, name):
    import torch
    from pyfwix_amd import ops
    z = load(name)
    g = grid_of(z)
    d = _dims(g)
    shape = (g.NS, g.NW, g.NY, g.NX)
    x, y = _t(z["x"], dev), _t(z["y"], dev)
    out = torch.zeros(shape, dtype=torch.complex64, device=dev)
    out2 = torch.zeros_like(out)
    if name == "fft2":
        op = ops.FFT2d(d)
        op.forward(False, x, out)
        op.adjoint(False, out2, y)
    elif name in ("pspi", "nsps"):
        tap = int(z["taper"]) if "taper" in z else 0
        op = ops.OneStep(ops.FWIX_PSPI if name == "pspi" else ops.FWIX_NSPS, d, nref=z["sref"].shape[1],
                         taperx=tap, tapery=tap)
        op.set_refs(_t(z["sref"], dev), _t(z["labels"], dev))
        op.set_depth(int(z["iz"]))
        op.forward(False, x, out)
        op.adjoint(False, out2, y)
    elif name == "scatter":
        op = ops.Scatter(d, taylor=int(z["taylor"]), eps_scat=0.04)
        op.set_slow(_t(z["slow"], dev))
        op.set_depth(int(z["iz"]))
        op.forward(False, x, out)
        op.adjoint(False, out2, y)
    elif name == "injection":
        op = ops.Injection(d, z["cx"], z["cy"], z["cz"], z["ids"])
        op.set_depth(int(z["iz"]))
        op.forward(False, _t(z["traces"], dev), out)
        tr = torch.zeros(z["traces"].shape, dtype=torch.complex64, device=dev)
        op.adjoint(False, tr, y)
        assert rel(out.cpu().numpy(), z["fwd"]) < 1e-6
        assert rel(tr.cpu().numpy(), z["adj"]) < 1e-6
        return
    elif name == "born":
        tap = int(z["taper"])
        b = ops.Born(d, nref=z["sref"].shape[1], taperx=tap, tapery=tap, taylor=1, eps_scat=0.04)
        b.set_refs(_t(z["sref"], dev), _t(z["labels"], dev))
        b.set_slow(_t(z["slow"], dev))
        b.set_sources(z["sx"], z["sy"], z["sz"], z["sid"])
        b.set_receivers(z["rx"], z["ry"], z["rz"], z["rid"])
        b.set_wavelet(_t(z["wavelet"], dev))
        d0 = torch.zeros(z["data0"].shape, dtype=torch.complex64, device=dev)
        b.model(d0)
        assert rel(d0.cpu().numpy(), z["data0"]) < TOL_WFLD
        dd = torch.zeros_like(d0)
        b.forward(False, _t(z["ds"], dev), dd)
        assert rel(dd.cpu().numpy(), z["fwd"]) < TOL_WFLD
        m = torch.zeros(z["ds"].shape, dtype=torch.complex64, device=dev)
        b.adjoint(False, m, _t(z["dd"], dev))
        assert rel(m.cpu().numpy(), z["adj"]) < TOL_WFLD
        img = torch.zeros(z["image"].shape, dtype=torch.complex64, device=dev)
        b.migrate(False, img, _t(z["dd"], dev))
        assert rel(img.cpu().numpy(), z["image"]) < TOL_WFLD
        return
    assert rel(out.cpu().numpy(), z["fwd"]) < TOL_WFLD
    assert rel(out2.cpu().numpy(), z["adj"]) < TOL_WFLD
