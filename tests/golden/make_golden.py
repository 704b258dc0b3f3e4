"""Write the committed golden fixtures (tests/golden/*.npz) from the FP64 oracle.

Small seeded cases of every operator on the path: inputs and expected
outputs.  CPU tests re-derive them from the oracle (regression of the
oracle); GPU tests compare the HIP path against them.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle as O  # noqa: E402
from fwix_testutil import crand, grid, smooth_slowness  # noqa: E402


def grid_dict(g):
    return {k: np.asarray(getattr(g, k)) for k in ("NX", "NY", "NW", "NS", "nx", "ny", "ox", "dx", "oy", "dy",
                                                     "ow", "dw", "oz", "dz", "nz")}


def make_cases():
    from pyfwix_amd.refsampler import select_refs
    out = {}
    rng = np.random.default_rng(2024)
    g = grid(32, 24, 3, 2, nx=24, ny=16, nz=4, ow=6.0, dw=3.0)
    slow = smooth_slowness(rng, g, imag=0.0).astype(np.complex64)
    sref, labels = select_refs(slow, 3, g.NX - g.nx, g.NY - g.ny)
    s64, lab = sref.astype(np.complex128), labels.astype(np.int64)
    shape = (g.NS, g.NW, g.NY, g.NX)
    x = crand(rng, shape).astype(np.complex64)
    y = crand(rng, shape).astype(np.complex64)
    X, Y = x.astype(np.complex128), y.astype(np.complex128)
    base = dict(grid_dict(g), slow=slow, sref=sref, labels=labels, x=x, y=y)
    out["fft2"] = dict(base, fwd=O.fft2(X), adj=O.ifft2(Y))
    tap = 5
    out["pspi"] = dict(base, taper=np.asarray(tap), iz=np.asarray(2),
                       fwd=O.pspi_forward(X, s64[2], lab[2], g, g.dz, 0.0, tap, tap),
                       adj=O.pspi_adjoint(Y, s64[2], lab[2], g, g.dz, 0.0, tap, tap))
    out["nsps"] = dict(base, iz=np.asarray(1), fwd=O.nsps_forward(X, s64[1], lab[1], g, g.dz, 0.0),
                       adj=O.nsps_adjoint(Y, s64[1], lab[1], g, g.dz, 0.0))
    sl64 = slow.astype(np.complex128)
    out["scatter"] = dict(base, iz=np.asarray(3), taylor=np.asarray(1),
                          fwd=O.scatter_forward(X, sl64[3], g, g.dz, 0.04, 1),
                          adj=O.scatter_adjoint(Y, sl64[3], g, g.dz, 0.04, 1))
    n = 25
    cx = rng.uniform(-10, (g.NX - 1) * g.dx, n).astype(np.float32)
    cy = rng.uniform(0, (g.NY - 1) * g.dy, n).astype(np.float32)
    cz = rng.uniform(0, 2.5 * g.dz, n).astype(np.float32)
    ids = (np.arange(n) % g.NS).astype(np.int32)
    tr = crand(rng, (n, g.NW)).astype(np.complex64)
    out["injection"] = dict(base, cx=cx, cy=cy, cz=cz, ids=ids, traces=tr, iz=np.asarray(1),
                            fwd=O.injection_forward(tr.astype(complex), cx, cy, cz, ids, g, 1, shape=shape),
                            adj=O.injection_adjoint(Y, cx, cy, cz, ids, g, 1))
    src = (np.array([95.0, 120.0], np.float32), np.array([75.0, 60.0], np.float32),
           np.array([4.0, 2.0], np.float32), np.array([0, 1], np.int32))
    rx, ry = np.meshgrid(np.arange(2, 22, 2) * g.dx, np.arange(2, 15, 3) * g.dy)
    rec = (rx.ravel().astype(np.float32), ry.ravel().astype(np.float32),
           np.full(rx.size, 7.0, np.float32), (np.arange(rx.size) % 2).astype(np.int32))
    wav = crand(rng, (2, g.NW)).astype(np.complex64)
    d0, bg = O.model_background(wav.astype(complex), src, rec, s64, lab, g, g.dz, 0.0, tap, tap)
    ds = (crand(rng, (g.nz, g.NW, g.ny, g.nx)) * 1e-8).astype(np.complex64)
    dd = crand(rng, (rx.size, g.NW)).astype(np.complex64)
    out["born"] = dict(base, taper=np.asarray(tap), sx=src[0], sy=src[1], sz=src[2], sid=src[3],
                       rx=rec[0], ry=rec[1], rz=rec[2], rid=rec[3], wavelet=wav, ds=ds, dd=dd, data0=d0,
                       fwd=O.born_forward(ds.astype(complex), bg, sl64, rec, s64, lab, g, g.dz, 0.0, tap, tap,
                                          0.04, 1),
                       adj=O.born_adjoint(dd.astype(complex), bg, sl64, rec, s64, lab, g, g.dz, 0.0, tap, tap,
                                          0.04, 1),
                       image=O.migrate(dd.astype(complex), bg, rec, s64, lab, g, g.dz, 0.0, tap, tap))
    return out


def main():
    for name, arrays in make_cases().items():
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **arrays)
        print("wrote", name, sum(a.nbytes for a in arrays.values()) // 1024, "KiB")


if __name__ == "__main__":
    main()
