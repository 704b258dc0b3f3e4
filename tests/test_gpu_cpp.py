"""The C++ operator layer (include/fwix_operators.hpp) on the GPU.

Runs the driver built from tests/cpp/test_operators.cpp: dot tests of every
operator class through its host forward/adjoint (1e-5, both add flags),
in-place vs out-of-place PSPI, and exceptions on invalid sizes; then checks
the PSPI forward it wrote against the FP64 oracle.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from fwix_testutil import TOL_WFLD, rel

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_operators")


def test_cpp_operator_layer(dev, tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/test_operators is not built (python -m pyfwix_amd.build)")
    res = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr)
    assert res.returncode == 0, res.stdout + res.stderr
    g = O.Grid(NX=32, NY=24, NW=3, NS=2, nx=24, ny=16, ox=0.0, dx=10.0, oy=0.0, dy=10.0, ow=5.0, dw=2.0,
               oz=0.0, dz=10.0, nz=4)
    m = np.fromfile(tmp_path / "pspi_in.bin", np.complex64).reshape(g.NS, g.NW, g.NY, g.NX)
    y = np.fromfile(tmp_path / "pspi_out.bin", np.complex64).reshape(m.shape)
    sref = np.fromfile(tmp_path / "sref.bin", np.complex64).reshape(g.nz, 2, g.NW)
    lab = np.fromfile(tmp_path / "labels.bin", np.uint8).reshape(g.nz, g.NW, g.NY, g.NX)
    ref = O.pspi_forward(m.astype(np.complex128), sref[2].astype(np.complex128), lab[2].astype(np.int64), g, g.dz,
                         0.0, 3, 2)
    assert rel(y, ref) < TOL_WFLD
