"""The compiled pyFWIX module loads without a GPU and exposes the reference's
classes (``fwix/src/pyFWIX.cpp:9-15``: init_operator + init_propagator) with
their methods; the stand-in SEP objects hand themselves over as ``.cppMode``."""
import numpy as np


def test_pyfwix_exposes_the_reference_classes():
    from pyfwix_amd import pyFWIX
    methods = {
        "cuFFT2d": ("forward", "adjoint"),
        "StreamingPropagator": ("forward",),
        "Propagator": ("forward", "get_compression_ratio"),
        "ExtendedBorn": ("forward", "adjoint"),
        "StreamingExtendedBorn": ("forward", "adjoint"),
        "PSPI": ("forward", "adjoint", "set_depth"),
        "NSPS": ("forward", "adjoint", "set_depth"),
        "Injection": ("forward", "adjoint", "set_depth", "set_coords"),
        "Downward": ("forward", "adjoint"),
        "Upward": ("forward", "adjoint"),
    }
    for cls, names in methods.items():
        c = getattr(pyFWIX, cls)
        for n in names:
            assert callable(getattr(c, n)), (cls, n)


def test_standins_hand_over_cppmode():
    from pyfwix_amd.sep import axis, complexVector, hypercube, paramObj
    h = hypercube([axis(4), axis(3)])
    v = complexVector(h, np.zeros((3, 4), np.complex64))
    p = paramObj(nref=2)
    assert h.cppMode is h and v.cppMode is v and p.cppMode is p
    assert v.getHyper().cppMode.getAxis(1).n == 4 and p.cppMode.getInt("nref") == 2
