"""pyFWIX: the reference's compiled front-end module (``fwix/src/pyFWIX.cpp``)
with the same classes -- cuFFT2d, StreamingPropagator, Propagator,
ExtendedBorn, StreamingExtendedBorn, PSPI, NSPS, Injection, Downward,
Upward -- and argument lists (``propagator/src/python/pyCudaWEM.cpp:22-242``,
``operator/src/python/pyCudaOperator.cpp:15-30``), implemented in C++
(``csrc/pyFWIX.cpp``, pybind11) over the HIP library.

The extension binds to the HIP runtime torch already loaded (one runtime per
process), so torch is imported before it.  There is no CPU fallback: a
missing extension raises ImportError.
"""
import torch  # noqa: F401  (loads the HIP runtime the extension binds to)

from . import _lib  # noqa: F401  (loads libpyfwix_hip.so against that runtime)

try:
    from ._pyFWIX import *  # noqa: F401,F403
    from ._pyFWIX import (NSPS, PSPI, Downward, ExtendedBorn, Injection, Propagator,  # noqa: F401
                          StreamingExtendedBorn, StreamingPropagator, Upward, cuFFT2d)
except ImportError as e:  # pragma: no cover - depends on the build
    raise ImportError(f"pyFWIX extension missing ({e}): build it with `python -m pyfwix_amd.build`") from e
