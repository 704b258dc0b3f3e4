"""CPU oracle for the one-way depth-stepping path (TEST INFRASTRUCTURE ONLY).

This package restates, in numpy, the arithmetic of pyfwix's PSPI/NSPS
phase-shift propagators, Born scattering, imaging condition, injection and
depth recursions, following the reference's own step formulas (every function
cites the reference file:line it follows).

It is the checker, never the thing measured or shipped: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product path (``pyfwix_amd``) never imports it and fails
loudly when its HIP library is missing.

Every function is run in float64 by default (``dtype=np.complex128``);
``bench.py`` times the same code in complex64 with multi-threaded
``scipy.fft`` as the reported CPU baseline.
"""
from .fwix_oracle import *  # noqa: F401,F403
