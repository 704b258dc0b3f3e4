"""pyfwix_amd: MI355X-native one-way wave-equation depth stepping (PSPI / NSPS,
Born scattering, imaging condition and their adjoints) behind pyfwix's
operator API.

The arithmetic lives in hand-written HIP kernels for gfx950
(``libpyfwix_hip.so``, C ABI in ``include/fwix_hip.h``); this package holds
the host-side mirror of the reference's Python front end.
"""
__version__ = "0.1.0"
