"""CPU checks of the C ABI boundary and the host-side logic (no GPU calls).

* libpyfwix_hip.so loads and exports every function include/fwix_hip.h declares,
  and the ctypes signature table covers them;
* argument validation errors are reported through the error convention
  (status -1 + fwix_last_error, RuntimeError in Python) without touching a GPU;
* the reference-slowness selector and the synthetic workloads behave as documented.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fwix_hip.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fwix_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from pyfwix_amd import _lib
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(_lib.lib, n)]
    assert not missing, missing
    assert set(names) <= set(_lib.EXPORTS), set(names) - set(_lib.EXPORTS)
    assert _lib.lib.fwix_version() >= 100


def test_error_convention_without_gpu():
    from pyfwix_amd import _lib
    d = _lib.make_dims(100, 64, 4, 1)  # 100 = 4 * 25: not a supported FFT size
    h = C.c_void_p()
    rc = _lib.lib.fwix_fft2d_create(C.byref(d), C.byref(h))
    assert rc == -1
    assert b"FFT sizes" in _lib.lib.fwix_last_error()
    with pytest.raises(RuntimeError, match="nref"):
        p = _lib.default_params(nref=0)
        _lib.create("fwix_onestep_create", 0, C.byref(_lib.make_dims(64, 64, 2, 1)), C.byref(p))
    with pytest.raises(RuntimeError, match="wrong kind|null"):
        _lib.call("fwix_onestep_set_depth", None, 0)
    p = _lib.default_params()
    assert (p.nref, p.taperx, p.tapery, p.taylor) == (1, 0, 0, 1)
    assert abs(p.eps_scat - 0.04) < 1e-7 and p.eps == 0.0


@pytest.mark.parametrize("n,ok", [(8, True), (12, True), (96, True), (512, True), (768, True), (1024, True),
                                  (100, False), (2048, False), (4, False), (36, False)])
def test_fft_sizes(n, ok):
    from pyfwix_amd import _lib
    assert _lib.fft_size_supported(n) == ok


def test_refsampler_quantiles_labels_and_padding():
    from pyfwix_amd.refsampler import RefSampler, select_refs_depth
    rng = np.random.default_rng(0)
    s = rng.uniform(1, 2, (2, 6, 5)) + 0j
    s[1] = s[0] * 1.5
    sref, lab = select_refs_depth(s, 3, padx=3, pady=2)
    assert sref.shape == (3, 2) and lab.shape == (2, 8, 8)
    for w in range(2):
        flat = np.sort(s[w].real.ravel())
        k = np.round((np.arange(3) + 0.5) / 3 * (flat.size - 1)).astype(int)
        assert np.allclose(sref[:, w].real, flat[k])
        d = np.abs(s[w].reshape(-1, 1) - sref[:, w][None, :])
        assert np.array_equal(lab[w, :6, :5].ravel(), d.argmin(axis=1))
        assert np.array_equal(lab[w, 6:, :5], np.repeat(lab[w, 5:6, :5], 2, axis=0))   # RefSampler.cpp:69-79
        assert np.array_equal(lab[w, :, 5:], np.repeat(lab[w, :, 4:5], 3, axis=1))     # RefSampler.cpp:80-86
    # frequency-invariant input: identical result for every frequency
    s2 = np.broadcast_to(s[:1], (4, 6, 5)).copy()
    r2, l2 = select_refs_depth(s2, 3)
    assert np.all(r2 == r2[:, :1]) and np.all(l2 == l2[:1])
    rs = RefSampler(s[None], 3, 3, 2)
    assert rs.get_ref_labels(0).dtype == np.int32 and rs.get_ref_slow(0, 1).shape == (2,)


def test_workload_c1_geometry():
    from pyfwix_amd.workloads import CONFIGS, born_workload, ricker_spectrum
    w = born_workload("c1", **CONFIGS["c1"])
    assert (w.NX, w.NY, w.NW, w.nz, w.nref) == (256, 256, 32, 100, 4)
    assert len(w.rec[0]) == 128 * 128 and len(w.src[0]) == 1
    v = 1.0 / np.sqrt(w.slow3)
    assert abs(v[0, 0, 0] - 1500.0) < 1e-6                       # corner: background only
    r2 = ((49 - 49.5) ** 2 + 2 * (63 - 63.5) ** 2) * 100.0          # node next to the model centre
    assert abs(v[49, 63, 63] - (1500 + 0.8 * 490 + 300 * np.exp(-r2 / (2 * 150.0 ** 2)))) < 1e-6
    assert np.count_nonzero(w.dslow3.reshape(w.nz, -1).any(axis=1)) == 3
    f = np.array([15.0])
    assert abs(abs(ricker_spectrum(f)[0]) - 2 / np.sqrt(np.pi) / 15 * np.exp(-1)) < 1e-12


def test_refsampler_kmeans_is_seeded_and_finds_clusters():
    from pyfwix_amd.refsampler import kmeans_1slice, select_refs
    rng = np.random.default_rng(5)
    true = np.array([2.0e-7, 3.0e-7, 4.4e-7]) * (1 - 0.01j)
    x = np.concatenate([c + 1e-9 * (rng.standard_normal(400) + 1j * rng.standard_normal(400)) for c in true])
    c = np.sort_complex(kmeans_1slice(x, 3, np.random.default_rng(1), eps=1e-12))
    assert np.max(np.abs(c - np.sort_complex(true))) < 2e-10
    s = np.broadcast_to(x.reshape(1, 1, 30, 40), (2, 3, 30, 40)).astype(np.complex64)
    r1, l1 = select_refs(s, 3, 5, 2, method="kmeans", seed=4)
    r2, l2 = select_refs(s, 3, 5, 2, method="kmeans", seed=4)
    assert np.array_equal(r1, r2) and np.array_equal(l1, l2)            # seeded: reproducible
    d = np.abs(s[0, 0].reshape(-1, 1) - r1[0, :, 0][None])
    assert np.array_equal(l1[0, 0, :30, :40].ravel(), d.argmin(axis=1))  # nearest-center labels
    assert np.array_equal(l1[0, 0, 30:, :40], np.repeat(l1[0, 0, 29:30, :40], 2, axis=0))


def test_column_budget_table():
    """Register-budget table of the column kernels: defaults, set/get, validation."""
    from pyfwix_amd import autotune
    assert autotune.get(0, 0, 512) == 6 and autotune.get(1, 1, 512) == 4 and autotune.get(0, 0, 1024) == 2
    assert autotune.get(1, 0, 1024) == 2 and autotune.get(1, 2, 768) == 4 and autotune.get(0, 1, 384) == 4
    assert autotune.get(1, 1, 384) == 2
    old = autotune.get(1, 2, 768)
    try:
        autotune.set_budget(1, 2, 768, 4)
        assert autotune.get(1, 2, 768) == 4 and autotune.table((768,))["768"]["fanin"][2] == 4
    finally:
        autotune.set_budget(1, 2, 768, old)
    for bad in ((2, 0, 512, 4), (0, 3, 512, 4), (0, 0, 500, 4), (0, 0, 512, 3)):
        with pytest.raises(RuntimeError, match="tune_columns"):
            autotune.set_budget(*bad)
    with pytest.raises(ValueError):
        autotune.get(0, 0, 500)
    # column layouts: two waves per column only for 1024-point columns
    assert autotune.get_layout(0, 0, 512) == 0 and autotune.get_layout(1, 2, 768) == 0
    old = autotune.get_layout(1, 1, 1024)
    try:
        autotune.set_layout(1, 1, 1024, 1 - old)
        assert autotune.get_layout(1, 1, 1024) == 1 - old and autotune.table((1024,))["1024"]["fanin_layout"][1] == 1 - old
    finally:
        autotune.set_layout(1, 1, 1024, old)
    for bad in ((0, 0, 512, 1), (2, 0, 1024, 1), (0, 3, 1024, 0), (0, 0, 1024, 2), (0, 0, 500, 0)):
        with pytest.raises(RuntimeError, match="tune_columns_layout"):
            autotune.set_layout(*bad)
    # scratch layout: quad blocks at 1024 points by default, row-major below
    from pyfwix_amd import _lib
    assert _lib.lib.fwix_tune_scratch_layout_get(1024) == 1 and _lib.lib.fwix_tune_scratch_layout_get(512) == 0
    assert _lib.lib.fwix_tune_scratch_layout_get(500) == -1
    for bad in ((500, 0), (512, 2), (1024, -1)):
        with pytest.raises(RuntimeError, match="tune_scratch_layout"):
            _lib.call("fwix_tune_scratch_layout_set", *bad)
