"""CPU checks of the Python front-end mirror: SEP stand-ins and host logic.

The operator classes themselves need the GPU (tests/test_gpu_frontend.py).
"""
import numpy as np
import pytest


def test_hypercube_axis_order_and_vectors():
    from pyfwix_amd.sep import axis, complexVector, hypercube
    h = hypercube([axis(6, 0.0, 10.0), axis(4, 5.0, 10.0), axis(3, 2.0, 1.5), axis(2)])
    assert h.getNdim() == 4 and h.getN123() == 144 and h.shape == (2, 3, 4, 6)
    assert h.getAxis(3).o == 2.0 and h.getAxis(1).d == 10.0
    h2 = hypercube(ns=[6, 4, 3, 2], os=[0, 5, 2, 0], ds=[10, 10, 1.5, 1])
    assert h2 == h and h.clone() == h
    v = complexVector(h).random(np.random.default_rng(0))
    w = v.clone()
    assert v.getNdArray().shape == (2, 3, 4, 6) and v.getNdArray().dtype == np.complex64
    assert w.getNdArray() is not v.getNdArray() and np.array_equal(w.arr, v.arr)
    # dot(a, b) = sum(conj(a) * b): <a, i a> = i |a|^2
    w.scale(1j)
    assert abs(v.dot(w) - 1j * v.norm() ** 2) < 1e-9 * v.norm() ** 2
    assert v.cloneSpace().norm() == 0.0
    with pytest.raises(ValueError):
        complexVector(h, np.zeros((2, 3, 4, 5), np.complex64))


def test_param_obj_defaults_and_required():
    from pyfwix_amd.sep import paramObj
    p = paramObj({"nref": "4", "eps": 0.1}, taperx=20)
    assert p.getInt("nref") == 4 and p.getInt("taperx", 0) == 20 and p.getInt("tapery", 0) == 0
    assert p.getFloat("eps", 0.0) == pytest.approx(0.1)
    with pytest.raises(KeyError):
        p.getInt("padx")


def test_dense_shot_ids_follow_the_propagator_map():
    from pyfwix_amd.CudaWEM import dense_shot_ids
    s, r, n = dense_shot_ids([31, 10, 21, 10], [10, 10, 21, 31, 31])
    assert n == 3 and s.tolist() == [2, 0, 1, 0] and r.tolist() == [0, 0, 1, 2, 2]
    with pytest.raises(RuntimeError, match="Shot ID not present"):
        dense_shot_ids([1, 2], [1, 3])


def test_streaming_split_rules():
    """Batch index maps as the reference's stream_unit_test.cpp:141-235 checks them."""
    from pyfwix_amd.streaming import ginsu_window, next_fft_size, split_frequencies, split_sources
    s_ids = [101, 102, 103, 101, 102]
    r_ids = [101, 102, 103, 101, 102, 103, 101, 102]
    b = split_sources(s_ids, r_ids, 2)
    assert [x[0].tolist() for x in b] == [[0, 1, 3, 4], [2]]
    assert [x[1].tolist() for x in b] == [[0, 1, 3, 4, 6, 7], [2, 5]]
    assert len(split_sources(s_ids, r_ids, 7)) == 3          # reduced to the unique sources
    starts, sizes = split_frequencies(10, 3)
    assert starts == [0, 4, 7] and sizes == [4, 3, 3]
    assert split_frequencies(2, 5) == ([0, 1], [1, 1])
    # window: extent of the points +- ginsu, floor/ceil to nodes, clamped; degenerate -> whole model
    assert ginsu_window([55.0, 130.0], [40.0, 41.0], 10.0, 5.0, 30, 20, 0.0, 10.0, 0.0, 10.0) == (4, 14, 3, 5)
    assert ginsu_window([55.0], [40.0], 0.0, 0.0, 30, 20, 0.0, 10.0, 0.0, 10.0) == (5, 6, 0, 19)  # y on a node
    assert ginsu_window([-100.0, 900.0], [0.0, 10.0], 0.0, 0.0, 30, 20, 0.0, 10.0, 0.0, 10.0)[:2] == (0, 29)
    assert (next_fft_size(41), next_fft_size(97), next_fft_size(8), next_fft_size(1000)) == (48, 128, 8, 1024)


def test_shot_batch_planning_and_references():
    """Shot-batch host logic: per-shot references equal the per-depth selector,
    launches group one padded grid, fit the budget and never raise a shot's
    checkpoint interval above what it needs alone."""
    from pyfwix_amd.batching import Shot, launch_bytes, min_ckpt, padded_size, plan_launches, shot_refs
    from pyfwix_amd.refsampler import select_refs_depth
    rng = np.random.default_rng(3)
    s3 = (1.0 / (1500.0 + 1000.0 * rng.random((5, 9, 11))) ** 2)
    sref, lab = shot_refs(s3, 3, 24, 16)
    for z in range(5):
        r1, l1 = select_refs_depth(s3[z][None].astype(np.complex64), 3, 24 - 11, 16 - 9)
        assert np.array_equal(sref[z], r1[:, 0]) and np.array_equal(lab[z], l1[0])
    assert (padded_size(96), padded_size(97), padded_size(300), padded_size(512)) == (192, 256, 768, 1024)
    shots = [Shot(n, n, np.arange(1, nw + 1, dtype=np.float32), np.zeros((1, 3)), np.zeros((1, nw)),
                  np.zeros((4, 3))) for n, nw in ((100, 50), (120, 200), (300, 80), (260, 120), (128, 60))]
    nz, budget = 250, 48 << 30
    plan = plan_launches(shots, nz, budget)
    assert sorted(i for idx, *_ in plan for i in idx) == list(range(len(shots)))
    for idx, NX, NY, k in plan:
        assert {padded_size(shots[i].nx) for i in idx} == {NX}
        P = sum(shots[i].nw for i in idx)
        nx = max(shots[i].nx for i in idx)
        assert launch_bytes(P, len(idx), NX, NY, nx, nx, nz, k) <= budget
        assert k <= max(min_ckpt(shots[i].nw, 1, NX, NY, shots[i].nx, shots[i].ny, nz, budget) for i in idx)
    # the three 256-padded shots share a launch; the 768-padded ones are their own group
    assert any(sorted(idx) == [0, 1, 4] for idx, *_ in plan)
    assert launch_bytes(10, 1, 64, 64, 32, 32, 100, 1) > launch_bytes(10, 1, 64, 64, 32, 32, 100, 4)


def test_fwi_start_model_is_the_smoothed_true_model():
    """The FWI start model (closed-form widening of the Gaussian anomalies, linear
    trend kept) equals scipy's Gaussian smoothing of the true velocity away from
    the edges."""
    from scipy.ndimage import gaussian_filter
    from pyfwix_amd.workloads import fwi_models
    nx, nz, sm = 64, 48, 3.0
    s_true, s_start = fwi_models(nx=nx, nz=nz, d=10.0, smooth=sm, anomalies=4, seed=1)
    v_true = 1.0 / np.sqrt(s_true.astype(np.float64))
    v_start = 1.0 / np.sqrt(s_start.astype(np.float64))
    ref = gaussian_filter(v_true, sm, mode="nearest", truncate=6.0)
    m = int(6 * sm)
    inner = (slice(m, nz - m), slice(m, nx - m), slice(m, nx - m))
    assert np.max(np.abs(v_start[inner] - ref[inner])) < 0.05      # m/s, on 1500-2000 m/s
    assert np.max(np.abs(v_start - v_true)) > 1.0                    # the anomalies did change
