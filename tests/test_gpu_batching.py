"""Shot batches on the GPU: shots of uneven aperture and frequency count laid
along the plane axis of one Born operator (per-plane frequencies, per-shot
slowness / references / model, ragged traces).

* each shot's data and gradient in a batch equal the shot run alone, and the
  shot run alone equals the regular operator (frequency-replicated inputs,
  dense traces), which the other GPU tests pin to the FP64 oracle;
* a batch with recomputed background segments equals one with every depth kept;
* the batched Born pair passes the dot-product test;
* a batch of one-way recursions (set_shot_planes on Downward) equals its shots
  run alone;
* torch.autograd through a batch operator (BornFunction) gives the batch's
  migration adjoint.
"""
import numpy as np
import pytest

from fwix_testutil import TOL_DOT, crand, rel

pytestmark = pytest.mark.gpu

NZ, D = 10, 10.0
PARAMS = {"nref": 3, "taper": 4, "eps": 0.0, "eps_scat": 0.04, "taylor": 1}


def _shots(rng):
    """Three shots of one padded grid (48 x 48): apertures 20, 24, 17 (x) by
    22, 17, 24 (y), 5 / 8 / 3 frequencies on different bands."""
    from pyfwix_amd.batching import Shot
    out = []
    for nx, ny, nw, f0 in ((20, 22, 5, 4.0), (24, 17, 8, 2.5), (17, 24, 3, 7.0)):
        freqs = (f0 + 1.5 * np.arange(nw)).astype(np.float32)
        src = np.array([[nx // 2 * D + 3.0, ny // 2 * D - 2.0, 0.0]], np.float32)
        gx, gy = np.meshgrid(np.arange(0, nx, 3) * D, np.arange(0, ny, 3) * D)
        rec = np.stack([gx.ravel(), gy.ravel(), np.full(gx.size, (NZ - 1) * D)], 1).astype(np.float32)
        wav = crand(rng, (1, nw)).astype(np.complex64)
        out.append(Shot(nx, ny, freqs, src, wav, rec))
    return out


def _models(rng, shots):
    slows, dss = [], []
    for s in shots:
        v = 1500.0 + 600.0 * rng.random((NZ, s.ny, s.nx))
        from scipy.ndimage import gaussian_filter
        v = gaussian_filter(v, 1.5)
        slows.append((1.0 / v ** 2).astype(np.complex64))
        dss.append((0.05 * slows[-1].real * rng.standard_normal(slows[-1].shape)).astype(np.complex64))
    return slows, dss


def _run(batch, slows, dss, data_in=None):
    batch.set_slowness(slows)
    batch.set_perturbation(dss)
    batch.model()
    data = batch.forward()
    grad = batch.adjoint(data if data_in is None else data_in)
    return data, grad


@pytest.mark.parametrize("chunk", [0, 3])
def test_batch_equals_shots_alone(dev, chunk):
    """chunk = 3: the batch's 16 planes run as launches of 3 planes, which
    straddle the shots' plane ranges (5 + 8 + 3)."""
    import torch
    from pyfwix_amd.batching import ShotBatch, padded_size
    rng = np.random.default_rng(4)
    shots = _shots(rng)
    assert len({padded_size(s.nx) for s in shots} | {padded_size(s.ny) for s in shots}) == 1
    slows, dss = _models(rng, shots)
    b = ShotBatch(shots, NZ, D, D, D, PARAMS, device=dev, planes_per_chunk=chunk)
    data, grad = _run(b, slows, dss)
    torch.cuda.synchronize()
    assert float(data.abs().max()) > 0
    for k, s in enumerate(shots):
        one = ShotBatch([s], NZ, D, D, D, PARAMS, device=dev)
        d1, g1 = _run(one, [slows[k]], [dss[k]])
        assert rel(b.shot_data(data, k).cpu().numpy(), one.shot_data(d1, 0).cpu().numpy()) < 1e-6
        assert rel(b.shot_grad(grad, k).cpu().numpy(), one.shot_grad(g1, 0).cpu().numpy()) < 1e-6


def test_single_shot_batch_equals_regular_operator(dev):
    """One shot through the shot-batch layout == the regular operator with the
    slowness and model replicated over frequency and the gradient summed."""
    import torch
    from pyfwix_amd import ops
    from pyfwix_amd.batching import ShotBatch, shot_refs
    rng = np.random.default_rng(5)
    s = _shots(rng)[1]
    slows, dss = _models(rng, [s])
    one = ShotBatch([s], NZ, D, D, D, PARAMS, device=dev)
    d1, g1 = _run(one, slows, dss)
    NX, NY, nw = one.NX, one.NY, s.nw
    dims = ops.make_dims(NX, NY, nw, 1, s.nx, s.ny, 0.0, D, 0.0, D, float(s.freqs[0]), 1.5, 0.0, D, NZ)
    born = ops.Born(dims, nref=PARAMS["nref"], taperx=PARAMS["taper"], tapery=PARAMS["taper"], eps_scat=0.04,
                    taylor=1)
    born.set_frequencies(s.freqs)
    sref, lab = shot_refs(slows[0], PARAMS["nref"], NX, NY)
    rep = lambda a: np.ascontiguousarray(np.repeat(a[:, None], nw, axis=1))  # noqa: E731
    born.set_refs(torch.from_numpy(np.ascontiguousarray(np.repeat(sref[:, :, None], nw, axis=2))).to(dev),
                  torch.from_numpy(rep(lab)).to(dev))
    slow = torch.from_numpy(rep(slows[0])).to(dev)
    born.set_slow(slow)
    born.set_sources(s.src[:, 0], s.src[:, 1], s.src[:, 2], np.zeros(1, np.int32))
    born.set_receivers(s.rec[:, 0], s.rec[:, 1], s.rec[:, 2], np.zeros(len(s.rec), np.int32))
    wav = torch.from_numpy(np.ascontiguousarray(s.wavelet)).to(dev)
    born.set_wavelet(wav)
    born.model()
    ds = torch.from_numpy(rep(dss[0])).to(dev)
    d = torch.zeros((len(s.rec), nw), dtype=torch.complex64, device=dev)
    born.forward(False, ds, d)
    g = torch.zeros_like(ds)
    born.adjoint(False, g, d)
    torch.cuda.synchronize()
    assert rel(one.shot_data(d1, 0).cpu().numpy(), d.cpu().numpy()) < 1e-6
    assert rel(one.shot_grad(g1, 0).cpu().numpy(), g.sum(1).cpu().numpy()) < 1e-5


def test_batch_recomputed_segments_and_dot(dev):
    import torch
    from pyfwix_amd.batching import ShotBatch
    rng = np.random.default_rng(6)
    shots = _shots(rng)
    slows, dss = _models(rng, shots)
    full = ShotBatch(shots, NZ, D, D, D, PARAMS, device=dev)
    d_full, g_full = _run(full, slows, dss)
    seg = ShotBatch(shots, NZ, D, D, D, PARAMS, ckpt_every=3, device=dev)
    d_seg, g_seg = _run(seg, slows, dss)
    torch.cuda.synchronize()
    # the stored run chains all its background steps; a recomputed segment's
    # first step runs its own row pass: equal up to float32 rounding order
    assert rel(d_seg.cpu().numpy(), d_full.cpu().numpy()) < 5e-6
    assert rel(g_seg.cpu().numpy(), g_full.cpu().numpy()) < 5e-6
    # dot test over the whole batched model space (nz, nshot, ny, nx)
    m = torch.from_numpy(crand(rng, tuple(full.dslow.shape)).astype(np.complex64) * 1e-8).to(dev)
    dd = torch.from_numpy(crand(rng, (full.nsamp,)).astype(np.complex64)).to(dev)
    full.dslow = m
    Bm = full.forward()
    BHd = full.adjoint(dd)
    torch.cuda.synchronize()
    lhs = complex(torch.vdot(Bm.cpu().to(torch.complex128).ravel(), dd.cpu().to(torch.complex128).ravel()))
    rhs = complex(torch.vdot(m.cpu().to(torch.complex128).ravel(), BHd.cpu().to(torch.complex128).ravel()))
    assert abs(lhs - rhs) / max(abs(lhs), abs(rhs)) < TOL_DOT


def test_batch_refuses_image_and_reflection_channels(dev):
    import torch
    from pyfwix_amd.batching import ShotBatch
    rng = np.random.default_rng(7)
    shots = _shots(rng)
    slows, dss = _models(rng, shots)
    b = ShotBatch(shots, NZ, D, D, D, PARAMS, device=dev)
    b.set_slowness(slows)
    b.model()
    data = torch.zeros(b.nsamp, dtype=torch.complex64, device=dev)
    img = torch.zeros((NZ, b.ny, b.nx), dtype=torch.complex64, device=dev)
    with pytest.raises(RuntimeError, match="shot batch"):
        b.op.migrate(False, img, data)
    with pytest.raises(RuntimeError, match="plane counts must add up"):
        b.op.set_shot_planes([1, 2])


def test_oneway_batch_equals_shots_alone(dev):
    import torch
    from pyfwix_amd import ops
    from pyfwix_amd.batching import shot_refs
    from pyfwix_amd._lib import FWIX_NSPS, FWIX_PSPI
    rng = np.random.default_rng(8)
    shots = _shots(rng)
    slows, _ = _models(rng, shots)
    NX = NY = 48
    nx, ny = max(s.nx for s in shots), max(s.ny for s in shots)
    P = sum(s.nw for s in shots)
    refs = [shot_refs(a, 3, NX, NY) for a in slows]
    freqs = np.concatenate([s.freqs for s in shots])
    u = [crand(rng, (s.nw, NY, NX)).astype(np.complex64) for s in shots]
    for method in (FWIX_PSPI, FWIX_NSPS):
        for direction in (0, 1):
            dims = ops.make_dims(NX, NY, P, 1, nx, ny, 0.0, D, 0.0, D, 1.0, 0.0, 0.0, D, NZ)
            ow = ops.OneWay(direction, method, dims, nref=3, taperx=4, tapery=4)
            ow.set_shot_planes([s.nw for s in shots])
            ow.set_frequencies(freqs)
            sref = torch.from_numpy(np.ascontiguousarray(np.stack([r[0] for r in refs], 2))).to(dev)
            lab = torch.from_numpy(np.ascontiguousarray(np.stack([r[1] for r in refs], 1))).to(dev)
            ow.set_refs(sref, lab)
            m = torch.from_numpy(np.concatenate(u)[None]).to(dev)
            out = torch.zeros_like(m)
            ow.forward(False, m, out)
            adj = torch.zeros_like(m)
            ow.adjoint(False, adj, m)
            o = 0
            for k, s in enumerate(shots):
                d1 = ops.make_dims(NX, NY, s.nw, 1, s.nx, s.ny, 0.0, D, 0.0, D, 1.0, 0.0, 0.0, D, NZ)
                one = ops.OneWay(direction, method, d1, nref=3, taperx=4, tapery=4)
                one.set_frequencies(s.freqs)
                rep_s = torch.from_numpy(np.ascontiguousarray(np.repeat(refs[k][0][:, :, None], s.nw, 2))).to(dev)
                rep_l = torch.from_numpy(np.ascontiguousarray(np.repeat(refs[k][1][:, None], s.nw, 1))).to(dev)
                one.set_refs(rep_s, rep_l)
                m1 = torch.from_numpy(u[k][None]).to(dev)
                o1 = torch.zeros_like(m1)
                one.forward(False, m1, o1)
                a1 = torch.zeros_like(m1)
                one.adjoint(False, a1, m1)
                torch.cuda.synchronize()
                assert rel(out[0, o:o + s.nw].cpu().numpy(), o1[0].cpu().numpy()) < 1e-6
                assert rel(adj[0, o:o + s.nw].cpu().numpy(), a1[0].cpu().numpy()) < 1e-6
                o += s.nw


def test_batch_equals_shots_alone_on_a_large_grid(dev):
    """Two shots on a 768 x 768 padded grid (apertures 300 x 320 and 330 x 290,
    2 and 3 frequencies), three depths."""
    import torch
    from scipy.ndimage import gaussian_filter
    from pyfwix_amd.batching import Shot, ShotBatch, padded_size
    rng = np.random.default_rng(14)
    nz = 3
    shots, slows, dss = [], [], []
    for nx, ny, nw, f0 in ((300, 320, 2, 6.0), (330, 290, 3, 9.0)):
        assert padded_size(nx) == padded_size(ny) == 768
        freqs = (f0 + 2.0 * np.arange(nw)).astype(np.float32)
        src = np.array([[nx // 2 * D, ny // 2 * D, 0.0]], np.float32)
        gx, gy = np.meshgrid(np.arange(0, nx, 25) * D, np.arange(0, ny, 25) * D)
        rec = np.stack([gx.ravel(), gy.ravel(), np.full(gx.size, (nz - 1) * D)], 1).astype(np.float32)
        shots.append(Shot(nx, ny, freqs, src, crand(rng, (1, nw)).astype(np.complex64), rec))
        v = gaussian_filter(1500.0 + 800.0 * rng.random((nz, ny, nx)), 4.0)
        slows.append((1.0 / v ** 2).astype(np.complex64))
        dss.append((0.05 * slows[-1].real * rng.standard_normal(slows[-1].shape)).astype(np.complex64))
    b = ShotBatch(shots, nz, D, D, D, PARAMS, device=dev)
    data, grad = _run(b, slows, dss)
    torch.cuda.synchronize()
    assert float(data.abs().max()) > 0
    for k, s in enumerate(shots):
        one = ShotBatch([s], nz, D, D, D, PARAMS, device=dev)
        d1, g1 = _run(one, [slows[k]], [dss[k]])
        assert rel(b.shot_data(data, k).cpu().numpy(), one.shot_data(d1, 0).cpu().numpy()) < 1e-6
        assert rel(b.shot_grad(grad, k).cpu().numpy(), one.shot_grad(g1, 0).cpu().numpy()) < 1e-6


def test_born_function_on_a_shot_batch(dev):
    """torch.autograd through a shot-batch operator: BornFunction returns the
    batch's ragged traces, and x.grad of 0.5 ||B ds - d||^2 is the batch's
    migration adjoint of the residual, (nz, nshot, ny, nx)."""
    import torch
    from pyfwix_amd.autograd import born
    from pyfwix_amd.batching import ShotBatch
    rng = np.random.default_rng(9)
    shots = _shots(rng)
    slows, dss = _models(rng, shots)
    b = ShotBatch(shots, NZ, D, D, D, PARAMS, device=dev)
    b.set_slowness(slows)
    b.set_perturbation(dss)
    b.model()
    d_ref = b.forward()
    x = b.dslow.clone().requires_grad_(True)
    assert x.shape == b.op.model_shape()
    d = born(b.op, x)
    assert d.shape == d_ref.shape
    assert torch.equal(d.detach(), d_ref)
    d_obs = torch.from_numpy(crand(rng, tuple(d.shape)).astype(np.complex64)).to(dev) * float(d_ref.abs().max())
    r = (d_ref - d_obs).contiguous()
    d.backward(r)
    g_ref = b.adjoint(r)
    torch.cuda.synchronize()
    assert float(g_ref.abs().max()) > 0
    assert rel(x.grad.cpu().numpy(), g_ref.cpu().numpy()) < 1e-6
    # through a loss: torch forms the residual from abs and pow with float32
    # rounding of its own, so the gradient agrees to a few float32 ulps
    x2 = b.dslow.clone().requires_grad_(True)
    loss = 0.5 * (born(b.op, x2) - d_obs).abs().pow(2).sum()
    loss.backward()
    torch.cuda.synchronize()
    assert rel(x2.grad.cpu().numpy(), g_ref.cpu().numpy()) < 1e-5
