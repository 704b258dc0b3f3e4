"""Captured depth recursions: a HIP graph of background modeling + Born
forward + Born adjoint (and of a NSPS Downward pass) replays to the eager
results, replay after replay, with other eager work and synchronisations in
between and with the inputs changed in place."""
import numpy as np
import pytest

from fwix_testutil import crand, rel

pytestmark = pytest.mark.gpu

NZ, D = 8, 10.0


def _born(dev, ckpt):
    """Two shots of one padded grid (32 x 32), 4 and 3 frequencies."""
    from scipy.ndimage import gaussian_filter
    from pyfwix_amd.batching import Shot, ShotBatch
    rng = np.random.default_rng(11)
    shots, slows, dss = [], [], []
    for nx, nw, f0 in ((16, 4, 5.0), (13, 3, 8.0)):
        freqs = (f0 + 2.0 * np.arange(nw)).astype(np.float32)
        src = np.array([[nx // 2 * D, nx // 2 * D, 0.0]], np.float32)
        gx, gy = np.meshgrid(np.arange(0, nx, 2) * D, np.arange(0, nx, 2) * D)
        rec = np.stack([gx.ravel(), gy.ravel(), np.full(gx.size, (NZ - 1) * D)], 1).astype(np.float32)
        shots.append(Shot(nx, nx, freqs, src, crand(rng, (1, nw)).astype(np.complex64), rec))
        v = gaussian_filter(1500.0 + 600.0 * rng.random((NZ, nx, nx)), 1.5)
        slows.append((1.0 / v ** 2).astype(np.complex64))
        dss.append((0.05 * slows[-1].real * rng.standard_normal(slows[-1].shape)).astype(np.complex64))
    params = {"nref": 3, "taper": 3, "eps": 0.0, "eps_scat": 0.04, "taylor": 1}
    b = ShotBatch(shots, NZ, D, D, D, params, ckpt_every=ckpt, device=dev)
    b.set_slowness(slows)
    b.set_perturbation(dss)
    return b, rng


@pytest.mark.parametrize("ckpt", [1, 3])
def test_born_graph_replays_match_eager(dev, ckpt):
    import torch
    from pyfwix_amd.graph import StepGraph
    b, rng = _born(dev, ckpt)
    data = torch.zeros(b.nsamp, dtype=torch.complex64, device=dev)
    grad = torch.zeros_like(b.dslow)

    def step():
        b.op.model()
        b.op.forward(False, b.dslow, data)
        b.op.adjoint(False, grad, data)

    step()
    torch.cuda.synchronize()
    d0, g0 = data.clone(), grad.clone()
    g = StepGraph(step)
    ds_other = torch.from_numpy(crand(rng, tuple(b.dslow.shape)).astype(np.complex64) * 1e-8).to(dev)
    ds_keep = b.dslow.clone()
    for k in range(3):
        data.zero_()
        grad.fill_(1.0)
        g.replay()
        torch.cuda.synchronize()
        assert rel(data.cpu().numpy(), d0.cpu().numpy()) < 1e-6
        assert rel(grad.cpu().numpy(), g0.cpu().numpy()) < 1e-6
        # eager work in between, with other inputs
        b.dslow.copy_(ds_other)
        step()
        torch.cuda.synchronize()
        d1 = data.clone()
        b.dslow.copy_(ds_keep)
    # inputs changed in place are seen by the graph
    b.dslow.copy_(ds_other)
    g.replay()
    torch.cuda.synchronize()
    assert rel(data.cpu().numpy(), d1.cpu().numpy()) < 1e-6


def test_oneway_graph_replays_match_eager(dev):
    import torch
    from pyfwix_amd import ops
    from pyfwix_amd.batching import shot_refs
    from pyfwix_amd.graph import StepGraph
    from pyfwix_amd._lib import FWIX_NSPS
    rng = np.random.default_rng(12)
    n, nw = 32, 5
    s3 = (1.0 / (1500.0 + 1500.0 * rng.random((NZ, n // 2, n // 2))) ** 2).astype(np.complex64)
    sref, lab = shot_refs(s3, 3, n, n)
    dims = ops.make_dims(n, n, nw, 1, n // 2, n // 2, 0.0, D, 0.0, D, 5.0, 2.0, 0.0, D, NZ)
    ow = ops.OneWay(0, FWIX_NSPS, dims, nref=3)
    ow.set_refs(torch.from_numpy(np.ascontiguousarray(np.repeat(sref[:, :, None], nw, 2))).to(dev),
                torch.from_numpy(np.ascontiguousarray(np.repeat(lab[:, None], nw, 1))).to(dev))
    m = torch.from_numpy(crand(rng, (1, nw, n, n)).astype(np.complex64)).to(dev)
    out = torch.zeros_like(m)
    adj = torch.zeros_like(m)

    def run():
        ow.forward(False, m, out)
        ow.adjoint(False, adj, out)

    run()
    torch.cuda.synchronize()
    o0, a0 = out.clone(), adj.clone()
    g = StepGraph(run)
    for _ in range(3):
        out.zero_()
        adj.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert rel(out.cpu().numpy(), o0.cpu().numpy()) < 1e-6
        assert rel(adj.cpu().numpy(), a0.cpu().numpy()) < 1e-6
        run()
        torch.cuda.synchronize()
