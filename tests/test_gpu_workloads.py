"""The batched (c4) and FWI-gradient (c5) workloads at their own shapes.

c4 (64 shots, apertures U{96..512}^2 padded to 2^a 3^b >= 2n, 50-200
frequencies over 3-40 Hz, nz 250): a multi-shot launch of the 160 GB launch
plan at full depth and frequency count equals its shots run alone and passes
the batched dot test; an oracle window (2 frequencies x 10 depths) on three
c4 shots batched together (float64 scattering arithmetic: 3 Hz is among
the frequencies) matches the FP64 oracle shot by shot.

c5 (512^2 x 400 padded to 1024^2, 128 frequencies over 2-40 Hz, 4 x 4
surface shots, checkpoints every 16): one shot's gradient slice at full
lateral size and frequency count over the first 48 depths with checkpoints
every 16 equals the store-everything gradient bit for bit; the linearized
misfit 0.5 |B ds - d|^2 of that slice agrees with central differences along
3 directions to 1e-3 through autograd; an oracle window (2 frequencies x 10
depths at 1024^2, 2 Hz in float64 scattering arithmetic and 20 Hz in
float32) matches the FP64 oracle.  The workload generators follow
``StreamingPropagator.cpp:68-170`` (shot batches) and ``fwix/workers.py:14-88``
(objective / gradient).
"""
import functools

import numpy as np
import pytest

import oracle as O
from fwix_testutil import TOL_DOT, TOL_WFLD, crand, rel

pytestmark = pytest.mark.gpu

BUDGET = 160 << 30


def _t(x, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def _c4():
    from pyfwix_amd.workloads import SHOT_BATCH_CONFIGS, uneven_shots
    cfg = SHOT_BATCH_CONFIGS["c4"]
    return cfg, uneven_shots(**cfg)


def _params(cfg):
    return {"nref": cfg["nref"], "taper": cfg["taper"], "eps": 0.0, "eps_scat": 0.04, "taylor": 1}


def test_c4_launch_equals_shots_alone_and_dot(dev):
    import torch
    from pyfwix_amd.batching import ShotBatch, plan_launches
    from pyfwix_amd.workloads import shot_models
    cfg, shots = _c4()
    nz, d = cfg["nz"], cfg["d"]
    plan = plan_launches(shots, nz, BUDGET)
    # the multi-shot launch of the smallest padded grid: the case batching is for
    idx, NX, NY, ck = min((p for p in plan if len(p[0]) > 1), key=lambda p: (p[1], -len(p[0])))
    print(f"c4 launch: shots {idx}, {NX}x{NY}, planes {sum(shots[i].nw for i in idx)}, ckpt {ck}", flush=True)
    group = [shots[i] for i in idx]
    models = [shot_models(s, nz, d) for s in group]
    slows = [m[0].astype(np.complex64) for m in models]
    dss = [m[1].astype(np.complex64) for m in models]
    b = ShotBatch(group, nz, d, d, d, _params(cfg), ckpt_every=ck, device=dev, NX=NX, NY=NY)
    b.set_slowness(slows)
    b.set_perturbation(dss)
    b.model()
    data = b.forward()
    grad = b.adjoint(data)
    torch.cuda.synchronize()
    assert float(data.abs().max()) > 0
    for k, s in enumerate(group):
        one = ShotBatch([s], nz, d, d, d, _params(cfg), ckpt_every=ck, device=dev, NX=NX, NY=NY)
        one.set_slowness([slows[k]])
        one.set_perturbation([dss[k]])
        one.model()
        d1 = one.forward()
        g1 = one.adjoint(d1)
        assert rel(b.shot_data(data, k).cpu().numpy(), one.shot_data(d1, 0).cpu().numpy()) < 1e-6
        assert rel(b.shot_grad(grad, k).cpu().numpy(), one.shot_grad(g1, 0).cpu().numpy()) < 1e-6
        del one, d1, g1
    # batched dot test: <B m, y> = <m, B^H y>
    rng = np.random.default_rng(44)
    m = torch.zeros_like(b.dslow)
    m.copy_(_t(crand(rng, tuple(m.shape)).astype(np.complex64), dev))
    for k, s in enumerate(group):  # zero outside each shot's aperture (its model region)
        m[:, k, s.ny:, :] = 0
        m[:, k, :, s.nx:] = 0
    y = _t(crand(rng, (b.nsamp,)).astype(np.complex64), dev)
    Bm = torch.zeros(b.nsamp, dtype=torch.complex64, device=dev)
    b.op.forward(False, m, Bm)
    Bty = torch.zeros_like(m)
    b.op.adjoint(False, Bty, y)
    lhs = torch.sum(Bm.conj().to(torch.complex128) * y.to(torch.complex128))
    rhs = torch.sum(m.conj().to(torch.complex128) * Bty.to(torch.complex128))
    assert abs((lhs / rhs).real.item() - 1.0) < TOL_DOT


def _oracle_shot(shot, slow3, ds3, ref, NX, NY, nz, d, taper, nref, resid):
    """FP64 oracle of one shot: (data (nrec, nw), gradient summed over frequency (nz, ny, nx))."""
    f = np.asarray(shot.freqs, np.float64)
    nw = len(f)
    g = O.Grid(NX=NX, NY=NY, NW=nw, NS=1, nx=shot.nx, ny=shot.ny, dx=d, dy=d, ow=float(f[0]),
               dw=float(f[1] - f[0]) if nw > 1 else 1.0, dz=d, nz=nz)
    sref = np.repeat(ref[0][:, :, None], nw, 2).astype(np.complex128)
    lab = np.repeat(ref[1][:, None], nw, 1).astype(np.int64)
    slow = np.repeat(slow3[:, None], nw, 1).astype(np.complex128)
    ds = np.repeat(ds3[:, None], nw, 1).astype(np.complex128)
    src = tuple(np.asarray(shot.src[:, k], np.float32) for k in range(3)) + (np.zeros(1, np.int32),)
    rec = tuple(np.asarray(shot.rec[:, k], np.float32) for k in range(3)) + (np.zeros(len(shot.rec), np.int32),)
    _, bg = O.model_background(shot.wavelet.astype(np.complex128), src, rec, sref, lab, g, d, 0.0, taper, taper)
    data = O.born_forward(ds, bg, slow, rec, sref, lab, g, d, 0.0, taper, taper, 0.04, 1)
    grad = O.born_adjoint(resid, bg, slow, rec, sref, lab, g, d, 0.0, taper, taper, 0.04, 1).sum(axis=1)
    return data, grad


def test_c4_window_three_shots_match_oracle(dev):
    import torch
    from pyfwix_amd.batching import Shot, ShotBatch, padded_size, shot_refs
    from pyfwix_amd.workloads import shot_models
    cfg, shots = _c4()
    d, nz, taper, nref = cfg["d"], 10, cfg["taper"], cfg["nref"]
    # three c4 shots of one padded grid, two of their frequencies each, the first 10 depths
    by = {}
    for s in shots:
        by.setdefault(padded_size(s.nx), []).append(s)
    N, grp = min(((n, g) for n, g in by.items() if len(g) >= 3), key=lambda t: t[0])
    group = []
    for s in grp[:3]:
        rec = s.rec.copy()
        rec[:, 2] = (nz - 1) * d
        group.append(Shot(s.nx, s.ny, s.freqs[[0, len(s.freqs) // 2]], s.src, s.wavelet[:, [0, len(s.freqs) // 2]],
                          rec[::7]))
    models = [shot_models(s, nz, d) for s in group]      # the workload's three flat reflectors, nz = 10
    slows = [m[0].astype(np.complex64) for m in models]
    dss = [m[1].astype(np.complex64) for m in models]
    rng = np.random.default_rng(45)
    b = ShotBatch(group, nz, d, d, d, _params(cfg), device=dev, NX=N, NY=N)
    # 3 Hz on a 10 m grid: the Taylor term weighs k by k^2 / (w^2 s) up to ~600, beyond float32 at
    # 1e-4 for white residual traces; the scattered-field passes run in float64 arithmetic
    b.op.set_precision(1)
    refs = [shot_refs(sl.real, nref, N, N) for sl in slows]
    b.set_slowness(slows, refs)
    b.set_perturbation(dss)
    b.model()
    data = b.forward()
    resid = [crand(rng, (len(s.rec), s.nw)).astype(np.complex64) for s in group]
    grad = b.adjoint(b.pack_data(resid))
    torch.cuda.synchronize()
    for k, s in enumerate(group):
        # the batch's interior grid is the largest aperture: the oracle runs each shot on the batch grid
        # with its slowness edge-replicated and its perturbation zero outside its aperture
        nx = b.nx
        s_pad = np.ascontiguousarray(b.slow[:, k].cpu().numpy())
        ds_pad = np.ascontiguousarray(b.dslow[:, k].cpu().numpy())
        sh = Shot(nx, nx, s.freqs, s.src, s.wavelet, s.rec)
        ref_d, ref_g = _oracle_shot(sh, s_pad, ds_pad, refs[k], N, N, nz, d, taper, nref,
                                    resid[k].astype(np.complex128))
        assert rel(b.shot_data(data, k).cpu().numpy(), ref_d) < TOL_WFLD
        assert rel(grad[:, k].cpu().numpy(), ref_g) < TOL_WFLD


@functools.lru_cache(maxsize=1)
def _c5_models():
    """(true, start) slowness of the whole c5 model (400 x 512 x 512)."""
    from pyfwix_amd.workloads import FWI_CONFIGS, fwi_models
    return fwi_models(**FWI_CONFIGS["c5"])


def _c5_slice(dev, nz, freqs=None, nrec_step=1, ckpt=16):
    """One c5 shot over the first nz depths of the start model, receivers on the slice's deepest plane."""
    from pyfwix_amd.batching import Shot, ShotBatch, padded_size, shot_refs
    from pyfwix_amd.workloads import FWI_CONFIGS, grid_shots
    cfg = FWI_CONFIGS["c5"]
    d = cfg["d"]
    shot = grid_shots(**cfg)[5]
    s_true, s_start = _c5_models()
    s_true, s_start = s_true[:nz], s_start[:nz]
    rec = shot.rec[::nrec_step].copy()
    rec[:, 2] = (nz - 1) * d
    if freqs is None:
        shot = Shot(shot.nx, shot.ny, shot.freqs, shot.src, shot.wavelet, rec)
    else:
        sel = [int(np.argmin(np.abs(shot.freqs - f))) for f in freqs]
        shot = Shot(shot.nx, shot.ny, shot.freqs[sel], shot.src, shot.wavelet[:, sel], rec)
    N = padded_size(cfg["nx"])
    params = {"nref": cfg["nref"], "taper": cfg["taper"], "eps": 0.0, "eps_scat": 0.04, "taylor": 1}
    b = ShotBatch([shot], nz, d, d, d, params, ckpt_every=ckpt, device=dev, NX=N, NY=N)
    refs = shot_refs(s_start, cfg["nref"], N, N)
    b.set_slowness([s_start.astype(np.complex64)], [refs])
    return cfg, shot, b, s_true, s_start, refs, N


def test_c5_gradient_checkpointing_is_exact(dev):
    """Checkpoints every 16 depths with recomputed segments == every depth stored, bit for bit."""
    import torch
    nz = 48
    grads = []
    for ck in (16, 1):
        cfg, shot, b, s_true, s_start, refs, N = _c5_slice(dev, nz, ckpt=ck)
        data = torch.zeros(b.nsamp, dtype=torch.complex64, device=dev)
        b.op.model(data)
        rng = np.random.default_rng(51)
        data.sub_(_t(crand(rng, (b.nsamp,)).astype(np.complex64), dev) * float(data.abs().max()))
        grads.append(b.adjoint(data).cpu())
        del b, data
        torch.cuda.empty_cache()
    assert torch.equal(grads[0], grads[1])
    assert float(grads[0].abs().max()) > 0


def test_c5_born_misfit_gradient_matches_finite_differences(dev):
    import torch
    from pyfwix_amd.autograd import born, born_misfit_gradient
    nz = 48
    cfg, shot, b, s_true, s_start, refs, N = _c5_slice(dev, nz)
    b.op.model()
    ds_true = torch.from_numpy(np.ascontiguousarray((s_true - s_start)[:, None].astype(np.complex64))).to(dev)
    with torch.no_grad():
        dobs = born(b.op, ds_true)
    x0 = 0.5 * ds_true
    loss, grad = born_misfit_gradient(b.op, x0, dobs)
    scale = float(torch.linalg.vector_norm(x0))

    def misfit(x):
        with torch.no_grad():
            r = born(b.op, x) - dobs
            return 0.5 * float(torch.sum(r.real.double() ** 2 + r.imag.double() ** 2))

    rng = np.random.default_rng(52)
    for i in range(3):
        v = torch.from_numpy(rng.standard_normal(tuple(x0.shape)).astype(np.float32)).to(dev).to(torch.complex64)
        v *= scale / float(torch.linalg.vector_norm(v))
        h = 0.25
        fd = (misfit(x0 + h * v) - misfit(x0 - h * v)) / (2 * h)
        an = float(torch.sum(grad.real.double() * v.real.double() + grad.imag.double() * v.imag.double()))
        print(f"c5 direction {i}: fd {fd:.6e} autograd {an:.6e}", flush=True)
        assert abs(fd - an) <= 1e-3 * abs(an), (fd, an)


@pytest.mark.parametrize("freqs,prec", [((2.0, 2.3), 1), ((20.0, 20.3), 0)], ids=["2Hz-float64", "20Hz-float32"])
def test_c5_window_matches_oracle(dev, freqs, prec):
    import torch
    nz = 10
    cfg, shot, b, s_true, s_start, refs, N = _c5_slice(dev, nz, freqs=freqs, nrec_step=9, ckpt=1)
    b.op.set_precision(prec)
    rng = np.random.default_rng(53)
    ds3 = (s_true - s_start).astype(np.complex64)
    b.set_perturbation([ds3])
    b.model()
    data = b.forward()
    resid = crand(rng, (len(shot.rec), shot.nw)).astype(np.complex64)
    grad = b.adjoint(b.pack_data([resid]))
    torch.cuda.synchronize()
    ref_d, ref_g = _oracle_shot(shot, s_start.astype(np.complex64), ds3, refs, N, N, nz, cfg["d"], cfg["taper"],
                                cfg["nref"], resid.astype(np.complex128))
    assert rel(b.shot_data(data, 0).cpu().numpy(), ref_d) < TOL_WFLD
    assert rel(grad[:, 0].cpu().numpy(), ref_g) < TOL_WFLD
