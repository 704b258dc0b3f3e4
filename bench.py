#!/usr/bin/env python3
"""Benchmark: Born forward + adjoint (migration) on one MI355X.

One step = background down-going modeling (every U_z kept in HBM) + Born
forward (linearized modeling of a smoothed random slowness perturbation) +
Born adjoint (migration of the Born data into the extended model), on the
256^2 x 200 grid padded to 512^2, 64 frequencies over 3-40 Hz, 4 reference
slownesses, 20-node taper.  The metric counts every executed wavefield depth
step (NX*NY*NW*NS points each) per second.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2] [--no-cpu]

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def _device_inputs(w, dev):
    """Upload a workload and expand the frequency-invariant arrays over NW on the GPU."""
    import torch
    nz, NW = w.nz, w.NW
    s3 = torch.from_numpy(w.slow3.astype(np.float32)).to(dev)
    slow = torch.empty((nz, NW, w.ny, w.nx), dtype=torch.complex64, device=dev)
    slow.real.copy_(s3[:, None].expand(nz, NW, w.ny, w.nx))
    slow.imag.zero_()
    d3 = torch.from_numpy(w.dslow3.astype(np.float32)).to(dev)
    ds = torch.empty_like(slow)
    ds.real.copy_(d3[:, None].expand(nz, NW, w.ny, w.nx))
    ds.imag.zero_()
    sref = torch.from_numpy(w.sref1).to(dev).expand(nz, w.nref, NW).contiguous()
    labels = torch.from_numpy(w.labels1).to(dev).expand(nz, NW, w.NY, w.NX).contiguous()
    wav = torch.from_numpy(w.wavelet).to(dev)
    del s3, d3
    return slow, ds, sref, labels, wav


def _dims(w):
    from pyfwix_amd._lib import make_dims
    return make_dims(w.NX, w.NY, w.NW, w.NS, w.nx, w.ny, 0.0, w.dx, 0.0, w.dy, w.ow, w.dw, 0.0, w.dz, w.nz)


def _step_alg_bytes(w, steps_bg, steps_fwd, steps_adj):
    """Compulsory HBM bytes of the executed depth steps (labels as uint8)."""
    P = w.points
    lab = w.NX * w.NY * w.NW
    m = w.nx * w.ny * w.NW
    pspi = 16 * P + lab + 8 * P          # read/write U, labels, checkpoint write
    fwd = 24 * P + 8 * m + lab            # read dU_z, U_z, ds_z; write dU_z+1
    adj = 24 * P + 16 * m + lab           # read lam, U_z; write lam; model slice RMW
    return steps_bg * pspi + steps_fwd * fwd + steps_adj * adj


def cpu_baseline(w, budget_s=30.0):
    """FP32 numpy/scipy restatement (the oracle in complex64, threaded FFTs) of one
    background step, one Born forward step and one Born adjoint step at the full
    slice size; returns Gpts*freq/s over those three depth steps."""
    import oracle as O
    cores = min(16, os.cpu_count() or 1)
    O.set_fft_workers(cores)
    g = O.Grid(NX=w.NX, NY=w.NY, NW=w.NW, NS=w.NS, nx=w.nx, ny=w.ny, dx=w.dx, dy=w.dy, ow=w.ow, dw=w.dw,
               dz=w.dz, nz=w.nz)
    z = w.nz // 2
    rng = np.random.default_rng(0)
    shape = (w.NS, w.NW, w.NY, w.NX)
    U = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    lam = U[:, ::-1].copy()
    sref = np.repeat(w.sref1[z], w.NW, axis=1)
    lab = np.repeat(w.labels1[z], w.NW, axis=0)
    slow_z = np.broadcast_to(w.slow3[z][None], (w.NW, w.ny, w.nx)).astype(np.complex64)
    ds_z = np.broadcast_to(w.dslow3[z][None], (w.NW, w.ny, w.nx)).astype(np.complex64)
    t0 = time.perf_counter()
    O.pspi_forward(U, sref, lab, g, w.dz, 0.0, w.taper, w.taper)                       # background step
    sc = O.scatter_forward(O.ic_forward(ds_z, U), slow_z, g, w.dz, 0.04, 1)          # Born forward step
    O.pspi_forward(U + sc, sref, lab, g, w.dz, 0.0, w.taper, w.taper)
    mu = O.pspi_adjoint(lam, sref, lab, g, w.dz, 0.0, w.taper, w.taper)               # Born adjoint step
    O.ic_adjoint(O.scatter_adjoint(mu, slow_z, g, w.dz, 0.04, 1), U, w.nx, w.ny)
    dt = time.perf_counter() - t0
    return {"value": round(3 * w.points / dt / 1e9, 4), "unit": "Gpts*freq/s", "cores": cores,
            "kind": "port",
            "sample": (f"one background PSPI step + one Born forward step + one Born adjoint step at "
                       f"depth {z} of the full {w.NX}x{w.NY}x{w.NW} slice, numpy/scipy.fft complex64, "
                       f"{dt:.1f} s")}


def plain_vs_fused(w, slow_dims, sref, labels, u, reps=5):
    """PSPI forward + adjoint depth step: fused path vs the plain per-operator path."""
    import torch
    from pyfwix_amd.ops import FWIX_PSPI, FWIX_PSPI_PLAIN, OneStep
    out = {}
    for name, method in (("fused", FWIX_PSPI), ("plain", FWIX_PSPI_PLAIN)):
        op = OneStep(method, slow_dims, nref=w.nref, taperx=w.taper, tapery=w.taper)
        op.set_refs(sref, labels)
        op.set_depth(w.nz // 2)
        for _ in range(2):
            op.forward(False, u, u)
            op.adjoint(False, u, u)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            op.forward(False, u, u)
            op.adjoint(False, u, u)
        torch.cuda.synchronize()
        out[name] = 2 * reps * w.points / (time.perf_counter() - t0) / 1e9
        del op
    return {"fused_gpts": round(out["fused"], 3), "plain_gpts": round(out["plain"], 3),
            "speedup": round(out["fused"] / out["plain"], 3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-path comparison")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); shots and "
              "frequency bands would shard across GPUs with one image all-reduce, which is not built here",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from pyfwix_amd import _lib
    from pyfwix_amd.ops import Born
    from pyfwix_amd.workloads import CONFIGS, born_workload

    dev = torch.device("cuda:0")
    cfg = CONFIGS[args.config]
    w = born_workload(args.config, **cfg)
    slow, ds, sref, labels, wav = _device_inputs(w, dev)
    dims = _dims(w)
    born = Born(dims, nref=w.nref, taperx=w.taper, tapery=w.taper, taylor=1, eps_scat=0.04, ckpt_every=1)
    born.set_refs(sref, labels)
    born.set_slow(slow)
    born.set_sources(*w.src)
    born.set_receivers(*w.rec)
    born.set_wavelet(wav)
    nrec = born.nrec
    data = torch.zeros((nrec, w.NW), dtype=torch.complex64, device=dev)
    grad = torch.zeros_like(ds)

    def step():
        born.model()
        born.forward(False, ds, data)
        born.adjoint(False, grad, data)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    s0 = born.steps
    _lib.timing_reset()
    _lib.timing_enable(1 << 16)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    _lib.timing_disable()
    passes = _lib.timing_collect()
    steps_exec = born.steps - s0
    per = steps_exec // args.steps
    value = steps_exec * w.points / wall / 1e9
    ms_step = wall / args.steps * 1e3

    dom_name, dom = max(passes.items(), key=lambda kv: kv[1]["ms"])
    achieved = dom["bytes"] / (dom["ms"] * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{args.config}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get(dom_name)
    nz1 = w.nz - 1
    alg_total = _step_alg_bytes(w, nz1, nz1, nz1) * args.steps
    res = {
        "metric": "wavefield Gpts*freq/s per depth step (fwd+adj)",
        "value": round(value, 4),
        "unit": "Gpts*freq/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex64",
        "data": "synthetic (seeded velocity model, Ricker source, smoothed random perturbation)",
        "config": {"workload": f"{args.config}: Born forward + adjoint (migration), background modeling "
                               f"included", "nx": w.nx, "ny": w.ny, "nz": w.nz, "NX": w.NX, "NY": w.NY,
                   "nw": w.NW, "shots": w.NS, "nref": w.nref, "taper": w.taper, "receivers": nrec,
                   "depth_steps_per_step": per, "checkpointing": "all depths in HBM"},
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "launches": dom["launches"], "avg_us": round(dom["ms"] * 1e3 / dom["launches"], 2)},
        "step_roofline": {"achieved": round(alg_total / wall / 1e9, 1), "unit": "GB/s",
                          "frac": round(alg_total / wall / 1e9 / HBM_PEAK_GBS, 4),
                          "bytes": "16P+labels per background step (+8P checkpoint), 24P+8m+labels per "
                                   "Born forward step, 24P+16m+labels per adjoint step"},
        "grad_s_per_shot": round(wall / args.steps / w.NS, 4),
        "passes": {k: {"ms": round(v["ms"], 3), "launches": v["launches"],
                       "GBs": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1)} for k, v in
                   sorted(passes.items(), key=lambda kv: -kv[1]["ms"])},
    }
    del born, grad, data
    if not args.no_plain:
        u = torch.randn((w.NS, w.NW, w.NY, w.NX), dtype=torch.complex64, device=dev)
        res["vs_plain"] = plain_vs_fused(w, dims, sref, labels, u)
        del u
    del slow, ds
    torch.cuda.empty_cache()
    res["cpu_baseline"] = None if args.no_cpu else cpu_baseline(w)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
