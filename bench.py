#!/usr/bin/env python3
"""Benchmark: the FWI slowness gradient on one MI355X (the largest workload
that fits one GPU), plus the other workloads of the path.

Default (--config c5): 512^2 x 400 model padded to 1024^2, 128 frequencies
over 2-40 Hz, 4 reference slownesses, 20-node taper, 16 shots on a 4x4
surface grid.  One step = the gradient of one shot: background modeling on
the start model with checkpoints every 16 depths, the L2 residual against
data modeled on the true model (before the timed region), and the Born
adjoint of the residual with segment recomputation.  Steps cycle through the
shots.  The metric counts the background and adjoint depth steps of every
timed shot slice (NX*NY*NW points each) per second; seconds per shot are
reported beside it (`grad_s_per_shot`).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c5] [--no-cpu]

Prints one JSON line.  Before the timed region the column kernels' register
budget is tuned for the padded size (`--no-autotune` keeps the defaults) and
a stream copy measures the achievable HBM bandwidth (`roofline.peak_measured`).

Other workloads (same JSON shape, their own metric where it differs):

    --config c2        Born forward + adjoint (migration) with background
                       modeling, 256^2 x 200 padded to 512^2, 64 frequencies
                       (`--graph` also times the step as one HIP graph)
    --config c1        the c1 grid through the same Born step
    --config c3        NSPS Downward forward + adjoint in the salt model,
                       beside the same recursion with PSPI steps
    --config c4        64 shots of uneven aperture / frequency count as shot
                       batches (`pyfwix_amd/batching.py`), vs one shot per
                       operator, plus a PSPI step sample vs the plain path

Profiling helpers (the summaries under profiles/ are made this way):

    --copy-probe       run a plain device copy of COPY_PROBE_BYTES first, so a
                       counter run can scale FETCH_SIZE / WRITE_SIZE to bytes
    --summarize-stats CSV OUT        kernel_stats.csv of `rocprofv3 --stats`
                                     -> markdown table OUT
    --summarize-traffic FETCH WRITE OUT   counter_collection.csv of two
                                     `rocprofv3 --pmc` runs -> markdown OUT,
                                     read back by the bench for `traffic`
    --summarize-valu SQ STATS OUT    counter_collection.csv of a `rocprofv3
                                     --pmc SQ_...` run + kernel_stats.csv of
                                     the same workload -> markdown OUT, read
                                     back by the bench for `valu_issue`
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
COPY_PROBE_BYTES = 2 << 30          # bytes read (and written) by the copy probe
COPY_PROBE_TAG = "fwix::k_copy16"   # the probe is the library's 16-byte stream copy (fwix_copy),
                                    # the same kernel and size as the bandwidth probe hbm_copy_gbs


def pass_kernel(name, w):
    """Kernel (demangled name) that a timed pass launches, for the c2-shaped PSPI/Born path
    (column kernels carry their register budget, row kernels the shot-map flag)."""
    from pyfwix_amd import autotune
    nx, ny = w.NX, w.NY
    def col(kind, mode):
        stem, wpe = ("k_cols_fanout", "k_cols_fanin")[kind], autotune.get(kind, mode, ny)
        if autotune.get_layout(kind, mode, ny) == 1:
            return f"fwix::{stem}_split<{mode}, {wpe}>"
        return f"fwix::{stem}<{ny}, {mode}, {wpe}>"
    return {"cols_fanout": col(0, 0), "cols_fanin": col(1, 1),
            "born_cols_kfac": col(0, 2), "born_cols_kfac_adj": col(1, 2),
            "rows_select": f"fwix::k_rows_select<{nx}, false>", "rows_fanout": f"fwix::k_rows_fanout<{nx}, false>",
            "born_rows_scatter": f"fwix::k_rows_scatter<{nx}, true, false>",
            "born_rows_inv_scatter_adj": f"fwix::k_rows_scatter_adj<{nx}, true, false>",
            "born_rows_adj_ic": f"fwix::k_rows_adj_out<{nx}, true, false>",
            "born_rows_ic": f"fwix::k_rows_ic_fwd<{nx}, false>"}.get(name)


def _csv_rows(path):
    import csv
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def _short(name):
    name = name.split("(")[0]
    return name[5:] if name.startswith("void ") else name


def summarize_stats(csv_path, out_path, command, top=16,
                    title="c2 Born forward + adjoint (512x512 padded, 64 frequencies, 200 depths)"):
    rows = _csv_rows(csv_path)
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    tot = sum(float(r["TotalDurationNs"]) for r in rows)
    lines = [f"# Kernel time, {title}", "",
             "Command (from the repository root):", "", "```", command, "```", "",
             f"Total kernel time {tot / 1e6:.1f} ms over {sum(int(r['Calls']) for r in rows)} dispatches.", "",
             "| kernel | calls | total ms | avg us | min us | max us | % |", "|---|---|---|---|---|---|---|"]
    for r in rows[:top]:
        lines.append(f"| `{_short(r['Name'])}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.2f} | "
                     f"{float(r['AverageNs']) / 1e3:.2f} | {float(r['MinNs']) / 1e3:.2f} | "
                     f"{float(r['MaxNs']) / 1e3:.2f} | {float(r['Percentage']):.1f} |")
    open(out_path, "w").write("\n".join(lines) + "\n")


def _counter_means(csv_path, counter):
    acc = {}
    for r in _csv_rows(csv_path):
        if r["Counter_Name"] != counter:
            continue
        k = _short(r["Kernel_Name"])
        n, v = acc.get(k, (0, 0.0))
        acc[k] = (n + 1, v + float(r["Counter_Value"]))
    return {k: (n, v / n) for k, (n, v) in acc.items()}


def summarize_traffic(fetch_csv, write_csv, out_path, command_fetch, command_write,
                      title="c2 Born forward + adjoint"):
    fetch = _counter_means(fetch_csv, "FETCH_SIZE")
    write = _counter_means(write_csv, "WRITE_SIZE")
    probe = [k for k in fetch if COPY_PROBE_TAG in k]
    if not probe or not any(COPY_PROBE_TAG in k for k in write):
        raise SystemExit(f"copy probe kernel ({COPY_PROBE_TAG}) not found in the counter runs")
    kib = COPY_PROBE_BYTES / 1024.0
    ff = kib / fetch[probe[0]][1]
    wf = kib / write[[k for k in write if COPY_PROBE_TAG in k][0]][1]
    lines = [f"# HBM traffic per dispatch, {title}", "",
             "Two counter runs (from the repository root):", "", "```", command_fetch, command_write, "```", "",
             f"Calibration: a plain copy of {COPY_PROBE_BYTES} bytes (`{COPY_PROBE_TAG}`, every dispatch of it in the "
             f"run copies that many) reads "
             f"{fetch[probe[0]][1]:.0f} KiB (FETCH_SIZE) and writes {write[[k for k in write if COPY_PROBE_TAG in k][0]][1]:.0f} KiB "
             f"(WRITE_SIZE); scale factors fetch {ff:.4f}, write {wf:.4f} turn the counters into bytes.", "",
             "| kernel | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | traffic bytes (scaled) |",
             "|---|---|---|---|---|"]
    for k, (n, f) in sorted(fetch.items(), key=lambda kv: -kv[1][1] * kv[1][0]):
        if k not in write or COPY_PROBE_TAG in k:
            continue
        wv = write[k][1]
        lines.append(f"| `{k}` | {n} | {f:.1f} | {wv:.1f} | {(f * ff + wv * wf) * 1024:.0f} |")
    open(out_path, "w").write("\n".join(lines) + "\n")


def summarize_valu(sq_csv, stats_csv, out_path, command, command_stats, title, top=12):
    """VALU issue table from a `rocprofv3 --pmc SQ_INSTS_VALU SQ_WAVES SQ_INSTS_VMEM_RD
    SQ_INSTS_VMEM_WR SQ_INSTS_LDS SQ_INSTS_SALU` run and the kernel times of a
    `--kernel-trace --stats` run of the same workload (read back by the bench
    for `valu_issue`)."""
    names = ("SQ_INSTS_VALU", "SQ_INSTS_LDS", "SQ_INSTS_VMEM_RD", "SQ_INSTS_VMEM_WR")
    cnt = {c: _counter_means(sq_csv, c) for c in names}
    times = {_short(r["Name"]): (float(r["AverageNs"]) / 1e3, float(r["TotalDurationNs"]))
             for r in _csv_rows(stats_csv)}
    lines = [f"# VALU issue, {title}", "", "Counter run and kernel-time run (from the repository root):", "",
             "```", command, command_stats, "```", "",
             f"Peak VALU issue of the code as built (non-packed FP32; packed FP32 is switched off): one wave64 "
             f"instruction per cycle per CU (4 SIMD16 units, 4 cycles each), 256 CUs x 2.4 GHz = "
             f"{VALU_PEAK_GINSTR} G wave-instructions/s. `issue frac` = VALU wave-instructions per dispatch / "
             f"(avg kernel time x {VALU_PEAK_GINSTR} G/s); transcendental instructions (sin, cos, sqrt, exp, rcp) "
             f"issue at a quarter of that rate, so cycle occupancy is somewhat higher than the fraction shown.", "",
             "| kernel | dispatches | VALU wave-instr / dispatch | LDS instr / dispatch | VMEM rd / wr per dispatch | "
             "avg us | issue frac |", "|---|---|---|---|---|---|---|"]
    rows = [k for k in cnt["SQ_INSTS_VALU"] if k in times]
    rows.sort(key=lambda k: -times[k][1])
    for k in rows[:top]:
        n, valu = cnt["SQ_INSTS_VALU"][k]
        avg_us = times[k][0]
        lds = cnt["SQ_INSTS_LDS"].get(k, (0, 0.0))[1]
        rd = cnt["SQ_INSTS_VMEM_RD"].get(k, (0, 0.0))[1]
        wr = cnt["SQ_INSTS_VMEM_WR"].get(k, (0, 0.0))[1]
        frac = valu / (avg_us * 1e-6) / (VALU_PEAK_GINSTR * 1e9)
        lines.append(f"| `{k}` | {n} | {valu:.0f} | {lds:.0f} | {rd:.0f} / {wr:.0f} | {avg_us:.2f} | {frac:.3f} |")
    open(out_path, "w").write("\n".join(lines) + "\n")


def _profile_row(path, kernel, ncells, col):
    """Cell `col` of the row of `kernel` in a profile table, or None.  A column
    kernel whose register budget differs from the profiled one matches the
    row of the same kernel at another budget."""
    if kernel is None or not os.path.exists(path):
        return None
    rows = {}
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == ncells and cells[0].startswith("`"):
            rows[cells[0].strip("`")] = cells[col]
    if kernel in rows:
        return float(rows[kernel].split("/")[0])
    if "k_cols_" in kernel:
        stem = kernel.rsplit(",", 1)[0]
        for k, v in rows.items():
            if k.rsplit(",", 1)[0] == stem:
                return float(v.split("/")[0])
    return None


def traffic_for(config, kernel):
    """Scaled HBM bytes per dispatch of `kernel` from profiles/traffic_<config>.md, or None."""
    return _profile_row(os.path.join(ROOT, "profiles", f"traffic_{config}.md"), kernel, 5, 4)


VALU_PEAK_GINSTR = 614.4   # wave64 VALU instructions per second (1e9): 256 CUs x 1 per cycle x 2.4 GHz


def valu_for(config, kernel):
    """VALU wave-instructions per dispatch of `kernel` from profiles/valu_<config>.md, or None."""
    return _profile_row(os.path.join(ROOT, "profiles", f"valu_{config}.md"), kernel, 7, 2)


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
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": round(3 * w.points / dt / 1e9, 4), "unit": "Gpts*freq/s", "cores": cores,
            "cpu_model": model, "host_cpus": os.cpu_count(),
            "kind": "port",
            "sample": (f"one background PSPI step + one Born forward step + one Born adjoint step at "
                       f"depth {z} of the full {w.NX}x{w.NY}x{w.NW} slice, numpy/scipy.fft complex64, "
                       f"{dt:.1f} s")}


def plain_vs_fused(w, slow_dims, sref, labels, u, reps=5, copy_gbs=None):
    """PSPI forward + adjoint depth step: fused path vs the plain per-operator path
    (with the plain kernels' per-pass bandwidth when copy_gbs is given)."""
    import torch
    from pyfwix_amd.ops import FWIX_PSPI, FWIX_PSPI_PLAIN, OneStep
    out = {}
    table = None
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
        if method == FWIX_PSPI_PLAIN and copy_gbs:
            def _once():
                op.forward(False, u, u)
                op.adjoint(False, u, u)
                torch.cuda.synchronize()
            table = plain_pass_table(copy_gbs, _once)
        del op
    return {"fused_gpts": round(out["fused"], 3), "plain_gpts": round(out["plain"], 3),
            "speedup": round(out["fused"] / out["plain"], 3), "plain_passes": table}


def hbm_copy_gbs(dev, nbytes=2 << 30, reps=10):
    """Achievable HBM bandwidth: the library's 16-byte stream copy of `nbytes`,
    read + write bytes over the mean HIP-event time of `reps` copies."""
    import torch
    from pyfwix_amd import _lib
    x = torch.empty(nbytes // 4, dtype=torch.float32, device=dev).fill_(1.0)
    y = torch.empty_like(x)
    st = torch.cuda.current_stream()
    for _ in range(2):
        _lib.call("fwix_copy", x.data_ptr(), y.data_ptr(), nbytes, st.cuda_stream)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(st)
    for _ in range(reps):
        _lib.call("fwix_copy", x.data_ptr(), y.data_ptr(), nbytes, st.cuda_stream)
    b.record(st)
    torch.cuda.synchronize()
    gbs = 2.0 * nbytes * reps / (a.elapsed_time(b) * 1e-3) / 1e9
    ok = bool(torch.equal(x[:1024], y[:1024]))
    del x, y
    torch.cuda.empty_cache()
    if not ok:
        raise RuntimeError("stream copy probe produced wrong data")
    return gbs


def autotune_sizes(sizes, refs=None, freqs=None):
    """Tune the column kernels' register budget at every padded size (outside
    the timed region), on a real depth's references when given; -> {size:
    {"chosen": ..., "us": {...}}, "seconds": t}."""
    from pyfwix_amd import autotune
    t0 = time.perf_counter()
    out = {}
    for n in sizes:
        r = autotune.tune_size(int(n), refs=refs, freqs=freqs)
        out[str(n)] = {"chosen": r["chosen"], "us": {k: v for k, v in r.items() if k != "chosen"}}
        print(f"[tune] {n}: {r['chosen']}", file=sys.stderr, flush=True)
    out["seconds"] = round(time.perf_counter() - t0, 2)
    return out


def bench_shot_batch(args):
    """Batched multi-shot modeling + migration (64 shots of uneven aperture and
    frequency count): every launch of the plan runs background modeling + Born
    forward + Born adjoint for its shots; the same shots one per operator; and
    a PSPI forward + adjoint step sample, batched fused vs per-shot plain."""
    import concurrent.futures as cf

    import torch
    from pyfwix_amd import _lib
    from pyfwix_amd.batching import ShotBatch, min_ckpt, padded_size, plan_launches, shot_refs
    from pyfwix_amd.ops import FWIX_PSPI, FWIX_PSPI_PLAIN, OneStep, make_dims
    from pyfwix_amd.workloads import SHOT_BATCH_CONFIGS, shot_models, uneven_shots

    cfg = SHOT_BATCH_CONFIGS[args.config]
    dev = torch.device("cuda:0")
    nz, d, nref, taper = cfg["nz"], cfg["d"], cfg["nref"], cfg["taper"]
    params = {"nref": nref, "taper": taper, "eps": 0.0, "eps_scat": 0.04, "taylor": 1}
    shots = uneven_shots(**cfg)
    if args.shots:
        shots = shots[:args.shots]
    budget = int(args.budget_gb * (1 << 30))
    plan = plan_launches(shots, nz, budget)
    t_prep = time.perf_counter()

    def prep(k):
        s = shots[k]
        s3, ds3 = shot_models(s, nz, d)
        return s3, ds3, shot_refs(s3, nref, padded_size(s.nx), padded_size(s.ny))

    with cf.ThreadPoolExecutor(8) as ex:
        models = list(ex.map(prep, range(len(shots))))
    t_prep = time.perf_counter() - t_prep
    print(f"[c4] {len(shots)} shots, {len(plan)} launches, host models {t_prep:.1f} s", file=sys.stderr, flush=True)
    nominal = 3 * (nz - 1)                     # background + forward + adjoint steps per shot slice
    pts = sum(s.nw * padded_size(s.nx) * padded_size(s.ny) for s in shots)

    def run(launches, timed_passes=False):
        total, steps = 0.0, 0
        if timed_passes:
            _lib.timing_reset()
            _lib.timing_select(None, max(1, args.pass_sample))
            _lib.timing_enable(1 << 15)
        for idx, NX, NY, k in launches:
            b = ShotBatch([shots[i] for i in idx], nz, d, d, d, params, ckpt_every=k, device=dev, NX=NX, NY=NY,
                          planes_per_chunk=args.chunk)
            b.set_slowness([models[i][0] for i in idx], [models[i][2] for i in idx])
            b.set_perturbation([models[i][1] for i in idx])
            torch.cuda.synchronize()
            s0 = b.op.steps
            t0 = time.perf_counter()
            b.model()
            data = b.forward()
            b.adjoint(data)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            total += dt
            steps += b.op.steps - s0
            print(f"[c4] launch {NX}x{NY} shots={len(idx)} planes={b.planes} ckpt={k}: {dt:.3f} s",
                  file=sys.stderr, flush=True)
            del b, data
            torch.cuda.empty_cache()
        passes = {}
        if timed_passes:
            _lib.timing_disable()
            passes = _lib.timing_collect()
            _lib.timing_select(None, 1)
        return total, steps, passes

    tuned = None if args.no_autotune else autotune_sizes(sorted({NX for _i, NX, _NY, _k in plan}))
    # warm-up: one small shot, a few depths
    w0 = ShotBatch([shots[0]], 4, d, d, d, params, device=dev)
    w0.set_slowness([models[0][0][:4]])
    w0.set_perturbation([models[0][1][:4]])
    w0.model()
    w0.adjoint(w0.forward())
    torch.cuda.synchronize()
    del w0
    t_b, steps_b, passes = run(plan, timed_passes=not args.no_pass_timing)
    print(f"[c4] batched {t_b:.2f} s", file=sys.stderr, flush=True)
    alone = [([i], padded_size(shots[i].nx), padded_size(shots[i].ny),
              min_ckpt(shots[i].nw, 1, padded_size(shots[i].nx), padded_size(shots[i].ny), shots[i].nx,
                       shots[i].ny, nz, budget)) for i in range(len(shots))]
    t_s, steps_s, _ = (None, None, None) if args.no_per_shot else run(alone)
    if t_s is not None:
        print(f"[c4] per shot {t_s:.2f} s", file=sys.stderr, flush=True)

    # PSPI forward + adjoint step sample (depths 0 and 1): per-shot plain vs batched fused
    def step_rate(launches, method, batched):
        total, npts = 0.0, 0
        for idx, NX, NY, _k in launches:
            sh = [shots[i] for i in idx]
            P = sum(s.nw for s in sh)
            dims = make_dims(NX, NY, P, 1, max(s.nx for s in sh), max(s.ny for s in sh), 0.0, d, 0.0, d,
                             1.0, 0.0, 0.0, d, 2)
            op = OneStep(method, dims, nref=nref, taperx=taper, tapery=taper)
            if batched:
                op.set_shot_planes([s.nw for s in sh])
                sref = np.stack([models[i][2][0][:2] for i in idx], 2)
                lab = np.stack([models[i][2][1][:2] for i in idx], 1)
            else:
                sref = np.repeat(models[idx[0]][2][0][:2, :, None], P, 2)
                lab = np.repeat(models[idx[0]][2][1][:2, None], P, 1)
            op.set_frequencies(np.concatenate([s.freqs for s in sh]))
            refs = (torch.from_numpy(np.ascontiguousarray(sref)).to(dev),
                    torch.from_numpy(np.ascontiguousarray(lab)).to(dev))
            op.set_refs(*refs)
            u = torch.randn((1, P, NY, NX), dtype=torch.complex64, device=dev)

            def once():
                for z in (0, 1):
                    op.set_depth(z)
                    op.forward(False, u, u)
                    op.adjoint(False, u, u)
            once()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            once()
            torch.cuda.synchronize()
            total += time.perf_counter() - t0
            npts += 4 * P * NX * NY
            del op, u, refs
        torch.cuda.empty_cache()
        return npts / total / 1e9

    sample = None
    if not args.no_plain:
        g_plain = step_rate(alone, FWIX_PSPI_PLAIN, False)
        g_batch = step_rate(plan, FWIX_PSPI, True)
        sample = {"plain_per_shot_gpts": round(g_plain, 3), "batched_fused_gpts": round(g_batch, 3),
                  "speedup": round(g_batch / g_plain, 3),
                  "sample": "PSPI forward + adjoint at depths 0 and 1 of every shot (4 steps per shot)"}
    roof = None
    if passes:
        est = {k: v["ms"] / v["launches"] * v["total_launches"] for k, v in passes.items()}
        dom_name = max(est, key=est.get)
        dom = passes[dom_name]
        avg_us = dom["ms"] * 1e3 / dom["launches"]
        achieved = dom["bytes"] / (dom["ms"] * 1e-3) / 1e9
        roof = {"bound": "hbm", "kernel": dom_name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "alg_bytes_per_launch": round(dom["bytes"] / dom["launches"]), "avg_us": round(avg_us, 2),
                "launches_timed": dom["launches"], "launches": dom["total_launches"],
                "note": "all padded grid sizes of the plan together"}
    classes = {}
    for idx, NX, NY, k in plan:
        c = classes.setdefault(f"{NX}x{NY}", {"shots": 0, "launches": 0, "planes": 0, "ckpt_every": set()})
        c["shots"] += len(idx)
        c["launches"] += 1
        c["planes"] += sum(shots[i].nw for i in idx)
        c["ckpt_every"].add(k)
    for c in classes.values():
        c["ckpt_every"] = sorted(c["ckpt_every"])
    value = nominal * pts / t_b / 1e9
    res = {
        "metric": "wavefield Gpts*freq/s per depth step (fwd+adj)",
        "value": round(value, 4),
        "unit": "Gpts*freq/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": round(t_b * 1e3, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex64",
        "data": "synthetic (seeded apertures / frequency counts, smooth velocity, three flat reflectors)",
        "config": {"workload": f"{args.config}: batched multi-shot Born modeling + migration (background "
                               f"modeling included), {len(shots)} shots",
                   "nz": nz, "aperture_nodes": [cfg["nmin"], cfg["nmax"]], "freqs_per_shot": [cfg["wmin"], cfg["wmax"]],
                   "band_hz": [cfg["fmin"], cfg["fmax"]], "nref": nref, "taper": taper,
                   "receivers": f"every {cfg['rec_step']}th node of the aperture on the deepest plane",
                   "hbm_budget_gb": args.budget_gb, "launches": len(plan), "size_classes": classes,
                   "points_per_depth": pts, "depth_steps_counted": f"{nominal} per shot slice (background, "
                   f"forward, adjoint); recomputed background steps not counted",
                   "depth_steps_executed_batched": steps_b},
        "roofline": roof,
        "per_shot": None if t_s is None else {
            "value": round(nominal * pts / t_s / 1e9, 4), "seconds": round(t_s, 3), "steps_executed": steps_s,
            "form": "the same fused kernels, one shot per operator and launch"},
        "batching_speedup": None if t_s is None else round(t_s / t_b, 3),
        "vs_plain_per_shot": sample,
        "column_budgets": tuned,
        "host_model_prep_s": round(t_prep, 1),
        "cpu_baseline": None,
    }
    print(json.dumps(res), flush=True)


def bench_nsps(args):
    """NSPS propagation (Downward forward + adjoint over every depth) in the
    layered + salt model, beside the same recursion with PSPI steps."""
    import torch
    from pyfwix_amd import _lib
    from pyfwix_amd.ops import FWIX_NSPS, FWIX_PSPI, OneWay
    from pyfwix_amd.workloads import NSPS_CONFIGS, nsps_workload, workload_dims

    dev = torch.device("cuda:0")
    t_prep = time.perf_counter()
    w = nsps_workload(args.config, **NSPS_CONFIGS[args.config])
    t_prep = time.perf_counter() - t_prep
    dims = workload_dims(w)
    sref = torch.from_numpy(w.sref1).to(dev).expand(w.nz, w.nref, w.NW).contiguous()
    labels = torch.from_numpy(w.labels1).to(dev).expand(w.nz, w.NW, w.NY, w.NX).contiguous()
    zt = w.nz // 2
    tuned = None if args.no_autotune else autotune_sizes(
        [w.NY], refs=(sref[zt:zt + 1].contiguous(), labels[zt:zt + 1].contiguous()),
        freqs=w.ow + w.dw * np.arange(w.NW))
    # the wavelet injected at the source node of every frequency plane
    u0 = torch.zeros((1, w.NW, w.NY, w.NX), dtype=torch.complex64, device=dev)
    u0[0, :, w.ny // 2, w.nx // 2] = torch.from_numpy(w.wavelet[0]).to(dev)
    out = {}
    passes = {}
    for name, method in (("nsps", FWIX_NSPS), ("pspi", FWIX_PSPI)):
        op = OneWay(0, method, dims, nref=w.nref, taperx=w.taper, tapery=w.taper)
        op.set_refs(sref, labels)
        d = torch.zeros_like(u0)
        m = torch.zeros_like(u0)
        op.forward(False, u0, d)
        op.adjoint(False, m, d)
        torch.cuda.synchronize()
        timed = name == "nsps" and not args.no_pass_timing
        if timed:
            _lib.timing_reset()
            _lib.timing_select(None, max(1, args.pass_sample))
            _lib.timing_enable(1 << 15)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            op.forward(False, u0, d)
            op.adjoint(False, m, d)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        if timed:
            _lib.timing_disable()
            passes = _lib.timing_collect()
            _lib.timing_select(None, 1)
        out[name] = (2 * w.nz * w.points * args.steps / wall / 1e9, wall / args.steps)
        del op, d, m
        torch.cuda.empty_cache()
    roof = None
    if passes:
        est = {k: v["ms"] / v["launches"] * v["total_launches"] for k, v in passes.items()}
        dom_name = max(est, key=est.get)
        dom = passes[dom_name]
        avg_us = dom["ms"] * 1e3 / dom["launches"]
        achieved = dom["bytes"] / dom["launches"] / (avg_us * 1e-6) / 1e9
        roof = {"bound": "hbm", "kernel": dom_name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "alg_bytes_per_launch": round(dom["bytes"] / dom["launches"]), "avg_us": round(avg_us, 2),
                "launches_timed": dom["launches"], "launches": dom["total_launches"]}
    res = {
        "metric": "wavefield Gpts*freq/s per depth step (fwd+adj)",
        "value": round(out["nsps"][0], 4),
        "unit": "Gpts*freq/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 1,
        "ms_per_step": round(out["nsps"][1] * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex64",
        "data": "synthetic (layered 1500-4500 m/s + salt ellipsoid + 3 % seeded fluctuation, Ricker point source)",
        "config": {"workload": f"{args.config}: NSPS Downward forward + adjoint over every depth, single shot",
                   "nx": w.nx, "ny": w.ny, "nz": w.nz, "NX": w.NX, "NY": w.NY, "nw": w.NW, "nref": w.nref,
                   "band_hz": [w.ow, w.ow + w.dw * (w.NW - 1)]},
        "roofline": roof,
        "pspi_same_recursion": {"value": round(out["pspi"][0], 4), "ms_per_step": round(out["pspi"][1] * 1e3, 3)},
        "passes": {k: {"avg_us": round(v["ms"] * 1e3 / v["launches"], 2),
                       "GBs": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1)} for k, v in passes.items()},
        "column_budgets": tuned,
        "host_model_prep_s": round(t_prep, 1),
        "cpu_baseline": None,
    }
    print(json.dumps(res), flush=True)


def _roofline(passes, config, kernel_of, copy_gbs, ev_us):
    """Roofline of the dominant pass (largest estimated total time) from the
    sampled per-pass event times; traffic / VALU from the committed counter
    summaries of the same config when they exist."""
    if not passes:
        return None
    est = {k: v["ms"] / v["launches"] * v["total_launches"] for k, v in passes.items()}
    dom_name = max(est, key=est.get)
    dom = passes[dom_name]
    avg_us = dom["ms"] * 1e3 / dom["launches"]
    per_launch = dom["bytes"] / dom["launches"]
    achieved = per_launch / (avg_us * 1e-6) / 1e9
    kern = kernel_of(dom_name)
    vi = valu_for(config, kern)
    return {"bound": "hbm", "kernel": dom_name, "kernel_symbol": kern, "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "peak_measured": round(copy_gbs, 1), "frac_measured": round(achieved / copy_gbs, 4),
            "traffic": traffic_for(config, kern),
            "traffic_source": f"profiles/traffic_{config}.md (two rocprofv3 --pmc runs of this config)",
            "valu_issue": None if vi is None else {
                "instr_per_launch": vi, "achieved": round(vi / avg_us / 1e3, 1), "peak": VALU_PEAK_GINSTR,
                "unit": "G wave-instr/s", "frac": round(vi / avg_us / 1e3 / VALU_PEAK_GINSTR, 4)},
            "alg_bytes_per_launch": round(per_launch), "avg_us": round(avg_us, 2),
            "launches_timed": dom["launches"], "launches": dom["total_launches"],
            "empty_event_pair_us": None if ev_us is None else round(ev_us, 2)}


def _passes_table(passes):
    return {k: {"ms_est": round(v["ms"] / v["launches"] * v["total_launches"], 3),
                "launches": v["total_launches"], "timed": v["launches"],
                "avg_us": round(v["ms"] * 1e3 / v["launches"], 2),
                "GBs": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1)} for k, v in
            sorted(passes.items(), key=lambda kv: -kv[1]["ms"] / kv[1]["launches"] * kv[1]["total_launches"])}


def plain_pass_table(copy_gbs, timed):
    """Per-pass bandwidth of the plain per-operator kernels recorded while
    `timed` ran (every launch timed with HIP events), against the copy rate."""
    from pyfwix_amd import _lib
    _lib.timing_reset()
    _lib.timing_select(None, 1)
    _lib.timing_enable(1 << 14)
    try:
        timed()
    finally:
        _lib.timing_disable()
    passes = _lib.timing_collect()
    _lib.timing_reset()
    out = {}
    for k, v in sorted(passes.items()):
        if not k.startswith("plain_") or not v["launches"]:
            continue
        gbs = v["bytes"] / (v["ms"] * 1e-3) / 1e9
        out[k] = {"avg_us": round(v["ms"] * 1e3 / v["launches"], 2), "GBs": round(gbs, 1),
                  "frac_of_copy": round(gbs / copy_gbs, 3), "launches": v["launches"]}
    return out


def cpu_baseline_fwi(N, nw, nx, taper, slow3, refs, d, freqs, budget_planes=16):
    """FP32 numpy/scipy restatement (the oracle in complex64, threaded FFTs) of
    one background PSPI step and one Born adjoint step of the FWI slice at its
    full lateral size on `budget_planes` of its frequencies; Gpts*freq/s."""
    import oracle as O
    cores = min(16, os.cpu_count() or 1)
    O.set_fft_workers(cores)
    nwp = min(nw, budget_planes)
    g = O.Grid(NX=N, NY=N, NW=nwp, NS=1, nx=nx, ny=nx, dx=d, dy=d, ow=float(freqs[0]),
               dw=float(freqs[1] - freqs[0]), dz=d, nz=slow3.shape[0])
    z = slow3.shape[0] // 2
    rng = np.random.default_rng(0)
    shape = (1, nwp, N, N)
    U = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    lam = U[:, ::-1].copy()
    sref = np.repeat(refs[0][z][:, None], nwp, axis=1).astype(np.complex64)
    lab = np.repeat(refs[1][z][None], nwp, axis=0)
    slow_z = np.broadcast_to(slow3[z][None], (nwp, nx, nx)).astype(np.complex64)
    t0 = time.perf_counter()
    O.pspi_forward(U, sref, lab, g, d, 0.0, taper, taper)                              # background step
    mu = O.pspi_adjoint(lam, sref, lab, g, d, 0.0, taper, taper)                       # Born adjoint step
    O.ic_adjoint(O.scatter_adjoint(mu, slow_z, g, d, 0.04, 1), U, nx, nx)
    dt = time.perf_counter() - t0
    return {"value": round(2 * nwp * N * N / dt / 1e9, 4), "unit": "Gpts*freq/s", "cores": cores,
            "host_cpus": os.cpu_count(), "kind": "port",
            "sample": (f"one background PSPI step + one Born adjoint step at depth {z} of one shot slice "
                       f"({N}x{N} padded, {nwp} of its {nw} frequencies), numpy/scipy.fft complex64, "
                       f"{dt:.1f} s")}


def bench_fwi(args):
    """FWI-style slowness gradient, one group of shots per step (default one
    shot): background modeling on the start model with checkpoints every
    `ckpt` depths, the L2 data residual against data modeled on the true model
    (before the timed region), and the Born adjoint of the residual with
    segment recomputation; the gradient is summed over shots.  Steps cycle
    through the shots.  The value counts the background and adjoint depth
    steps of every timed shot slice (recomputed steps not counted); the
    seconds per shot are reported beside it."""
    import torch
    from pyfwix_amd import _lib
    from pyfwix_amd.batching import ShotBatch, padded_size, shot_refs
    from pyfwix_amd.ops import FWIX_PSPI, FWIX_PSPI_PLAIN, OneStep, make_dims
    from pyfwix_amd.workloads import FWI_CONFIGS, fwi_models, grid_shots

    cfg = FWI_CONFIGS[args.config]
    dev = torch.device("cuda:0")
    nz, d, nref, taper, ck = cfg["nz"], cfg["d"], cfg["nref"], cfg["taper"], cfg["ckpt"]
    nx = cfg["nx"]
    params = {"nref": nref, "taper": taper, "eps": 0.0, "eps_scat": 0.04, "taylor": 1}
    t_prep = time.perf_counter()
    shots = grid_shots(**cfg)
    if args.shots:
        shots = shots[:args.shots]
    s_true, s_start = fwi_models(**cfg)
    N = padded_size(nx)
    refs_true = shot_refs(s_true, nref, N, N)
    refs_start = shot_refs(s_start, nref, N, N)
    t_prep = time.perf_counter() - t_prep
    per = max(1, min(args.shots_per_step, len(shots)))
    groups = [shots[i:i + per] for i in range(0, len(shots) - per + 1, per)]
    nw = shots[0].nw
    print(f"[fwi] {len(shots)} shots, {per} per step, {len(groups)} groups, host models {t_prep:.1f} s",
          file=sys.stderr, flush=True)
    copy_gbs = hbm_copy_gbs(dev)
    zt = nz // 2
    # tune on the middle depth's references over every plane of one step
    npl = nw * per
    tuned = None if args.no_autotune else autotune_sizes(
        [N], refs=(torch.from_numpy(np.ascontiguousarray(np.repeat(refs_start[0][zt:zt + 1, :, None], npl, 2))).to(dev),
                   torch.from_numpy(refs_start[1][zt:zt + 1, None]).to(dev).expand(1, npl, N, N).contiguous()),
        freqs=np.tile(shots[0].freqs, per))
    b = ShotBatch(groups[0], nz, d, d, d, params, ckpt_every=ck, device=dev, NX=N, NY=N,
                  planes_per_chunk=args.chunk)
    # observed data on the true model (outside the timed region)
    t_obs = time.perf_counter()
    b.set_slowness([s_true] * per, [refs_true] * per)
    observed = []
    for grp in groups:
        b.set_shots(grp)
        dobs = torch.zeros(b.nsamp, dtype=torch.complex64, device=dev)
        b.op.model(dobs)
        observed.append(dobs)
    torch.cuda.synchronize()
    t_obs = time.perf_counter() - t_obs
    b.set_slowness([s_start] * per, [refs_start] * per)
    grad = torch.zeros((nz, nx, nx), dtype=torch.complex64, device=dev)
    data = torch.zeros(b.nsamp, dtype=torch.complex64, device=dev)
    g = torch.zeros((nz, per, b.ny, b.nx), dtype=torch.complex64, device=dev)
    misfit = torch.zeros((), dtype=torch.float32, device=dev)
    k_step = [0]

    def step():
        k = k_step[0] % len(groups)
        k_step[0] += 1
        b.set_shots(groups[k])
        b.op.model(data)                          # background on the start model, data recorded
        data.sub_(observed[k])                    # L2 residual
        misfit.add_(0.5 * torch.linalg.vector_norm(data) ** 2)
        b.op.adjoint(False, g, data)              # gradient of 0.5 |d - d_obs|^2 (Born adjoint)
        grad.add_(g.sum(1))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    grad.zero_()
    misfit.zero_()
    s0 = b.op.steps
    timed = not args.no_pass_timing
    if timed:
        _lib.timing_reset()
        _lib.timing_select(None, max(1, args.pass_sample))
        _lib.timing_enable(1 << 15)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    passes, ev_us = {}, None
    if timed:
        _lib.timing_disable()
        passes = _lib.timing_collect()
        _lib.timing_select(None, 1)
        ev_us = _lib.timing_overhead(torch.cuda.current_stream().cuda_stream, 200)
    steps_exec = b.op.steps - s0
    counted = 2 * (nz - 1) * args.steps           # background + adjoint steps of each timed shot slice
    P = b.points
    value = counted * P / wall / 1e9

    class _W:  # pass -> kernel symbol: a shot batch maps planes to shots (row kernels' MAP flag)
        NX = NY = N

    def kern(name):
        k = (pass_kernel(name, _W) or "").replace(", false>", ", true>") or None
        # the background steps are chained: their select pass also leaves the next step's spectrum
        return f"fwix::k_rows_select<{N}, true, true>" if name == "rows_select" else k
    roof = _roofline(passes, args.config, kern, copy_gbs, ev_us)
    lab = N * N * b.planes
    m = nx * nx * b.planes
    alg = (nz - 1) * args.steps * ((16 * P + lab) + (24 * P + 16 * m + lab)) + (nz // ck) * 8 * P * args.steps
    res = {
        "metric": "wavefield Gpts*freq/s per depth step (fwd+adj)",
        "value": round(value, 4),
        "unit": "Gpts*freq/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex64",
        "data": "synthetic (smooth background + seeded Gaussian anomalies; start model smoothed in closed form; "
                "observed data modeled on the true model)",
        "config": {"workload": f"{args.config}: FWI slowness gradient (background modeling on the start model "
                               f"with checkpoints every {ck} depths, L2 residual, Born adjoint with segment "
                               f"recomputation), {per} shot(s) per step cycling over {len(groups) * per} shots",
                   "nx": nx, "ny": nx, "nz": nz, "NX": N, "NY": N, "nw": nw, "band_hz": [cfg["fmin"], cfg["fmax"]],
                   "nref": nref, "taper": taper, "ckpt_every": ck, "shots_per_step": per,
                   "receivers": f"every {cfg['rec_step']}th node on the deepest plane",
                   "planes_per_chunk": args.chunk or "auto"},
        "grad_s_per_shot": round(wall / args.steps / per, 4),
        "depth_steps": {"counted_per_step": 2 * (nz - 1), "executed_per_step": steps_exec // args.steps,
                        "note": "background and adjoint steps counted; recomputed background segments not"},
        "roofline": roof,
        "step_roofline": {"achieved": round(alg / wall / 1e9, 1), "unit": "GB/s",
                          "frac": round(alg / wall / 1e9 / HBM_PEAK_GBS, 4),
                          "bytes": "16P+labels per background step, 24P+16m+labels per adjoint step, "
                                   "8P per checkpoint write"},
        "passes": _passes_table(passes),
        "misfit": float(misfit),
        "grad_norm": float(torch.linalg.vector_norm(grad)),
        "column_budgets": tuned,
        "host_model_prep_s": round(t_prep, 1),
        "observed_data_s": round(t_obs, 1),
    }
    del b, data, g
    torch.cuda.empty_cache()
    if not args.no_plain:
        # the gradient of one shot, fused vs reference-granularity operators, on the
        # regular operator layout (frequency-replicated slowness and labels: 107 GB
        # and 54 GB at the full depth, so the first `zw` depths are sampled)
        from pyfwix_amd.ops import Born
        sh = shots[0]
        zw = min(nz, 3 * ck)
        dims = make_dims(N, N, nw, 1, nx, nx, 0.0, d, 0.0, d, float(sh.freqs[0]), 1.0, 0.0, d, zw)
        born = Born(dims, nref=nref, taperx=taper, tapery=taper, eps_scat=0.04, taylor=1, ckpt_every=ck)
        s3 = torch.from_numpy(s_start[:zw]).to(dev)
        slow = torch.zeros((zw, nw, nx, nx), dtype=torch.complex64, device=dev)
        slow.real.copy_(s3[:, None].expand(zw, nw, nx, nx))
        del s3
        sref = torch.from_numpy(refs_start[0][:zw]).to(dev)[:, :, None].expand(zw, nref, nw).contiguous()
        labels = torch.from_numpy(refs_start[1][:zw]).to(dev)[:, None].expand(zw, nw, N, N).contiguous()
        sh_rec = sh.rec.copy()
        sh_rec[:, 2] = (zw - 1) * d                    # receivers on the window's deepest plane
        wav1 = torch.from_numpy(np.ascontiguousarray(sh.wavelet)).to(dev)
        data1 = torch.zeros((len(sh.rec), nw), dtype=torch.complex64, device=dev)
        g1 = torch.zeros_like(slow)
        t_one = {}
        for name, method in (("fused", FWIX_PSPI), ("plain", FWIX_PSPI_PLAIN)):
            born.set_method(method)
            born.set_frequencies(sh.freqs)
            born.set_refs(sref, labels)
            born.set_slow(slow)
            born.set_sources(sh.src[:, 0], sh.src[:, 1], sh.src[:, 2], np.zeros(1, np.int32))
            born.set_receivers(sh_rec[:, 0], sh_rec[:, 1], sh_rec[:, 2], np.zeros(len(sh_rec), np.int32))
            born.set_wavelet(wav1)
            born.model(data1)                          # warm
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            born.model(data1)
            born.adjoint(False, g1, data1)
            torch.cuda.synchronize()
            t_one[name] = time.perf_counter() - t0
            print(f"[fwi] one shot, {name}: {t_one[name]:.3f} s", file=sys.stderr, flush=True)
        res["vs_plain_shot"] = {
            "fused_s": round(t_one["fused"], 3), "plain_s": round(t_one["plain"], 3),
            "speedup": round(t_one["plain"] / t_one["fused"], 3), "depths": zw,
            "plain_s_per_shot_est": round(res["grad_s_per_shot"] * t_one["plain"] / t_one["fused"], 3),
            "sample": f"background modeling + Born adjoint of shot 0 alone over the first {zw} depths on the "
                      f"regular operator layout, fused steps vs imaging condition / scattering / PSPI as separate "
                      f"operators; plain_s_per_shot_est scales the measured s/shot by that ratio"}
        del born, slow, sref, labels, wav1, data1, g1
        torch.cuda.empty_cache()
        # PSPI forward + adjoint step, fused vs the plain per-operator kernels, on one shot's planes
        rates = {}
        dims = make_dims(N, N, nw, 1, nx, nx, 0.0, d, 0.0, d, 1.0, 0.0, 0.0, d, 2)
        sref = torch.from_numpy(np.ascontiguousarray(np.repeat(refs_start[0][:2, :, None], nw, 2))).to(dev)
        lab_t = torch.from_numpy(np.ascontiguousarray(np.repeat(refs_start[1][:2, None], nw, 1))).to(dev)
        u = torch.randn((1, nw, N, N), dtype=torch.complex64, device=dev)
        plain_table = None
        for name, method in (("fused", FWIX_PSPI), ("plain", FWIX_PSPI_PLAIN)):
            op = OneStep(method, dims, nref=nref, taperx=taper, tapery=taper)
            op.set_frequencies(shots[0].freqs)
            op.set_refs(sref, lab_t)
            op.set_depth(1)
            op.forward(False, u, u)
            op.adjoint(False, u, u)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(3):
                op.forward(False, u, u)
                op.adjoint(False, u, u)
            torch.cuda.synchronize()
            rates[name] = 6 * nw * N * N / (time.perf_counter() - t0) / 1e9
            if method == FWIX_PSPI_PLAIN:
                def _once():
                    op.forward(False, u, u)
                    op.adjoint(False, u, u)
                    torch.cuda.synchronize()
                plain_table = plain_pass_table(copy_gbs, _once)
            del op
        res["vs_plain"] = {"fused_gpts": round(rates["fused"], 3), "plain_gpts": round(rates["plain"], 3),
                           "speedup": round(rates["fused"] / rates["plain"], 3),
                           "sample": f"PSPI forward + adjoint at one depth of one shot ({N}x{N}x{nw}), 3 repetitions",
                           "plain_passes": plain_table}
        del u, sref, lab_t
        torch.cuda.empty_cache()
    res["cpu_baseline"] = None if args.no_cpu else cpu_baseline_fwi(
        N, nw, nx, taper, s_start, refs_start, d, shots[0].freqs)
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c5")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-path comparison")
    ap.add_argument("--copy-probe", action="store_true", help="run the counter calibration copy first")
    ap.add_argument("--pass-sample", type=int, default=8,
                    help="time every n-th launch of each pass with HIP events inside the timed region")
    ap.add_argument("--chunk", type=int, default=0,
                    help="planes per launch of the fused passes (0: the library's choice)")
    ap.add_argument("--no-pass-timing", action="store_true",
                    help="no event markers at all (counter-collection runs)")
    ap.add_argument("--graph", action="store_true",
                    help="also time the step captured as one HIP graph (reported as `graph`)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default column-kernel register budgets (no tuning before the timed region)")
    ap.add_argument("--shots", type=int, default=0, help="shot-batch configs: first n shots only (0: all)")
    ap.add_argument("--budget-gb", type=float, default=160.0, help="shot-batch configs: HBM per launch")
    ap.add_argument("--no-per-shot", action="store_true", help="shot-batch configs: skip the one-shot-per-launch leg")
    ap.add_argument("--shots-per-step", type=int, default=1, help="c5: shots per gradient step (one launch)")
    ap.add_argument("--summarize-stats", nargs=2, metavar=("CSV", "OUT"))
    ap.add_argument("--summarize-traffic", nargs=3, metavar=("FETCH_CSV", "WRITE_CSV", "OUT"))
    ap.add_argument("--summarize-valu", nargs=3, metavar=("SQ_CSV", "STATS_CSV", "OUT"))
    ap.add_argument("--command", action="append", default=[], help="profiling command recorded in a summary")
    ap.add_argument("--title", default=None, help="workload named in a summary's title")
    args = ap.parse_args()
    if os.environ.get("ROCPROF_COUNTER_COLLECTION", "").strip().lower() in ("1", "true", "yes", "on"):
        # Under rocprofv3 --pmc the profiler intercepts the dispatch queue and
        # serializes kernels with its own packets; the barrier packets that
        # hipEventRecord puts in the queue for the per-pass timers are not
        # handled there ("A barrier packet in the queue!", then
        # HSA_STATUS_ERROR_INVALID_PACKET_FORMAT), so counter runs record no events.
        args.no_pass_timing = True
    if args.summarize_stats:
        kw = {"title": args.title} if args.title else {}
        summarize_stats(*args.summarize_stats, command=" ".join(args.command), **kw)
        return
    if args.summarize_traffic:
        cmds = args.command + ["", ""]
        kw = {"title": args.title} if args.title else {}
        summarize_traffic(*args.summarize_traffic, command_fetch=cmds[0], command_write=cmds[1], **kw)
        return
    if args.summarize_valu:
        cmds = args.command + ["", ""]
        summarize_valu(*args.summarize_valu, command=cmds[0], command_stats=cmds[1],
                       title=args.title or args.config)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); shots and "
              "frequency bands would shard across GPUs with one image all-reduce, which is not built here",
              file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        # a plain copy of known size with ordinary loads and stores, like the
        # pass kernels': counter runs scale FETCH_SIZE / WRITE_SIZE by it
        import torch
        from pyfwix_amd import _lib
        x = torch.ones(COPY_PROBE_BYTES // 4, dtype=torch.float32, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(2):
            _lib.call("fwix_copy", x.data_ptr(), y.data_ptr(), COPY_PROBE_BYTES, torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        del x, y
        torch.cuda.empty_cache()
    from pyfwix_amd.workloads import FWI_CONFIGS, NSPS_CONFIGS, SHOT_BATCH_CONFIGS
    if args.config in NSPS_CONFIGS:
        bench_nsps(args)
        return
    if args.config in SHOT_BATCH_CONFIGS:
        bench_shot_batch(args)
        return
    if args.config in FWI_CONFIGS:
        bench_fwi(args)
        return

    import torch
    from pyfwix_amd import _lib
    from pyfwix_amd.ops import Born
    from pyfwix_amd.workloads import CONFIGS, born_workload, device_inputs, workload_dims

    dev = torch.device("cuda:0")
    cfg = CONFIGS[args.config]
    w = born_workload(args.config, **cfg)
    copy_gbs = hbm_copy_gbs(dev)
    slow, ds, sref, labels, wav = device_inputs(w, dev)
    zt = w.nz // 2          # tune on the references of the middle depth
    tuned = None if args.no_autotune else autotune_sizes(
        [w.NY], refs=(sref[zt:zt + 1].contiguous(), labels[zt:zt + 1].contiguous()),
        freqs=w.ow + w.dw * np.arange(w.NW))
    dims = workload_dims(w)
    born = Born(dims, nref=w.nref, taperx=w.taper, tapery=w.taper, taylor=1, eps_scat=0.04, ckpt_every=1,
                planes_per_chunk=args.chunk)
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
    timed = not args.no_pass_timing
    if timed:
        _lib.timing_reset()
        _lib.timing_select(None, max(1, args.pass_sample))
        _lib.timing_enable(1 << 15)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    passes = {}
    ev_us = None
    if timed:
        _lib.timing_disable()
        passes = _lib.timing_collect()
        _lib.timing_select(None, 1)
        ev_us = _lib.timing_overhead(torch.cuda.current_stream().cuda_stream, 200)
    steps_exec = born.steps - s0
    per = steps_exec // args.steps
    value = steps_exec * w.points / wall / 1e9
    ms_step = wall / args.steps * 1e3
    graph = None
    if args.graph:
        from pyfwix_amd.graph import StepGraph
        g = StepGraph(step)
        g.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            g.replay()
        torch.cuda.synchronize()
        gw = time.perf_counter() - t0
        graph = {"value": round(per * args.steps * w.points / gw / 1e9, 4), "ms_per_step": round(gw / args.steps * 1e3, 3),
                 "form": "background + forward + adjoint captured once (torch.cuda.CUDAGraph on the library's launches), "
                         "replayed per step"}
        del g

    roof = None
    if passes:
        # estimated total time of each pass: mean recorded launch x all launches
        est = {k: v["ms"] / v["launches"] * v["total_launches"] for k, v in passes.items()}
        dom_name = max(est, key=est.get)
        dom = passes[dom_name]
        avg_us = dom["ms"] * 1e3 / dom["launches"]
        per_launch = dom["bytes"] / dom["launches"]
        achieved = per_launch / (avg_us * 1e-6) / 1e9
        vi = valu_for(args.config, pass_kernel(dom_name, w))
        roof = {"bound": "hbm", "kernel": dom_name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "peak_measured": round(copy_gbs, 1), "frac_measured": round(achieved / copy_gbs, 4),
                "traffic": traffic_for(args.config, pass_kernel(dom_name, w)),
                "valu_issue": None if vi is None else {
                    "instr_per_launch": vi, "achieved": round(vi / avg_us / 1e3, 1), "peak": VALU_PEAK_GINSTR,
                    "unit": "G wave-instr/s", "frac": round(vi / avg_us / 1e3 / VALU_PEAK_GINSTR, 4)},
                "alg_bytes_per_launch": round(per_launch), "avg_us": round(avg_us, 2),
                "launches_timed": dom["launches"], "launches": dom["total_launches"],
                "empty_event_pair_us": round(ev_us, 2)}
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
                   "receiver_depth_m": float(w.rec[2][0]), "depth_steps_per_step": per,
                   "planes_per_chunk": args.chunk or "auto", "checkpointing": "all depths in HBM"},
        "roofline": roof,
        "graph": graph,
        "column_budgets": tuned,
        "step_roofline": {"achieved": round(alg_total / wall / 1e9, 1), "unit": "GB/s",
                          "frac": round(alg_total / wall / 1e9 / HBM_PEAK_GBS, 4),
                          "bytes": "16P+labels per background step (+8P checkpoint), 24P+8m+labels per "
                                   "Born forward step, 24P+16m+labels per adjoint step"},
        "grad_s_per_shot": round(wall / args.steps / w.NS, 4),
        "passes": {k: {"ms_est": round(v["ms"] / v["launches"] * v["total_launches"], 3),
                       "launches": v["total_launches"], "timed": v["launches"],
                       "avg_us": round(v["ms"] * 1e3 / v["launches"], 2),
                       "GBs": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1)} for k, v in
                   sorted(passes.items(), key=lambda kv: -kv[1]["ms"] / kv[1]["launches"] * kv[1]["total_launches"])},
    }
    if not args.no_plain:
        # the same Born step through separate operators, one shot per launch
        from pyfwix_amd.ops import FWIX_PSPI, FWIX_PSPI_PLAIN
        born.set_method(FWIX_PSPI_PLAIN)
        step()
        torch.cuda.synchronize()
        s1 = born.steps
        t0 = time.perf_counter()
        step()
        torch.cuda.synchronize()
        pw = time.perf_counter() - t0
        res["vs_plain_born"] = {"plain_gpts": round((born.steps - s1) * w.points / pw / 1e9, 3),
                                "plain_ms_per_step": round(pw * 1e3, 1), "speedup": round(pw / (wall / args.steps), 3),
                                "form": "imaging condition, scattering operator and PSPI step as separate operators "
                                        "(taper, FFT, per-reference phase shift / IFFT / select), one shot per launch"}
        born.set_method(FWIX_PSPI)
    del born, grad, data
    if not args.no_plain:
        u = torch.randn((w.NS, w.NW, w.NY, w.NX), dtype=torch.complex64, device=dev)
        res["vs_plain"] = plain_vs_fused(w, dims, sref, labels, u, copy_gbs=copy_gbs)
        del u
    del slow, ds
    torch.cuda.empty_cache()
    res["cpu_baseline"] = None if args.no_cpu else cpu_baseline(w)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
