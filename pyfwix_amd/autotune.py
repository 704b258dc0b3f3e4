"""Autotuner of the fused column kernels.

Every fan-out / fan-in column kernel of the library is compiled for three
register budgets, 2, 4 or 6 waves per SIMD (256, 128 or 80 VGPRs).  More
resident waves hide more of the LDS and HBM latency these transforms wait on,
but a tight budget can spill, and the workgroup's LDS tile caps residency
anyway; which budget wins depends on the column length and the transform.
:func:`tune_size` times each budget on real launches at one padded grid size
(a PSPI forward + adjoint step for the phase-shift transforms, modes 0 / 1,
and a scattering forward + adjoint for the scattering-term transforms, mode
2) and keeps the fastest in the library's process-wide table; results are
bit-identical across budgets (only register allocation changes).
:func:`save` / :func:`load` keep a tuned table as JSON.

    from pyfwix_amd import autotune
    autotune.tune_size(768)           # once per padded size, before the run
"""
from __future__ import annotations

import json

import numpy as np

from . import _lib

__all__ = ["CANDIDATES", "get", "set_budget", "get_layout", "set_layout", "table", "tune_size", "save", "load"]

CANDIDATES = (2, 4, 6)
FANOUT, FANIN = 0, 1


def get(kind: int, mode: int, n: int) -> int:
    """Waves-per-SIMD budget in use for (kind, mode, column length n)."""
    v = _lib.lib.fwix_tune_columns_get(int(kind), int(mode), int(n))
    if v < 0:
        raise ValueError(f"no column kernel for kind={kind} mode={mode} n={n}")
    return v


def set_budget(kind: int, mode: int, n: int, wpe: int) -> None:
    _lib.call("fwix_tune_columns_set", int(kind), int(mode), int(n), int(wpe))


def get_layout(kind: int, mode: int, n: int) -> int:
    """Layout of the column transforms for (kind, mode, column length n): 0 one
    wave per column, 1 two waves per column (n = 1024 only)."""
    v = _lib.lib.fwix_tune_columns_layout_get(int(kind), int(mode), int(n))
    if v < 0:
        raise ValueError(f"no column kernel for kind={kind} mode={mode} n={n}")
    return v


def set_layout(kind: int, mode: int, n: int, layout: int) -> None:
    _lib.call("fwix_tune_columns_layout_set", int(kind), int(mode), int(n), int(layout))


def table(sizes=(96, 128, 192, 256, 384, 512, 768, 1024)) -> dict:
    """{"n": {"fanout": [m0, m1, m2], "fanin": [m0, m1, m2], "fanout_layout": [...],
    "fanin_layout": [...]}} for the given sizes."""
    return {str(n): {"fanout": [get(FANOUT, m, n) for m in range(3)],
                     "fanin": [get(FANIN, m, n) for m in range(3)],
                     "fanout_layout": [get_layout(FANOUT, m, n) for m in range(3)],
                     "fanin_layout": [get_layout(FANIN, m, n) for m in range(3)]} for n in sizes}


def save(path: str, sizes=(96, 128, 192, 256, 384, 512, 768, 1024)) -> None:
    with open(path, "w") as f:
        json.dump(table(sizes), f, indent=1)


def load(path: str) -> None:
    with open(path) as f:
        t = json.load(f)
    for n, v in t.items():
        for m in range(3):
            set_budget(FANOUT, m, int(n), v["fanout"][m])
            set_budget(FANIN, m, int(n), v["fanin"][m])
            set_layout(FANOUT, m, int(n), v.get("fanout_layout", [0, 0, 0])[m])
            set_layout(FANIN, m, int(n), v.get("fanin_layout", [0, 0, 0])[m])


def _time_passes(run, names, reps):
    """Mean microseconds per launch of each named pass over ``reps`` runs (events on every launch)."""
    import torch
    run()
    torch.cuda.synchronize()
    _lib.timing_reset()
    _lib.timing_select(None, 1)
    _lib.timing_enable(1 << 14)
    for _ in range(reps):
        run()
    torch.cuda.synchronize()
    _lib.timing_disable()
    p = _lib.timing_collect()
    _lib.timing_reset()
    return {k: p[k]["ms"] * 1e3 / p[k]["launches"] for k in names if k in p and p[k]["launches"]}


def tune_size(n: int, planes: int | None = None, nref: int = 4, taylor: int = 1, reps: int = 5,
              candidates=CANDIDATES, device=None, seed: int = 0, margin: float = 0.03, refs=None,
              freqs=None) -> dict:
    """Tune the budgets of every column kernel at column length ``n`` (an n x n
    padded grid with ``planes`` planes, default about 16 M points) and set the
    fastest, keeping the budget in use unless another beats it by ``margin``
    (launch-time noise is a few per cent).  ``refs`` = (sref (1, nref, planes),
    labels (1, planes, n, n)) device tensors and ``freqs`` (planes values, Hz)
    tune on a real depth of a workload (which references are present in which
    rows changes what the fan-out / fan-in transforms do); by default the
    labels are random.  At n = 1024 the two-wave layout is timed beside the
    one-wave one.  Returns {"fanout": {config: us}, "fanin": {...},
    "scatter_fanout": {...}, "scatter_fanin": {...}, "chosen": {...}}, config
    "4" for one wave per column at 4 waves per SIMD, "split4" for two."""
    import torch
    from .ops import FWIX_PSPI, OneStep, Scatter, make_dims
    dev = torch.device(device or "cuda:0")
    if refs is not None:
        nref, planes = int(refs[0].shape[1]), int(refs[0].shape[2])
    planes = planes or max(8, min(256, (16 << 20) // (n * n)))
    rng = np.random.default_rng(seed)
    dims = make_dims(n, n, planes, 1, n // 2, n // 2, 0.0, 10.0, 0.0, 10.0, 3.0, 0.5, 0.0, 10.0, 1)
    step = OneStep(FWIX_PSPI, dims, nref=nref, taperx=min(20, n // 4), tapery=min(20, n // 4))
    if freqs is not None:
        step.set_frequencies(freqs)
    if refs is not None:
        sref, lab = refs
    else:
        sref = torch.from_numpy((1.0 / (1500.0 + 1500.0 * rng.random((1, nref, 1))) ** 2).astype(np.complex64))
        sref = sref.to(dev).expand(1, nref, planes).contiguous()
        lab = torch.from_numpy(rng.integers(0, nref, (1, 1, n, n)).astype(np.uint8)).to(dev)
        lab = lab.expand(1, planes, n, n).contiguous()
    step.set_refs(sref, lab)
    step.set_depth(0)
    scat = Scatter(dims, taylor=taylor)  # its frequency axis stays the regular one of `dims`
    s = torch.from_numpy((1.0 / (1500.0 + 1500.0 * rng.random((1, 1, n // 2, n // 2))) ** 2).astype(np.complex64))
    slow = s.to(dev).expand(1, planes, n // 2, n // 2).contiguous()
    scat.set_slow(slow)
    scat.set_depth(0)
    u = torch.randn((1, planes, n, n), dtype=torch.complex64, device=dev)
    v = torch.empty_like(u)
    w = torch.empty_like(u)

    def pspi():
        step.forward(False, u, v)
        step.adjoint(False, w, v)

    def scatter():
        scat.forward(False, u, v)
        scat.adjoint(False, w, v)

    # configurations: register budget, and at n = 1024 also the two-wave layout
    # (keys "4" = one wave per column at 4 waves per SIMD, "split4" = two waves)
    configs = [(0, w) for w in candidates]
    if n == 1024:
        configs += [(1, w) for w in candidates]
    key = lambda lw: (f"split{lw[1]}" if lw[0] else str(lw[1]))  # noqa: E731
    before = {(k, m): (get_layout(k, m, n), get(k, m, n)) for k in (FANOUT, FANIN) for m in range(3)}

    def apply(k, m, lw):
        set_layout(k, m, n, lw[0])
        set_budget(k, m, n, lw[1])

    out = {"fanout": {}, "fanin": {}, "scatter_fanout": {}, "scatter_fanin": {}}
    try:
        for lw in configs:
            for k in (FANOUT, FANIN):
                for m in range(3):
                    apply(k, m, lw)
            t = _time_passes(pspi, ("cols_fanout", "cols_fanin"), reps)
            out["fanout"][key(lw)] = t.get("cols_fanout")
            out["fanin"][key(lw)] = t.get("cols_fanin")
            if taylor >= 1:
                t = _time_passes(scatter, ("scatter_cols", "scatter_cols_adj"), reps)
                out["scatter_fanout"][key(lw)] = t.get("scatter_cols")
                out["scatter_fanin"][key(lw)] = t.get("scatter_cols_adj")
    except Exception:
        for (k, m), lw in before.items():
            apply(k, m, lw)
        raise

    def best(d, k, m):
        """The configuration in use unless another is faster by more than ``margin``."""
        cur = before[(k, m)]
        if not d or any(x is None for x in d.values()) or key(cur) not in d:
            return cur
        w = min(configs, key=lambda lw: d[key(lw)])
        return w if d[key(w)] < (1.0 - margin) * d[key(cur)] else cur

    chosen = {"fanout": best(out["fanout"], FANOUT, 0), "fanin": best(out["fanin"], FANIN, 1),
              "scatter_fanout": best(out["scatter_fanout"], FANOUT, 2),
              "scatter_fanin": best(out["scatter_fanin"], FANIN, 2)}
    for m in (0, 1):
        apply(FANOUT, m, chosen["fanout"])
        apply(FANIN, m, chosen["fanin"])
    apply(FANOUT, 2, chosen["scatter_fanout"])
    apply(FANIN, 2, chosen["scatter_fanin"])
    out = {k: {b: (None if t is None else round(t, 2)) for b, t in v.items()} for k, v in out.items()}
    out["chosen"] = {k: key(v) for k, v in chosen.items()}
    del step, scat, u, v, w, slow
    torch.cuda.empty_cache()
    return out
