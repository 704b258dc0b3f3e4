"""Column-kernel register budgets on the GPU: every budget gives the same
results (only register allocation differs), and the autotuner picks one of
the candidates from timed launches and leaves it in the table."""
import numpy as np
import pytest

from fwix_testutil import crand, rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [48, 384])
def test_budgets_give_identical_steps(dev, n):
    import torch
    from pyfwix_amd import autotune
    from pyfwix_amd.ops import FWIX_NSPS, FWIX_PSPI, OneStep, Scatter, make_dims
    rng = np.random.default_rng(n)
    nw = 6
    dims = make_dims(n, n, nw, 1, n // 2, n // 2, 0.0, 10.0, 0.0, 10.0, 4.0, 1.5, 0.0, 10.0, 1)
    sref = torch.from_numpy((1.0 / (1500.0 + 1500.0 * rng.random((1, 3, 1))) ** 2).astype(np.complex64))
    sref = sref.to(dev).expand(1, 3, nw).contiguous()
    lab = torch.from_numpy(rng.integers(0, 3, (1, nw, n, n)).astype(np.uint8)).to(dev)
    slow = torch.from_numpy((1.0 / (1500.0 + 1500.0 * rng.random((1, nw, n // 2, n // 2))) ** 2)
                            .astype(np.complex64)).to(dev)
    u = torch.from_numpy(crand(rng, (1, nw, n, n)).astype(np.complex64)).to(dev)
    before = {(k, m): autotune.get(k, m, n) for k in (0, 1) for m in range(3)}
    outs = {}
    try:
        for wpe in autotune.CANDIDATES:
            for k in (0, 1):
                for m in range(3):
                    autotune.set_budget(k, m, n, wpe)
            res = []
            for method in (FWIX_PSPI, FWIX_NSPS):
                op = OneStep(method, dims, nref=3, taperx=4, tapery=4)
                op.set_refs(sref, lab)
                op.set_depth(0)
                f = torch.zeros_like(u)
                a = torch.zeros_like(u)
                op.forward(False, u, f)
                op.adjoint(False, a, u)
                res += [f, a]
            sc = Scatter(dims, taylor=2)
            sc.set_slow(slow)
            f = torch.zeros_like(u)
            a = torch.zeros_like(u)
            sc.forward(False, u, f)
            sc.adjoint(False, a, u)
            res += [f, a]
            torch.cuda.synchronize()
            outs[wpe] = [r.cpu().numpy() for r in res]
    finally:
        for (k, m), wpe in before.items():
            autotune.set_budget(k, m, n, wpe)
    for wpe in autotune.CANDIDATES[1:]:
        for x, y in zip(outs[wpe], outs[autotune.CANDIDATES[0]]):
            assert rel(x, y) < 1e-6


def test_tune_size_sets_a_candidate(dev):
    from pyfwix_amd import autotune
    n = 64
    before = {(k, m): autotune.get(k, m, n) for k in (0, 1) for m in range(3)}
    try:
        r = autotune.tune_size(n, planes=16, reps=2)
        for key in ("fanout", "fanin", "scatter_fanout", "scatter_fanin"):
            assert set(r[key]) == {str(w) for w in autotune.CANDIDATES}
            assert all(t is not None and t > 0 for t in r[key].values())
            assert r["chosen"][key] in {str(w) for w in autotune.CANDIDATES}
            # the fastest, or the budget in use when nothing beats it by the margin
            assert r[key][r["chosen"][key]] == min(r[key].values()) or r["chosen"][key] == str(before[
                (0 if "fanout" in key else 1, 2 if "scatter" in key else (0 if key == "fanout" else 1))])
        assert autotune.get(0, 0, n) == int(r["chosen"]["fanout"])
        assert autotune.get(1, 2, n) == int(r["chosen"]["scatter_fanin"])
    finally:
        for (k, m), wpe in before.items():
            autotune.set_budget(k, m, n, wpe)


@pytest.mark.parametrize("nref", [4, 40])
def test_1024_column_layouts_match_oracle(dev, nref):
    """1024-point columns: one wave per column and the two-wave radix-2 split
    give the oracle's steps and scattering terms, and each other's to float
    rounding (references absent from rows and planes, more references than
    the split kernels stage in LDS, every budget the split is built for)."""
    import torch
    import oracle as O
    from fwix_testutil import TOL_WFLD, dims_of, grid, smooth_slowness
    from pyfwix_amd import autotune
    from pyfwix_amd.ops import FWIX_NSPS, FWIX_PSPI, OneStep, Scatter
    g = grid(48, 1024, 3, 1, nx=40, ny=1000, nz=2)
    rng = np.random.default_rng(1024)
    sref = (1.0 / rng.uniform(1500, 3000, (2, nref, g.NW)) ** 2).astype(np.complex64)
    labels = np.zeros((2, g.NW, g.NY, g.NX), np.uint8)
    labels[1, 0] = rng.integers(0, nref, (g.NY, g.NX))
    labels[1, 1, :400] = nref - 1                           # rows with the last reference alone
    labels[1, 2] = rng.choice([0, nref // 2], (g.NY, g.NX))  # most references absent
    slow = smooth_slowness(rng, g, imag=0.01)
    shape = (g.NS, g.NW, g.NY, g.NX)
    x, y = crand(rng, shape), crand(rng, shape)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a.astype(np.complex64))).to(dev)  # noqa: E731
    s64, lab = sref[1].astype(np.complex128), labels[1].astype(np.int64)
    refs = {}
    for name, method in (("pspi", FWIX_PSPI), ("nsps", FWIX_NSPS)):
        tap = 4 if name == "pspi" else 0
        fwd = O.pspi_forward if name == "pspi" else O.nsps_forward
        adj = O.pspi_adjoint if name == "pspi" else O.nsps_adjoint
        kw = dict(tx=tap, ty=tap) if name == "pspi" else {}
        refs[name] = (fwd(x, s64, lab, g, g.dz, 0.0, **kw), adj(y, s64, lab, g, g.dz, 0.0, **kw))
    sl2 = slow[1].astype(np.complex64).astype(np.complex128)
    refs["scatter"] = (O.scatter_forward(x, sl2, g, g.dz, 0.04, 2), O.scatter_adjoint(y, sl2, g, g.dz, 0.04, 2))
    before = {(k, m): (autotune.get_layout(k, m, 1024), autotune.get(k, m, 1024)) for k in (0, 1) for m in range(3)}
    outs = {}
    try:
        for lw in [(0, 2), (1, 2), (1, 4)]:
            for k in (0, 1):
                for m in range(3):
                    autotune.set_layout(k, m, 1024, lw[0])
                    autotune.set_budget(k, m, 1024, lw[1])
            res = {}
            for name, method in (("pspi", FWIX_PSPI), ("nsps", FWIX_NSPS)):
                op = OneStep(method, dims_of(g), nref=nref, taperx=4 if name == "pspi" else 0,
                             tapery=4 if name == "pspi" else 0)
                op.set_refs(t(sref), torch.from_numpy(labels).to(dev))
                op.set_depth(1)
                f, a = torch.zeros(shape, dtype=torch.complex64, device=dev), torch.zeros(shape, dtype=torch.complex64,
                                                                                        device=dev)
                op.forward(False, t(x), f)
                op.adjoint(False, a, t(y))
                res[name] = (f.cpu().numpy(), a.cpu().numpy())
            sc = Scatter(dims_of(g), taylor=2, eps_scat=0.04)
            sc.set_slow(t(slow))
            sc.set_depth(1)
            f, a = torch.zeros(shape, dtype=torch.complex64, device=dev), torch.zeros(shape, dtype=torch.complex64,
                                                                                    device=dev)
            sc.forward(False, t(x), f)
            sc.adjoint(False, a, t(y))
            res["scatter"] = (f.cpu().numpy(), a.cpu().numpy())
            outs[lw] = res
    finally:
        for (k, m), lw in before.items():
            autotune.set_layout(k, m, 1024, lw[0])
            autotune.set_budget(k, m, 1024, lw[1])
    for lw, res in outs.items():
        for name, (f, a) in res.items():
            assert rel(f, refs[name][0]) < TOL_WFLD, (lw, name, "forward")
            assert rel(a, refs[name][1]) < TOL_WFLD, (lw, name, "adjoint")
            assert rel(f, outs[(0, 2)][name][0]) < 2e-6, (lw, name, "forward vs one wave")
            assert rel(a, outs[(0, 2)][name][1]) < 2e-6, (lw, name, "adjoint vs one wave")


@pytest.mark.parametrize("NX,NY", [(64, 48), (48, 1024)])
def test_scratch_layouts_give_identical_steps(dev, NX, NY):
    """The fused passes' scratch in quad blocks or row-major: the same PSPI /
    NSPS steps and scattering terms bit for bit (only memory placement
    differs)."""
    import torch
    from fwix_testutil import dims_of, grid, smooth_slowness
    from pyfwix_amd import _lib
    from pyfwix_amd.ops import FWIX_NSPS, FWIX_PSPI, OneStep, Scatter
    g = grid(NX, NY, 3, 1, nx=NX - 8, ny=NY - 8, nz=2)
    rng = np.random.default_rng(NX + NY)
    nref = 3
    sref = (1.0 / rng.uniform(1500, 3000, (2, nref, g.NW)) ** 2).astype(np.complex64)
    labels = rng.integers(0, nref, (2, g.NW, g.NY, g.NX)).astype(np.uint8)
    slow = smooth_slowness(rng, g, imag=0.01)
    shape = (g.NS, g.NW, g.NY, g.NX)
    x = torch.from_numpy(crand(rng, shape).astype(np.complex64)).to(dev)
    before = _lib.lib.fwix_tune_scratch_layout_get(NY)
    outs = {}
    try:
        for quad in (0, 1):
            _lib.call("fwix_tune_scratch_layout_set", NY, quad)
            res = []
            for method in (FWIX_PSPI, FWIX_NSPS):
                op = OneStep(method, dims_of(g), nref=nref, taperx=4, tapery=4)
                op.set_refs(torch.from_numpy(sref).to(dev), torch.from_numpy(labels).to(dev))
                op.set_depth(1)
                f, a = torch.zeros_like(x), torch.zeros_like(x)
                op.forward(False, x, f)
                op.adjoint(False, a, x)
                res += [f, a]
            sc = Scatter(dims_of(g), taylor=2, eps_scat=0.04)
            sc.set_slow(torch.from_numpy(slow.astype(np.complex64)).to(dev))
            sc.set_depth(1)
            f, a = torch.zeros_like(x), torch.zeros_like(x)
            sc.forward(False, x, f)
            sc.adjoint(False, a, x)
            res += [f, a]
            torch.cuda.synchronize()
            outs[quad] = [r.cpu().numpy() for r in res]
    finally:
        _lib.call("fwix_tune_scratch_layout_set", NY, before)
    for p, q in zip(outs[0], outs[1]):
        assert np.abs(p).max() > 0
        assert np.array_equal(p, q)
