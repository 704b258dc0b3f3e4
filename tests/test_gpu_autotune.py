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
            assert r["chosen"][key] in autotune.CANDIDATES
            # the fastest, or the budget in use when nothing beats it by the margin
            assert r[key][str(r["chosen"][key])] == min(r[key].values()) or r["chosen"][key] == before[
                (0 if "fanout" in key else 1, 2 if "scatter" in key else (0 if key == "fanout" else 1))]
        assert autotune.get(0, 0, n) == r["chosen"]["fanout"] and autotune.get(1, 2, n) == r["chosen"]["scatter_fanin"]
    finally:
        for (k, m), wpe in before.items():
            autotune.set_budget(k, m, n, wpe)
