"""HIP-graph capture of whole depth recursions.

The reference launches every operator of every depth step from the host
(``propagator/src/OneWay.cpp:22-79``, ``ExtendedBorn.cpp:152-208``); a
depth loop is thousands of launches.  Everything the library launches inside
``Born.model/forward/adjoint``, ``OneWay.forward/adjoint`` and the one-step
operators goes to the caller's current stream with no host synchronisation,
no allocation and no memset node (zero fills are kernels), so the whole
recursion can be captured once and replayed as one graph launch.

    g = StepGraph(lambda: (born.model(), born.forward(False, ds, d), born.adjoint(False, grad, d)))
    g.replay()          # same kernels, same buffers, same results

Inputs are read from, and outputs written to, the buffers that were bound at
capture time: change their contents in place (``copy_``), not the tensors.
Pass timing records nothing inside a graph.  Host-side statistics
(``Born.steps``) advance only during the capture.
"""
from __future__ import annotations

__all__ = ["StepGraph"]


class StepGraph:
    def __init__(self, fn, warmup: int = 1):
        import torch
        self.fn = fn
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            for _ in range(warmup):
                fn()
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            fn()
        torch.cuda.synchronize()

    def replay(self) -> None:
        self.graph.replay()
