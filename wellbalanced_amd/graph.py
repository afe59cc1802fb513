"""HIP-graph capture of whole steps.

A model's ``advance(k)`` (CFL reduction, device dt, both Heun stages, halo fills)
enqueues only kernels whose arguments are fixed device pointers -- dt and the time
live on the device and counters are cleared by kernels, not memsets -- so it can be
captured once into a HIP graph (``torch.cuda.CUDAGraph`` over hipGraph) and replayed:
one graph launch per k steps instead of ~10 kernel launches per step.
"""

from __future__ import annotations

import torch


class StepGraph:
    """``steps`` captured steps of ``model``; `replay` runs them again."""

    def __init__(self, model, steps: int = 1) -> None:
        self.model = model
        self.steps = int(steps)
        side = torch.cuda.Stream(device=model.device)
        side.wait_stream(torch.cuda.current_stream(model.device))
        with torch.cuda.stream(side):
            # a first eager run on the capture stream loads every kernel
            saved = [b.clone() for b in model._buf], model._t.clone(), model._dt.clone()
            model.advance(1)
            for b, s in zip(model._buf, saved[0]):
                b.copy_(s)
            model._t.copy_(saved[1])
            model._dt.copy_(saved[2])
        torch.cuda.current_stream(model.device).wait_stream(side)
        torch.cuda.synchronize(model.device)
        model.steps -= 1
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            model.advance(self.steps)
        model.steps -= self.steps  # capture does not run the steps

    def replay(self, times: int = 1) -> None:
        for _ in range(times):
            self.graph.replay()
            self.model.steps += self.steps
