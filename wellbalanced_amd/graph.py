"""HIP-graph capture of whole steps (``wb_graph_*`` of the C ABI).

A model's ``advance(k)`` (CFL reduction, device dt, both Heun stages, halo fills)
enqueues only this library's kernels, whose arguments are fixed device pointers --
dt and the time live on the device and counters are cleared by kernels, not
memsets -- so it is captured once into a HIP graph on a side stream and replayed:
one graph launch per k steps instead of ~10 kernel launches per step.

The graph holds the device addresses of the model's buffers and tables at capture.
A model bumps its ``generation`` whenever it replaces them (a regrid, a new tiling,
the autotuner's choice); `StepGraph.replay` then refuses to launch.
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib


class StepGraph:
    """``steps`` captured steps of ``model``; `replay` runs them again."""

    def __init__(self, model, steps: int = 1) -> None:
        self.model = model
        self.steps = int(steps)
        self._graph = None
        dev = model.device
        side = torch.cuda.Stream(device=dev)
        side.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(side):
            # a first eager run on the capture stream loads every kernel
            saved = [b.clone() for b in model._buf], model._t.clone(), model._dt.clone()
            model.advance(1)
            for b, s in zip(model._buf, saved[0]):
                b.copy_(s)
            model._t.copy_(saved[1])
            model._dt.copy_(saved[2])
        torch.cuda.synchronize(dev)
        model.steps -= 1
        handle = ctypes.c_void_p()
        with torch.cuda.stream(side):
            _lib.call("wb_graph_capture_begin", side.cuda_stream)
            try:
                model.advance(self.steps)
            finally:
                _lib.call("wb_graph_capture_end", side.cuda_stream, ctypes.addressof(handle))
        self._graph = handle.value
        model.steps -= self.steps  # capture does not run the steps
        self._side = side
        self._generation = getattr(model, "generation", 0)

    def replay(self, times: int = 1) -> None:
        """Launch the captured steps ``times`` times on the current stream; raises,
        before any launch, if the model replaced its buffers or tables since."""
        if getattr(self.model, "generation", 0) != self._generation:
            raise RuntimeError("the model's buffers or tables changed since the capture "
                               "(regrid, tiling or autotuning): capture a new StepGraph")
        stream = _lib.stream_handle(self.model.device)
        _lib.call("wb_graph_launch", self._graph, int(times), stream)
        self.model.steps += self.steps * int(times)

    def __del__(self) -> None:
        if getattr(self, "_graph", None):
            try:
                _lib.call("wb_graph_destroy", self._graph)
            except Exception:
                pass
            self._graph = None
