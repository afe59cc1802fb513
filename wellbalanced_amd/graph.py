"""HIP-graph capture of whole steps (``wb_graph_*`` of the C ABI).

A model's ``advance(k)`` (CFL reduction, device dt, both Heun stages, halo fills)
enqueues only this library's kernels, whose arguments are fixed device pointers --
dt and the time live on the device and counters are cleared by kernels, not
memsets -- so it is captured once into a HIP graph on a side stream and replayed:
one graph launch per k steps instead of ~10 kernel launches per step.

The graph holds the device addresses of the model's buffers and tables at capture.
A model bumps its ``generation`` whenever it replaces them (a regrid, a new tiling,
the autotuner's choice); `StepGraph.replay` then refuses to launch.

`LockstepGraph` captures several models stepping together at the first one's
stable step (a precision comparison over many steps without a time offset).
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib


def _snapshot(model):
    return [b.clone() for b in model._buf], model._t.clone(), model._dt.clone()


def _restore(model, saved) -> None:
    for b, s in zip(model._buf, saved[0]):
        b.copy_(s)
    model._t.copy_(saved[1])
    model._dt.copy_(saved[2])


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
            saved = _snapshot(model)
            model.advance(1)
            _restore(model, saved)
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


class LockstepGraph:
    """``steps`` captured steps of ``models`` (one device) in lockstep: per step the
    first model's CFL reduction sets its device dt and advances its time, then every
    model runs its Heun stages with that dt.  The followers' own dt is untouched (the
    captured stages read the first model's); `replay` sets their time to the first
    model's after the launch."""

    def __init__(self, models, steps: int = 1) -> None:
        self.models = list(models)
        if len(self.models) < 2:
            raise ValueError("a lockstep graph needs two models or more; use StepGraph")
        lead = self.models[0]
        dev = lead.device
        if any(m.device != dev for m in self.models):
            raise ValueError("the models of a lockstep graph share one device")
        for m in self.models:
            if hasattr(m, "_check_ready"):
                m._check_ready()
        self.steps = int(steps)
        self._graph = None
        side = torch.cuda.Stream(device=dev)
        side.wait_stream(torch.cuda.current_stream(dev))
        counts = [m.steps for m in self.models]
        with torch.cuda.stream(side):
            saved = [_snapshot(m) for m in self.models]
            self._enqueue(1)
            for m, s in zip(self.models, saved):
                _restore(m, s)
        torch.cuda.synchronize(dev)
        handle = ctypes.c_void_p()
        with torch.cuda.stream(side):
            _lib.call("wb_graph_capture_begin", side.cuda_stream)
            try:
                self._enqueue(self.steps)
            finally:
                _lib.call("wb_graph_capture_end", side.cuda_stream, ctypes.addressof(handle))
        self._graph = handle.value
        for m, c in zip(self.models, counts):
            m.steps = c
        self._side = side
        self._generation = [getattr(m, "generation", 0) for m in self.models]

    def _enqueue(self, steps: int) -> None:
        lead, followers = self.models[0], self.models[1:]
        stream = _lib.stream_handle(lead.device)
        own = [m._dt for m in followers]
        try:
            for m in followers:
                m._dt = lead._dt
            for _ in range(steps):
                lead.advance(1)
                for m in followers:
                    m._heun(stream)
        finally:
            for m, dt in zip(followers, own):
                m._dt = dt

    def replay(self, times: int = 1) -> None:
        """Launch the captured steps ``times`` times on the current stream; raises,
        before any launch, if a model replaced its buffers or tables since."""
        if [getattr(m, "generation", 0) for m in self.models] != self._generation:
            raise RuntimeError("a model's buffers or tables changed since the capture "
                               "(regrid, tiling or autotuning): capture a new LockstepGraph")
        lead = self.models[0]
        stream = _lib.stream_handle(lead.device)
        _lib.call("wb_graph_launch", self._graph, int(times), stream)
        for m in self.models:
            m.steps += self.steps * int(times)
        for m in self.models[1:]:
            m._t.copy_(lead._t)

    def __del__(self) -> None:
        if getattr(self, "_graph", None):
            try:
                _lib.call("wb_graph_destroy", self._graph)
            except Exception:
                pass
            self._graph = None
