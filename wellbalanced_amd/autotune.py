"""Autotuner of the fused stage's launch shape.

The fused stages have free parameters that do not change the results (the FP64
path gives the same bits for any of them): the cells per workgroup along the
contiguous axis (``tile``: 30, 62, 126 or 254 in 2D -> 64, 64, 128 or 256 threads, the
small tiles and short segments giving a small grid enough workgroups to fill the chip;
30, 62 or 126 on a hierarchy, whose small patches fill a 30-row tile better) and the
cells each workgroup marches along x (``nxs``), for the refined 2D hierarchy the stage
kernel (``kernel``: a workgroup per 16 x 16 tile, or the persistent marching stage and
its number of workgroups), for the mixed 3D stages the waves per SIMD the register
allocation targets (``waves``), and for the 3D block hierarchy whether blocks with
every neighbour read them in place instead of filled ghosts (``direct``).  `tune`
times each candidate on the model's own
state (state restored afterwards) for `steps` steps, or for more steps until the run
lasts MIN_SECONDS (on small grids a step is ~0.1 ms, below the host clock's noise),
and keeps the fastest per step; the choice is cached per (model kind, shape,
precision) for the process.
"""

from __future__ import annotations

import itertools
import time

import torch

_CACHE: dict[tuple, dict] = {}
MIN_SECONDS = 0.005

DEFAULTS_2D = {"tile": [30, 62, 126, 254], "nxs": [4, 8, 16, 32, 64]}
DEFAULTS_3D = {"nxs": [8, 16, 32, 64]}
DEFAULTS_3D_MIXED = {"nxs": [8, 16, 32, 64], "waves": [2, 3]}
DEFAULTS_AMR = {"tile": [30, 62, 126], "nxs": [4, 8, 16, 32], "workgroups": [512, 1024, 2048]}


def _candidates(defaults: dict) -> list[dict]:
    """The tiled stage, then every marching launch shape."""
    names = list(defaults)
    return [{"kernel": "tiles"}] + [dict(zip(names, v), kernel="march")
                                    for v in itertools.product(*(defaults[n] for n in names))]
DEFAULTS_AMR3D_MIXED = {"waves": [2, 3], "direct": [True, False]}
DEFAULTS_AMR3D = {"direct": [True, False]}


def _key(model) -> tuple:
    shape = getattr(model, "n", None) or (getattr(model, "nx", None), getattr(model, "ny", None))
    if hasattr(model, "layouts"):
        shape = tuple((l.nx, l.ny) for l in model.layouts)
    if hasattr(model, "nblocks"):
        shape = (model.nblocks,)
    return (type(model).__name__, tuple(shape) if not isinstance(shape, tuple) else shape,
            model.precision)


def _apply(model, choice: dict) -> None:
    if hasattr(model, "set_tiling"):
        model.set_tiling(**choice)
    else:
        for k, v in choice.items():
            setattr(model, k, v)
    # a captured graph holds the old launch shape (wellbalanced_amd.graph)
    model.generation = getattr(model, "generation", 0) + 1


def tune(model, space: dict | list | None = None, steps: int = 3) -> dict:
    """Pick the fastest launch shape for ``model``; returns and applies it.  ``space``:
    a dict of parameter values (every combination is tried) or a list of choices."""
    key = _key(model)
    if key in _CACHE:
        _apply(model, _CACHE[key])
        return _CACHE[key]
    if space is None:
        if hasattr(model, "nblocks"):
            space = DEFAULTS_AMR3D_MIXED if model.precision == "mixed" else DEFAULTS_AMR3D
        elif hasattr(model, "layouts"):
            space = _candidates(DEFAULTS_AMR)
        elif hasattr(model, "n"):
            space = DEFAULTS_3D if model.precision == "fp64" else DEFAULTS_3D_MIXED
        elif getattr(model, "kernels", "fused") == "fused":
            space = _candidates(DEFAULTS_2D)
        else:
            space = DEFAULTS_2D
    if not space:
        _CACHE[key] = {}
        return {}
    saved = [b.clone() for b in model._buf], model._t.clone(), model.steps
    best, best_t = None, float("inf")
    if isinstance(space, dict):  # the product of every parameter's values
        names = list(space)
        space = [dict(zip(names, v)) for v in itertools.product(*(space[n] for n in names))]
    for choice in space:
        _apply(model, choice)
        model.advance(1)
        torch.cuda.synchronize(model.device)
        reps = steps
        while True:
            t0 = time.perf_counter()
            model.advance(reps)
            torch.cuda.synchronize(model.device)
            elapsed = time.perf_counter() - t0
            if elapsed >= MIN_SECONDS or reps >= 64 * steps:
                break
            reps *= 4
        elapsed /= reps
        if elapsed < best_t:
            best, best_t = choice, elapsed
        for b, s in zip(model._buf, saved[0]):
            b.copy_(s)
        model._t.copy_(saved[1])
        model.steps = saved[2]
    _apply(model, best)
    _CACHE[key] = best
    return best
