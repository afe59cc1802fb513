"""Refinement indicators of the 3D step: the Löhner estimator and the learned CNN.

The reference tags cells with |J| times the cell size (``mhd3d.py:771-801``,
`MHD3D.current_tags` here).  This module adds the second-derivative estimator of
Löhner (1987) in FLASH's normalised form (``wb_lohner3d``): max over density and
pressure, flag where it exceeds 0.8.  It is the label source of the learned
indicator (`wellbalanced_amd.cnn`) and its fallback.
"""

from __future__ import annotations

import torch

from . import _lib

CUTOFF = 0.8
EPS = 0.01


def lohner(model, eps: float = EPS) -> torch.Tensor:
    """The estimate of every cell of an `MHD3D` model, ``(nx, ny, nz)`` FP64 on the device."""
    nx, ny, nz = model.n
    X, Y, Z = model.dims
    stream = _lib.stream_handle(model.device)
    pmask = sum(int(p) << a for a, p in enumerate(model.periodic))
    _lib.call("wb_mhd3d_fill_halos", _lib.ptr(model._buf[0]), nx, ny, nz, pmask, stream)
    scratch = torch.empty(2 * X * Y * Z, dtype=torch.float64, device=model.device)
    out = torch.empty((nx, ny, nz), dtype=torch.float64, device=model.device)
    _lib.call("wb_lohner3d", _lib.ptr(model._buf[0]), nx, ny, nz, *model.spacing, *model.mean_field,
              model.gamma, eps, _lib.ptr(scratch), _lib.ptr(out), stream)
    return out


def flags(estimate: torch.Tensor, cutoff: float = CUTOFF) -> torch.Tensor:
    return estimate > cutoff


def block_flags(cell_flags: torch.Tensor, block: int = 16) -> torch.Tensor:
    """A block of ``block``^3 cells is flagged if any of its cells is."""
    nx, ny, nz = (s // block for s in cell_flags.shape)
    f = cell_flags[: nx * block, : ny * block, : nz * block]
    f = f.reshape(nx, block, ny, block, nz, block)
    return f.any(dim=5).any(dim=3).any(dim=1)
