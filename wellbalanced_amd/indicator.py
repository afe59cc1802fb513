"""Refinement indicators of the 3D step: the Löhner estimator and the learned CNN.

The reference tags cells with |J| times the cell size (``mhd3d.py:771-801``).  This
module adds the second-derivative estimator of Löhner (1987) in FLASH's normalised
form (``wb_lohner3d``): max over density and pressure, flag where it exceeds 0.8.
It is the label source of the learned indicator (`wellbalanced_amd.cnn`) and its
fallback.

The learned indicator is trained to look ahead: from the state at step s it
predicts the Löhner flags of step s + lead, so that a hierarchy regridded every
``lead`` steps still covers the features that move in between.
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


def refine(model, net=None, cutoff: float = CUTOFF) -> torch.Tensor:
    """Cell flags of ``model``: the learned indicator's (logit > 0) when ``net`` is
    given, else the Löhner estimator's."""
    if net is None:
        return flags(lohner(model), cutoff)
    return net.predict(model) > 0.0


def lookahead_samples(model, net, snapshots: int, every: int, lead: int, cutoff: float = CUTOFF):
    """Advance ``model`` and collect ``snapshots`` triples (step, features at the step,
    Löhner estimate at the step, Löhner estimate ``lead`` steps later), one every
    ``every`` steps, the first after ``every`` steps."""
    wanted = {every * (i + 1) for i in range(snapshots)}
    need = wanted | {s + lead for s in wanted}
    feats, ests = {}, {}
    for s in range(1, max(need) + 1):
        model.advance(1)
        if s in wanted:
            feats[s] = net.features(model)
        if s in need:
            ests[s] = lohner(model)
    return [(s, feats[s], ests[s], ests[s + lead]) for s in sorted(wanted)]


def recall_at_count(score: torch.Tensor, target: torch.Tensor, count: int) -> float:
    """The fraction of ``target`` cells among the ``count`` highest-scoring cells."""
    idx = torch.topk(score.reshape(-1), count).indices
    return float(target.reshape(-1)[idx].float().sum() / target.float().sum())


def train_lookahead(problem, *, snapshots: int = 10, every: int = 4, lead: int = 8,
                    iterations: int = 150, lr: float = 3e-3, seed: int = 0, cutoff: float = CUTOFF):
    """The learned indicator trained on the GPU on a uniform mixed run of ``problem``:
    the features of ``snapshots`` states, one every ``every`` steps, against the Löhner
    flags ``lead`` steps later.  Returns (network, losses, training seconds)."""
    import time

    from . import cnn
    from .mhd3d import MHD3D

    uni = MHD3D.from_problem(problem, kernels="fused", precision="mixed")
    uni.check_interval = 0
    net = cnn.Indicator(uni.n, lr=lr, seed=seed)
    uni.advance(every)
    net.fit_normalisation(uni)
    samples = lookahead_samples(uni, net, snapshots=snapshots, every=every, lead=lead,
                                cutoff=cutoff)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses = net.fit([(x, (later > cutoff).float().reshape(-1).contiguous())
                      for _, x, _, later in samples], iterations)
    torch.cuda.synchronize()
    return net, losses, time.perf_counter() - t0


def block_hierarchy(n0: int, flags1, flags2, block: int = 16):
    """A 3-level hierarchy of ``block``^3 blocks over an ``n0``^3 base: level 1 over the
    base blocks holding a cell of ``flags1``, level 2 over the level-1 blocks holding a
    base cell of ``flags2`` (each parent's 8 children), the nesting added."""
    import numpy as np

    from .amr3d_maps import Hierarchy3D

    h = Hierarchy3D(n0, block=block)
    h.refine_flagged(0, np.asarray(flags1, bool))
    flags2 = np.asarray(flags2, bool)
    if len(h.levels) > 1 and flags2.any():
        half = block // 2  # a level-1 block covers half a base block per axis
        parents = {tuple(int(v) for v in c // half) for c in np.argwhere(flags2)}
        parents = [p for p in parents if p in h.levels[1]]
        if parents:
            h.refine(1, parents)
    return h
