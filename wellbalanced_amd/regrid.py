"""Regridding a periodic 2D hierarchy on refinement tags: the host side of
``MHD2D.regrid`` (mhd.py:1043-1056), which calls patchwork's ``AMRHierarchy.regrid``
and then moves the fields onto the new patches.

patchwork is not part of the reference tree, so the clustering is this module's own:

* the tags of each level are widened by ``buffer`` cells (periodically) and clustered
  into boxes by the Berger-Rigoutsos algorithm (Berger and Rigoutsos 1991, IEEE SMC
  21:1278): shrink a box to its tags; accept it when the tagged fraction reaches
  ``efficiency`` and no side exceeds ``max_patch_size`` fine cells; otherwise cut it
  at a hole of the tag signature nearest the middle, else at the strongest sign change
  of the signature's second difference, else in the middle of its longest side;
* levels are clustered finest first, and the footprint of every new patch, widened by
  ``NEST`` cells, is added to the tags of the level below, so each new patch nests in
  the new level below with a margin (its ghost cells then interpolate from cells of
  that level, not from its halo).

The state moves by tables (`transfer_tables`) that the device applies level by level,
coarse to fine: cells and nodes that a patch of the same level held before are
copied; new cells take the conservative linear prolongation of their coarse cell
(minmod slopes: the four children sum to four times the coarse cell, so the integrals
are kept) and new A_z nodes the bilinear interpolation of the coarse nodes (the field
stays divergence-free), as the reference's docstring states (mhd.py:91-94).
"""

from __future__ import annotations

import numpy as np

from . import amr_maps
from .hierarchy import AMRHierarchy

HALO = amr_maps.HALO
#: Cells of the new level below around every new patch's footprint.
NEST = 2


# -- clustering -------------------------------------------------------------------------
def dilate(flags: np.ndarray, width: int) -> np.ndarray:
    """``flags`` widened by ``width`` cells along each axis (a square), periodically."""
    out = flags.copy()
    for a in range(flags.ndim):
        grown = out.copy()
        for s in range(1, width + 1):
            grown |= np.roll(out, s, axis=a) | np.roll(out, -s, axis=a)
        out = grown
    return out


def _signature(sub: np.ndarray, axis: int) -> np.ndarray:
    """The tag count of every slab of ``sub`` normal to ``axis``."""
    return sub.sum(axis=tuple(a for a in range(sub.ndim) if a != axis))


def _cut(sub: np.ndarray, min_size: int):
    """(axis, position) to split a tag block at, or None (see the module)."""
    sides = sub.shape
    order = sorted(range(sub.ndim), key=lambda a: -sides[a])  # the longest side first
    for axis in order:  # holes
        sig = _signature(sub, axis)
        n = sides[axis]
        holes = [k for k in range(min_size, n - min_size + 1) if k < n and sig[k] == 0]
        if holes:
            k = min(holes, key=lambda k: abs(2 * k - n))
            return axis, k
    best = None
    for axis in order:  # inflections of the signature
        sig = _signature(sub, axis).astype(np.int64)
        n = sides[axis]
        if n < 4:
            continue
        lap = sig[:-2] - 2 * sig[1:-1] + sig[2:]  # at cells 1 .. n-2
        for k in range(1, len(lap)):
            if lap[k - 1] * lap[k] < 0:
                pos = k + 1  # between cells k and k+1
                if min_size <= pos <= n - min_size:
                    strength = abs(lap[k] - lap[k - 1])
                    if best is None or strength > best[0]:
                        best = (strength, axis, pos)
    if best is not None:
        return best[1], best[2]
    axis = order[0]
    if sides[axis] >= 2 * min_size:
        return axis, sides[axis] // 2
    return None


def cluster(flags: np.ndarray, *, efficiency: float = 0.7, max_size: int | None = None,
            min_size: int = 1) -> list[tuple]:
    """Boxes ``((i0, i1), (j0, j1)[, (k0, k1)])`` covering every flagged cell (see the
    module); 2D or 3D."""
    flags = np.asarray(flags, bool)
    boxes = []
    todo = [tuple((0, n) for n in flags.shape)]
    while todo:
        box = todo.pop()
        sub = flags[tuple(slice(lo, hi) for lo, hi in box)]
        if not sub.any():
            continue
        shrunk = []
        for a, (lo, _) in enumerate(box):
            hit = np.flatnonzero(_signature(sub, a))
            shrunk.append((lo + int(hit[0]), lo + int(hit[-1]) + 1))
        box = tuple(shrunk)
        sub = flags[tuple(slice(lo, hi) for lo, hi in box)]
        big = max_size is not None and any(hi - lo > max_size for lo, hi in box)
        if sub.mean() >= efficiency and not big:
            boxes.append(box)
            continue
        cut = _cut(sub, min_size)
        if cut is None:
            if not big:
                boxes.append(box)
                continue
            axis = int(np.argmax(sub.shape))  # too long for a patch: halve it
            cut = (axis, sub.shape[axis] // 2)
        axis, k = cut
        lo, hi = box[axis]
        first = tuple((lo, lo + k) if a == axis else b for a, b in enumerate(box))
        second = tuple((lo + k, hi) if a == axis else b for a, b in enumerate(box))
        todo += [first, second]
    return sorted(boxes)


def regrid(hierarchy: AMRHierarchy, tags: dict, *, max_patch_size: int | None = None,
           buffer: int = 1, efficiency: float = 0.7, nest: int = NEST) -> AMRHierarchy:
    """A new hierarchy over the same domain whose level ``L + 1`` covers the tagged
    cells of level ``L`` (``tags``: level -> bool array of the level's shape); 2D or 3D,
    periodic."""
    if not all(hierarchy.periodic):
        raise NotImplementedError("regridding runs on periodic hierarchies")
    flags = {}
    for L, t in tags.items():
        t = np.asarray(t, bool)
        if t.shape != hierarchy.level_shape(L):
            raise ValueError(f"the tags of level {L} have shape {t.shape}, "
                             f"not {hierarchy.level_shape(L)}")
        if t.any():
            flags[L] = dilate(t, buffer) if buffer > 0 else t.copy()
    top = max(flags) + 1 if flags else 0
    boxes: dict[int, list] = {}
    fine_max = None if max_patch_size is None else max(1, int(max_patch_size) // 2)
    ndim = hierarchy.ndim
    for L in range(top - 1, -1, -1):
        shape = hierarchy.level_shape(L)
        f = flags.get(L, np.zeros(shape, bool)).copy()
        if L + 2 in boxes:  # the new level L + 2 nests in the new level L + 1
            fine = np.zeros(hierarchy.level_shape(L + 1), bool)
            for box in boxes[L + 2]:
                fine[tuple(slice(lo // 2, hi // 2) for lo, hi in box)] = True
            fine = dilate(fine, nest)
            split = [v for n in shape for v in (n, 2)]
            f |= fine.reshape(split).any(axis=tuple(range(1, 2 * ndim, 2)))
        if not f.any():
            continue
        found = cluster(f, efficiency=efficiency, max_size=fine_max)
        boxes[L + 1] = [tuple((2 * lo, 2 * hi) for lo, hi in box) for box in found]
    out = AMRHierarchy(hierarchy.base_shape, hierarchy.bounds, hierarchy.periodic)
    for L in sorted(boxes):
        for box in boxes[L]:
            out.add_patch(L, box)
    return out


# -- moving the state ---------------------------------------------------------------------
def transfer_tables(old: AMRHierarchy, new: AMRHierarchy, new_maps) -> dict:
    """Per new level: ``copy`` (cells: new offset, old offset, new plane, old plane;
    nodes: new offset, old offset) of what the old hierarchy held on the same level,
    ``cells`` (new offset, new plane, coarse cell offset, coarse row length, coarse
    plane, side x, side y; side -1 / +1: the lower / upper child) and ``nodes`` (new
    offset, coarse base node offset, coarse node row length, kind x, kind y; kind 1:
    the mean with the next coarse node) of the rest.  Whole levels at once."""
    old_layouts, _ = amr_maps.layouts_of(old)
    new_layouts = new_maps.layouts
    boxes = amr_maps.box_array(new_layouts)
    g = amr_maps._patch_arrays(new_layouts)
    Y = g["ny"] + 2 * HALO
    plane = (g["nx"] + 2 * HALO) * Y
    A0 = g["off"] + amr_maps.NFIELDS * plane
    if old_layouts:
        go = amr_maps._patch_arrays(old_layouts)
        Yo = go["ny"] + 2 * HALO
        planeo = (go["nx"] + 2 * HALO) * Yo
        A0o = go["off"] + amr_maps.NFIELDS * planeo
    tables = {}
    for L in range(new.max_level + 1):
        shape = new.level_shape(L)
        if L <= old.max_level:
            cell_own, node_own = amr_maps._owner_tables(old_layouts, L, shape)
        else:
            cell_own = np.full(shape, -1, np.int64)
            node_own = np.full(shape, -1, np.int64)
        coarse_owner = (amr_maps._owner_tables(new_layouts, L - 1, new.level_shape(L - 1))[0]
                        if L > 0 else None)
        ids = np.flatnonzero(g["level"] == L)
        # cells: every interior cell of the level's new patches
        k, a, b = amr_maps._points(ids, g["nx"][ids], g["ny"][ids])
        a, b = a + HALO, b + HALO
        I, J = g["i0"][k] + a - HALO, g["j0"][k] + b - HALO
        src = cell_own[I, J]
        have = src >= 0
        q = src[have]
        copy_cells = (g["off"][k[have]] + a[have] * Y[k[have]] + b[have],
                      go["off"][q] + (I[have] - go["i0"][q] + HALO) * Yo[q]
                      + (J[have] - go["j0"][q] + HALO) if len(q) else np.zeros(0, np.int64),
                      plane[k[have]], planeo[q] if len(q) else np.zeros(0, np.int64))
        miss = ~have
        if miss.any():
            if L == 0:
                raise ValueError("level 0 must cover the domain")
            km, fi, fj = k[miss], I[miss], J[miss]
            q, qa, qbb = amr_maps._coarse_container(new_layouts, L - 1, fi // 2 - 1, fj // 2 - 1,
                                                    3, new, cells=True, owner=coarse_owner,
                                                    boxes=boxes)
            cells = (g["off"][km] + a[miss] * Y[km] + b[miss], plane[km],
                     g["off"][q] + (qa + 1) * Y[q] + (qbb + 1), Y[q], plane[q],
                     np.where(fi % 2 == 1, 1, -1).astype(np.int64),
                     np.where(fj % 2 == 1, 1, -1).astype(np.int64))
        else:
            cells = tuple(np.zeros(0, np.int64) for _ in range(7))
        # nodes, the patch's boundary nodes included (periodic: node n is node 0)
        k, a, b = amr_maps._points(ids, g["nx"][ids] + 1, g["ny"][ids] + 1)
        a, b = a + HALO, b + HALO
        I = (g["i0"][k] + a - HALO) % shape[0]
        J = (g["j0"][k] + b - HALO) % shape[1]
        src = node_own[I, J]
        have = src >= 0
        q = src[have]
        copy_nodes = (A0[k[have]] + a[have] * (Y[k[have]] + 1) + b[have],
                      A0o[q] + ((I[have] - go["i0"][q]) % shape[0] + HALO) * (Yo[q] + 1)
                      + (J[have] - go["j0"][q]) % shape[1] + HALO if len(q)
                      else np.zeros(0, np.int64))
        miss = ~have
        if miss.any():
            if L == 0:
                raise ValueError("level 0 must cover the domain")
            km = k[miss]
            fi = g["i0"][km] + a[miss] - HALO  # unwrapped: the container tries periodic images
            fj = g["j0"][km] + b[miss] - HALO
            q, qa, qbb = amr_maps._coarse_container(new_layouts, L - 1, fi // 2, fj // 2, 2, new,
                                                    cells=False, owner=coarse_owner, boxes=boxes)
            nodes = (A0[km] + a[miss] * (Y[km] + 1) + b[miss], A0[q] + qa * (Y[q] + 1) + qbb,
                     Y[q] + 1, (fi % 2).astype(np.int64), (fj % 2).astype(np.int64))
        else:
            nodes = tuple(np.zeros(0, np.int64) for _ in range(5))
        tables[L] = {"copy_cells": copy_cells, "copy_nodes": copy_nodes, "cells": cells,
                     "nodes": nodes}
    return tables
