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


def _cut(sub: np.ndarray, min_size: int):
    """(axis, position) to split a tag block at, or None (see the module)."""
    sides = sub.shape
    order = sorted(range(2), key=lambda a: -sides[a])  # the longest side first
    for axis in order:  # holes
        sig = sub.sum(axis=1 - axis)
        n = sides[axis]
        holes = [k for k in range(min_size, n - min_size + 1) if k < n and sig[k] == 0]
        if holes:
            k = min(holes, key=lambda k: abs(2 * k - n))
            return axis, k
    best = None
    for axis in order:  # inflections of the signature
        sig = sub.sum(axis=1 - axis).astype(np.int64)
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
    """Boxes ``((i0, i1), (j0, j1))`` covering every flagged cell (see the module)."""
    flags = np.asarray(flags, bool)
    boxes = []
    todo = [((0, flags.shape[0]), (0, flags.shape[1]))]
    while todo:
        (i0, i1), (j0, j1) = todo.pop()
        sub = flags[i0:i1, j0:j1]
        if not sub.any():
            continue
        rows = np.flatnonzero(sub.any(axis=1))
        cols = np.flatnonzero(sub.any(axis=0))
        i0, i1 = i0 + int(rows[0]), i0 + int(rows[-1]) + 1
        j0, j1 = j0 + int(cols[0]), j0 + int(cols[-1]) + 1
        sub = flags[i0:i1, j0:j1]
        big = max_size is not None and (i1 - i0 > max_size or j1 - j0 > max_size)
        if sub.mean() >= efficiency and not big:
            boxes.append(((i0, i1), (j0, j1)))
            continue
        cut = _cut(sub, min_size)
        if cut is None:
            if not big:
                boxes.append(((i0, i1), (j0, j1)))
                continue
            axis = int(np.argmax(sub.shape))  # too long for a patch: halve it
            cut = (axis, sub.shape[axis] // 2)
        axis, k = cut
        if axis == 0:
            todo += [((i0, i0 + k), (j0, j1)), ((i0 + k, i1), (j0, j1))]
        else:
            todo += [((i0, i1), (j0, j0 + k)), ((i0, i1), (j0 + k, j1))]
    return sorted(boxes)


def regrid(hierarchy: AMRHierarchy, tags: dict, *, max_patch_size: int | None = None,
           buffer: int = 1, efficiency: float = 0.7, nest: int = NEST) -> AMRHierarchy:
    """A new hierarchy over the same domain whose level ``L + 1`` covers the tagged
    cells of level ``L`` (``tags``: level -> bool array of the level's shape)."""
    if hierarchy.ndim != 2 or not all(hierarchy.periodic):
        raise NotImplementedError("regridding runs on periodic 2D hierarchies")
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
    for L in range(top - 1, -1, -1):
        shape = hierarchy.level_shape(L)
        f = flags.get(L, np.zeros(shape, bool)).copy()
        if L + 2 in boxes:  # the new level L + 2 nests in the new level L + 1
            fine = np.zeros(hierarchy.level_shape(L + 1), bool)
            for (i0, i1), (j0, j1) in boxes[L + 2]:
                fine[i0 // 2 : i1 // 2, j0 // 2 : j1 // 2] = True
            fine = dilate(fine, nest)
            f |= fine.reshape(shape[0], 2, shape[1], 2).any(axis=(1, 3))
        if not f.any():
            continue
        found = cluster(f, efficiency=efficiency, max_size=fine_max)
        boxes[L + 1] = [((2 * i0, 2 * i1), (2 * j0, 2 * j1)) for (i0, i1), (j0, j1) in found]
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
    the mean with the next coarse node) of the rest."""
    old_layouts, _ = amr_maps.layouts_of(old)
    new_layouts = new_maps.layouts
    boxes = amr_maps.box_array(new_layouts)
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
        cd, cs, cpd, cps, nd, ns = [], [], [], [], [], []
        pd, pp, pc0, pcy, pcp, psx, psy = [], [], [], [], [], [], []
        qd, qb, qs, qkx, qky = [], [], [], [], []
        for lay in new_layouts:
            if lay.level != L:
                continue
            (i0, i1), (j0, j1) = lay.box
            a, b = np.meshgrid(np.arange(HALO, HALO + lay.nx), np.arange(HALO, HALO + lay.ny),
                               indexing="ij")
            a, b = a.ravel(), b.ravel()
            I, J = i0 + a - HALO, j0 + b - HALO
            src = cell_own[I, J]
            for q in np.unique(src[src >= 0]):
                sel = src == q
                ql = old_layouts[q]
                cd.append(lay.cell(a[sel], b[sel]))
                cs.append(ql.cell(I[sel] - ql.box[0][0] + HALO, J[sel] - ql.box[1][0] + HALO))
                cpd.append(np.full(sel.sum(), lay.plane, np.int64))
                cps.append(np.full(sel.sum(), ql.plane, np.int64))
            miss = src < 0
            if miss.any():
                if L == 0:
                    raise ValueError("level 0 must cover the domain")
                fi, fj = I[miss], J[miss]
                q, qa, qbb = amr_maps._coarse_container(new_layouts, L - 1, fi // 2 - 1,
                                                        fj // 2 - 1, 3, new, cells=True,
                                                        owner=coarse_owner, boxes=boxes)
                for qq in np.unique(q):
                    sel = q == qq
                    ql = new_layouts[qq]
                    pd.append(lay.cell(a[miss][sel], b[miss][sel]))
                    pp.append(np.full(sel.sum(), lay.plane, np.int64))
                    pc0.append(ql.cell(qa[sel] + 1, qbb[sel] + 1))
                    pcy.append(np.full(sel.sum(), ql.Y, np.int64))
                    pcp.append(np.full(sel.sum(), ql.plane, np.int64))
                    psx.append(np.where(fi[sel] % 2 == 1, 1, -1).astype(np.int64))
                    psy.append(np.where(fj[sel] % 2 == 1, 1, -1).astype(np.int64))
            # nodes, the patch's boundary nodes included (periodic: node n is node 0)
            a, b = np.meshgrid(np.arange(HALO, HALO + lay.nx + 1),
                               np.arange(HALO, HALO + lay.ny + 1), indexing="ij")
            a, b = a.ravel(), b.ravel()
            I, J = (i0 + a - HALO) % shape[0], (j0 + b - HALO) % shape[1]
            src = node_own[I, J]
            for q in np.unique(src[src >= 0]):
                sel = src == q
                ql = old_layouts[q]
                nd.append(lay.node(a[sel], b[sel]))
                ns.append(ql.node((I[sel] - ql.box[0][0]) % shape[0] + HALO,
                                  (J[sel] - ql.box[1][0]) % shape[1] + HALO))
            miss = src < 0
            if miss.any():
                if L == 0:
                    raise ValueError("level 0 must cover the domain")
                fi = i0 + a[miss] - HALO  # unwrapped: the container tries periodic images
                fj = j0 + b[miss] - HALO
                q, qa, qbb = amr_maps._coarse_container(new_layouts, L - 1, fi // 2, fj // 2, 2,
                                                        new, cells=False, owner=coarse_owner,
                                                        boxes=boxes)
                for qq in np.unique(q):
                    sel = q == qq
                    ql = new_layouts[qq]
                    qd.append(lay.node(a[miss][sel], b[miss][sel]))
                    qb.append(ql.node(qa[sel], qbb[sel]))
                    qs.append(np.full(sel.sum(), ql.Y + 1, np.int64))
                    qkx.append((fi[sel] % 2).astype(np.int64))
                    qky.append((fj[sel] % 2).astype(np.int64))
        cat = amr_maps._cat
        tables[L] = {
            "copy_cells": (cat(cd), cat(cs), cat(cpd), cat(cps)),
            "copy_nodes": (cat(nd), cat(ns)),
            "cells": (cat(pd), cat(pp), cat(pc0), cat(pcy), cat(pcp), cat(psx), cat(psy)),
            "nodes": (cat(qd), cat(qb), cat(qs), cat(qkx), cat(qky)),
        }
    return tables
