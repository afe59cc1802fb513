"""Host-side geometry of a refined 3D patch hierarchy: the gather tables the kernels of
`wellbalanced_amd.patches3d` run.

The reference runs ``MHD3D`` on patchwork hierarchies whose patches are boxes of any
size (``mhd3d.py:448-473,635-671,726-754``): the cells and the edge potential of every
patch take their ghosts from the patches of their level, else by interpolation from the
level below (patchwork's ``exchange_halos``; linear for the conserved cells,
``interpolation="quadratic"`` for the potential); ``FluxRegister`` refluxes the
coarse-fine faces; ``restrict`` sets a coarse edge to the mean of the two fine edges
along it.  patchwork is not available (``uv.lock:3139``), so its rules are this module's
own choices, the 3D analogues of the 2D ones (`wellbalanced_amd.amr_maps`):

* every field's element is owned by the first patch of its level that holds it (a
  node-centred axis holds its boundary): ghosts and shared boundary elements copy from
  that owner;
* the remaining ghosts interpolate from the level below, per axis: a conserved cell
  linearly between the two nearest coarse cells (weights 1/4, 3/4); the potential
  quadratically: along a node-centred axis the coarse node itself (even fine index) or
  the quadratic through coarse nodes m - 1, m, m + 1 (weights -1/8, 3/4, 3/8) midway
  between m and m + 1, along its cell-centred axis the quadratic through the coarse cell
  and its two neighbours at the fine cell's centre (weights 5/32, 30/32, -3/32 for the
  lower child, mirrored for the upper); the weights multiply, x slowest;
* a gathered value is ``w_0 v_0 + w_1 v_1 + ...`` summed in that order (a copy is one
  term of weight 1);
* average-down: the eight children, ``sum 0.125 v`` in x-slowest order; restriction of
  a coarse edge: ``0.5 e_lo + 0.5 e_hi`` of the two fine edges along it;
* the flux register: per uncovered coarse cell next to a finer patch, its coarse-fine
  faces in the order x-high, x-low, y-high, y-low, z-high, z-low, each trading the
  coarse flux for the mean of the four fine ones, ``((f00 + f01) + (f10 + f11)) / 4``.

Domains are periodic on every axis (the refined set-ups this path runs).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

HALO = 2
NCELL = 5
#: Per field: the node-centred axes (the potential A_c is cell-centred along c).
CENTRING = {"cells": (False, False, False), "ax": (False, True, True),
            "ay": (True, False, True), "az": (True, True, False)}
FIELDS = ("cells", "ax", "ay", "az")


@dataclass
class PatchLayout3D:
    level: int
    box: tuple
    offset: int

    @property
    def n(self) -> tuple[int, int, int]:
        return tuple(hi - lo for lo, hi in self.box)

    @property
    def dims(self) -> tuple[int, int, int]:
        return tuple(m + 2 * HALO for m in self.n)

    def extent(self, name: str) -> tuple[int, int, int]:
        """Haloed shape of a field (one component)."""
        return tuple(d + int(node) for d, node in zip(self.dims, CENTRING[name]))

    def base(self, name: str) -> int:
        """Offset of the field's first element (component 0 for the cells)."""
        X, Y, Z = self.dims
        off = self.offset
        if name == "cells":
            return off
        off += NCELL * X * Y * Z
        for other in ("ax", "ay"):
            if name == other:
                return off
            off += int(np.prod(self.extent(other)))
        return off

    @property
    def size(self) -> int:
        X, Y, Z = self.dims
        return NCELL * X * Y * Z + sum(int(np.prod(self.extent(f))) for f in ("ax", "ay", "az"))

    def index(self, name: str, a, b, c) -> np.ndarray:
        """Offsets of haloed local elements (a, b, c) of a field (cells: component 0)."""
        e = self.extent(name)
        return (self.base(name) + (np.asarray(a, np.int64) * e[1] + np.asarray(b, np.int64)) * e[2]
                + np.asarray(c, np.int64))


def layouts_of(hierarchy) -> tuple[list[PatchLayout3D], int]:
    out, offset = [], 0
    for p in hierarchy.patches:
        lay = PatchLayout3D(p.level, tuple(tuple(b) for b in p.box), offset)
        out.append(lay)
        offset += lay.size
    return out, offset


@dataclass
class Stencils:
    """A gather: ``dst[i] = sum_t w[t] * src_buf[src[t]]`` over terms ``start[i] ..
    start[i] + count[i]``."""

    dst: np.ndarray
    start: np.ndarray
    count: np.ndarray
    src: np.ndarray
    w: np.ndarray

    @staticmethod
    def build(parts) -> "Stencils":
        """From a list of (dst array, list of (src array, weight array)) blocks, each
        block's terms interleaved per destination in list order."""
        dsts, starts, counts, srcs, ws = [], [], [], [], []
        total = 0
        for dst, terms in parts:
            n, k = len(dst), len(terms)
            if n == 0:
                continue
            dsts.append(np.asarray(dst, np.int64))
            starts.append(total + k * np.arange(n, dtype=np.int64))
            counts.append(np.full(n, k, np.int32))
            srcs.append(np.stack([np.broadcast_to(np.asarray(s, np.int64), n) for s, _ in terms], 1).ravel())
            ws.append(np.stack([np.broadcast_to(np.asarray(w, np.float64), n) for _, w in terms], 1).ravel())
            total += n * k
        if not dsts:
            z = np.zeros(0, np.int64)
            return Stencils(z, z, np.zeros(0, np.int32), z, np.zeros(0))
        return Stencils(np.concatenate(dsts), np.concatenate(starts), np.concatenate(counts),
                        np.concatenate(srcs), np.concatenate(ws))

    def __len__(self) -> int:
        return len(self.dst)


@dataclass
class Maps3D:
    layouts: list
    total: int
    fill: dict = field(default_factory=dict)        # level -> Stencils (ghosts, shared boundaries)
    coarsen: dict = field(default_factory=dict)     # coarse level -> Stencils (cells + edges)
    restrict: dict = field(default_factory=dict)    # coarse level -> Stencils (edges only)
    reflux: list = field(default_factory=list)      # per refluxed cell: (patch, a, b, c, faces)
    covered: list = field(default_factory=list)     # per patch: bool (n) mask of covered cells


def _owners(layouts, level, shape, name):
    """The first patch of ``level`` holding each element of field ``name`` (-1: none),
    on the periodic index space of the level (a node index n is node 0)."""
    nodes = CENTRING[name]
    own = np.full(shape, -1, np.int64)
    for k, lay in enumerate(layouts):
        if lay.level != level:
            continue
        idx = [np.arange(lo, hi + int(node)) % n for (lo, hi), node, n in zip(lay.box, nodes, shape)]
        sub = own[np.ix_(*idx)]
        own[np.ix_(*idx)] = np.where(sub < 0, k, sub)
    return own


def _local(lay, shape, I):
    """Haloed local coordinates in patch ``lay`` of wrapped level coordinates I (3, N)."""
    return [(I[a] - lay.box[a][0]) % shape[a] + HALO for a in range(3)]


def _axis_rule(kind, f):
    """(coarse index arrays, weights) per term along one axis for fine indices f."""
    f = np.asarray(f, np.int64)
    if kind == "linear":
        c0 = np.floor_divide(f - 1, 2)
        w1 = np.where(f % 2 == 0, 0.75, 0.25)
        return [(c0, 1.0 - w1), (c0 + 1, w1)]
    if kind == "node":
        odd = f % 2 == 1
        m = np.floor_divide(f - 1, 2)
        # even: the coarse node f / 2 (a one-term rule padded with zero weights)
        return [(np.where(odd, m - 1, f // 2), np.where(odd, -0.125, 1.0)),
                (np.where(odd, m, f // 2), np.where(odd, 0.75, 0.0)),
                (np.where(odd, m + 1, f // 2), np.where(odd, 0.375, 0.0))]
    # "cell-quadratic": the coarse cell c = f // 2 and its neighbours at the fine centre
    c = np.floor_divide(f, 2)
    lower = f % 2 == 0
    return [(c - 1, np.where(lower, 5.0 / 32.0, -3.0 / 32.0)), (c, np.full(f.shape, 30.0 / 32.0)),
            (c + 1, np.where(lower, -3.0 / 32.0, 5.0 / 32.0))]


def _cell_plane(lay) -> int:
    return int(np.prod(lay.extent("cells")))


def _interp_terms(layouts, coarse_owner, name, fine, kinds):
    """Terms (component-0 src offsets, weights, cell planes of the holding patches) of the
    interpolation of fine elements ``fine`` (3, N, unwrapped coordinates) of field
    ``name`` from the level below (owner map ``coarse_owner``), x slowest, with the terms
    of zero weight for every element left out."""
    shape_c = coarse_owner.shape
    rules = [_axis_rule(kind, fine[a]) for a, kind in enumerate(kinds)]
    terms = []
    for tx in rules[0]:
        for ty in rules[1]:
            for tz in rules[2]:
                w = (tx[1] * ty[1]) * tz[1]
                if np.all(w == 0.0):
                    continue
                I = np.stack([tx[0] % shape_c[0], ty[0] % shape_c[1], tz[0] % shape_c[2]])
                q = coarse_owner[I[0], I[1], I[2]]
                if (q < 0).any():
                    raise ValueError("a fine patch is not properly nested in the level below")
                off = np.empty(len(q), np.int64)
                plane = np.empty(len(q), np.int64)
                for k in np.unique(q):
                    sel = q == k
                    lay = layouts[k]
                    off[sel] = lay.index(name, *_local(lay, shape_c, I[:, sel]))
                    plane[sel] = _cell_plane(lay)
                terms.append((off, w, plane))
    return terms


def _interp_parts(layouts, coarse_owner, name, fine, dst, dst_plane, kinds):
    """Stencil blocks of interpolated elements, one per pattern of nonzero weights (the
    node rule of an even index has one term along that axis), per component."""
    rules = [_axis_rule(kind, fine[a]) for a, kind in enumerate(kinds)]
    pattern = sum((np.stack([r[1] != 0.0 for r in rule]).sum(0)) * 10**a
                  for a, rule in enumerate(rules))
    comps = NCELL if name == "cells" else 1
    parts = []
    for pat in np.unique(pattern):
        sel = pattern == pat
        terms = _interp_terms(layouts, coarse_owner, name, fine[:, sel], kinds)
        for comp in range(comps):
            parts.append((dst[sel] + comp * dst_plane,
                          [(off + comp * plane, w) for off, w, plane in terms]))
    return parts


def build(hierarchy) -> Maps3D:
    if hierarchy.ndim != 3 or not all(hierarchy.periodic):
        raise NotImplementedError("refined 3D patch hierarchies run on periodic domains")
    layouts, total = layouts_of(hierarchy)
    maps = Maps3D(layouts, total)
    nlev = hierarchy.max_level + 1
    owners = {(L, name): _owners(layouts, L, hierarchy.level_shape(L), name)
              for L in range(nlev) for name in FIELDS}
    if (owners[0, "cells"] < 0).any():
        raise ValueError("level 0 must cover the domain")
    for L in range(nlev):
        shape = hierarchy.level_shape(L)
        parts = []
        for k, lay in enumerate(layouts):
            if lay.level != L:
                continue
            for name in FIELDS:
                nodes = CENTRING[name]
                ext = lay.extent(name)
                P = np.indices(ext).reshape(3, -1).astype(np.int64)
                inner = np.ones(P.shape[1], bool)
                edge = np.zeros(P.shape[1], bool)
                for a in range(3):
                    hi = HALO + lay.n[a] + (1 if nodes[a] else 0)
                    inner &= (P[a] >= HALO) & (P[a] < hi)
                    if nodes[a]:
                        edge |= (P[a] == HALO) | (P[a] == hi - 1)
                edge &= inner
                todo = ~inner | edge
                P, edge = P[:, todo], edge[todo]
                fine = P + np.array([lo for lo, _ in lay.box])[:, None] - HALO
                I = np.stack([fine[a] % shape[a] for a in range(3)])
                own = owners[L, name][I[0], I[1], I[2]]
                dst_all = lay.index(name, *P)
                comps = NCELL if name == "cells" else 1
                plane = _cell_plane(lay)
                have = own >= 0
                src = np.empty(P.shape[1], np.int64)
                src_plane = np.empty(P.shape[1], np.int64)
                local = np.zeros_like(P)
                for q in np.unique(own[have]):
                    sel = own == q
                    loc = _local(layouts[q], shape, I[:, sel])
                    local[:, sel] = np.stack(loc)
                    src[sel] = layouts[q].index(name, *loc)
                    src_plane[sel] = _cell_plane(layouts[q])
                same = have & (own == k) & np.all(local == P, axis=0)
                copy = have & ~same
                for comp in range(comps):
                    parts.append((dst_all[copy] + comp * plane,
                                  [(src[copy] + comp * src_plane[copy], 1.0)]))
                miss = ~have
                if miss.any():
                    if L == 0 or edge[miss].any():
                        raise ValueError("level 0 must cover the domain")
                    kinds = (["linear"] * 3 if name == "cells" else
                             ["node" if nodes[a] else "cell-quadratic" for a in range(3)])
                    parts += _interp_parts(layouts, owners[L - 1, name], name, fine[:, miss],
                                           dst_all[miss], plane, kinds)
        maps.fill[L] = Stencils.build(parts)
    for L in range(nlev - 1):
        maps.coarsen[L], maps.restrict[L] = _coarsen_tables(layouts, owners, hierarchy, L)
    maps.reflux = _reflux_cells(layouts, owners, hierarchy)
    for k, lay in enumerate(layouts):
        cov = np.zeros(lay.n, bool)
        if lay.level + 1 < nlev:
            fine = owners[lay.level + 1, "cells"]
            idx = np.indices(lay.n).reshape(3, -1) + np.array([lo for lo, _ in lay.box])[:, None]
            hit = np.zeros(idx.shape[1], bool)
            for d in np.ndindex(2, 2, 2):
                f = [(2 * idx[a] + d[a]) for a in range(3)]
                hit |= fine[f[0], f[1], f[2]] >= 0
            cov = hit.reshape(lay.n)
        maps.covered.append(cov)
    return maps


def _coarsen_tables(layouts, owners, hierarchy, L):
    """Level L from level L + 1: covered cells := the mean of their 8 children, and the
    coarse edges over fine edges := the mean of the two fine edges along them (every
    coarse patch's own elements)."""
    shape_f = hierarchy.level_shape(L + 1)
    cell_parts, edge_parts = [], []
    fine_cells = owners[L + 1, "cells"]
    for k, lay in enumerate(layouts):
        if lay.level != L:
            continue
        plane = _cell_plane(lay)
        C = np.indices(lay.n).reshape(3, -1) + np.array([lo for lo, _ in lay.box])[:, None]
        child = [np.stack([2 * C[a] + d[a] for a in range(3)]) for d in np.ndindex(2, 2, 2)]
        qs = np.stack([fine_cells[c[0], c[1], c[2]] for c in child])
        full = np.all(qs >= 0, axis=0)
        if full.any():
            lo = np.array([b0 for b0, _ in lay.box])[:, None]
            dst = lay.index("cells", *(C[:, full] - lo + HALO))
            terms = []
            for c, q in zip(child, qs):
                c, q = c[:, full], q[full]
                off = np.empty(len(q), np.int64)
                pf = np.empty(len(q), np.int64)
                for p in np.unique(q):
                    sel = q == p
                    off[sel] = layouts[p].index("cells", *_local(layouts[p], shape_f, c[:, sel]))
                    pf[sel] = _cell_plane(layouts[p])
                terms.append((off, 0.125, pf))
            for comp in range(NCELL):
                cell_parts.append((dst + comp * plane,
                                   [(off + comp * pf, w) for off, w, pf in terms]))
        for c_axis, name in enumerate(("ax", "ay", "az")):
            nodes = CENTRING[name]
            ext = [lay.n[a] + int(nodes[a]) for a in range(3)]
            Q = np.indices(ext).reshape(3, -1) + np.array([lo for lo, _ in lay.box])[:, None]
            ends = []
            for s in (0, 1):
                F = np.stack([(2 * Q[a] + (s if a == c_axis else 0)) % shape_f[a] for a in range(3)])
                q = owners[L + 1, name][F[0], F[1], F[2]]
                ends.append((F, q))
            ok = (ends[0][1] >= 0) & (ends[1][1] >= 0)
            if not ok.any():
                continue
            dst = lay.index(name, *(Q[:, ok] - np.array([lo for lo, _ in lay.box])[:, None] + HALO))
            terms = []
            for F, q in ends:
                F, q = F[:, ok], q[ok]
                off = np.empty(len(q), np.int64)
                for p in np.unique(q):
                    sel = q == p
                    off[sel] = layouts[p].index(name, *_local(layouts[p], shape_f, F[:, sel]))
                terms.append((off, 0.5))
            edge_parts.append((dst, terms))
    return Stencils.build(cell_parts + edge_parts), Stencils.build(edge_parts)


def _reflux_cells(layouts, owners, hierarchy):
    """Per uncovered coarse cell next to a finer patch: (coarse patch, haloed local cell,
    level, faces); a face is (axis, side, coarse face (patch, local a, b, c: the low face
    of that local cell), four fine faces (patch, local face) in the order (p, q) =
    (0,0), (0,1), (1,0), (1,1) over the two other axes in increasing order)."""
    out = []
    nlev = hierarchy.max_level + 1
    for L in range(nlev - 1):
        shape_c = hierarchy.level_shape(L)
        shape_f = hierarchy.level_shape(L + 1)
        fine = owners[L + 1, "cells"]
        cov = np.zeros(shape_c, bool)
        for d in np.ndindex(2, 2, 2):
            cov |= fine[d[0]::2, d[1]::2, d[2]::2] >= 0
        for k, lay in enumerate(layouts):
            if lay.level != L:
                continue
            lo = np.array([b[0] for b in lay.box])
            for C in np.argwhere(~cov[tuple(slice(b0, b1) for b0, b1 in lay.box)]) + lo:
                faces = []
                for axis in range(3):
                    for side, step in ((0, 1), (1, -1)):  # high face, low face
                        nb = C.copy()
                        nb[axis] = (nb[axis] + step) % shape_c[axis]
                        if not cov[tuple(nb)]:
                            continue
                        a1, a2 = [a for a in range(3) if a != axis]
                        loc = C - lo + HALO
                        cface = loc.copy()
                        if side == 0:
                            cface[axis] += 1
                        fines = []
                        for p in (0, 1):
                            for q in (0, 1):
                                F = 2 * C.copy()
                                F[a1] += p
                                F[a2] += q
                                F[axis] = 2 * C[axis] + 2 if side == 0 else 2 * C[axis] - 1
                                F = F % np.array(shape_f)
                                pf = int(fine[tuple(F)])
                                fl = np.array(_local(layouts[pf], shape_f, F[:, None]))[:, 0]
                                if side == 1:  # the fine cell's high face
                                    fl[axis] += 1
                                fines.append((pf, *fl))
                        faces.append((axis, side, (k, *cface), fines))
                if faces:
                    out.append((k, *(C - lo + HALO), L, faces))
    return out


def pack_state(hierarchy, layouts, total, gamma, mean_field, state) -> np.ndarray:
    """The flat FP64 buffer of every patch's haloed state (``set_state`` on each patch,
    ``mhd3d.py:388-440``), patch after patch as in ``layouts``."""
    from .initial import host_state_3d

    buf = np.zeros(total)
    for lay, patch in zip(layouts, hierarchy.patches):
        cons, pots = host_state_3d(hierarchy, patch, gamma, mean_field, **state)
        buf[lay.offset : lay.offset + cons.size] = cons.ravel()
        for name, pot in zip(("ax", "ay", "az"), pots):
            b = lay.base(name)
            buf[b : b + pot.size] = pot.ravel()
    return buf


def _regrid_rule(node, f):
    """Per axis (coarse indices, weights) of a new edge of a regrid: along the edge the
    coarse cell holding it, across it the coarse node (even index) or the mean of the
    two around it (odd)."""
    f = np.asarray(f, np.int64)
    if not node:
        return [(np.floor_divide(f, 2), np.ones(f.shape))]
    odd = f % 2 == 1
    m = np.floor_divide(f, 2)
    return [(m, np.where(odd, 0.5, 1.0)), (m + 1, np.where(odd, 0.5, 0.0))]


def transfer_tables(old_hierarchy, old_maps, new_hierarchy, new_maps) -> dict:
    """The move of a regrid (``MHD3D.regrid``, mhd3d.py:803-821), per new level: the
    cells and edges the old hierarchy held on the same level (a gather from the old
    buffer), the conservative minmod prolongation of the new cells from the new level
    below (rows of `wb_patches3d_prolong`), and the new edges interpolated from the new
    level below (along the edge the coarse edge holding it, across it linearly), a
    gather in the new buffer."""
    old_layouts, new_layouts = old_maps.layouts, new_maps.layouts
    nlev = new_hierarchy.max_level + 1
    out = {}
    for L in range(nlev):
        shape = new_hierarchy.level_shape(L)
        old_own = {name: (_owners(old_layouts, L, shape, name) if L <= old_hierarchy.max_level
                          else np.full(shape, -1, np.int64)) for name in FIELDS}
        coarse_own = ({name: _owners(new_layouts, L - 1, new_hierarchy.level_shape(L - 1), name)
                       for name in FIELDS} if L > 0 else None)
        copies, prolong, interp = [], [], []
        for lay in new_layouts:
            if lay.level != L:
                continue
            lo = np.array([b for b, _ in lay.box])[:, None]
            plane = _cell_plane(lay)
            for name in FIELDS:
                nodes = CENTRING[name]
                ext = [lay.n[a] + int(nodes[a]) for a in range(3)]
                P = np.indices(ext).reshape(3, -1).astype(np.int64) + HALO
                fine = P - HALO + lo
                I = np.stack([fine[a] % shape[a] for a in range(3)])
                own = old_own[name][I[0], I[1], I[2]]
                dst = lay.index(name, *P)
                have = own >= 0
                src = np.empty(P.shape[1], np.int64)
                src_plane = np.empty(P.shape[1], np.int64)
                for q in np.unique(own[have]):
                    sel = own == q
                    src[sel] = old_layouts[q].index(name, *_local(old_layouts[q], shape, I[:, sel]))
                    src_plane[sel] = _cell_plane(old_layouts[q])
                comps = NCELL if name == "cells" else 1
                for comp in range(comps):
                    copies.append((dst[have] + comp * plane, [(src[have] + comp * src_plane[have], 1.0)]))
                miss = ~have
                if not miss.any():
                    continue
                if L == 0:
                    raise ValueError("level 0 must cover the domain")
                shape_c = new_hierarchy.level_shape(L - 1)
                if name == "cells":
                    C = np.stack([np.floor_divide(fine[a][miss], 2) % shape_c[a] for a in range(3)])
                    q = coarse_own["cells"][C[0], C[1], C[2]]
                    if (q < 0).any():
                        raise ValueError("a new patch is not properly nested in the level below")
                    rows = np.zeros((int(miss.sum()), 9), np.int64)
                    rows[:, 0] = dst[miss]
                    rows[:, 1] = plane
                    for p in np.unique(q):
                        sel = q == p
                        cl = new_layouts[p]
                        _, Yc, Zc = cl.dims
                        rows[sel, 2] = cl.index("cells", *_local(cl, shape_c, C[:, sel]))
                        rows[sel, 3] = Yc * Zc
                        rows[sel, 4] = Zc
                        rows[sel, 5] = _cell_plane(cl)
                    for a in range(3):
                        rows[:, 6 + a] = np.where(fine[a][miss] % 2 == 1, 1, -1)
                    prolong.append(rows)
                    continue
                F = fine[:, miss]
                rules = [_regrid_rule(nodes[a], F[a]) for a in range(3)]
                pattern = sum((np.stack([r[1] != 0.0 for r in rule]).sum(0)) * 10**a
                              for a, rule in enumerate(rules))
                for pat in np.unique(pattern):
                    sel = pattern == pat
                    terms = []
                    for tx in rules[0]:
                        for ty in rules[1]:
                            for tz in rules[2]:
                                w = ((tx[1] * ty[1]) * tz[1])[sel]
                                if np.all(w == 0.0):
                                    continue
                                Ic = np.stack([tx[0][sel] % shape_c[0], ty[0][sel] % shape_c[1],
                                               tz[0][sel] % shape_c[2]])
                                qq = coarse_own[name][Ic[0], Ic[1], Ic[2]]
                                if (qq < 0).any():
                                    raise ValueError("a new patch is not properly nested")
                                off = np.empty(len(qq), np.int64)
                                for p in np.unique(qq):
                                    s2 = qq == p
                                    off[s2] = new_layouts[p].index(name, *_local(new_layouts[p], shape_c, Ic[:, s2]))
                                terms.append((off, w))
                    interp.append((dst[miss][sel], terms))
        out[L] = {"copies": Stencils.build(copies),
                  "prolong": np.concatenate(prolong) if prolong else np.zeros((0, 9), np.int64),
                  "interp": Stencils.build(interp)}
    return out
