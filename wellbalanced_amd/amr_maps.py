"""Host-side geometry of a 2D patch hierarchy: the gather tables the AMR kernels run.

The reference lets ``patchwork`` (not available here) fill the halos of refined
patches (same-level copies, bilinear interpolation of the conserved fields and
quadratic of A_z from the level below, ``mhd.py:744-777``, ``CHANGELOG.md:10-13``),
reflux coarse-fine faces (``FluxRegister.update/reflux``, ``mhd.py:911-919,989-990``)
and restrict A_z (``mhd.py:981,991``).  Here each of those becomes a flat table of
integer offsets into one FP64 buffer that holds every patch's state (6 haloed cell
planes then the haloed nodal A_z, patch after patch), built once per hierarchy with
NumPy and consumed by gather kernels:

* ``copy``: ghost cells/nodes covered by a same-level patch (or its periodic image)
  and shared boundary nodes, taken from their canonical owner (the first patch that
  holds them);
* ``cell_interp``: the remaining ghost cells, bilinear from the 2x2 nearest coarse
  cells (weights 1/4, 3/4 per axis) for each of the 6 fields;
* ``node_interp``: the remaining ghost nodes, per axis the coarse node itself
  (even fine index) or, midway between coarse nodes m and m + 1 (odd index), the
  quadratic through nodes m - 1, m, m + 1 (weights -1/8, 3/4, 3/8);
* ``reflux``: per uncovered coarse cell next to a finer patch, up to four faces
  whose coarse flux (of both Heun stages) is traded for the mean of the two fine
  fluxes once per step;
* ``average``: covered coarse cells, the mean of their four fine children;
* ``restrict``: coarse nodes that coincide with fine nodes take the fine value.

On a non-periodic axis the ghosts outside the domain are left to the outflow or
wall rules (`wb_amr2d_domain_faces`, after each level's fill), interpolation stencils
that reach outside the domain read them from a coarse patch's halo, and the nodes of
a level run 0 .. n along that axis (on a periodic axis node n is node 0).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

HALO = 2
NFIELDS = 6


@dataclass
class PatchLayout:
    level: int
    box: tuple  # ((i0, i1), (j0, j1)) in level cells
    offset: int  # first double of the patch's state in the buffer

    @property
    def nx(self) -> int:
        return self.box[0][1] - self.box[0][0]

    @property
    def ny(self) -> int:
        return self.box[1][1] - self.box[1][0]

    @property
    def X(self) -> int:
        return self.nx + 2 * HALO

    @property
    def Y(self) -> int:
        return self.ny + 2 * HALO

    @property
    def plane(self) -> int:
        return self.X * self.Y

    @property
    def size(self) -> int:
        return NFIELDS * self.plane + (self.X + 1) * (self.Y + 1)

    def cell(self, a, b):
        """Offset of local haloed cell (a, b) in field 0."""
        return self.offset + np.asarray(a, np.int64) * self.Y + np.asarray(b, np.int64)

    def node(self, a, b):
        """Offset of local haloed node (a, b) of A_z."""
        return (self.offset + NFIELDS * self.plane + np.asarray(a, np.int64) * (self.Y + 1)
                + np.asarray(b, np.int64))


@dataclass
class Maps:
    """The gather tables of one hierarchy (all int64 offsets into the state buffer)."""

    layouts: list
    total: int
    # per level >= 0, ghost fill in this order: copies, cell interpolation, node interpolation
    copy_cells: dict = field(default_factory=dict)      # level -> (dst, src, plane_dst, plane_src)
    copy_nodes: dict = field(default_factory=dict)      # level -> (dst, src)
    cell_interp: dict = field(default_factory=dict)     # level -> (dst, plane_dst, c00, cy, plane_c, wx, wy)
    node_interp: dict = field(default_factory=dict)     # level -> (dst, base, sx, sy, kx, ky)
    reflux: dict = field(default_factory=dict)          # level -> arrays (coarse level)
    average: dict = field(default_factory=dict)         # level -> (coarse, plane_c, fine, fine_Y, plane_f)
    restrict: dict = field(default_factory=dict)        # level -> (coarse, fine)
    domain_faces: dict = field(default_factory=dict)    # level -> rows (patch, faces)


def layouts_of(hierarchy) -> tuple[list[PatchLayout], int]:
    out, offset = [], 0
    for p in hierarchy.patches:
        lay = PatchLayout(p.level, tuple(tuple(b) for b in p.box), offset)
        out.append(lay)
        offset += lay.size
    return out, offset


def _owner_tables(layouts, level, shape, periodic=(True, True)):
    """Per level: the index of the first patch owning each cell (-1 if none) and node
    (nodes 0 .. n - 1 along a periodic axis, where node n is node 0; 0 .. n along a
    non-periodic one)."""
    nx, ny = shape
    cells = np.full((nx, ny), -1, np.int64)
    nodes = np.full((nx + (not periodic[0]), ny + (not periodic[1])), -1, np.int64)
    for k, lay in enumerate(layouts):
        if lay.level != level:
            continue
        (i0, i1), (j0, j1) = lay.box
        view = cells[i0:i1, j0:j1]
        view[view < 0] = k
        ii = np.arange(i0, i1 + 1) % nx if periodic[0] else np.arange(i0, i1 + 1)
        jj = np.arange(j0, j1 + 1) % ny if periodic[1] else np.arange(j0, j1 + 1)
        sub = nodes[np.ix_(ii, jj)]
        sub[sub < 0] = k
        nodes[np.ix_(ii, jj)] = sub
    return cells, nodes


def _wrap(v, n, periodic):
    """Level coordinates ``v`` on an axis of ``n`` cells: the periodic image in 0 .. n - 1,
    or unchanged on a non-periodic axis."""
    return v % n if periodic else v


def _patch_arrays(layouts) -> dict:
    """Per-patch geometry as arrays (index = patch number)."""
    return {
        "off": np.array([lay.offset for lay in layouts], np.int64),
        "i0": np.array([lay.box[0][0] for lay in layouts], np.int64),
        "j0": np.array([lay.box[1][0] for lay in layouts], np.int64),
        "nx": np.array([lay.nx for lay in layouts], np.int64),
        "ny": np.array([lay.ny for lay in layouts], np.int64),
        "level": np.array([lay.level for lay in layouts], np.int64),
    }


def _points(ids, ext_a, ext_b):
    """Every local point (a, b) of the grids ext_a[k] x ext_b[k] of the patches ``ids``:
    (patch, a, b) flattened, row-major per patch."""
    sizes = ext_a * ext_b
    k = np.repeat(ids, sizes)
    t = np.arange(int(sizes.sum()), dtype=np.int64) - np.repeat(np.cumsum(sizes) - sizes, sizes)
    eb = np.repeat(ext_b, sizes)
    return k, t // eb, t % eb


def build(hierarchy) -> Maps:
    """All gather tables for a hierarchy (whole levels at once)."""
    per = tuple(bool(p) for p in hierarchy.periodic)
    layouts, total = layouts_of(hierarchy)
    maps = Maps(layouts=layouts, total=total)
    nlev = hierarchy.max_level + 1
    owners = [_owner_tables(layouts, L, hierarchy.level_shape(L), per) for L in range(nlev)]
    boxes = box_array(layouts)
    g = _patch_arrays(layouts)
    Y = g["ny"] + 2 * HALO
    plane = (g["nx"] + 2 * HALO) * Y
    YN = Y + 1
    A0 = g["off"] + NFIELDS * plane  # first A_z node of each patch

    def cell(k, a, b):
        return g["off"][k] + a * Y[k] + b

    def node(k, a, b):
        return A0[k] + a * YN[k] + b

    for L in range(nlev):
        shape = hierarchy.level_shape(L)
        cell_own, node_own = owners[L]
        ids = np.flatnonzero(g["level"] == L)
        # ghost cells
        k, a, b = _points(ids, g["nx"][ids] + 2 * HALO, g["ny"][ids] + 2 * HALO)
        ghost = ~((a >= HALO) & (a < HALO + g["nx"][k]) & (b >= HALO) & (b < HALO + g["ny"][k]))
        k, a, b = k[ghost], a[ghost], b[ghost]
        fi = g["i0"][k] + a - HALO  # unwrapped level coordinates
        fj = g["j0"][k] + b - HALO
        # ghosts outside the domain on a non-periodic axis: the boundary rules fill them
        inside = ((per[0] | ((fi >= 0) & (fi < shape[0])))
                  & (per[1] | ((fj >= 0) & (fj < shape[1]))))
        k, a, b, fi, fj = k[inside], a[inside], b[inside], fi[inside], fj[inside]
        I, J = _wrap(fi, shape[0], per[0]), _wrap(fj, shape[1], per[1])
        src = cell_own[I, J]
        have = src >= 0
        q = src[have]
        maps.copy_cells[L] = (
            cell(k[have], a[have], b[have]),
            cell(q, I[have] - g["i0"][q] + HALO, J[have] - g["j0"][q] + HALO),
            plane[k[have]], plane[q])
        miss = ~have
        if miss.any() and L == 0:
            raise ValueError("level 0 must cover the domain")
        if miss.any():
            mk, ma, mb, mfi, mfj = k[miss], a[miss], b[miss], fi[miss], fj[miss]
            ci0 = np.floor_divide(mfi - 1, 2)  # lower coarse cell of the stencil
            cj0 = np.floor_divide(mfj - 1, 2)
            q, qa, qb = _coarse_container(layouts, L - 1, ci0, cj0, 2, hierarchy, cells=True,
                                          owner=owners[L - 1][0], boxes=boxes, periodic=per)
            maps.cell_interp[L] = (
                cell(mk, ma, mb), plane[mk], cell(q, qa, qb), Y[q], plane[q],
                np.where(mfi % 2 == 0, 0.75, 0.25), np.where(mfj % 2 == 0, 0.75, 0.25))
        else:
            maps.cell_interp[L] = tuple([np.zeros(0, np.int64)] * 5 + [np.zeros(0)] * 2)
        # ghost nodes and shared boundary nodes
        k, a, b = _points(ids, g["nx"][ids] + 2 * HALO + 1, g["ny"][ids] + 2 * HALO + 1)
        nxk, nyk = g["nx"][k], g["ny"][k]
        inner = (a >= HALO) & (a <= HALO + nxk) & (b >= HALO) & (b <= HALO + nyk)
        boundary = inner & ((a == HALO) | (a == HALO + nxk) | (b == HALO) | (b == HALO + nyk))
        keep = ~inner | boundary
        k, a, b, isb = k[keep], a[keep], b[keep], boundary[keep]
        fi = g["i0"][k] + a - HALO
        fj = g["j0"][k] + b - HALO
        inside = ((per[0] | ((fi >= 0) & (fi <= shape[0])))
                  & (per[1] | ((fj >= 0) & (fj <= shape[1]))))
        k, a, b, isb, fi, fj = (x[inside] for x in (k, a, b, isb, fi, fj))
        I, J = _wrap(fi, shape[0], per[0]), _wrap(fj, shape[1], per[1])
        src = node_own[I, J]
        # a boundary node keeps its value unless an earlier patch owns it
        take = (src >= 0) & ~(isb & (src >= k))
        q = src[take]
        nd = [node(k[take], a[take], b[take])]
        ns = [node(q, _wrap(I[take] - g["i0"][q], shape[0], per[0]) + HALO,
                   _wrap(J[take] - g["j0"][q], shape[1], per[1]) + HALO)]
        # the periodic duplicate node of a patch that spans the level is its own node 0
        dup = isb & (src == k)
        sa = _wrap(I[dup] - g["i0"][k[dup]], shape[0], per[0]) + HALO
        sb = _wrap(J[dup] - g["j0"][k[dup]], shape[1], per[1]) + HALO
        moved = (sa != a[dup]) | (sb != b[dup])
        kd = k[dup][moved]
        nd.append(node(kd, a[dup][moved], b[dup][moved]))
        ns.append(node(kd, sa[moved], sb[moved]))
        maps.copy_nodes[L] = (np.concatenate(nd), np.concatenate(ns))
        miss = (src < 0) & ~isb
        if miss.any() and L == 0:
            raise ValueError("level 0 must cover the domain")
        if miss.any():
            mk, ma, mb, mfi, mfj = k[miss], a[miss], b[miss], fi[miss], fj[miss]
            # per axis: even -> coarse node fi/2 (stencil base fi/2, kind 0);
            # odd -> quadratic on nodes m-1..m+1, m = (fi-1)/2 (base m-1, kind 1)
            kx = (mfi % 2 == 1).astype(np.int64)
            ky = (mfj % 2 == 1).astype(np.int64)
            bx = np.where(kx == 1, (mfi - 1) // 2 - 1, mfi // 2)
            by = np.where(ky == 1, (mfj - 1) // 2 - 1, mfj // 2)
            q, qa, qb = _coarse_container(layouts, L - 1, bx, by, 3, hierarchy, cells=False,
                                          owner=owners[L - 1][0], boxes=boxes, periodic=per)
            maps.node_interp[L] = (node(mk, ma, mb), node(q, qa, qb), YN[q],
                                   np.ones(q.size, np.int64), kx, ky)
        else:
            maps.node_interp[L] = tuple([np.zeros(0, np.int64)] * 6)
    for L in range(1, nlev):
        maps.reflux[L - 1] = _reflux_table(layouts, owners, L, hierarchy)
        maps.average[L - 1] = _average_table(layouts, owners, L, hierarchy)
        maps.restrict[L - 1] = _restrict_table(layouts, owners, L, hierarchy)
    maps.domain_faces = _domain_face_rows(layouts, hierarchy)
    return maps


def _domain_face_rows(layouts, hierarchy) -> dict:
    """Per level: the patches on a non-periodic domain face, as rows {patch, faces
    (1 x-low, 2 x-high, 4 y-low, 8 y-high)} (`wb_amr2d_domain_faces` adds the walls)."""
    out = {}
    for L in range(hierarchy.max_level + 1):
        shape = hierarchy.level_shape(L)
        rows = []
        for k, lay in enumerate(layouts):
            if lay.level != L:
                continue
            faces = 0
            for axis in range(2):
                if hierarchy.periodic[axis]:
                    continue
                if lay.box[axis][0] == 0:
                    faces |= 1 << (2 * axis)
                if lay.box[axis][1] == shape[axis]:
                    faces |= 2 << (2 * axis)
            if faces:
                rows.append((k, faces))
        out[L] = np.array(rows, np.int64).reshape(-1, 2)
    return out


def _cat(parts, dtype=np.int64):
    if not parts:
        return np.zeros(0, dtype)
    return np.concatenate(parts).astype(dtype)


def box_array(layouts) -> np.ndarray:
    """(npatch, 4) int64: i0, i1, j0, j1 of every patch."""
    return np.array([[*lay.box[0], *lay.box[1]] for lay in layouts], np.int64).reshape(-1, 4)


def _coarse_container(layouts, level, ci, cj, width, hierarchy, cells, owner=None, boxes=None,
                      periodic=(True, True)):
    """For stencils whose lower corner is (ci, cj) and extent ``width`` per axis, a
    patch of ``level`` whose haloed range holds the whole stencil, and the stencil's
    local corner there (the smallest non-negative periodic image; on a non-periodic
    axis the stencil may reach into the halo outside the domain).  The candidates are
    the owners of the stencil's corners and centre (``owner``: the level's cell-owner
    map, as `_owner_tables` gives it); stencils none of them holds try every patch."""
    shape = hierarchy.level_shape(level)
    per = tuple(bool(p) for p in periodic)

    def local(v, lo, n, p):
        return (v - lo + HALO) % n if p else v - lo + HALO

    def probe(v, n, p):  # a cell of the level to ask for its owner
        return v % n if p else np.clip(v, 0, n - 1)
    ci = np.asarray(ci, np.int64)
    cj = np.asarray(cj, np.int64)
    n = len(ci)
    q = np.full(n, -1, np.int64)
    qa = np.zeros(n, np.int64)
    qb = np.zeros(n, np.int64)
    ext = 0 if cells else 1
    box = box_array(layouts) if boxes is None else boxes
    if owner is None:
        owner = _owner_tables(layouts, level, shape)[0]
    half = width // 2
    for di, dj in ((0, 0), (width - 1, width - 1), (0, width - 1), (width - 1, 0), (half, half)):
        todo = np.flatnonzero(q < 0)
        if todo.size == 0:
            break
        k = owner[probe(ci[todo] + di, shape[0], per[0]), probe(cj[todo] + dj, shape[1], per[1])]
        have = k >= 0
        todo, k = todo[have], k[have]
        i0, i1, j0, j1 = box[k, 0], box[k, 1], box[k, 2], box[k, 3]
        la = local(ci[todo], i0, shape[0], per[0])
        lb = local(cj[todo], j0, shape[1], per[1])
        fit = ((la >= 0) & (lb >= 0) & (la + width <= i1 - i0 + 2 * HALO + ext)
               & (lb + width <= j1 - j0 + 2 * HALO + ext))
        q[todo[fit]], qa[todo[fit]], qb[todo[fit]] = k[fit], la[fit], lb[fit]
    rest = np.flatnonzero(q < 0)
    if rest.size:
        for k, lay in enumerate(layouts):
            if lay.level != level:
                continue
            (i0, i1), (j0, j1) = lay.box
            hi_a = (i1 - i0) + 2 * HALO + ext
            hi_b = (j1 - j0) + 2 * HALO + ext
            la = local(ci[rest], i0, shape[0], per[0])
            lb = local(cj[rest], j0, shape[1], per[1])
            ok = (q[rest] < 0) & (la >= 0) & (lb >= 0) & (la + width <= hi_a) & (lb + width <= hi_b)
            q[rest[ok]], qa[rest[ok]], qb[rest[ok]] = k, la[ok], lb[ok]
    if (q < 0).any():
        raise ValueError("a fine patch is not properly nested in the level below")
    return q, qa, qb


def _reflux_table(layouts, owners, L, hierarchy):
    """Uncovered coarse cells of level L-1 next to level-L patches: per cell up to four
    faces (x-lo, x-hi, y-lo, y-hi), each with the coarse face and its two fine faces."""
    shape_c = hierarchy.level_shape(L - 1)
    shape_f = hierarchy.level_shape(L)
    fine_cells, _ = owners[L]
    coarse_cells, _ = owners[L - 1]
    covered = np.zeros(shape_c, bool)
    cov = fine_cells >= 0
    covered |= cov[0::2, 0::2] | cov[1::2, 0::2] | cov[0::2, 1::2] | cov[1::2, 1::2]
    entries = {}
    for axis in (0, 1):
        for side in (0, 1):  # side 0: the fine patch is on the high side of the coarse cell
            # coarse cells whose neighbour across the face along `axis` is covered
            nb = np.roll(covered, -1 if side == 0 else 1, axis=axis)
            if not hierarchy.periodic[axis]:  # no neighbour across the domain face
                edge = [slice(None), slice(None)]
                edge[axis] = -1 if side == 0 else 0
                nb[tuple(edge)] = False
            cells = np.argwhere(~covered & nb)
            for ci, cj in cells:
                key = (ci, cj)
                entries.setdefault(key, []).append((axis, side))
    rows = []
    for (ci, cj), faces in entries.items():
        q = coarse_cells[ci, cj]
        ql = layouts[q]
        for axis, side in faces:
            # fine cells just across the face, and the fine face index
            if axis == 0:
                fi = (2 * ci + 2) % shape_f[0] if side == 0 else (2 * ci - 1) % shape_f[0]
                fjs = [(2 * cj) % shape_f[1], (2 * cj + 1) % shape_f[1]]
                fis = [fi, fi]
            else:
                fj = (2 * cj + 2) % shape_f[1] if side == 0 else (2 * cj - 1) % shape_f[1]
                fis = [(2 * ci) % shape_f[0], (2 * ci + 1) % shape_f[0]]
                fjs = [fj, fj]
            fine = []
            for fi, fj in zip(fis, fjs):
                p = fine_cells[fi, fj]
                pl = layouts[p]
                # the fine face lies on the fine cell's low side (side 0) or high side (1)
                fa = fi - pl.box[0][0] + HALO
                fb = fj - pl.box[1][0] + HALO
                if axis == 0 and side == 1:
                    fa += 1
                if axis == 1 and side == 1:
                    fb += 1
                fine.append((p, fa, fb))
            ca = ci - ql.box[0][0] + HALO
            cb = cj - ql.box[1][0] + HALO
            face_a = ca + (1 if (axis == 0 and side == 0) else 0)
            face_b = cb + (1 if (axis == 1 and side == 0) else 0)
            rows.append((q, ca, cb, axis, side, face_a, face_b, *fine[0], *fine[1]))
    if not rows:
        return np.zeros((0, 13), np.int64)
    rows.sort(key=lambda r: (r[0], r[1], r[2], r[3], r[4]))
    return np.array(rows, np.int64)


def _average_table(layouts, owners, L, hierarchy):
    """Covered coarse cells of level L-1 with the offsets of their four fine children
    (the level's patches at once)."""
    coarse_cells, _ = owners[L - 1]
    g = _patch_arrays(layouts)
    ids = np.flatnonzero(g["level"] == L)
    if len(ids) == 0:
        return np.zeros((0, 5), np.int64)
    if ((g["i0"][ids] | g["nx"][ids] | g["j0"][ids] | g["ny"][ids]) % 2).any():
        raise ValueError("fine patches are aligned to the coarse cells")
    k, a, b = _points(ids, g["nx"][ids] // 2, g["ny"][ids] // 2)
    ci, cj = g["i0"][k] // 2 + a, g["j0"][k] // 2 + b
    q = coarse_cells[ci, cj]
    if (q < 0).any():
        raise ValueError("a fine patch lies outside the level below")
    Y = g["ny"] + 2 * HALO
    plane = (g["nx"] + 2 * HALO) * Y
    return np.stack([
        g["off"][q] + (ci - g["i0"][q] + HALO) * Y[q] + (cj - g["j0"][q] + HALO),
        plane[q],
        g["off"][k] + (2 * ci - g["i0"][k] + HALO) * Y[k] + (2 * cj - g["j0"][k] + HALO),
        Y[k], plane[k]], axis=1).astype(np.int64)


def _restrict_table(layouts, owners, L, hierarchy):
    """Every coarse copy of a node that coincides with a fine node, and the fine
    node's canonical owner."""
    shape_c = hierarchy.level_shape(L - 1)
    shape_f = hierarchy.level_shape(L)
    _, fine_nodes = owners[L]
    rows = []
    for q, ql in enumerate(layouts):
        if ql.level != L - 1:
            continue
        (i0, i1), (j0, j1) = ql.box
        ci = np.arange(i0, i1 + 1)
        cj = np.arange(j0, j1 + 1)
        CI, CJ = np.meshgrid(ci, cj, indexing="ij")
        per = hierarchy.periodic
        FI = _wrap(2 * CI, shape_f[0], per[0])
        FJ = _wrap(2 * CJ, shape_f[1], per[1])
        own = fine_nodes[FI, FJ]
        sel = own >= 0
        for p in np.unique(own[sel]):
            s = sel & (own == p)
            pl = layouts[p]
            fa = _wrap(FI[s] - pl.box[0][0], shape_f[0], per[0]) + HALO
            fb = _wrap(FJ[s] - pl.box[1][0], shape_f[1], per[1]) + HALO
            ca = CI[s] - i0 + HALO
            cb = CJ[s] - j0 + HALO
            rows.append(np.stack([ql.node(ca, cb), pl.node(fa, fb)], axis=1))
    del shape_c
    if not rows:
        return np.zeros((0, 2), np.int64)
    return np.concatenate(rows).astype(np.int64)


def random_patches(hierarchy, level: int, region, sizes=(16, 128), seed: int = 0) -> None:
    """Tile ``region`` ((i0, i1), (j0, j1)) of ``level`` with patches whose sides are
    drawn uniformly in ``sizes`` (inclusive): column strips of random widths, each cut
    into random heights.  Boxes are aligned to even indices (coarse cells)."""
    rng = np.random.default_rng(seed)
    (i0, i1), (j0, j1) = region
    lo, hi = sizes

    def cuts(a, b):
        out, x = [], a
        while x < b:
            w = int(rng.integers(lo, hi + 1))
            w -= w % 2
            if b - x - w < lo:
                w = b - x
            out.append((x, min(x + w, b)))
            x += w
        return out

    for xa, xb in cuts(i0, i1):
        for ya, yb in cuts(j0, j1):
            hierarchy.add_patch(level, ((xa, xb), (ya, yb)))


def pack_state(hierarchy, layouts, total, gamma, mean_field, state) -> np.ndarray:
    """The flat FP64 buffer of every patch's haloed state (``set_state`` on each patch,
    ``mhd.py:683-737``), patch after patch as in ``layouts``."""
    from .initial import host_state_2d

    buf = np.zeros(total)
    for lay, patch in zip(layouts, hierarchy.patches):
        cons, az = host_state_2d(hierarchy, patch, gamma, mean_field, **state)
        buf[lay.offset : lay.offset + 6 * lay.plane] = cons.ravel()
        buf[lay.offset + 6 * lay.plane : lay.offset + lay.size] = az.ravel()
    return buf
