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
  (even fine index) or the 4-point midpoint rule (-1, 9, 9, -1)/16 (odd index);
* ``reflux``: per uncovered coarse cell next to a finer patch, up to four faces
  whose coarse flux is replaced by the mean of the two fine fluxes;
* ``average``: covered coarse cells, the mean of their four fine children;
* ``restrict``: coarse nodes that coincide with fine nodes take the fine value.

Domains are periodic on both axes (the refined set-ups this path runs).
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


def layouts_of(hierarchy) -> tuple[list[PatchLayout], int]:
    out, offset = [], 0
    for p in hierarchy.patches:
        lay = PatchLayout(p.level, tuple(tuple(b) for b in p.box), offset)
        out.append(lay)
        offset += lay.size
    return out, offset


def _owner_tables(layouts, level, shape):
    """Per level: the index of the first patch owning each cell (-1 if none) and node."""
    nx, ny = shape
    cells = np.full((nx, ny), -1, np.int64)
    nodes = np.full((nx, ny), -1, np.int64)  # periodic: node n is node 0
    for k, lay in enumerate(layouts):
        if lay.level != level:
            continue
        (i0, i1), (j0, j1) = lay.box
        view = cells[i0:i1, j0:j1]
        view[view < 0] = k
        ii = np.arange(i0, i1 + 1) % nx
        jj = np.arange(j0, j1 + 1) % ny
        sub = nodes[np.ix_(ii, jj)]
        sub[sub < 0] = k
        nodes[np.ix_(ii, jj)] = sub
    return cells, nodes


def build(hierarchy) -> Maps:
    """All gather tables for a hierarchy whose axes are periodic."""
    if not all(hierarchy.periodic):
        raise NotImplementedError("refined 2D hierarchies run on periodic domains")
    layouts, total = layouts_of(hierarchy)
    maps = Maps(layouts=layouts, total=total)
    nlev = hierarchy.max_level + 1
    owners = [_owner_tables(layouts, L, hierarchy.level_shape(L)) for L in range(nlev)]
    for L in range(nlev):
        shape = hierarchy.level_shape(L)
        cell_own, node_own = owners[L]
        cd, cs, cpd, cps = [], [], [], []
        nd, ns = [], []
        id_, ipd, ic00, icy, ipc, iwx, iwy = [], [], [], [], [], [], []
        nid, nbase, nsx, nsy, nkx, nky = [], [], [], [], [], []
        for k, lay in enumerate(layouts):
            if lay.level != L:
                continue
            (i0, i1), (j0, j1) = lay.box
            # ghost cells
            a, b = np.meshgrid(np.arange(lay.X), np.arange(lay.Y), indexing="ij")
            ghost = ~((a >= HALO) & (a < HALO + lay.nx) & (b >= HALO) & (b < HALO + lay.ny))
            a, b = a[ghost], b[ghost]
            I = (i0 + a - HALO) % shape[0]
            J = (j0 + b - HALO) % shape[1]
            src = cell_own[I, J]
            have = src >= 0
            for q in np.unique(src[have]):
                sel = have & (src == q)
                ql = layouts[q]
                cd.append(lay.cell(a[sel], b[sel]))
                cs.append(ql.cell(I[sel] - ql.box[0][0] + HALO, J[sel] - ql.box[1][0] + HALO))
                cpd.append(np.full(sel.sum(), lay.plane, np.int64))
                cps.append(np.full(sel.sum(), ql.plane, np.int64))
            miss = ~have
            if miss.any():
                if L == 0:
                    raise ValueError("level 0 must cover the domain")
                ia, ib = a[miss], b[miss]
                fi = i0 + ia - HALO  # unwrapped fine cell coordinates
                fj = j0 + ib - HALO
                ci0 = np.floor_divide(fi - 1, 2)  # lower coarse cell of the stencil
                cj0 = np.floor_divide(fj - 1, 2)
                wx1 = np.where(fi % 2 == 0, 0.75, 0.25)
                wy1 = np.where(fj % 2 == 0, 0.75, 0.25)
                q, qa, qb = _coarse_container(layouts, L - 1, ci0, cj0, 2, hierarchy, cells=True)
                for qq in np.unique(q):
                    sel = q == qq
                    ql = layouts[qq]
                    id_.append(lay.cell(ia[sel], ib[sel]))
                    ipd.append(np.full(sel.sum(), lay.plane, np.int64))
                    ic00.append(ql.cell(qa[sel], qb[sel]))
                    icy.append(np.full(sel.sum(), ql.Y, np.int64))
                    ipc.append(np.full(sel.sum(), ql.plane, np.int64))
                    iwx.append(wx1[sel])
                    iwy.append(wy1[sel])
            # ghost nodes and shared boundary nodes
            a, b = np.meshgrid(np.arange(lay.X + 1), np.arange(lay.Y + 1), indexing="ij")
            inner = (a >= HALO) & (a <= HALO + lay.nx) & (b >= HALO) & (b <= HALO + lay.ny)
            boundary = inner & ((a == HALO) | (a == HALO + lay.nx) | (b == HALO) | (b == HALO + lay.ny))
            a_all, b_all = a[~inner | boundary], b[~inner | boundary]
            isb = boundary[~inner | boundary]
            I = (i0 + a_all - HALO) % shape[0]
            J = (j0 + b_all - HALO) % shape[1]
            src = node_own[I, J]
            # a boundary node keeps its value unless an earlier patch owns it
            take = (src >= 0) & ~(isb & (src >= k))
            # the periodic duplicate node of a patch that spans the level is its own node 0
            for q in np.unique(src[take]):
                sel = take & (src == q)
                ql = layouts[q]
                qa = (I[sel] - ql.box[0][0]) % shape[0] + HALO
                qb = (J[sel] - ql.box[1][0]) % shape[1] + HALO
                nd.append(lay.node(a_all[sel], b_all[sel]))
                ns.append(ql.node(qa, qb))
            dup = isb & (src == k)
            if dup.any():
                # this patch owns the node; if it spans the level it may own it twice
                # (node n and node 0 of a periodic axis): keep index HALO as the source
                da, db = a_all[dup], b_all[dup]
                sa = (I[dup] - i0) % shape[0] + HALO
                sb = (J[dup] - j0) % shape[1] + HALO
                moved = (sa != da) | (sb != db)
                if moved.any():
                    nd.append(lay.node(da[moved], db[moved]))
                    ns.append(lay.node(sa[moved], sb[moved]))
            miss = (src < 0) & ~isb
            if miss.any():
                if L == 0:
                    raise ValueError("level 0 must cover the domain")
                ia, ib = a_all[miss], b_all[miss]
                fi = i0 + ia - HALO
                fj = j0 + ib - HALO
                # per axis: even -> coarse node fi/2 (stencil base fi/2, kind 0);
                # odd -> 4-point rule on nodes m-1..m+2, m = (fi-1)/2 (base m-1, kind 1)
                kx = (fi % 2 == 1).astype(np.int64)
                ky = (fj % 2 == 1).astype(np.int64)
                bx = np.where(kx == 1, (fi - 1) // 2 - 1, fi // 2)
                by = np.where(ky == 1, (fj - 1) // 2 - 1, fj // 2)
                q, qa, qb = _coarse_container(layouts, L - 1, bx, by, 4, hierarchy, cells=False)
                for qq in np.unique(q):
                    sel = q == qq
                    ql = layouts[qq]
                    nid.append(lay.node(ia[sel], ib[sel]))
                    nbase.append(ql.node(qa[sel], qb[sel]))
                    nsx.append(np.full(sel.sum(), ql.Y + 1, np.int64))
                    nsy.append(np.full(sel.sum(), 1, np.int64))
                    nkx.append(kx[sel])
                    nky.append(ky[sel])
        cat = _cat
        maps.copy_cells[L] = (cat(cd), cat(cs), cat(cpd), cat(cps))
        maps.copy_nodes[L] = (cat(nd), cat(ns))
        maps.cell_interp[L] = (cat(id_), cat(ipd), cat(ic00), cat(icy), cat(ipc),
                               cat(iwx, np.float64), cat(iwy, np.float64))
        maps.node_interp[L] = (cat(nid), cat(nbase), cat(nsx), cat(nsy), cat(nkx), cat(nky))
    for L in range(1, nlev):
        maps.reflux[L - 1] = _reflux_table(layouts, owners, L, hierarchy)
        maps.average[L - 1] = _average_table(layouts, owners, L, hierarchy)
        maps.restrict[L - 1] = _restrict_table(layouts, owners, L, hierarchy)
    return maps


def _cat(parts, dtype=np.int64):
    if not parts:
        return np.zeros(0, dtype)
    return np.concatenate(parts).astype(dtype)


def _coarse_container(layouts, level, ci, cj, width, hierarchy, cells):
    """For stencils whose lower corner is (ci, cj) and extent ``width`` per axis, a
    patch of ``level`` whose haloed range holds the whole stencil, and the stencil's
    local corner there.  Periodic images are tried."""
    shape = hierarchy.level_shape(level)
    n = len(ci)
    q = np.full(n, -1, np.int64)
    qa = np.zeros(n, np.int64)
    qb = np.zeros(n, np.int64)
    for k, lay in enumerate(layouts):
        if lay.level != level:
            continue
        (i0, i1), (j0, j1) = lay.box
        hi_a = (i1 - i0) + 2 * HALO - (0 if cells else -1)  # haloed extent
        hi_b = (j1 - j0) + 2 * HALO - (0 if cells else -1)
        for sx in (-shape[0], 0, shape[0]):
            for sy in (-shape[1], 0, shape[1]):
                la = ci + sx - i0 + HALO
                lb = cj + sy - j0 + HALO
                ok = (q < 0) & (la >= 0) & (la + width <= hi_a) & (lb >= 0) & (lb + width <= hi_b)
                q[ok], qa[ok], qb[ok] = k, la[ok], lb[ok]
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
    """Covered coarse cells of level L-1 with the offsets of their four fine children."""
    fine_cells, _ = owners[L]
    coarse_cells, _ = owners[L - 1]
    rows = []
    for p, pl in enumerate(layouts):
        if pl.level != L:
            continue
        (i0, i1), (j0, j1) = pl.box
        for ci in range(i0 // 2, (i1 + 1) // 2):
            for cj in range(j0 // 2, (j1 + 1) // 2):
                if not (i0 <= 2 * ci and 2 * ci + 1 < i1 and j0 <= 2 * cj and 2 * cj + 1 < j1):
                    raise ValueError("fine patches are aligned to the coarse cells")
                q = coarse_cells[ci, cj]
                if q < 0:
                    raise ValueError("a fine patch lies outside the level below")
                ql = layouts[q]
                rows.append((ql.cell(ci - ql.box[0][0] + HALO, cj - ql.box[1][0] + HALO), ql.plane,
                             pl.cell(2 * ci - i0 + HALO, 2 * cj - j0 + HALO), pl.Y, pl.plane))
    return np.array(rows, np.int64).reshape(-1, 5)


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
        FI = (2 * CI) % shape_f[0]
        FJ = (2 * CJ) % shape_f[1]
        own = fine_nodes[FI, FJ]
        sel = own >= 0
        for p in np.unique(own[sel]):
            s = sel & (own == p)
            pl = layouts[p]
            fa = (FI[s] - pl.box[0][0]) % shape_f[0] + HALO
            fb = (FJ[s] - pl.box[1][0]) % shape_f[1] + HALO
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
