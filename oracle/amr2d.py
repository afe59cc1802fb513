"""The refined 2D step on a patch hierarchy, in NumPy FP64 (oracle only).

Follows ``MHD2D.step`` (``mhd.py:965-994``) on every patch of a hierarchy:

1. stage 1 (``_advance(dt, 0.5 dt)``, ``mhd.py:889-927``): the halos of every level,
   the fluxes of every patch and the update ``u - dt div F`` of every patch cell,
   covered coarse cells included, and ``A_z -= dt E_z``; the flux register records
   ``0.5 dt`` times the fluxes of the coarse-fine faces;
2. ``hierarchy.restrict(A_z)`` (``mhd.py:981``): coarse nodes on fine nodes take
   the fine value;
3. stage 2 on that state, recorded the same way;
4. the Heun average ``(U2 + U0) * 0.5`` (``mhd.py:983-988``);
5. ``register.reflux`` (``mhd.py:989-990``): each uncovered coarse cell next to a
   finer patch trades the coarse flux of such a face, from both stages, for the mean
   of the two fine fluxes over it; then the covered coarse cells take the mean of
   their four children (Berger and Colella's synchronisation; the reference's test
   states "the fine patch and the coarse level carry the same solution where both
   exist", ``tests/unit/test_mhd.py:188``);
6. ``hierarchy.restrict(A_z)`` (``mhd.py:991``).

The geometry is this module's own, computed from the patch boxes alone: which
patch owns a cell or node of a level, which ghosts copy from a patch of their level
and which interpolate from the level below, which coarse faces are coarse-fine
faces.  Arithmetic choices where patchwork's (absent, ``uv.lock:3139``) cannot be
seen -- all parity-unpinned beyond the invariants and known answers the tests check:

* ghost cells missing on their level: bilinear in the 2x2 nearest coarse cells,
  weights 1/4 and 3/4 per axis, ``wy0 (wx0 c00 + wx1 c10) + wy1 (wx0 c01 + wx1 c11)``;
* ghost nodes of ``A_z`` missing on their level: per axis the coarse node itself
  (even fine index) or the midpoint rule `NODE_RULE` over coarse nodes around it
  (odd index), rows then columns, each a running sum from 0 in stencil order;
* the register: per coarse cell the faces in the order x-high, x-low, y-high,
  y-low, each adding ``dt/h (F_c - mean)`` (high face) or ``dt/h (mean - F_c)`` (low
  face), ``mean = 0.5 (F_f0 + F_f1)``; stage 1 keeps ``R = 0.5 d_1``, the reflux
  adds ``R + 0.5 d_2`` to the averaged cell;
* average-down ``((c00 + c10) + (c01 + c11)) * 0.25``;
* on a non-periodic axis each level's patches take the reference's outflow or wall
  rules (``mhd.py:744-777``) right after that level's exchange, so the level above
  interpolates from current ghost values.
"""

from __future__ import annotations

import numpy as np

from . import boundaries, kernels

HALO = 2
XR = (0, 1, 2, 3, 4, 6)
YR = (0, 2, 1, 3, 4, 6)
#: Rows zeroed at a wall face (``mhd.py:136``).
WALL_ROWS = [0, 2, 3, 4, 6]

#: The rule for an A_z node midway between coarse nodes m and m+1 (odd fine index):
#: (first coarse node relative to m, weights).
NODE_RULES = {
    "cubic": (-1, (-0.0625, 0.5625, 0.5625, -0.0625)),
    "quadratic-lower": (-1, (-0.125, 0.75, 0.375)),
    "quadratic-upper": (0, (0.375, 0.75, -0.125)),
    "bilinear": (0, (0.5, 0.5)),
}
NODE_RULE = "quadratic-lower"


class Geometry2D:
    """Owners of the cells and nodes of every level, from the boxes alone."""

    def __init__(self, patches, base_shape, periodic=(True, True)):
        self.patches = [(int(L), tuple(tuple(int(v) for v in ax) for ax in box))
                        for L, box in patches]
        self.base = tuple(int(n) for n in base_shape)
        self.periodic = tuple(bool(p) for p in periodic)
        self.nlev = max(L for L, _ in self.patches) + 1
        self.cell_owner, self.node_owner = [], []
        for L in range(self.nlev):
            n, m = self.shape(L)
            cells = np.full((n, m), -1, np.int64)
            # node (I, J) for I in 0..n (n is node 0 on a periodic axis)
            nodes = np.full((n + 1, m + 1), -1, np.int64)
            for k, (lev, ((i0, i1), (j0, j1))) in enumerate(self.patches):
                if lev != L:
                    continue
                for i in range(i0, i1):
                    for j in range(j0, j1):
                        if cells[i, j] < 0:
                            cells[i, j] = k
                for i in range(i0, i1 + 1):
                    for j in range(j0, j1 + 1):
                        I, J = self.wrap_node(L, i, j)
                        if nodes[I, J] < 0:
                            nodes[I, J] = k
            self.cell_owner.append(cells)
            self.node_owner.append(nodes)
        if (self.cell_owner[0] < 0).any():
            raise ValueError("level 0 must cover the domain")

    def shape(self, L):
        return (self.base[0] << L, self.base[1] << L)

    def box(self, k):
        return self.patches[k][1]

    def level(self, k):
        return self.patches[k][0]

    def inside(self, L, i, j):
        """Whether cell (i, j) of level L lies in the domain (periodic axes wrap)."""
        n, m = self.shape(L)
        return ((self.periodic[0] or 0 <= i < n) and (self.periodic[1] or 0 <= j < m))

    def wrap_cell(self, L, i, j):
        n, m = self.shape(L)
        return (i % n if self.periodic[0] else i), (j % m if self.periodic[1] else j)

    def wrap_node(self, L, i, j):
        n, m = self.shape(L)
        return (i % n if self.periodic[0] else i), (j % m if self.periodic[1] else j)

    def cell_of(self, L, i, j):
        """(patch, local haloed a, b) of the patch owning cell (i, j) of level L, or None."""
        I, J = self.wrap_cell(L, i, j)
        n, m = self.shape(L)
        if not (0 <= I < n and 0 <= J < m):
            return None
        k = self.cell_owner[L][I, J]
        if k < 0:
            return None
        (i0, _), (j0, _) = self.box(k)
        return int(k), I - i0 + HALO, J - j0 + HALO

    def node_of(self, L, i, j):
        """(patch, local haloed a, b) of the first patch holding node (i, j), or None."""
        I, J = self.wrap_node(L, i, j)
        n, m = self.shape(L)
        if not (0 <= I <= n and 0 <= J <= m):
            return None
        k = self.node_owner[L][I, J]
        if k < 0:
            return None
        (i0, _), (j0, _) = self.box(k)
        a = (I - i0) % n + HALO if self.periodic[0] else I - i0 + HALO
        b = (J - j0) % m + HALO if self.periodic[1] else J - j0 + HALO
        return int(k), a, b

    def covered(self, L):
        """Cells of level L under a patch of level L + 1."""
        if L + 1 >= self.nlev:
            return np.zeros(self.shape(L), bool)
        f = self.cell_owner[L + 1] >= 0
        return f[0::2, 0::2] | f[1::2, 0::2] | f[0::2, 1::2] | f[1::2, 1::2]


def layout(patches):
    """Offsets of the flat state buffer: per patch 6 haloed cell planes, then A_z."""
    offs, total = [], 0
    for _, ((i0, i1), (j0, j1)) in patches:
        X, Y = i1 - i0 + 2 * HALO, j1 - j0 + 2 * HALO
        offs.append(total)
        total += 6 * X * Y + (X + 1) * (Y + 1)
    return offs, total


def unpack(buf, patches):
    """Per patch ``(cells (6, X, Y), az (X+1, Y+1))`` copies of a flat buffer."""
    offs, _ = layout(patches)
    out = []
    for off, (_, ((i0, i1), (j0, j1))) in zip(offs, patches):
        X, Y = i1 - i0 + 2 * HALO, j1 - j0 + 2 * HALO
        cells = np.array(buf[off : off + 6 * X * Y]).reshape(6, X, Y)
        az = np.array(buf[off + 6 * X * Y : off + 6 * X * Y + (X + 1) * (Y + 1)]).reshape(X + 1, Y + 1)
        out.append((cells, az))
    return out


def pack(state, patches):
    offs, total = layout(patches)
    buf = np.zeros(total)
    for off, (cells, az) in zip(offs, state):
        buf[off : off + cells.size] = cells.ravel()
        buf[off + cells.size : off + cells.size + az.size] = az.ravel()
    return buf


def _node_value(get, L1, fi, fj):
    """A_z at fine node (fi, fj) of level L1 + 1 from the nodes of level L1 (``get``)."""
    def axis_rule(f):
        if f % 2 == 0:
            return f // 2, (1.0,), False
        start, w = NODE_RULES[NODE_RULE]
        return (f - 1) // 2 + start, w, True

    bx, wx, kx = axis_rule(fi)
    by, wy, ky = axis_rule(fj)
    total = 0.0
    for j, wj in enumerate(wy):
        row = 0.0
        for i, wi in enumerate(wx):
            row = row + (wi if kx else 1.0) * get(L1, bx + i, by + j)
        total = total + (wj if ky else 1.0) * row
    return total


class OracleAMR2D:
    """Heun steps of a 2D patch hierarchy, every patch in its own arrays."""

    def __init__(self, patches, base_shape, bounds, buf, *, periodic=(True, True),
                 walls=(False, False), gamma=5.0 / 3.0, mean_field=(0.0, 0.0),
                 limiter="minmod", cfl=0.4, resistivity=0.0, conduction=(0.0, 0.0)):
        self.geo = Geometry2D(patches, base_shape, periodic)
        self.patches = self.geo.patches
        self.bounds = tuple(tuple(float(v) for v in b) for b in bounds)
        self.walls = tuple(bool(w) for w in walls)
        self.state = unpack(np.asarray(buf, np.float64), self.patches)
        self.gamma = float(gamma)
        self.b0 = tuple(float(b) for b in mean_field)
        self.code = kernels.LIMITER_CODES.index(limiter)
        self.cfl = float(cfl)
        self.resistivity = float(resistivity)
        self.conduction = (float(conduction[0]), float(conduction[1]))
        self.nlev = self.geo.nlev
        self.time = 0.0
        self.steps = 0
        self._plan()

    @classmethod
    def from_hierarchy(cls, hierarchy, buf, **options):
        patches = [(p.level, p.box) for p in hierarchy.patches]
        return cls(patches, hierarchy.base_shape, hierarchy.bounds, buf,
                   periodic=hierarchy.periodic, **options)

    # -- geometry --------------------------------------------------------------------
    def spacing(self, L):
        n, m = self.geo.shape(L)
        (x0, x1), (y0, y1) = self.bounds
        return (x1 - x0) / n, (y1 - y0) / m

    def _plan(self):
        """Per patch: the ghost cells and nodes (copied or interpolated), the
        boundary nodes an earlier patch owns, and the coarse-fine faces."""
        geo = self.geo
        self.cell_plan, self.node_plan = [], []
        for k, (L, ((i0, i1), (j0, j1))) in enumerate(self.patches):
            nx, ny = i1 - i0, j1 - j0
            X, Y = nx + 2 * HALO, ny + 2 * HALO
            cells = []
            for a in range(X):
                for b in range(Y):
                    if HALO <= a < HALO + nx and HALO <= b < HALO + ny:
                        continue
                    fi, fj = i0 + a - HALO, j0 + b - HALO
                    if not geo.inside(L, fi, fj):
                        continue  # a domain-face ghost: the boundary rules fill it
                    src = geo.cell_of(L, fi, fj)
                    if src is not None:
                        cells.append((a, b, "copy", src))
                        continue
                    if L == 0:
                        raise ValueError("level 0 must cover the domain")
                    cells.append((a, b, "interp", (fi, fj)))
            nodes = []
            for a in range(X + 1):
                for b in range(Y + 1):
                    inner = HALO <= a <= HALO + nx and HALO <= b <= HALO + ny
                    edge = inner and (a in (HALO, HALO + nx) or b in (HALO, HALO + ny))
                    if inner and not edge:
                        continue
                    fi, fj = i0 + a - HALO, j0 + b - HALO
                    n, m = geo.shape(L)
                    if not ((geo.periodic[0] or 0 <= fi <= n) and (geo.periodic[1] or 0 <= fj <= m)):
                        continue
                    src = geo.node_of(L, fi, fj)
                    if edge:
                        if src is not None and src != (k, a, b):
                            nodes.append((a, b, "copy", src))
                        continue
                    if src is not None:
                        nodes.append((a, b, "copy", src))
                        continue
                    if L == 0:
                        raise ValueError("level 0 must cover the domain")
                    nodes.append((a, b, "interp", (fi, fj)))
            self.cell_plan.append(cells)
            self.node_plan.append(nodes)
        # coarse-fine faces of every level pair: per uncovered coarse cell next to a
        # finer patch, its faces in the order x-high, x-low, y-high, y-low
        self.reflux_cells = []
        for L in range(self.nlev - 1):
            cov = geo.covered(L)
            n, m = geo.shape(L)
            for k, (lev, ((i0, i1), (j0, j1))) in enumerate(self.patches):
                if lev != L:
                    continue
                for i in range(i0, i1):
                    for j in range(j0, j1):
                        if cov[i, j]:
                            continue
                        faces = []
                        for axis, side, di, dj in ((0, 0, 1, 0), (0, 1, -1, 0),
                                                   (1, 0, 0, 1), (1, 1, 0, -1)):
                            ni, nj = i + di, j + dj
                            if not geo.inside(L, ni, nj):
                                continue
                            ni, nj = geo.wrap_cell(L, ni, nj)
                            if cov[ni, nj]:
                                faces.append(self._fine_faces(L, k, i, j, axis, side))
                        if faces:
                            self.reflux_cells.append((k, i - i0 + HALO, j - j0 + HALO, L, faces))

    def _fine_faces(self, L, k, i, j, axis, side):
        """The coarse face of cell (i, j) of patch k and the two fine faces over it,
        each as (patch, axis, local face index a, b): an x-face a is the low face of
        local cell a."""
        geo = self.geo
        (i0, _), (j0, _) = geo.box(k)
        ca, cb = i - i0 + HALO, j - j0 + HALO
        coarse = (k, ca + (1 if axis == 0 and side == 0 else 0),
                  cb + (1 if axis == 1 and side == 0 else 0))
        fine = []
        for t in (0, 1):
            if axis == 0:
                fi = 2 * i + 2 if side == 0 else 2 * i - 1
                fj = 2 * j + t
            else:
                fi = 2 * i + t
                fj = 2 * j + 2 if side == 0 else 2 * j - 1
            p, fa, fb = geo.cell_of(L + 1, fi, fj)
            if side == 1:  # the fine cell's high face
                fa, fb = (fa + 1, fb) if axis == 0 else (fa, fb + 1)
            fine.append((p, fa, fb))
        return axis, side, coarse, fine

    # -- halos -----------------------------------------------------------------------
    def _cell_value(self, state, L, i, j):
        """The 6 conserved values of cell (i, j) of level L from its owner (a ghost of a
        patch touching a non-periodic domain face outside the domain)."""
        src = self.geo.cell_of(L, i, j)
        if src is None:
            src = self._boundary_holder(L, i, j, cells=True)
        k, a, b = src
        return state[k][0][:, a, b]

    def _node_get(self, state):
        def get(L, i, j):
            src = self.geo.node_of(L, i, j)
            if src is None:
                src = self._boundary_holder(L, i, j, cells=False)
            k, a, b = src
            return state[k][1][a, b]
        return get

    def _boundary_holder(self, L, i, j, cells):
        """A patch of level L whose halo holds the point (i, j) outside the domain on a
        non-periodic axis (its value set there by the boundary rules)."""
        n, m = self.geo.shape(L)
        top = (n, m) if cells else (n + 1, m + 1)
        outside = any(not self.geo.periodic[a] and not 0 <= v < top[a]
                      for a, v in enumerate((i, j)))
        if outside:
            for k, (lev, ((i0, i1), (j0, j1))) in enumerate(self.patches):
                if lev != L:
                    continue
                a, b = i - i0 + HALO, j - j0 + HALO
                ext = 0 if cells else 1
                if 0 <= a < i1 - i0 + 2 * HALO + ext and 0 <= b < j1 - j0 + 2 * HALO + ext:
                    return k, a, b
        raise ValueError(f"a fine patch is not properly nested in level {L}")

    def fill(self, state):
        """Halos of every level, coarse to fine (``mhd.py:744-777``)."""
        for L in range(self.nlev):
            self._fill_level(state, L)

    def _fill_level(self, state, L):
        get = self._node_get(state)
        ks = [k for k, (lev, _) in enumerate(self.patches) if lev == L]
        for k in ks:  # same-level copies first (they read interiors only)
            cells, az = state[k]
            for a, b, kind, src in self.cell_plan[k]:
                if kind == "copy":
                    q, qa, qb = src
                    cells[:, a, b] = state[q][0][:, qa, qb]
            for a, b, kind, src in self.node_plan[k]:
                if kind == "copy":
                    q, qa, qb = src
                    az[a, b] = state[q][1][qa, qb]
        for k in ks:  # then the interpolation from the level below
            cells, az = state[k]
            for a, b, kind, src in self.cell_plan[k]:
                if kind != "interp":
                    continue
                fi, fj = src
                ci, cj = (fi - 1) // 2, (fj - 1) // 2
                wx1 = 0.75 if fi % 2 == 0 else 0.25
                wy1 = 0.75 if fj % 2 == 0 else 0.25
                wx0, wy0 = 1.0 - wx1, 1.0 - wy1
                c00 = self._cell_value(state, L - 1, ci, cj)
                c10 = self._cell_value(state, L - 1, ci + 1, cj)
                c01 = self._cell_value(state, L - 1, ci, cj + 1)
                c11 = self._cell_value(state, L - 1, ci + 1, cj + 1)
                cells[:, a, b] = wy0 * (wx0 * c00 + wx1 * c10) + wy1 * (wx0 * c01 + wx1 * c11)
            for a, b, kind, src in self.node_plan[k]:
                if kind == "interp":
                    az[a, b] = _node_value(get, L - 1, *src)
        for k in ks:
            self._domain_faces(state, k)

    def _domain_faces(self, state, k):
        """Outflow or wall ghosts of a patch on the domain faces (``mhd.py:748-777``)."""
        L, box = self.patches[k]
        shape = self.geo.shape(L)
        cells, az = state[k]
        for axis in range(2):
            if self.geo.periodic[axis]:
                continue
            lower = box[axis][0] == 0
            upper = box[axis][1] == shape[axis]
            if not (lower or upper):
                continue
            if self.walls[axis]:
                normal = 1 + axis
                for f in range(6):
                    boundaries.wall_cells(cells[f], axis, lower, upper, -1.0 if f == normal else 1.0)
                boundaries.wall_nodes(az, axis, lower, upper)
            else:
                for f in range(6):
                    boundaries.outflow_cells(cells[f], axis, lower, upper)
                boundaries.outflow_nodes(az, axis, lower, upper)

    # -- the step --------------------------------------------------------------------
    def patch_fluxes(self, state, k):
        L, box = self.patches[k]
        cells, az = state[k]
        dx, dy = self.spacing(L)
        prim, bxf, byf, ecc, bad = kernels.primitive(*cells, az, dx, dy, *self.b0, self.gamma)
        if bad:
            raise FloatingPointError("non-positive density or pressure")
        fx = kernels.sweep(prim, bxf, 0, self.code, self.gamma)
        fy = kernels.sweep(prim, byf, 1, self.code, self.gamma)
        ez = kernels.ct_contact(fx[5], fy[5], fx[0], fy[0], ecc)
        if max(self.conduction) > 0.0:  # mhd.py:868-871
            kernels.conductive(prim, bxf, byf, dx, dy, *self.conduction, fx, fy)
        if self.resistivity:
            kernels.resistive(prim, bxf, byf, dx, dy, self.resistivity, fx, fy, ez)
        shape = self.geo.shape(L)
        for axis in range(2):  # the walls of the domain (mhd.py:873-886)
            if self.geo.periodic[axis] or not self.walls[axis]:
                continue
            f = fx if axis == 0 else fy
            lower, upper = box[axis][0] == 0, box[axis][1] == shape[axis]
            n = box[axis][1] - box[axis][0]
            # the sweep's face 0 is the patch's low boundary face, face n its high one
            for on, face, node in ((lower, 0, 0), (upper, n, -1)):
                if not on:
                    continue
                if axis == 0:
                    ez[node, :] = 0.0
                    f[WALL_ROWS, face, :] = 0.0
                else:
                    ez[:, node] = 0.0
                    f[WALL_ROWS, :, face] = 0.0
        return fx, fy, ez

    def stage(self, src, dt, start=None):
        """One Euler stage of every patch from ``src`` (its halos filled here); with
        ``start`` the Heun average ``(U2 + U0) * 0.5``.  Returns the new state and the
        fluxes."""
        self.fill(src)
        out, fluxes = [], []
        for k, (L, ((i0, i1), (j0, j1))) in enumerate(self.patches):
            fx, fy, ez = self.patch_fluxes(src, k)
            fluxes.append((fx, fy))
            cs, azs = src[k]
            cd, azd = cs.copy(), azs.copy()
            dx, dy = self.spacing(L)
            dtdx, dtdy = dt / dx, dt / dy
            h, nx, ny = HALO, i1 - i0, j1 - j0
            for f in range(6):
                gx = fx[XR[f], :, h : h + ny]
                gy = fy[YR[f], h : h + nx, :]
                u = cs[f, h : h + nx, h : h + ny]
                v = u - (dtdx * (gx[1:] - gx[:-1]) + dtdy * (gy[:, 1:] - gy[:, :-1]))
                if start is not None:
                    v = (v + start[k][0][f, h : h + nx, h : h + ny]) * 0.5
                cd[f, h : h + nx, h : h + ny] = v
            a = azs[h : h + nx + 1, h : h + ny + 1] - dt * ez
            if start is not None:
                a = (a + start[k][1][h : h + nx + 1, h : h + ny + 1]) * 0.5
            azd[h : h + nx + 1, h : h + ny + 1] = a
            out.append((cd, azd))
        return out, fluxes

    def register(self, fluxes, dt):
        """Per coarse-fine cell the correction ``d`` of one stage's fluxes (6 fields)."""
        out = []
        for k, ca, cb, L, faces in self.reflux_cells:
            dx, dy = self.spacing(L)
            delta = np.zeros(6)
            for axis, side, coarse, fine in faces:
                dtd = dt / dx if axis == 0 else dt / dy
                rows = XR if axis == 0 else YR
                for f in range(6):
                    fc = _face(fluxes, axis, rows[f], coarse)
                    mean = 0.5 * (_face(fluxes, axis, rows[f], fine[0])
                                  + _face(fluxes, axis, rows[f], fine[1]))
                    d = dtd * (fc - mean) if side == 0 else dtd * (mean - fc)
                    delta[f] = delta[f] + d
            out.append(delta)
        return out

    def restrict(self, state):
        """Coarse nodes on fine nodes take the fine value, finest level pair first
        (``mhd.py:981,991``)."""
        for L in range(self.nlev - 2, -1, -1):
            for q, (lev, ((i0, i1), (j0, j1))) in enumerate(self.patches):
                if lev != L:
                    continue
                for i in range(i0, i1 + 1):
                    for j in range(j0, j1 + 1):
                        src = self.geo.node_of(L + 1, 2 * i, 2 * j)
                        if src is None:
                            continue
                        p, fa, fb = src
                        state[q][1][i - i0 + HALO, j - j0 + HALO] = state[p][1][fa, fb]

    def average_down(self, state):
        """Covered coarse cells take the mean of their four children, finest first."""
        for L in range(self.nlev - 2, -1, -1):
            for p, (lev, ((i0, i1), (j0, j1))) in enumerate(self.patches):
                if lev != L + 1:
                    continue
                fine = state[p][0]
                for ci in range(i0 // 2, i1 // 2):
                    for cj in range(j0 // 2, j1 // 2):
                        q, qa, qb = self.geo.cell_of(L, ci, cj)
                        a, b = 2 * ci - i0 + HALO, 2 * cj - j0 + HALO
                        state[q][0][:, qa, qb] = ((fine[:, a, b] + fine[:, a + 1, b])
                                                  + (fine[:, a, b + 1] + fine[:, a + 1, b + 1])) * 0.25

    def stable_step(self):
        """``mhd.py:931-963`` over the patches: the fast-wave limit of each and, with a
        diffusivity, its diffusion limit with the patch's smallest density."""
        smallest = np.inf
        for k, (L, _) in enumerate(self.patches):
            cells, az = self.state[k]
            dx, dy = self.spacing(L)
            r, rho_min = kernels.rate(*cells, az, HALO, dx, dy, *self.b0, self.gamma)
            smallest = min(smallest, 1.0 / r)
            diffusivity = max(self.resistivity,
                              (self.gamma - 1.0) * max(self.conduction) / float(rho_min))
            if diffusivity > 0.0:
                smallest = min(smallest, 0.5 / (diffusivity * (dx**-2 + dy**-2)))
        return self.cfl * float(smallest)

    def step(self, dt=None):
        """One Heun step in the reference's order (see the module)."""
        if dt is None:
            dt = self.stable_step()
        start = self.state
        one, fluxes = self.stage(start, dt)
        kept = [0.5 * d for d in self.register(fluxes, dt)]
        self.restrict(one)
        two, fluxes = self.stage(one, dt, start=start)
        for (k, ca, cb, _, _), r, d in zip(self.reflux_cells, kept, self.register(fluxes, dt)):
            cells = two[k][0]
            for f in range(6):
                cells[f, ca, cb] = cells[f, ca, cb] + (r[f] + 0.5 * d[f])
        self.average_down(two)
        self.restrict(two)
        self.state = two
        self.time += dt
        self.steps += 1
        return dt

    def run(self, duration, max_steps=1_000_000):
        """``mhd.py:996-1008``: the last step shortened to land on the end."""
        end = self.time + duration
        taken = 0
        while self.time < end * (1.0 - 1e-14) and taken < max_steps:
            self.step(min(self.stable_step(), end - self.time))
            taken += 1
        return taken

    # -- diagnostics -----------------------------------------------------------------
    @property
    def buf(self):
        return pack(self.state, self.patches)

    def leaf_mask(self, k):
        L, ((i0, i1), (j0, j1)) = self.patches[k]
        return ~self.geo.covered(L)[i0:i1, j0:j1]

    def leaf_totals(self):
        """Integrals of the 6 conserved fields over the leaf cells."""
        totals = np.zeros(6)
        for k, (L, _) in enumerate(self.patches):
            cells = self.state[k][0][:, HALO:-HALO, HALO:-HALO]
            dx, dy = self.spacing(L)
            totals += (cells * self.leaf_mask(k)).sum(axis=(1, 2)) * (dx * dy)
        return totals

    def leaf_values(self):
        """Cell centres, levels and primitive variables of the leaf cells, patch by patch
        (``mhd.py:1059-1112``; the cell field is the mean of its two faces)."""
        out = {name: [] for name in ("x", "y", "level", "density", "pressure",
                                     "magnetic_field_x", "magnetic_field_y", "magnetic_field_z")}
        (x0, _), (y0, _) = self.bounds
        for k, (L, ((i0, i1), (j0, j1))) in enumerate(self.patches):
            cells, az = self.state[k]
            dx, dy = self.spacing(L)
            a = az[HALO:-HALO, HALO:-HALO]
            bxf = (a[:, 1:] - a[:, :-1]) / dy + self.b0[0]
            byf = -(a[1:, :] - a[:-1, :]) / dx + self.b0[1]
            bx, by = 0.5 * (bxf[:-1] + bxf[1:]), 0.5 * (byf[:, :-1] + byf[:, 1:])
            c = cells[:, HALO:-HALO, HALO:-HALO]
            rho = c[0]
            u, v, w = c[1] / rho, c[2] / rho, c[3] / rho
            p = (self.gamma - 1.0) * (c[4] - 0.5 * rho * (u**2 + v**2 + w**2)
                                      - 0.5 * (bx**2 + by**2 + c[5] ** 2))
            mask = self.leaf_mask(k)
            ii, jj = np.meshgrid(np.arange(i0, i1), np.arange(j0, j1), indexing="ij")
            out["x"].append((x0 + (ii + 0.5) * dx)[mask])
            out["y"].append((y0 + (jj + 0.5) * dy)[mask])
            out["level"].append(np.full(int(mask.sum()), L))
            for name, value in (("density", rho), ("pressure", p), ("magnetic_field_x", bx),
                                ("magnetic_field_y", by), ("magnetic_field_z", c[5])):
                out[name].append(value[mask])
        return {name: np.concatenate(v) for name, v in out.items()}

    def divergence(self):
        """Largest |div B| times the cell size over every patch (``mhd.py:1114-1131``)."""
        largest = 0.0
        for k, (L, _) in enumerate(self.patches):
            dx, dy = self.spacing(L)
            a = self.state[k][1][HALO:-HALO, HALO:-HALO]
            bxf = (a[:, 1:] - a[:, :-1]) / dy + self.b0[0]
            byf = -(a[1:, :] - a[:-1, :]) / dx + self.b0[1]
            div = (bxf[1:] - bxf[:-1]) / dx + (byf[:, 1:] - byf[:, :-1]) / dy
            largest = max(largest, float(np.max(np.abs(div))) * min(dx, dy))
        return largest


def _face(fluxes, axis, row, where):
    """The flux of local face (a, b) of patch k: sweep arrays hold face k between
    cells k + 1 and k + 2 along the swept axis."""
    k, fa, fb = where
    flux = fluxes[k][axis]
    if axis == 0:
        return flux[row, fa - 2, fb]
    return flux[row, fa, fb - 2]


def _minmod(a, b):
    return np.where(a * b > 0.0, np.where(np.abs(a) < np.abs(b), a, b), 0.0)


def regrid_state(old: OracleAMR2D, new_patches) -> OracleAMR2D:
    """The oracle of the move of a regrid (``mhd.py:1043-1056``) onto ``new_patches``:
    level by level, coarse to fine, the cells and nodes the old hierarchy held on the
    same level are copied; a new cell takes the conservative prolongation of its
    coarse cell with minmod slopes (``c0 + s_x slope_x / 4 + s_y slope_y / 4``, s = -1
    for the lower child, +1 for the upper), a new node the bilinear interpolation of
    the coarse nodes (``mhd.py:91-94``); the level's halos are filled; then the covered
    cells are averaged down and A_z restricted."""
    opts = dict(periodic=old.geo.periodic, walls=old.walls, gamma=old.gamma, mean_field=old.b0,
                limiter=kernels.LIMITER_CODES[old.code], cfl=old.cfl,
                resistivity=old.resistivity, conduction=old.conduction)
    new = OracleAMR2D(new_patches, old.geo.base, old.bounds,
                      np.zeros(layout(new_patches)[1]), **opts)
    ogeo, ngeo = old.geo, new.geo
    get = new._node_get(new.state)
    for L in range(new.nlev):
        ks = [k for k, (lev, _) in enumerate(new.patches) if lev == L]
        for k in ks:
            cells, az = new.state[k]
            (i0, i1), (j0, j1) = new.patches[k][1]
            for i in range(i0, i1):
                for j in range(j0, j1):
                    a, b = i - i0 + HALO, j - j0 + HALO
                    src = ogeo.cell_of(L, i, j) if L < ogeo.nlev else None
                    if src is not None:
                        q, qa, qb = src
                        cells[:, a, b] = old.state[q][0][:, qa, qb]
                        continue
                    q, qa, qb = ngeo.cell_of(L - 1, i // 2, j // 2)
                    c = new.state[q][0]
                    c0 = c[:, qa, qb]
                    slx = _minmod(c[:, qa + 1, qb] - c0, c0 - c[:, qa - 1, qb])
                    sly = _minmod(c[:, qa, qb + 1] - c0, c0 - c[:, qa, qb - 1])
                    hx = 0.25 if i % 2 == 1 else -0.25
                    hy = 0.25 if j % 2 == 1 else -0.25
                    cells[:, a, b] = (c0 + hx * slx) + hy * sly
            for i in range(i0, i1 + 1):
                for j in range(j0, j1 + 1):
                    a, b = i - i0 + HALO, j - j0 + HALO
                    src = ogeo.node_of(L, i, j) if L < ogeo.nlev else None
                    if src is not None:
                        q, qa, qb = src
                        az[a, b] = old.state[q][1][qa, qb]
                        continue
                    ci, cj = i // 2, j // 2
                    a0 = get(L - 1, ci, cj)
                    if i % 2 and j % 2:
                        az[a, b] = 0.25 * ((a0 + get(L - 1, ci + 1, cj))
                                           + (get(L - 1, ci, cj + 1) + get(L - 1, ci + 1, cj + 1)))
                    elif i % 2:
                        az[a, b] = 0.5 * (a0 + get(L - 1, ci + 1, cj))
                    elif j % 2:
                        az[a, b] = 0.5 * (a0 + get(L - 1, ci, cj + 1))
                    else:
                        az[a, b] = a0
        new._fill_level(new.state, L)
    new.average_down(new.state)
    new.restrict(new.state)
    new.time = old.time
    return new
