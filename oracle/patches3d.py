"""The refined 3D step on a hierarchy of patches of any size, in NumPy FP64 (oracle
only: test infrastructure, never imported by the product).

Follows ``MHD3D.step`` (``mhd3d.py:726-754``) on every patch, in the reference's order:
stage 1 on every patch cell (covered coarse cells included) with the flux register
recording half of each coarse-fine correction; the potential restricted; stage 2, its
Heun average ``(U2 + U0) * 0.5``; one reflux; the covered cells averaged; the
potential restricted.  A stage is the halo fill of every level, coarse to fine, and
per patch ``Oracle3D.fluxes`` (``mhd3d.py:604-633``) with the resistive terms added
into the sweeps and edges (``mhd3d.py:561-602``) and the conservative update.

Its geometry is its own, from the boxes alone: the owner of an element is the first
patch of its level holding it (a node-centred axis holds its ends), the level's values
are read from there.  The arithmetic choices (parity-unpinned beyond the invariants and
known answers the tests check, since patchwork is absent):

* ghosts with no owner on their level interpolate from the level below, the terms'
  weights multiplied per axis (x slowest) and summed from the first term on: conserved
  cells linearly (1/4, 3/4); the potential along a node-centred axis by the coarse node
  (even index) or the quadratic (-1/8, 3/4, 3/8) through nodes m - 1, m, m + 1 (odd),
  along its cell-centred axis by the quadratic through the coarse cell and its
  neighbours at the fine centre (5/32, 30/32, -3/32 for the lower child, mirrored);
* average-down ``sum 0.125 v`` of the eight children (x slowest); restriction
  ``0.5 e_lo + 0.5 e_hi``;
* the register: per coarse cell its coarse-fine faces x-high, x-low, y-high, y-low,
  z-high, z-low, each adding ``dt/h (F_c - m)`` (high) or ``dt/h (m - F_c)`` (low), ``m =
  ((f00 + f01) + (f10 + f11)) * 0.25``; stage 1 keeps ``0.5 d``, the reflux adds
  ``kept + 0.5 d``;
* regrid: cells the old level held are copied, new cells take the minmod prolongation
  ``((c0 + sx/4 s_x) + sy/4 s_y) + sz/4 s_z``, new edges along the edge the coarse edge
  holding them and across it the coarse node or the mean of the two around it.
"""

from __future__ import annotations

import numpy as np

from .mhd3d import Oracle3D, resistive3d

HALO = 2
#: Node-centred axes of each field (cells, A_x, A_y, A_z).
NODES = {"cells": (0, 0, 0), "ax": (0, 1, 1), "ay": (1, 0, 1), "az": (1, 1, 0)}
FIELDS = ("cells", "ax", "ay", "az")
ROWS = ((0, 1, 2, 3, 4), (0, 3, 1, 2, 4), (0, 2, 3, 1, 4))


def _sizes(box):
    n = [hi - lo for lo, hi in box]
    X, Y, Z = (m + 2 * HALO for m in n)
    return n, (X, Y, Z), [(X, Y + 1, Z + 1), (X + 1, Y, Z + 1), (X + 1, Y + 1, Z)]


def layout(patches):
    offs, total = [], 0
    for _, box in patches:
        _, (X, Y, Z), pots = _sizes(box)
        offs.append(total)
        total += 5 * X * Y * Z + sum(int(np.prod(s)) for s in pots)
    return offs, total


def unpack(buf, patches):
    """Per patch ``[cells (5, X, Y, Z), A_x, A_y, A_z]`` copies of a flat buffer."""
    offs, _ = layout(patches)
    out = []
    for off, (_, box) in zip(offs, patches):
        _, (X, Y, Z), pots = _sizes(box)
        arrays = [np.array(buf[off : off + 5 * X * Y * Z]).reshape(5, X, Y, Z)]
        o = off + 5 * X * Y * Z
        for s in pots:
            m = int(np.prod(s))
            arrays.append(np.array(buf[o : o + m]).reshape(s))
            o += m
        out.append(arrays)
    return out


def pack(state, patches):
    offs, total = layout(patches)
    buf = np.zeros(total)
    for off, arrays in zip(offs, state):
        o = off
        for a in arrays:
            buf[o : o + a.size] = a.ravel()
            o += a.size
    return buf


def _field(arrays, name):
    """The haloed arrays of a field: the 5 cell planes stacked, or one potential."""
    return arrays[FIELDS.index(name)]


class OraclePatches3D:
    """Heun steps of a periodic refined 3D hierarchy of patches of any size."""

    def __init__(self, patches, base_shape, bounds, buf, *, gamma=5.0 / 3.0,
                 mean_field=(0.0, 0.0, 0.0), limiter="vanleer", cfl=0.3, resistivity=0.0):
        self.patches = [(int(L), tuple(tuple(int(v) for v in b) for b in box)) for L, box in patches]
        self.base = tuple(int(n) for n in base_shape)
        self.bounds = tuple(tuple(float(v) for v in b) for b in bounds)
        self.state = unpack(np.asarray(buf, np.float64), self.patches)
        self.gamma = float(gamma)
        self.b0 = tuple(float(b) for b in mean_field)
        self.limiter = limiter
        self.cfl = float(cfl)
        self.resistivity = float(resistivity)
        self.nlev = max(L for L, _ in self.patches) + 1
        self.time = 0.0
        self.owner = {(L, f): self._owners(L, f) for L in range(self.nlev) for f in FIELDS}
        if (self.owner[0, "cells"] < 0).any():
            raise ValueError("level 0 must cover the domain")

    @classmethod
    def from_hierarchy(cls, hierarchy, buf, **options):
        return cls([(p.level, p.box) for p in hierarchy.patches], hierarchy.base_shape,
                   hierarchy.bounds, buf, **options)

    # -- geometry --------------------------------------------------------------------
    def shape(self, L):
        return tuple(n << L for n in self.base)

    def spacing(self, L):
        return tuple((hi - lo) / n for (lo, hi), n in zip(self.bounds, self.shape(L)))

    def _owners(self, L, name):
        shape = self.shape(L)
        own = np.full(shape, -1, np.int64)
        for k in range(len(self.patches) - 1, -1, -1):  # the first patch wins
            lev, box = self.patches[k]
            if lev != L:
                continue
            idx = [np.arange(lo, hi + nd) % n for (lo, hi), nd, n in zip(box, NODES[name], shape)]
            own[np.ix_(*idx)] = k
        return own

    def _level_values(self, state, L, name):
        """The owners' values of a field on level L's index space (NaN: no owner); the
        cells as (5, ...) arrays."""
        shape = self.shape(L)
        comps = 5 if name == "cells" else 1
        out = np.full((comps, *shape), np.nan)
        for k in range(len(self.patches) - 1, -1, -1):
            lev, box = self.patches[k]
            if lev != L:
                continue
            arr = _field(state[k], name).reshape(comps, *_field(state[k], name).shape[-3:])
            # a patch spanning a periodic axis holds node n as node 0: its own node 0 counts
            idx = [(np.arange(lo, hi + nd) % n)[:n] for (lo, hi), nd, n in zip(box, NODES[name], shape)]
            sl = (slice(None),) + tuple(slice(HALO, HALO + len(i)) for i in idx)
            out[(slice(None),) + np.ix_(*idx)] = arr[sl]
        return out

    # -- halos -----------------------------------------------------------------------
    @staticmethod
    def _rule(kind, f):
        """Per fine index f along one axis: the list of (coarse index, weight) terms."""
        if kind == "linear":
            c0 = (f - 1) // 2
            w1 = 0.75 if f % 2 == 0 else 0.25
            return [(c0, 1.0 - w1), (c0 + 1, w1)]
        if kind == "node":
            if f % 2 == 0:
                return [(f // 2, 1.0)]
            m = (f - 1) // 2
            return [(m - 1, -0.125), (m, 0.75), (m + 1, 0.375)]
        if kind == "regrid-cell":
            return [(f // 2, 1.0)]
        if kind == "regrid-node":
            if f % 2 == 0:
                return [(f // 2, 1.0)]
            return [(f // 2, 0.5), (f // 2 + 1, 0.5)]
        c = f // 2
        lower = f % 2 == 0
        return [(c - 1, 5.0 / 32.0 if lower else -3.0 / 32.0), (c, 30.0 / 32.0),
                (c + 1, -3.0 / 32.0 if lower else 5.0 / 32.0)]

    def _interpolate(self, coarse, shape_c, kinds, fine):
        """The stencil value at fine coordinates ``fine`` from the level array below."""
        rules = [self._rule(kinds[a], int(fine[a])) for a in range(3)]
        value = None
        for cx, wx in rules[0]:
            for cy, wy in rules[1]:
                for cz, wz in rules[2]:
                    w = (wx * wy) * wz
                    v = coarse[:, cx % shape_c[0], cy % shape_c[1], cz % shape_c[2]]
                    if np.isnan(v).any():
                        raise ValueError("a fine patch is not properly nested")
                    value = w * v if value is None else value + w * v
        return value

    def fill_level(self, state, L):
        shape = self.shape(L)
        for name in FIELDS:
            here = self._level_values(state, L, name)
            below = self._level_values(state, L - 1, name) if L > 0 else None
            kinds = (["linear"] * 3 if name == "cells" else
                     ["node" if nd else "cell-quadratic" for nd in NODES[name]])
            for k, (lev, box) in enumerate(self.patches):
                if lev != L:
                    continue
                arr = _field(state[k], name)
                arr = arr.reshape(-1, *arr.shape[-3:])
                nd = NODES[name]
                for pos in np.ndindex(*arr.shape[1:]):
                    inner = all(HALO <= pos[a] < HALO + box[a][1] - box[a][0] + nd[a] for a in range(3))
                    edge = inner and any(nd[a] and pos[a] in (HALO, HALO + box[a][1] - box[a][0])
                                         for a in range(3))
                    if inner and not edge:
                        continue
                    fine = [box[a][0] + pos[a] - HALO for a in range(3)]
                    I = tuple(fine[a] % shape[a] for a in range(3))
                    if self.owner[L, name][I] >= 0:
                        if edge and self.owner[L, name][I] == k:
                            # the patch's own element unless it is the periodic image of
                            # another of its elements
                            own = tuple((I[a] - box[a][0]) % shape[a] + HALO for a in range(3))
                            if own == pos:
                                continue
                        arr[(slice(None), *pos)] = here[(slice(None), *I)]
                        continue
                    arr[(slice(None), *pos)] = self._interpolate(below, self.shape(L - 1), kinds, fine)

    def fill(self, state):
        for L in range(self.nlev):
            self.fill_level(state, L)

    # -- the step ------------------------------------------------------------------------
    def _patch(self, state, k):
        cells, ax, ay, az = state[k]
        return Oracle3D(cells, [ax, ay, az], self.spacing(self.patches[k][0]), gamma=self.gamma,
                        mean_field=self.b0, limiter=self.limiter, cfl=self.cfl)

    def stage(self, src, dt, start=None):
        """Every patch's Euler stage from ``src`` (its halos filled here); with ``start``
        the Heun average.  Returns the new state and every patch's sweeps."""
        self.fill(src)
        out, sweeps_all = [], []
        h = HALO
        for k, (L, box) in enumerate(self.patches):
            o = self._patch(src, k)
            sweeps, edges = o.fluxes()
            n = o.n
            block = []
            for s in range(3):
                idx = [slice(None)] + [slice(h, h + n[a]) if a != s else slice(None) for a in range(3)]
                block.append(sweeps[s][tuple(idx)])
            if self.resistivity:  # mhd3d.py:631-632
                energy = [np.array(block[s][4]) for s in range(3)]
                resistive3d(o._faces, n, o.spacing, self.resistivity, edges, energy)
                for s in range(3):
                    idx = [4] + [slice(h, h + n[a]) if a != s else slice(None) for a in range(3)]
                    sweeps[s][tuple(idx)] = energy[s]
                    block[s] = block[s].copy()
                    block[s][4] = energy[s]
            sweeps_all.append(sweeps)
            rates = [dt / d for d in o.spacing]
            cells = src[k][0].copy()
            for f in range(5):
                div = None
                for s in range(3):
                    row = 0 if f == 0 else (4 if f == 4 else 1 + (f - 1 - s) % 3)
                    term = rates[s] * np.diff(block[s][row], axis=s)
                    div = term if div is None else div + term
                inner = (f, slice(h, h + n[0]), slice(h, h + n[1]), slice(h, h + n[2]))
                v = src[k][0][inner] - div
                if start is not None:
                    v = (v + start[k][0][inner]) * 0.5
                cells[inner] = v
            pots = []
            for c in range(3):
                p = src[k][1 + c].copy()
                idx = tuple(slice(h, h + n[a] + (a != c)) for a in range(3))
                v = p[idx] - dt * edges[c]
                if start is not None:
                    v = (v + start[k][1 + c][idx]) * 0.5
                p[idx] = v
                pots.append(p)
            out.append([cells, *pots])
        return out, sweeps_all

    def _coarse_fine_cells(self):
        """Per uncovered coarse cell next to a finer patch: (patch, haloed cell, level,
        faces), a face (axis, side, coarse (patch, low-face cell), four fine ones)."""
        cells = []
        for L in range(self.nlev - 1):
            shape, fshape = self.shape(L), self.shape(L + 1)
            fine = self.owner[L + 1, "cells"]
            cov = np.zeros(shape, bool)
            for d in np.ndindex(2, 2, 2):
                cov |= fine[d[0]::2, d[1]::2, d[2]::2] >= 0
            for k, (lev, box) in enumerate(self.patches):
                if lev != L:
                    continue
                for C in np.ndindex(*[hi - lo for lo, hi in box]):
                    G = [box[a][0] + C[a] for a in range(3)]
                    if cov[tuple(G)]:
                        continue
                    faces = []
                    for axis in range(3):
                        for side, step in ((0, 1), (1, -1)):
                            nb = list(G)
                            nb[axis] = (nb[axis] + step) % shape[axis]
                            if not cov[tuple(nb)]:
                                continue
                            a1, a2 = [a for a in range(3) if a != axis]
                            loc = [C[a] + HALO for a in range(3)]
                            cf = list(loc)
                            if side == 0:
                                cf[axis] += 1
                            fines = []
                            for p in (0, 1):
                                for q in (0, 1):
                                    F = [2 * G[a] for a in range(3)]
                                    F[a1] += p
                                    F[a2] += q
                                    F[axis] = 2 * G[axis] + 2 if side == 0 else 2 * G[axis] - 1
                                    F = [F[a] % fshape[a] for a in range(3)]
                                    pf = int(fine[tuple(F)])
                                    fb = self.patches[pf][1]
                                    fl = [(F[a] - fb[a][0]) % fshape[a] + HALO for a in range(3)]
                                    if side == 1:
                                        fl[axis] += 1
                                    fines.append((pf, fl))
                            faces.append((axis, side, (k, cf), fines))
                    if faces:
                        cells.append((k, tuple(c + HALO for c in C), L, faces))
        return cells

    @staticmethod
    def _flux(sweeps, axis, row, where):
        k, loc = where
        idx = list(loc)
        idx[axis] -= 2  # the sweep's face f lies between cells f + 1 and f + 2
        return sweeps[k][axis][(row, *idx)]

    def register(self, sweeps, dt):
        out = []
        for k, cell, L, faces in self.reflux_cells:
            h = self.spacing(L)
            d = np.zeros(5)
            for axis, side, coarse, fines in faces:
                dtd = dt / h[axis]
                for fld in range(5):
                    r = ROWS[axis][fld]
                    fc = self._flux(sweeps, axis, r, coarse)
                    f00, f01, f10, f11 = (self._flux(sweeps, axis, r, f) for f in fines)
                    m = ((f00 + f01) + (f10 + f11)) * 0.25
                    d[fld] = d[fld] + (dtd * (fc - m) if side == 0 else dtd * (m - fc))
            out.append(d)
        return out

    def coarsen(self, state, cells=True):
        """Finest level pair first: covered cells := the mean of their children
        (``cells``), coarse edges over fine ones := the mean of the two along them."""
        for L in range(self.nlev - 2, -1, -1):
            fshape = self.shape(L + 1)
            for name in FIELDS if cells else FIELDS[1:]:
                fine = self._level_values(state, L + 1, name)
                for k, (lev, box) in enumerate(self.patches):
                    if lev != L:
                        continue
                    arr = _field(state[k], name)
                    arr = arr.reshape(-1, *arr.shape[-3:])
                    nd = NODES[name]
                    ext = [box[a][1] - box[a][0] + nd[a] for a in range(3)]
                    for C in np.ndindex(*ext):
                        G = [box[a][0] + C[a] for a in range(3)]
                        if name == "cells":
                            terms = [[(2 * G[a] + d[a]) % fshape[a] for a in range(3)]
                                     for d in np.ndindex(2, 2, 2)]
                            weight = 0.125
                        else:
                            c = "xyz".index(name[1])
                            terms = [[(2 * G[a] + (s if a == c else 0)) % fshape[a] for a in range(3)]
                                     for s in (0, 1)]
                            weight = 0.5
                        vals = [fine[(slice(None), *t)] for t in terms]
                        if any(np.isnan(v).any() for v in vals):
                            continue
                        value = weight * vals[0]
                        for v in vals[1:]:
                            value = value + weight * v
                        arr[(slice(None), *[C[a] + HALO for a in range(3)])] = value

    def stable_step(self):
        """mhd3d.py:674-724 over the patches."""
        return self.cfl * min(self._patch(self.state, k).smallest_time()
                              for k in range(len(self.patches)))

    def step(self, dt=None):
        if dt is None:
            dt = self.stable_step()
        if not hasattr(self, "reflux_cells"):
            self.reflux_cells = self._coarse_fine_cells()
        start = self.state
        one, sweeps = self.stage(start, dt)
        kept = [0.5 * d for d in self.register(sweeps, dt)]
        self.coarsen(one, cells=False)
        two, sweeps = self.stage(one, dt, start=start)
        for (k, cell, _, _), r, d in zip(self.reflux_cells, kept, self.register(sweeps, dt)):
            for fld in range(5):
                two[k][0][(fld, *cell)] = two[k][0][(fld, *cell)] + (r[fld] + 0.5 * d[fld])
        self.coarsen(two, cells=True)
        self.state = two
        self.time += dt
        return dt

    def run(self, duration):
        end = self.time + duration
        while self.time < end * (1.0 - 1e-14):
            self.step(min(self.stable_step(), end - self.time))

    # -- diagnostics -------------------------------------------------------------------
    @property
    def buf(self):
        return pack(self.state, self.patches)

    def leaf_totals(self):
        tot = np.zeros(5)
        for k, (L, box) in enumerate(self.patches):
            inner = self.state[k][0][:, HALO:-HALO, HALO:-HALO, HALO:-HALO]
            mask = np.ones(inner.shape[1:], bool)
            if L + 1 < self.nlev:
                fine = self.owner[L + 1, "cells"]
                for C in np.ndindex(*mask.shape):
                    G = [box[a][0] + C[a] for a in range(3)]
                    mask[C] = not any(fine[tuple(2 * G[a] + d[a] for a in range(3))] >= 0
                                      for d in np.ndindex(2, 2, 2))
            tot += (inner * mask).sum(axis=(1, 2, 3)) * float(np.prod(self.spacing(L)))
        return tot


def regrid_state(old: OraclePatches3D, new_patches) -> OraclePatches3D:
    """The oracle of ``MHD3D.regrid``'s move onto ``new_patches`` (see the module)."""
    new = OraclePatches3D(new_patches, old.base, old.bounds, np.zeros(layout(new_patches)[1]),
                          gamma=old.gamma, mean_field=old.b0, limiter=old.limiter, cfl=old.cfl,
                          resistivity=old.resistivity)
    for L in range(new.nlev):
        shape = new.shape(L)
        for name in FIELDS:
            have = old._level_values(old.state, L, name) if L < old.nlev else None
            below = new._level_values(new.state, L - 1, name) if L > 0 else None
            nd = NODES[name]
            for k, (lev, box) in enumerate(new.patches):
                if lev != L:
                    continue
                arr = _field(new.state[k], name)
                arr = arr.reshape(-1, *arr.shape[-3:])
                ext = [box[a][1] - box[a][0] + nd[a] for a in range(3)]
                for C in np.ndindex(*ext):
                    fine = [box[a][0] + C[a] for a in range(3)]
                    I = tuple(fine[a] % shape[a] for a in range(3))
                    pos = (slice(None), *[C[a] + HALO for a in range(3)])
                    if have is not None and not np.isnan(have[(slice(None), *I)]).any():
                        arr[pos] = have[(slice(None), *I)]
                        continue
                    if name == "cells":
                        arr[pos] = _prolong(new, k, L, fine)
                        continue
                    kinds = ["regrid-node" if nd[a] else "regrid-cell" for a in range(3)]
                    arr[pos] = new._interpolate(below, new.shape(L - 1), kinds, fine)
        new.fill_level(new.state, L)
    new.coarsen(new.state, cells=True)
    new.time = old.time
    return new


def _minmod(a, b):
    return np.where(a * b > 0.0, np.where(np.abs(a) < np.abs(b), a, b), 0.0)


def _prolong(new, k, L, fine):
    """A new cell from the coarse cell holding it: its value and minmod slopes, read
    from the haloed arrays of the coarse cell's owner (halos filled)."""
    shape_c = new.shape(L - 1)
    C = tuple((fine[a] // 2) % shape_c[a] for a in range(3))
    q = int(new.owner[L - 1, "cells"][C])
    box = new.patches[q][1]
    loc = [(C[a] - box[a][0]) % shape_c[a] + HALO for a in range(3)]
    arr = new.state[q][0]
    c0 = arr[(slice(None), *loc)]
    value = c0
    for a in range(3):
        up, dn = list(loc), list(loc)
        up[a] += 1
        dn[a] -= 1
        slope = _minmod(arr[(slice(None), *up)] - c0, c0 - arr[(slice(None), *dn)])
        value = value + (0.25 if fine[a] % 2 == 1 else -0.25) * slope
    return value
