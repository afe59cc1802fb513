"""The step on a refined 3D hierarchy of equal blocks in NumPy (oracle only).

Restates this package's device algorithm for ``MHD3D.step`` on a patchwork
hierarchy (``mhd3d.py:635-671,726-754``) on the geometry and tables of
``wellbalanced_amd.amr3d_maps``: ghost fill (copies from the block of the same level
holding a cell or edge, else interpolation from the level below), the stage of every
block (``Oracle3D.fluxes`` on the haloed block and this package's update form), the
flux register of both stages (the corrections of the six directions summed per coarse
cell; stage 1 keeps half), the potential restricted between the stages, and after the
Heun average one reflux, the average-down of covered cells and the restriction of the
potential (``mhd3d.py:739-751``).  patchwork's own operation order is not available,
so these rules are parity unpinned; the invariants in the tests pin them.
"""

from __future__ import annotations

import numpy as np

from .mhd3d import Oracle3D, cell_fields, face_fields, fast_speed_np

HALO = 2


class OracleAMR3D:
    def __init__(self, hierarchy, tables, buf, *, gamma=5.0 / 3.0, mean_field=(0.0, 0.0, 0.0),
                 limiter="minmod", cfl=0.3):
        self.h = hierarchy
        self.t = tables
        self.buf = np.array(buf, np.float64, copy=True)
        self.gamma = float(gamma)
        self.b0 = tuple(float(b) for b in mean_field)
        self.limiter = limiter
        self.cfl = float(cfl)
        self.B = hierarchy.block
        self.X = self.B + 4
        self.nlev = len(hierarchy.levels)
        self.time = 0.0
        self._plan()

    # -- block layout ----------------------------------------------------------------------
    def cell_index(self, blk, f, i, j, k):
        X = self.X
        return self.t.blocks["off"][blk] + f * X**3 + (i * X + j) * X + k

    def edge_index(self, blk, c, i, j, k):
        X, XN = self.X, self.X + 1
        base = self.t.blocks["off"][blk] + 5 * X**3
        ne = X * XN * XN
        if c == 0:
            return base + (i * XN + j) * XN + k
        if c == 1:
            return base + ne + (i * X + j) * XN + k
        return base + 2 * ne + (i * XN + j) * X + k

    def block_of(self, L, g):
        """Block ids of level L holding level coordinates g (3, N) canonically."""
        nb = self.t.level_nb[L]
        b = g // self.B
        return self.t.lookup[self.t.level_base[L] + (b[0] * nb + b[1]) * nb + b[2]]

    def edge_master(self, L, c, g):
        """The master copy of edges (component c) at level coordinates g (3, N): the
        block holding them in its own range, else on its upper boundary (local index B on
        a node axis where g is a multiple of B), first in the fixed order; ids -1 where
        the level has none, and the master's local coordinates."""
        B, nb = self.B, self.t.level_nb[L]
        N = g.shape[1]
        ids = np.full(N, -1, np.int64)
        loc = np.zeros((3, N), np.int64)
        for s in range(8):
            sh = [(s >> (2 - a)) & 1 for a in range(3)]
            ok = np.ones(N, bool)
            blk = np.empty((3, N), np.int64)
            lc = np.empty((3, N), np.int64)
            for a in range(3):
                if sh[a]:
                    ok &= (a != c) & (g[a] % B == 0)
                    blk[a], lc[a] = (g[a] // B - 1) % nb, B + 2
                else:
                    blk[a], lc[a] = g[a] // B, g[a] % B + 2
            cand = self.t.lookup[self.t.level_base[L] + (blk[0] * nb + blk[1]) * nb + blk[2]]
            take = ok & (cand >= 0) & (ids < 0)
            ids[take] = cand[take]
            loc[:, take] = lc[:, take]
        return ids, loc

    # -- the ghost-fill plan ---------------------------------------------------------------
    def _plan(self):
        """Per block, the positions filled by a copy and by interpolation (taps in the
        kernel's order, padded with zero weights)."""
        B, X = self.B, self.X
        copies_d, copies_s = [], []
        interp_d, interp_s, interp_w = [], [], []
        for blk, row in enumerate(self.t.blocks):
            L = int(row["level"])
            n = self.t.level_nb[L] * B
            bo = np.array([row["bx"], row["by"], row["bz"]], np.int64)[:, None] * B
            # cells
            p = np.array(np.unravel_index(np.arange(X**3), (X, X, X)), np.int64)
            inner = np.all((p >= 2) & (p < 2 + B), axis=0)
            p = p[:, ~inner]
            g = (bo + p - 2) % n
            owner = self.block_of(L, g)
            have = owner >= 0
            for f in range(5):
                copies_d.append(self.cell_index(blk, f, *p[:, have]))
                copies_s.append(self.cell_index(owner[have], f, *(g[:, have] % B + 2)))
            if (~have).any():
                q, gq = p[:, ~have], g[:, ~have]
                nc = self.t.level_nb[L - 1] * B
                c0 = np.where(gq & 1, gq >> 1, (gq >> 1) - 1)
                tt = np.where(gq & 1, 0.25, 0.75)
                w = np.stack([1.0 - tt, tt])  # (2, 3, N)
                for f in range(5):
                    S, W = [], []
                    for x in range(2):
                        for y in range(2):
                            for z in range(2):
                                G = (c0 + np.array([x, y, z])[:, None]) % nc
                                cb = self.block_of(L - 1, G)
                                S.append(self.cell_index(cb, f, *(G % B + 2)))
                                W.append((w[x, 0] * w[y, 1]) * w[z, 2])
                    interp_d.append(self.cell_index(blk, f, *q))
                    interp_s.append(np.stack(S, 1))
                    interp_w.append(np.stack(W, 1))
            # edges
            for c in range(3):
                ext = tuple(X if a == c else X + 1 for a in range(3))
                p = np.array(np.unravel_index(np.arange(np.prod(ext)), ext), np.int64)
                own = np.all((p >= 2) & (p < 2 + B), axis=0)
                p = p[:, ~own]
                g = (bo + p - 2) % n
                master, mloc = self.edge_master(L, c, g)
                keep = (master == blk) & np.all(mloc == p, axis=0)
                cp = (master >= 0) & ~keep
                copies_d.append(self.edge_index(blk, c, *p[:, cp]))
                copies_s.append(self.edge_index(master[cp], c, *mloc[:, cp]))
                todo = master < 0
                if not todo.any():
                    continue
                q, gq = p[:, todo], g[:, todo]
                nc = self.t.level_nb[L - 1] * B
                N = q.shape[1]
                # per axis: (offsets, weights) lists of 2 (padded)
                offs = np.zeros((3, 2, N), np.int64)
                wts = np.zeros((3, 2, N))
                ntap = np.zeros((3, N), np.int64)
                for a in range(3):
                    m = gq[a] >> 1
                    odd = (gq[a] & 1).astype(bool)
                    if a == c:
                        tt = np.where(odd, 0.25, 0.75)
                        base = np.where(odd, m, m - 1)
                        offs[a, 0], offs[a, 1] = base, base + 1
                        wts[a, 0], wts[a, 1] = 1.0 - tt, tt
                        ntap[a] = 2
                    else:
                        offs[a, 0], offs[a, 1] = m, m + 1
                        wts[a, 0] = np.where(odd, 0.5, 1.0)
                        wts[a, 1] = np.where(odd, 0.5, 0.0)
                        ntap[a] = np.where(odd, 2, 1)
                # taps in the kernel's order, for each position its own tap counts;
                # positions group by their (nt0, nt1, nt2) pattern
                for pattern in {tuple(v) for v in ntap.T}:
                    sel = np.all(ntap.T == np.array(pattern), axis=1)
                    S, W = [], []
                    for x in range(pattern[0]):
                        for y in range(pattern[1]):
                            for z in range(pattern[2]):
                                G = np.stack([offs[0, x, sel], offs[1, y, sel], offs[2, z, sel]]) % nc
                                cb, cl = self.edge_master(L - 1, c, G)
                                assert (cb >= 0).all(), "outside the level below"
                                S.append(self.edge_index(cb, c, *cl))
                                W.append((wts[0, x, sel] * wts[1, y, sel]) * wts[2, z, sel])
                    interp_d.append(self.edge_index(blk, c, *q[:, sel]))
                    interp_s.append(np.stack(S, 1))
                    interp_w.append(np.stack(W, 1))
        self.copy_d = np.concatenate(copies_d)
        self.copy_s = np.concatenate(copies_s)
        self.interp = list(zip(interp_d, interp_s, interp_w))

    def fill(self, buf):
        vals = buf[self.copy_s]
        ivals = []
        for d, s, w in self.interp:
            acc = np.zeros(len(d))
            for t in range(s.shape[1]):
                acc = acc + w[:, t] * buf[s[:, t]]
            ivals.append(acc)
        buf[self.copy_d] = vals
        for (d, _, _), v in zip(self.interp, ivals):
            buf[d] = v

    # -- per block views -----------------------------------------------------------------------
    def views(self, buf, blk):
        X = self.X
        off = int(self.t.blocks["off"][blk])
        cells = buf[off : off + 5 * X**3].reshape(5, X, X, X)
        o = off + 5 * X**3
        pots = []
        for s in [(X, X + 1, X + 1), (X + 1, X, X + 1), (X + 1, X + 1, X)]:
            m = int(np.prod(s))
            pots.append(buf[o : o + m].reshape(s))
            o += m
        return cells, pots

    def spacing(self, blk):
        r = self.t.blocks[blk]
        return (float(r["dx"]), float(r["dy"]), float(r["dz"]))

    # -- the stage ------------------------------------------------------------------------------
    def _oracle(self, buf, blk):
        cells, pots = self.views(buf, blk)
        return Oracle3D(cells, pots, self.spacing(blk), gamma=self.gamma, mean_field=self.b0,
                        limiter=self.limiter, cfl=self.cfl)

    def stage(self, src, dst, dt, heun):
        self.fill(src)
        h, B = HALO, self.B
        sweeps = {}
        for blk in range(len(self.t.blocks)):
            o = self._oracle(src, blk)
            sw, edges = o.fluxes()
            sweeps[blk] = sw
            rates = [dt / s for s in o.spacing]
            cells, pots = self.views(src, blk)
            ocells, opots = self.views(dst, blk)
            inner = (slice(h, h + B),) * 3
            for field in range(5):
                div = None
                for s in range(3):
                    row = 0 if field == 0 else (4 if field == 4 else 1 + (field - 1 - s) % 3)
                    idx = [slice(None)] * 3
                    for k in range(3):
                        if k != s:
                            idx[k] = slice(h, h + B)
                    f = sw[s][row][tuple(idx)]
                    term = rates[s] * np.diff(f, axis=s)
                    div = term if div is None else div + term
                v = cells[field][inner] - div
                if heun:
                    v = (v + ocells[field][inner]) * 0.5
                ocells[field][inner] = v
            for c in range(3):
                idx = tuple(slice(h, h + B + (k != c)) for k in range(3))
                v = pots[c][idx] - dt * edges[c]
                if heun:
                    v = (v + opots[c][idx]) * 0.5
                opots[c][idx] = v
        acc = {}
        for d in range(6):
            self.reflux(sweeps, acc, dt, d // 2, d % 2)
        if not heun:  # the register keeps half the first stage's correction
            self.kept = {key: 0.5 * v for key, v in acc.items()}
        else:  # the reflux after the Heun average (mhd3d.py:745-750)
            for (coarse, c), v in acc.items():
                cells, _ = self.views(dst, coarse)
                r = self.kept[coarse, c]
                for fld in range(5):
                    cells[fld][c] = cells[fld][c] + (r[fld] + 0.5 * v[fld])
        for L in range(self.nlev - 2, -1, -1):
            self.coarsen(dst, L, cells=heun)

    def reflux(self, sweeps, acc, dt, axis, side):
        """Add the corrections of the coarse-fine faces of one direction to ``acc``
        (per coarse block and haloed cell: 5 values, summed in direction order)."""
        B = self.B
        a1, a2 = [a for a in range(3) if a != axis]
        xrow, yrow, zrow = (0, 1, 2, 3, 4), (0, 3, 1, 2, 4), (0, 2, 3, 1, 4)
        rows = (xrow, yrow, zrow)[axis]
        for item in self.t.reflux[2 * axis + side]:
            fine, coarse = int(item["fine"]), int(item["coarse"])
            origin = np.array([item["ci"], item["cj"], item["ck"]]) + 2
            dtd = dt / self.spacing(coarse)[axis]
            for u in range(8):
                for v in range(8):
                    c = origin.copy()
                    c[a1] += u
                    c[a2] += v
                    cf = c.copy()
                    if side == 0:
                        cf[axis] += 1
                    idx = list(cf)
                    idx[axis] -= 2  # the sweep's face index of the low face of cell cf
                    fc = sweeps[coarse][axis][(slice(None), *idx)]
                    ff = {}
                    for p in range(2):
                        for q in range(2):
                            fi = [0, 0, 0]
                            fi[axis] = 0 if side == 0 else B
                            fi[a1] = 2 + 2 * u + p
                            fi[a2] = 2 + 2 * v + q
                            ff[p, q] = sweeps[fine][axis][(slice(None), *fi)]
                    fm = ((ff[0, 0] + ff[0, 1]) + (ff[1, 0] + ff[1, 1])) * 0.25
                    total = acc.setdefault((coarse, tuple(c)), np.zeros(5))
                    for fld in range(5):
                        r = rows[fld]
                        d = dtd * (fc[r] - fm[r]) if side == 0 else dtd * (fm[r] - fc[r])
                        total[fld] = total[fld] + d

    def coarsen(self, buf, L, cells=True):
        """The covered cells := the mean of their children (``cells``) and the coarse
        edges := the mean of the two fine edges along them."""
        B = self.B
        level = self.t.blocks["level"]
        for row in self.t.children:
            if level[row[0]] != L:
                continue
            pc, pp = self.views(buf, row[0])
            kids = [self.views(buf, k) for k in row[1:]]
            # cells: the mean of 8 children
            for oct_ in range(8 if cells else 0):
                ox, oy, oz = (oct_ >> 2) & 1, (oct_ >> 1) & 1, oct_ & 1
                kc = kids[oct_][0]
                f = lambda a, b, c: kc[:, 2 + a : 2 + B : 2, 2 + b : 2 + B : 2, 2 + c : 2 + B : 2]  # noqa: E731
                s0 = (f(0, 0, 0) + f(0, 0, 1)) + (f(0, 1, 0) + f(0, 1, 1))
                s1 = (f(1, 0, 0) + f(1, 0, 1)) + (f(1, 1, 0) + f(1, 1, 1))
                sl = (slice(None), slice(2 + 8 * ox, 10 + 8 * ox), slice(2 + 8 * oy, 10 + 8 * oy),
                      slice(2 + 8 * oz, 10 + 8 * oz))
                pc[sl] = (s0 + s1) * 0.125
            # edges: the mean of the two fine edges, written to every coarse copy
            r = self.t.blocks[row[0]]
            nb = self.t.level_nb[L]
            bo = np.array([r["bx"], r["by"], r["bz"]]) * B
            for c in range(3):
                ext = tuple(B if a == c else B + 1 for a in range(3))
                q = np.array(np.unravel_index(np.arange(np.prod(ext)), ext))
                vals = []
                for s in (0, 1):
                    fc = 2 * q + (np.arange(3)[:, None] == c) * s
                    oc = np.where(fc >= 2 * B, 1, fc // B)
                    loc = fc - B * oc + 2
                    kid = (oc[0] * 2 + oc[1]) * 2 + oc[2]
                    v = np.empty(q.shape[1])
                    for k in range(8):
                        m = kid == k
                        v[m] = kids[k][1][c][loc[0, m], loc[1, m], loc[2, m]]
                    vals.append(v)
                value = 0.5 * (vals[0] + vals[1])
                g = (bo[:, None] + q) % (nb * B)
                for sh in np.ndindex(2, 2, 2):
                    blk = np.empty_like(g)
                    loc = np.empty_like(g)
                    ok = np.ones(q.shape[1], bool)
                    for a in range(3):
                        if sh[a] == 0:
                            blk[a], loc[a] = g[a] // B, g[a] % B + 2
                        else:
                            ok &= (a != c) & (g[a] % B == 0)
                            blk[a], loc[a] = (g[a] // B - 1) % nb, B + 2
                    ids = self.t.lookup[self.t.level_base[L] + (blk[0] * nb + blk[1]) * nb + blk[2]]
                    ok &= ids >= 0
                    for i in np.flatnonzero(ok):
                        buf[self.edge_index(ids[i], c, loc[0, i], loc[1, i], loc[2, i])] = value[i]

    def prepare(self):
        """As the model does after loading a state: coarsen, finest level first, then
        fill every ghost."""
        self.kept = {}
        for L in range(self.nlev - 2, -1, -1):
            self.coarsen(self.buf, L)
        self.fill(self.buf)

    # -- the step -----------------------------------------------------------------------------
    def block_rate(self, buf, blk):
        h, B = HALO, self.B
        cells, pots = self.views(buf, blk)
        spacing = self.spacing(blk)
        local = [p[tuple(slice(h, h + B + (k != c)) for k in range(3))] for c, p in enumerate(pots)]
        field = cell_fields(face_fields(local, spacing, self.b0))
        rho, mx, my, mz, energy = cells[:, h : h + B, h : h + B, h : h + B]
        vel = [mx / rho, my / rho, mz / rho]
        p = (self.gamma - 1.0) * (
            energy - 0.5 * rho * (vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2])
            - 0.5 * (field[0] * field[0] + field[1] * field[1] + field[2] * field[2]))
        rate = None
        for a in range(3):
            term = (np.abs(vel[a]) + fast_speed_np(rho, p, field[a], field[(a + 1) % 3],
                                                   field[(a + 2) % 3], self.gamma)) / spacing[a]
            rate = term if rate is None else rate + term
        return float(np.max(rate))

    def stable_step(self):
        largest = max(self.block_rate(self.buf, k) for k in range(len(self.t.blocks)))
        return self.cfl * (1.0 / largest)

    def step(self, dt=None):
        if dt is None:
            dt = self.stable_step()
        other = self.buf.copy()
        self.stage(self.buf, other, dt, False)
        self.stage(other, self.buf, dt, True)
        self.time += dt
        return dt

    def leaf_totals(self):
        cov = np.zeros(len(self.t.blocks), bool)
        if len(self.t.children):
            cov[self.t.children[:, 0]] = True
        h, B = HALO, self.B
        tot = np.zeros(5)
        for k in range(len(self.t.blocks)):
            if cov[k]:
                continue
            cells, _ = self.views(self.buf, k)
            sp = self.spacing(k)
            tot += cells[:, h : h + B, h : h + B, h : h + B].sum(axis=(1, 2, 3)) * (sp[0] * sp[1] * sp[2])
        return tot
