"""The refined 2D step on a patch hierarchy, in NumPy FP64 (oracle only).

Follows ``MHD2D.step`` (``mhd.py:965-994``) on every patch with the same halo fill,
fluxes and update as ``oracle.mhd2d``, plus the hierarchy operations the reference
delegates to patchwork: halo fill from same-level patches and the level below,
refluxing of coarse cells next to finer patches (the coarse face flux replaced by
the mean of the fine face fluxes), the average of covered coarse cells over their
fine children, and the restriction of A_z (``mhd.py:981,991``).  The geometry (which
cell is copied or interpolated from where) comes from the product's gather tables
(``wellbalanced_amd.amr_maps``); the arithmetic is restated here.  Parity with the
reference is unpinned beyond invariants (patchwork is not available): conservation,
composite div B, and a refined level covering the domain equal to a uniform run.
"""

from __future__ import annotations

import numpy as np

from . import kernels

HALO = 2
XR = (0, 1, 2, 3, 4, 6)
YR = (0, 2, 1, 3, 4, 6)


def node_interp_value(buf, base, sx, sy, kx, ky):
    """The 4-point node rule for one entry (the arithmetic order the kernel uses)."""
    w4 = (-0.0625, 0.5625, 0.5625, -0.0625)
    nx_ = 4 if kx else 1
    ny_ = 4 if ky else 1
    total = 0.0
    for j in range(ny_):
        row = 0.0
        for i in range(nx_):
            wi = w4[i] if kx else 1.0
            row = row + wi * buf[base + i * sx + j * sy]
        wj = w4[j] if ky else 1.0
        total = total + wj * row
    return total


class OracleAMR2D:
    """Heun steps of a periodic 2D hierarchy held in one flat buffer."""

    def __init__(self, maps, hierarchy, buf, *, gamma=5.0 / 3.0, mean_field=(0.0, 0.0),
                 limiter="minmod", cfl=0.4, resistivity=0.0, conduction=(0.0, 0.0)):
        self.maps = maps
        self.h = hierarchy
        self.buf = np.array(buf, dtype=np.float64, copy=True)
        self.gamma = float(gamma)
        self.b0 = tuple(float(b) for b in mean_field)
        self.code = kernels.LIMITER_CODES.index(limiter)
        self.cfl = float(cfl)
        self.resistivity = float(resistivity)
        self.conduction = (float(conduction[0]), float(conduction[1]))
        self.nlev = hierarchy.max_level + 1
        self.time = 0.0

    def views(self, buf, lay):
        cells = buf[lay.offset : lay.offset + 6 * lay.plane].reshape(6, lay.X, lay.Y)
        az = buf[lay.offset + 6 * lay.plane : lay.offset + lay.size].reshape(lay.X + 1, lay.Y + 1)
        return cells, az

    def fill(self, buf):
        for L in range(self.nlev):
            fill_ghosts_exact(buf, self.maps, L)

    def patch_fluxes(self, buf, lay):
        cells, az = self.views(buf, lay)
        dx, dy = self.h.spacing(lay.level)
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
        return fx, fy, ez

    def stage(self, src, dst, dt, heun):
        self.fill(src)
        fluxes = []
        for lay in self.maps.layouts:
            fx, fy, ez = self.patch_fluxes(src, lay)
            fluxes.append((fx, fy))
            cs, azs = self.views(src, lay)
            cd, azd = self.views(dst, lay)
            dx, dy = self.h.spacing(lay.level)
            dtdx, dtdy = dt / dx, dt / dy
            h, nx, ny = HALO, lay.nx, lay.ny
            for f in range(6):
                gx = fx[XR[f], :, h : h + ny]
                gy = fy[YR[f], h : h + nx, :]
                u = cs[f, h : h + nx, h : h + ny]
                v = u - (dtdx * (gx[1:] - gx[:-1]) + dtdy * (gy[:, 1:] - gy[:, :-1]))
                if heun:
                    v = (v + cd[f, h : h + nx, h : h + ny]) * 0.5
                cd[f, h : h + nx, h : h + ny] = v
            a = azs[h : h + nx + 1, h : h + ny + 1] - dt * ez
            if heun:
                a = (a + azd[h : h + nx + 1, h : h + ny + 1]) * 0.5
            azd[h : h + nx + 1, h : h + ny + 1] = a
        for L in range(self.nlev - 2, -1, -1):
            self.reflux(dst, fluxes, L, dt, heun)
            self.average(dst, L)
            self.restrict(dst, L)

    def reflux(self, dst, fluxes, level, dt, heun):
        rows = self.maps.reflux.get(level)
        if rows is None or len(rows) == 0:
            return
        lays = self.maps.layouts
        dx, dy = self.h.spacing(level)
        dtd = (dt / dx, dt / dy)
        # group the faces of each coarse cell in table order
        i = 0
        while i < len(rows):
            q, ca, cb = rows[i, 0], rows[i, 1], rows[i, 2]
            j = i
            delta = np.zeros(6)
            while j < len(rows) and tuple(rows[j, :3]) == (q, ca, cb):
                axis, side, fa, fb = rows[j, 3], rows[j, 4], rows[j, 5], rows[j, 6]
                fc = fluxes[q][axis]
                fine = []
                for p, pa, pb in (rows[j, 7:10], rows[j, 10:13]):
                    fine.append(fluxes[p][axis])
                for f in range(6):
                    r = (XR if axis == 0 else YR)[f]
                    cflux = _face(fc, axis, r, fa, fb)
                    f0 = _face(fine[0], axis, r, rows[j, 8], rows[j, 9])
                    f1 = _face(fine[1], axis, r, rows[j, 11], rows[j, 12])
                    mean = 0.5 * (f0 + f1)
                    # side 0: the face is the cell's high face; side 1: its low face
                    d = dtd[axis] * (cflux - mean) if side == 0 else dtd[axis] * (mean - cflux)
                    delta[f] = delta[f] + d
                j += 1
            lay = lays[q]
            cells, _ = self.views(dst, lay)
            for f in range(6):
                cells[f, ca, cb] = cells[f, ca, cb] + (0.5 * delta[f] if heun else delta[f])
            i = j

    def average(self, dst, level):
        rows = self.maps.average.get(level)
        if rows is None:
            return
        for coarse, pc, fine, fy, pf in rows:
            for f in range(6):
                b = fine + f * pf
                dst[coarse + f * pc] = ((dst[b] + dst[b + fy]) + (dst[b + 1] + dst[b + fy + 1])) * 0.25

    def restrict(self, dst, level):
        rows = self.maps.restrict.get(level)
        if rows is None or len(rows) == 0:
            return
        dst[rows[:, 0]] = dst[rows[:, 1]]

    def rate(self):
        largest = 0.0
        for lay in self.maps.layouts:
            cells, az = self.views(self.buf, lay)
            dx, dy = self.h.spacing(lay.level)
            r, _ = kernels.rate(*cells, az, HALO, dx, dy, *self.b0, self.gamma)
            largest = max(largest, r)
        return largest

    def stable_step(self):
        """``mhd.py:931-963`` over the patches: the fast-wave limit of each and, with a
        diffusivity, its diffusion limit with the patch's smallest density."""
        smallest = np.inf
        for lay in self.maps.layouts:
            cells, az = self.views(self.buf, lay)
            dx, dy = self.h.spacing(lay.level)
            r, rho_min = kernels.rate(*cells, az, HALO, dx, dy, *self.b0, self.gamma)
            smallest = min(smallest, 1.0 / r)
            diffusivity = max(self.resistivity,
                              (self.gamma - 1.0) * max(self.conduction) / float(rho_min))
            if diffusivity > 0.0:
                smallest = min(smallest, 0.5 / (diffusivity * (dx**-2 + dy**-2)))
        return self.cfl * float(smallest)

    def step(self, dt=None):
        if dt is None:
            dt = self.stable_step()
        other = self.buf.copy()
        self.stage(self.buf, other, dt, False)
        self.stage(other, self.buf, dt, True)
        self.time += dt
        return dt

    def leaf_totals(self):
        """Integrals over the leaf cells (covered coarse cells excluded)."""
        covered = {}
        for L in range(self.nlev - 1):
            for coarse, pc, fine, fy, pf in self.maps.average[L]:
                covered[int(coarse)] = True
        totals = np.zeros(6)
        for lay in self.maps.layouts:
            cells, _ = self.views(self.buf, lay)
            dx, dy = self.h.spacing(lay.level)
            inner = cells[:, HALO : HALO + lay.nx, HALO : HALO + lay.ny]
            a, b = np.meshgrid(np.arange(lay.nx) + HALO, np.arange(lay.ny) + HALO, indexing="ij")
            offs = lay.cell(a, b)
            mask = np.array([int(o) not in covered for o in offs.ravel()]).reshape(offs.shape)
            totals += (inner * mask).sum(axis=(1, 2)) * (dx * dy)
        return totals


def _face(flux, axis, row, fa, fb):
    """The flux of local face (fa, fb) (x-face fa = low face of cell fa) from a sweep
    array whose face k lies between cells k+1 and k+2."""
    if axis == 0:
        return flux[row, fa - 2, fb]
    return flux[row, fa, fb - 2]


def fill_ghosts_exact(buf: np.ndarray, maps, level: int) -> None:
    """Halo fill with the arithmetic order of the GPU gather kernels."""
    dst, src, pd, ps = maps.copy_cells[level]
    for f in range(6):
        buf[dst + f * pd] = buf[src + f * ps]
    d, s = maps.copy_nodes[level]
    buf[d] = buf[s]
    dst, pd, c00, cy, pc, wx1, wy1 = maps.cell_interp[level]
    wx0, wy0 = 1.0 - wx1, 1.0 - wy1
    for f in range(6):
        b = c00 + f * pc
        buf[dst + f * pd] = (wy0 * (wx0 * buf[b] + wx1 * buf[b + cy])
                             + wy1 * (wx0 * buf[b + 1] + wx1 * buf[b + cy + 1]))
    dst, base, sx, sy, kx, ky = maps.node_interp[level]
    for n in range(len(dst)):
        buf[dst[n]] = node_interp_value(buf, base[n], sx[n], sy[n], kx[n], ky[n])


def _minmod(a, b):
    return np.where(a * b > 0.0, np.where(np.abs(a) < np.abs(b), a, b), 0.0)


def regrid_state(old_buf, new_maps, tables):
    """The new hierarchy's buffer after a regrid, from the old buffer and the transfer
    tables of ``wellbalanced_amd.regrid.transfer_tables``: per level, coarse to fine,
    the same-level copies, the conservative minmod prolongation of the new cells, the
    bilinear new nodes and the ghost fill; then average-down and restriction."""
    buf = np.zeros(new_maps.total)
    nlev = max(lay.level for lay in new_maps.layouts) + 1
    for L in range(nlev):
        t = tables[L]
        cd, cs, cpd, cps = t["copy_cells"]
        for f in range(6):
            buf[cd + f * cpd] = old_buf[cs + f * cps]
        nd, ns = t["copy_nodes"]
        buf[nd] = old_buf[ns]
        d, pd, c, cy, pc, sx, sy = t["cells"]
        hx, hy = 0.25 * sx.astype(np.float64), 0.25 * sy.astype(np.float64)
        for f in range(6):
            q = c + f * pc
            c0 = buf[q]
            slx = _minmod(buf[q + cy] - c0, c0 - buf[q - cy])
            sly = _minmod(buf[q + 1] - c0, c0 - buf[q - 1])
            buf[d + f * pd] = (c0 + hx * slx) + hy * sly
        qd, qb, qs, kx, ky = t["nodes"]
        last = len(buf) - 1
        a0 = buf[qb]
        ax = buf[np.minimum(qb + qs, last)]
        ay = buf[np.minimum(qb + 1, last)]
        axy = buf[np.minimum(qb + qs + 1, last)]
        both = 0.25 * ((a0 + ax) + (ay + axy))
        buf[qd] = np.where((kx == 1) & (ky == 1), both,
                           np.where(kx == 1, 0.5 * (a0 + ax), np.where(ky == 1, 0.5 * (a0 + ay), a0)))
        fill_ghosts_exact(buf, new_maps, L)
    return buf
