"""The reference's ``MHD3D.step`` on one uniform patch, in NumPy FP64 (oracle only).

State layout (the same as the product's), halo 2 on every side, ``X = nx+4``:
``cons (5, X, Y, Z)`` = density, momentum x/y/z, energy; the edge potential
``ax (X, Y+1, Z+1)``, ``ay (X+1, Y, Z+1)``, ``az (X+1, Y+1, Z)``.  Follows
``mhd3d.py:361-386`` (face and cell fields), ``:448-473`` (halos), ``:475-509``
(primitive), ``:511-528`` (sweeps), ``:530-559`` (edge field), ``:604-671``
(fluxes, stage), ``:674-724`` (stable step) and ``:726-754`` (Heun step).
"""

from __future__ import annotations

import numpy as np

from . import boundaries, kernels

HALO = 2


def cyclic(axis):
    """``mhd3d.py:82-84``."""
    return axis, (axis + 1) % 3, (axis + 2) % 3


def face_fields(potential, spacing, b0):
    """``B_a = B0_a + d_b A_c - d_c A_b`` (``mhd3d.py:361-378``)."""
    faces = []
    for a in range(3):
        _, b, c = cyclic(a)
        faces.append(
            np.diff(potential[c], axis=b) / spacing[b]
            - np.diff(potential[b], axis=c) / spacing[c]
            + b0[a]
        )
    return faces


def cell_fields(faces):
    """Mean of the two faces of each cell along its axis (``mhd3d.py:380-386``)."""
    out = []
    for a, f in enumerate(faces):
        lo = np.take(f, np.arange(f.shape[a] - 1), axis=a)
        hi = np.take(f, np.arange(1, f.shape[a]), axis=a)
        out.append(0.5 * (lo + hi))
    return out


def _along(array, axis, start, stop):
    index = [slice(None)] * array.ndim
    index[axis] = slice(start, stop)
    return array[tuple(index)]


def _pair_mean(array, axis, start, count):
    """Means of points k and k+1 along ``axis``, k = start .. start+count-1 (``mhd3d.py:97-102``)."""
    return 0.5 * (_along(array, axis, start, start + count) + _along(array, axis, start + 1, start + 1 + count))


def _field_at_face(face, normal, node, shape):
    """A haloed face field at the block's faces normal to ``normal``: the mean of the four
    faces around each, cells along ``normal`` first, then nodes along ``node``
    (``mhd3d.py:105-117``)."""
    h = HALO
    face = _pair_mean(face, normal, h - 1, shape[normal] + 1)
    face = _pair_mean(face, node, h, shape[node])
    other = 3 - normal - node
    return _along(face, other, h, h + shape[other])


def resistive3d(faces, shape, spacing, eta, edges, energy):
    """``eta J`` into the edge fields and ``eta J x B`` into the energy fluxes of the block
    faces, in place (``mhd3d.py:561-602``): ``J_c = d_a B_b - d_b B_a`` on the block's
    c-edges from the haloed face fields; on a face normal to s the energy flux gains
    ``eta (J_b B_c - J_c B_b)``, J the mean of the face's two edges along each direction,
    B the mean of the four faces around it."""
    h = HALO
    current = []
    for c in range(3):
        _, a, b = cyclic(c)
        parts = []
        for axis, field, sign in ((a, faces[b], 1.0), (b, faces[a], -1.0)):
            derivative = np.diff(field, axis=axis) / spacing[axis]
            index = [slice(h, h + shape[i]) for i in range(3)]
            index[a] = slice(h, h + shape[a] + 1)
            index[b] = slice(h, h + shape[b] + 1)
            index[axis] = slice(h - 1, h + shape[axis])
            parts.append(sign * derivative[tuple(index)])
        current.append(parts[0] + parts[1])
    for c in range(3):
        edges[c] += eta * current[c]
    for s in range(3):
        _, b, c = cyclic(s)
        jb = _pair_mean(current[b], c, 0, shape[c])
        jc = _pair_mean(current[c], b, 0, shape[b])
        energy[s] += eta * (jb * _field_at_face(faces[c], s, c, shape)
                            - jc * _field_at_face(faces[b], s, b, shape))


def fast_speed_np(rho, p, bn, bt1, bt2, gamma):
    """The NumPy fast speed of ``mhd3d.py:109-114`` (``bn**2`` grouped)."""
    a2 = gamma * p / rho
    b2 = (bn * bn + bt1 * bt1 + bt2 * bt2) / rho
    root = np.sqrt(np.maximum((a2 + b2) * (a2 + b2) - 4.0 * a2 * (bn * bn) / rho, 0.0))
    return np.sqrt(0.5 * (a2 + b2 + root))


class Oracle3D:
    """One patch covering a uniform 3D level."""

    def __init__(self, cons, potential, spacing, *, gamma=5.0 / 3.0,
                 mean_field=(0.0, 0.0, 0.0), limiter="vanleer", cfl=0.3,
                 periodic=(True, True, True), resistivity=0.0):
        self.cons = np.array(cons, dtype=np.float64, copy=True)
        self.pot = [np.array(p, dtype=np.float64, copy=True) for p in potential]
        self.n = tuple(s - 2 * HALO for s in self.cons.shape[1:])
        X, Y, Z = self.cons.shape[1:]
        assert self.pot[0].shape == (X, Y + 1, Z + 1)
        assert self.pot[1].shape == (X + 1, Y, Z + 1)
        assert self.pot[2].shape == (X + 1, Y + 1, Z)
        self.spacing = tuple(float(s) for s in spacing)
        self.gamma = float(gamma)
        self.b0 = tuple(float(b) for b in mean_field)
        self.code = kernels.LIMITER_CODES.index(limiter)
        self.cfl = float(cfl)
        self.periodic = tuple(bool(p) for p in periodic)
        self.resistivity = float(resistivity)
        self.time = 0.0
        self.steps = 0

    def fill_halos(self):
        centring = [tuple(a == k for k in range(3)) for a in range(3)]
        boundaries.fill(
            list(self.cons), list(zip(self.pot, centring)), self.n, self.periodic,
            (False, False, False), (1, 2, 3),
        )

    def primitive(self):
        """``mhd3d.py:475-509``: prim of the haloed block, faces; checks the inner block."""
        faces = face_fields(self.pot, self.spacing, self.b0)
        bx, by, bz = cell_fields(faces)
        rho, mx, my, mz, energy = self.cons
        with np.errstate(all="ignore"):
            u, v, w = mx / rho, my / rho, mz / rho
            p = (self.gamma - 1.0) * (
                energy - 0.5 * rho * (u * u + v * v + w * w) - 0.5 * (bx * bx + by * by + bz * bz)
            )
        prim = np.stack((rho, u, v, w, p, bx, by, bz))
        h = HALO
        inner = prim[:, h:-h, h:-h, h:-h]
        if not (np.all(np.isfinite(inner)) and np.all(inner[0] > 0) and np.all(inner[4] > 0)):
            raise FloatingPointError(
                f"non-positive density or pressure on the level-0 patch at t = {self.time}"
            )
        return prim, faces

    def fluxes(self):
        """Sweeps and CT-contact edge field (``mhd3d.py:511-559,604-633``)."""
        prim, faces = self.primitive()
        sweeps = [kernels.sweep(prim, faces[s], s, self.code, self.gamma) for s in range(3)]
        h = HALO
        edges = []
        for c in range(3):
            _, a, b = cyclic(c)
            cell = prim[5 + a] * prim[1 + b] - prim[5 + b] * prim[1 + a]
            order = (c, a, b)
            fa = np.transpose(sweeps[a][[0, 5]], (0, *(1 + k for k in order)))
            fb = np.transpose(sweeps[b][[0, 6]], (0, *(1 + k for k in order)))
            cellp = np.transpose(cell, order)
            nc = self.n[c]
            sl = slice(h, h + nc)
            e = kernels.ct_contact(fa[1, sl], fb[1, sl], fa[0, sl], fb[0, sl], cellp[sl])
            edges.append(np.transpose(e, np.argsort(order)))
        self._faces = faces
        return sweeps, edges

    def advance(self, dt):
        """One forward-Euler stage (``mhd3d.py:635-671``) with this package's update form."""
        self.fill_halos()
        sweeps, edges = self.fluxes()
        h = HALO
        nx, ny, nz = self.n
        rates = [dt / s for s in self.spacing]
        block = []
        for s in range(3):
            idx = [slice(None)] + [slice(h, h + self.n[k]) if k != s else slice(None) for k in range(3)]
            block.append(sweeps[s][tuple(idx)])
        if self.resistivity:  # mhd3d.py:631-632
            energy = [np.array(block[s][4]) for s in range(3)]
            resistive3d(self._faces, self.n, self.spacing, self.resistivity, edges, energy)
            for s in range(3):
                block[s] = block[s].copy()
                block[s][4] = energy[s]
        # rows of each axis' flux for each conserved field (``mhd3d.py:620-631``)
        for field in range(5):
            div = None
            for s in range(3):
                if field == 0:
                    row = 0
                elif field == 4:
                    row = 4
                else:
                    row = 1 + (field - 1 - s) % 3
                f = block[s][row]
                d = np.diff(f, axis=s)
                term = rates[s] * d
                div = term if div is None else div + term
            u = self.cons[field, h : h + nx, h : h + ny, h : h + nz]
            u[...] = u - div
        for c in range(3):
            idx = tuple(slice(h, h + self.n[k] + (k != c)) for k in range(3))
            a = self.pot[c][idx]
            a[...] = a - dt * edges[c]

    def stable_step(self):
        """``mhd3d.py:674-724``."""
        return self.cfl * self.smallest_time()

    def smallest_time(self):
        """The smallest of 1 / (largest summed fast-wave rate) and, with a resistivity, the
        diffusion limit: the stable step before the factor ``cfl``."""
        h = HALO
        nx, ny, nz = self.n
        local = []
        for c in range(3):
            idx = tuple(slice(h, h + self.n[k] + (k != c)) for k in range(3))
            local.append(self.pot[c][idx])
        field = cell_fields(face_fields(local, self.spacing, self.b0))
        rho, mx, my, mz, energy = self.cons[:, h : h + nx, h : h + ny, h : h + nz]
        vel = [mx / rho, my / rho, mz / rho]
        p = (self.gamma - 1.0) * (
            energy
            - 0.5 * rho * (vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2])
            - 0.5 * (field[0] * field[0] + field[1] * field[1] + field[2] * field[2])
        )
        rate = None
        for a in range(3):
            term = (
                np.abs(vel[a])
                + fast_speed_np(rho, p, field[a], field[(a + 1) % 3], field[(a + 2) % 3], self.gamma)
            ) / self.spacing[a]
            rate = term if rate is None else rate + term
        smallest = 1.0 / float(np.max(rate))
        if self.resistivity:  # mhd3d.py:717-721
            smallest = min(smallest, 0.5 / (self.resistivity * sum(d**-2 for d in self.spacing)))
        return smallest

    def step(self, dt=None):
        """Heun (SSP-RK2) step (``mhd3d.py:726-754``)."""
        if dt is None:
            dt = self.stable_step()
        h = HALO
        cells = (slice(None),) + tuple(slice(h, h + n) for n in self.n)
        pidx = [tuple(slice(h, h + self.n[k] + (k != c)) for k in range(3)) for c in range(3)]
        start_c = self.cons[cells].copy()
        start_p = [self.pot[c][pidx[c]].copy() for c in range(3)]
        self.advance(dt)
        self.advance(dt)
        c = self.cons[cells]
        c += start_c
        c *= 0.5
        for k in range(3):
            a = self.pot[k][pidx[k]]
            a += start_p[k]
            a *= 0.5
        self.time += dt
        self.steps += 1
        return dt

    def interior(self):
        h = HALO
        return self.cons[(slice(None),) + tuple(slice(h, h + n) for n in self.n)]

    def totals(self):
        sx, sy, sz = self.spacing
        return self.interior().sum(axis=(1, 2, 3)) * (sx * sy * sz)

    def divergence(self):
        """``mhd3d.py:823-841``."""
        h = HALO
        local = []
        for c in range(3):
            idx = tuple(slice(h, h + self.n[k] + (k != c)) for k in range(3))
            local.append(self.pot[c][idx])
        faces = face_fields(local, self.spacing, self.b0)
        div = sum(np.diff(f, axis=a) / self.spacing[a] for a, f in enumerate(faces))
        return float(np.max(np.abs(div))) * min(self.spacing)
