"""The reference's ``MHD2D.step`` on one uniform patch, in NumPy FP64 (oracle only).

State layout (the same as the product's): ``cons (6, nx+4, ny+4)`` = density,
momentum x/y/z, energy, B_z with a halo of 2, and ``az (nx+5, ny+5)`` on the
nodes.  Follows ``mhd.py:744-777`` (halos), ``:800-887`` (primitive, sweeps,
CT-contact, walls), ``:889-927`` (stage), ``:930-963`` (stable step) and
``:965-994`` (Heun step).
"""

from __future__ import annotations

import numpy as np

from . import boundaries, kernels

HALO = 2
#: Flux rows of each conserved field in the x- and y-sweeps (``mhd.py:134``).
FLUX_ROWS = ((0, 1, 2, 3, 4, 6), (0, 2, 1, 3, 4, 6))
#: Rows zeroed at a wall face (``mhd.py:136``).
WALL_ROWS = [0, 2, 3, 4, 6]


class Oracle2D:
    """One patch covering a uniform 2D level."""

    def __init__(self, cons, az, spacing, *, gamma=5.0 / 3.0, mean_field=(0.0, 0.0),
                 limiter="vanleer", cfl=0.4, periodic=(True, True), walls=(False, False),
                 resistivity=0.0, conduction=(0.0, 0.0)):
        self.cons = np.array(cons, dtype=np.float64, copy=True)
        self.az = np.array(az, dtype=np.float64, copy=True)
        self.nx = self.cons.shape[1] - 2 * HALO
        self.ny = self.cons.shape[2] - 2 * HALO
        assert self.az.shape == (self.nx + 5, self.ny + 5)
        self.dx, self.dy = (float(s) for s in spacing)
        self.gamma = float(gamma)
        self.b0 = (float(mean_field[0]), float(mean_field[1]))
        self.code = kernels.LIMITER_CODES.index(limiter)
        self.cfl = float(cfl)
        self.periodic = tuple(bool(p) for p in periodic)
        self.walls = tuple(bool(w) for w in walls)
        self.resistivity = float(resistivity)
        self.conduction = (float(conduction[0]), float(conduction[1]))
        self.time = 0.0
        self.steps = 0

    # -- pieces --------------------------------------------------------------
    def fill_halos(self):
        boundaries.fill(
            list(self.cons), [(self.az, (False, False))], (self.nx, self.ny),
            self.periodic, self.walls, (1, 2),
        )

    def fluxes(self):
        """``_patch_fluxes`` (``mhd.py:844-887``): fx, fy, ez of the haloed block."""
        prim, bxf, byf, ecc, bad = kernels.primitive(
            *self.cons, self.az, self.dx, self.dy, *self.b0, self.gamma
        )
        if bad:
            raise FloatingPointError(
                f"non-positive density or pressure on the level-0 patch at t = {self.time}"
            )
        fx = kernels.sweep(prim, bxf, 0, self.code, self.gamma)
        fy = kernels.sweep(prim, byf, 1, self.code, self.gamma)
        ez = kernels.ct_contact(fx[5], fy[5], fx[0], fy[0], ecc)
        # mhd.py:868-871: the heat flux, then the resistive terms, before the walls
        if max(self.conduction) > 0.0:
            kernels.conductive(prim, bxf, byf, self.dx, self.dy, *self.conduction, fx, fy)
        if self.resistivity:
            kernels.resistive(prim, bxf, byf, self.dx, self.dy, self.resistivity, fx, fy, ez)
        if self.walls[0]:
            ez[0, :] = 0.0
            ez[-1, :] = 0.0
            fx[WALL_ROWS, 0, :] = 0.0
            fx[WALL_ROWS, -1, :] = 0.0
        if self.walls[1]:
            ez[:, 0] = 0.0
            ez[:, -1] = 0.0
            fy[WALL_ROWS, :, 0] = 0.0
            fy[WALL_ROWS, :, -1] = 0.0
        return fx, fy, ez

    def advance(self, dt):
        """One forward-Euler stage (``mhd.py:889-927``) with this package's update form."""
        self.fill_halos()
        fx, fy, ez = self.fluxes()
        h, nx, ny = HALO, self.nx, self.ny
        dtdx, dtdy = dt / self.dx, dt / self.dy
        for field, (xr, yr) in enumerate(zip(*FLUX_ROWS)):
            gx = fx[xr, :, h : h + ny]
            gy = fy[yr, h : h + nx, :]
            u = self.cons[field, h : h + nx, h : h + ny]
            u[...] = u - (dtdx * (gx[1:] - gx[:-1]) + dtdy * (gy[:, 1:] - gy[:, :-1]))
        a = self.az[h : h + nx + 1, h : h + ny + 1]
        a[...] = a - dt * ez

    def stable_step(self):
        """``mhd.py:930-963`` with ``mhd_rate``: the fast-wave limit and, with a
        diffusivity ``D = max(eta, (gamma - 1) kappa_max / rho_min)``, the explicit
        diffusion limit ``0.5 / (D (dx^-2 + dy^-2))``."""
        largest, rho_min = kernels.rate(
            *self.cons, self.az, HALO, self.dx, self.dy, *self.b0, self.gamma
        )
        smallest = 1.0 / largest
        diffusivity = max(self.resistivity,
                          (self.gamma - 1.0) * max(self.conduction) / float(rho_min))
        if diffusivity > 0.0:
            smallest = min(smallest, 0.5 / (diffusivity * (self.dx**-2 + self.dy**-2)))
        return self.cfl * smallest

    def step(self, dt=None):
        """Heun (SSP-RK2) step (``mhd.py:965-994``); returns ``dt``."""
        if dt is None:
            dt = self.stable_step()
        h, nx, ny = HALO, self.nx, self.ny
        cells = (slice(None), slice(h, h + nx), slice(h, h + ny))
        nodes = (slice(h, h + nx + 1), slice(h, h + ny + 1))
        start_c = self.cons[cells].copy()
        start_a = self.az[nodes].copy()
        self.advance(dt)
        self.advance(dt)
        c = self.cons[cells]
        c += start_c
        c *= 0.5
        a = self.az[nodes]
        a += start_a
        a *= 0.5
        self.time += dt
        self.steps += 1
        return dt

    def run(self, duration, max_steps=1_000_000):
        """``mhd.py:996-1008``."""
        end = self.time + duration
        taken = 0
        while self.time < end * (1.0 - 1e-14) and taken < max_steps:
            dt = min(self.stable_step(), end - self.time)
            self.step(dt)
            taken += 1
        return taken

    # -- diagnostics ---------------------------------------------------------
    def interior(self):
        h = HALO
        return self.cons[:, h : h + self.nx, h : h + self.ny]

    def nodes(self):
        h = HALO
        return self.az[h : h + self.nx + 1, h : h + self.ny + 1]

    def totals(self):
        return self.interior().sum(axis=(1, 2)) * (self.dx * self.dy)

    def divergence(self):
        """``mhd.py:1114-1131``."""
        az = self.nodes()
        bxf = (az[:, 1:] - az[:, :-1]) / self.dy + self.b0[0]
        byf = -(az[1:, :] - az[:-1, :]) / self.dx + self.b0[1]
        div = (bxf[1:] - bxf[:-1]) / self.dx + (byf[:, 1:] - byf[:, :-1]) / self.dy
        return float(np.max(np.abs(div))) * min(self.dx, self.dy)
