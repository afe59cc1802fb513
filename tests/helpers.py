"""Shared test helpers: oracle models built from the same host state the GPU gets."""

from __future__ import annotations

import numpy as np

from oracle.mhd2d import Oracle2D
from oracle.mhd3d import Oracle3D
from wellbalanced_amd import initial


def options2d(problem, **over):
    opts = {**problem.options, **over}
    walls = [False, False]
    for name, kind in (opts.get("boundaries") or {}).items():
        walls["xy".index(name)] = kind == "wall"
    return dict(
        gamma=opts.get("gamma", 5.0 / 3.0),
        mean_field=opts.get("mean_field", (0.0, 0.0)),
        limiter=opts.get("limiter", "vanleer"),
        cfl=opts.get("cfl", 0.4),
        walls=tuple(walls),
        resistivity=opts.get("resistivity", 0.0),
        conduction=tuple(opts.get("conduction", (0.0, 0.0))),
    )


def oracle2d(problem, **over) -> Oracle2D:
    h = problem.hierarchy
    o = options2d(problem, **over)
    cons, az = initial.host_state_2d(h, h.patches[0], o["gamma"], o["mean_field"], **problem.state)
    return Oracle2D(cons, az, h.spacing(0), periodic=h.periodic, **o)


def options3d(problem, **over):
    opts = {**problem.options, **over}
    return dict(
        gamma=opts.get("gamma", 5.0 / 3.0),
        mean_field=opts.get("mean_field", (0.0, 0.0, 0.0)),
        limiter=opts.get("limiter", "vanleer"),
        cfl=opts.get("cfl", 0.3),
        resistivity=opts.get("resistivity", 0.0),
    )


def oracle3d(problem, **over) -> Oracle3D:
    h = problem.hierarchy
    o = options3d(problem, **over)
    cons, pots = initial.host_state_3d(h, h.patches[0], o["gamma"], o["mean_field"], **problem.state)
    return Oracle3D(cons, pots, h.spacing(0), periodic=h.periodic, **o)


def block(seed=1, shape=(13, 9)):
    """The seeded primitive block of the reference's kernel test (test_mhd_kernels.py:17-33)."""
    rng = np.random.default_rng(seed)
    nx, ny = shape
    prim = rng.normal(size=(8, nx, ny))
    prim[0] = rng.uniform(0.5, 2.0, shape)
    prim[4] = rng.uniform(0.1, 1.0, shape)
    prim[1, :, 6] += 30.0
    prim[2, 8, :] -= 30.0
    prim[4, 5, 3] = 1.0e-3
    bxf = rng.normal(size=(nx + 1, ny))
    byf = rng.normal(size=(nx, ny + 1))
    bxf[4, 2] = byf[2, 4] = 0.0
    return prim, bxf, byf


def conserved(seed=3, shape=(13, 9)):
    """Conserved fields and nodal A_z with a positive pressure (test_mhd_kernels.py:95-105)."""
    rng = np.random.default_rng(seed)
    nx, ny = shape
    rho = rng.uniform(0.5, 2.0, shape)
    momenta = [rng.normal(size=shape) * rho for _ in range(3)]
    bz = rng.normal(size=shape)
    az = 0.1 * rng.normal(size=(nx + 1, ny + 1))
    energy = 20.0 + rng.uniform(size=shape)
    return [rho, *momenta, energy, bz], az


def state_of(fields):
    """Primitive rows (rho, u, v, w, p, bx, by, bz) to conserved rows, gamma 5/3."""
    g = 5.0 / 3.0
    rho, u, v, w, p, bx, by, bz = fields
    energy = p / (g - 1.0) + 0.5 * rho * (u * u + v * v + w * w) + 0.5 * (bx * bx + by * by + bz * bz)
    return rho, rho * u, rho * v, rho * w, energy


def amr_oracle(h, state, *, gamma=5.0 / 3.0, mean_field=(0.0, 0.0), **options):
    """The refined 2D oracle of hierarchy ``h`` from the host state the GPU gets, the
    covered cells averaged and A_z restricted as ``set_state`` does (mhd.py:683-737)."""
    from oracle.amr2d import OracleAMR2D, pack

    patches = [(p.level, p.box) for p in h.patches]
    per = [initial.host_state_2d(h, p, gamma, mean_field, **state) for p in h.patches]
    o = OracleAMR2D(patches, h.base_shape, h.bounds, pack(per, patches), periodic=h.periodic,
                    gamma=gamma, mean_field=mean_field, **options)
    o.average_down(o.state)
    o.restrict(o.state)
    return o


def composite_mismatch(o) -> float:
    """The largest difference of a coarse A_z node and the fine node on it (every
    coarse patch's own nodes), over the oracle's patches."""
    worst = 0.0
    geo = o.geo
    for q, (L, ((i0, i1), (j0, j1))) in enumerate(o.patches):
        if L + 1 >= o.nlev:
            continue
        for i in range(i0, i1 + 1):
            for j in range(j0, j1 + 1):
                src = geo.node_of(L + 1, 2 * i, 2 * j)
                if src is None:
                    continue
                p, fa, fb = src
                worst = max(worst, abs(o.state[q][1][i - i0 + 2, j - j0 + 2] - o.state[p][1][fa, fb]))
    return worst
