"""Refined 2D hierarchies on bounded domains: outflow and wall faces on every level
(``MHD2D._fill_halos``, ``mhd.py:744-777``; the wall fluxes, ``mhd.py:873-886``).
The oracle's invariants on the CPU, and the GPU path against the oracle (FP64: bit
for bit)."""

import numpy as np
import pytest

from tests.helpers import amr_oracle, composite_mismatch, oracle2d
from wellbalanced_amd import initial
from wellbalanced_amd.hierarchy import AMRHierarchy

B0 = (0.3, 0.0)


def _state():
    """A bump of density and pressure near the lower-left corner moving into it; A_z is
    zero along y = 0 and y = 1, so the walls start with no normal field."""
    bump = lambda x, y: np.exp(-((x - 0.22) ** 2 + (y - 0.2) ** 2) / 0.01)  # noqa: E731
    return dict(
        density=lambda x, y: 1.0 + 0.4 * bump(x, y),
        velocity=(-0.4, -0.3, 0.1),
        pressure=lambda x, y: 1.0 + 0.5 * bump(x, y),
        vector_potential=lambda x, y: 0.025 * np.cos(2 * np.pi * x) * (1.0 - np.cos(2 * np.pi * y)),
        bz=lambda x, y: 0.1 * np.cos(np.pi * x),
    )


def _hierarchy(periodic_x=False, levels=3):
    """32^2 base on the unit square; level-1 patches on the corners and one face, a
    level-2 patch in the lower-left corner (nested across the periodic seam when x is
    periodic)."""
    h = AMRHierarchy((32, 32), ((0.0, 1.0), (0.0, 1.0)), periodic=(periodic_x, False))
    h.add_patch(1, ((0, 24), (0, 24)))
    h.add_patch(1, ((24, 40), (0, 16)))
    h.add_patch(1, ((40, 64), (36, 64)))
    h.add_patch(1, ((56, 64), (0, 24)))
    if levels > 2:
        h.add_patch(2, ((0, 32), (0, 20)))
    return h


def _boundaries(periodic_x):
    return {"y": "wall"} if periodic_x else {"x": "outflow", "y": "wall"}


def _oracle(h, periodic_x, state=None, limiter="minmod"):
    walls = (False, True)
    return amr_oracle(h, state or _state(), mean_field=B0, limiter=limiter, walls=walls)


def _wall_rows(o):
    """A_z on the wall node rows (y = 0 and y = 1) of every patch touching them."""
    out = []
    for (L, box), (_, az) in zip(o.patches, o.state):
        n = 32 << L
        if box[1][0] == 0:
            out.append(az[2:-2, 2].copy())
        if box[1][1] == n:
            out.append(az[2:-2, -3].copy())
    return out


# -- the oracle ----------------------------------------------------------------------------
def test_a_level_covering_a_bounded_domain_is_the_uniform_run():
    """Four level-1 patches over the whole domain, outflow in x and walls in y, evolve
    bit for bit as one uniform 32^2 grid with those faces: the domain-face rules on
    patches (corners included) are the uniform grid's."""
    h = AMRHierarchy((16, 16), ((0.0, 1.0), (0.0, 1.0)), periodic=(False, False))
    for box in [((0, 16), (0, 16)), ((16, 32), (0, 16)), ((0, 16), (16, 32)), ((16, 32), (16, 32))]:
        h.add_patch(1, box)
    hu = AMRHierarchy((32, 32), ((0.0, 1.0), (0.0, 1.0)), periodic=(False, False))
    opts = {"limiter": "minmod", "mean_field": B0, "boundaries": {"x": "outflow", "y": "wall"}}
    uniform = oracle2d(initial.Problem(hu, opts, _state()))
    amr = _oracle(h, False)
    for _ in range(4):
        dt = uniform.stable_step()
        uniform.step(dt)
        amr.step(dt)
    for (L, ((i0, i1), (j0, j1))), (cells, az) in list(zip(amr.patches, amr.state))[1:]:
        np.testing.assert_array_equal(cells[:, 2:-2, 2:-2], uniform.interior()[:, i0:i1, j0:j1])
        np.testing.assert_array_equal(az[2:-2, 2:-2], uniform.nodes()[i0 : i1 + 1, j0 : j1 + 1])


def test_walls_on_every_level_conserve_and_keep_the_wall_field():
    """Periodic in x, walls in y: mass, x- and z-momentum, energy and B_z of the leaf
    cells kept to 1e-13, E_z zero along the walls (A_z on the wall rows unchanged, so no
    field crosses them), composite div B zero."""
    h = _hierarchy(periodic_x=True)
    amr = _oracle(h, True)
    before = amr.leaf_totals()
    rows = _wall_rows(amr)
    for _ in range(6):
        amr.step()
    after = amr.leaf_totals()
    for f in (0, 1, 3, 4, 5):
        assert abs(after[f] - before[f]) < 1e-13 * max(1.0, abs(before[f])), f
    for a, b in zip(rows, _wall_rows(amr)):
        np.testing.assert_array_equal(a, b)
    assert composite_mismatch(amr) == 0.0
    assert amr.divergence() < 1e-14


def test_outflow_and_walls_on_refined_corners_keep_div_b_zero():
    h = _hierarchy(periodic_x=False)
    amr = _oracle(h, False)
    rows = _wall_rows(amr)
    for _ in range(6):
        amr.step()
    assert composite_mismatch(amr) == 0.0
    assert amr.divergence() < 1e-14
    for a, b in zip(rows, _wall_rows(amr)):
        np.testing.assert_array_equal(a, b)


def _emulate_fill(maps, buf, walls):
    """The device fill of every level on the host: the gather tables as k_fill_level
    reads them (copies, bilinear cells, quadratic nodes), then the domain-face rows
    with the boundary rules."""
    from oracle import boundaries
    from oracle.amr2d import pack, unpack

    patches = [(lay.level, lay.box) for lay in maps.layouts]
    for L in sorted(maps.copy_cells):
        cd, cs, cpd, cps = maps.copy_cells[L]
        vals = [buf[cs + f * cps] for f in range(6)]
        for f in range(6):
            buf[cd + f * cpd] = vals[f]
        nd, ns = maps.copy_nodes[L]
        buf[nd] = buf[ns]
        dst, pd, c00, cy, pc, wx1, wy1 = maps.cell_interp[L]
        wx0, wy0 = 1.0 - wx1, 1.0 - wy1
        for f in range(6):
            q = c00 + f * pc
            buf[dst + f * pd] = (wy0 * (wx0 * buf[q] + wx1 * buf[q + cy])
                                 + wy1 * (wx0 * buf[q + 1] + wx1 * buf[q + cy + 1]))
        jd, jb, jsx, jsy, jkx, jky = maps.node_interp[L]
        w3 = (-0.125, 0.75, 0.375)
        for i in range(len(jd)):
            total = 0.0
            for j in range(3 if jky[i] else 1):
                row = 0.0
                for k in range(3 if jkx[i] else 1):
                    row = row + (w3[k] if jkx[i] else 1.0) * buf[jb[i] + k * jsx[i] + j * jsy[i]]
                total = total + (w3[j] if jky[i] else 1.0) * row
            buf[jd[i]] = total
        state = unpack(buf, patches)
        for k, faces in maps.domain_faces[L]:
            cells, az = state[k]
            for axis in range(2):
                lower, upper = bool(faces & (1 << 2 * axis)), bool(faces & (2 << 2 * axis))
                if not (lower or upper):
                    continue
                if walls[axis]:
                    for f in range(6):
                        boundaries.wall_cells(cells[f], axis, lower, upper,
                                              -1.0 if f == 1 + axis else 1.0)
                    boundaries.wall_nodes(az, axis, lower, upper)
                else:
                    for f in range(6):
                        boundaries.outflow_cells(cells[f], axis, lower, upper)
                    boundaries.outflow_nodes(az, axis, lower, upper)
        buf[:] = pack(state, patches)
    return buf


@pytest.mark.parametrize("periodic_x", [False, True])
def test_the_fill_tables_fill_the_ghosts_as_the_oracle(periodic_x):
    """The host-built gather tables of a bounded hierarchy (ghosts outside the domain
    left to the boundary rules, stencils reaching into a coarse patch's halo outside
    the domain, nodes 0 .. n on a non-periodic axis), applied level by level as the
    device applies them, give every ghost cell and node the oracle's value."""
    from oracle.amr2d import pack
    from wellbalanced_amd import amr_maps

    h = _hierarchy(periodic_x)
    amr = _oracle(h, periodic_x)
    maps = amr_maps.build(h)
    assert sum(len(r) for r in maps.domain_faces.values()) > 0
    got = _emulate_fill(maps, amr.buf.copy(), (False, True))
    amr.fill(amr.state)
    np.testing.assert_array_equal(got, pack(amr.state, amr.patches))


# -- GPU ----------------------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("periodic_x", [False, True])
@pytest.mark.parametrize("kernel,tile,nxs", [("march", 30, 8), ("march", 62, 16),
                                             ("tiles", 62, 16)])
def test_gpu_bounded_hierarchy_is_bitwise_the_oracle(gpu, periodic_x, kernel, tile, nxs):
    from wellbalanced_amd.amr import AMRMHD2D

    h = _hierarchy(periodic_x)
    amr = _oracle(h, periodic_x)
    model = AMRMHD2D(h, limiter="minmod", mean_field=B0, boundaries=_boundaries(periodic_x),
                     tile=tile, nxs=nxs, kernel=kernel)
    model.set_state(**_state())
    np.testing.assert_array_equal(model.buffer(), amr.buf)
    for _ in range(4):
        dt = amr.stable_step()
        assert model.stable_step() == dt
        model.step(dt)
        amr.step(dt)
    from tests.test_amr import _assert_interiors_equal

    _assert_interiors_equal(model.buffer(), amr)


@pytest.mark.gpu
def test_gpu_bounded_hierarchy_mixed_conserves_and_stays_near_fp64(gpu):
    """MHD2D on a refined hierarchy with walls returns the refined model; the mixed
    path at the FP64 step keeps the wall-conserved totals and div B."""
    from wellbalanced_amd import MHD2D
    from wellbalanced_amd.amr import AMRMHD2D

    h = _hierarchy(periodic_x=True)
    models = [MHD2D(h, limiter="minmod", mean_field=B0, boundaries={"y": "wall"},
                    precision=p) for p in ("fp64", "mixed")]
    assert all(isinstance(m, AMRMHD2D) for m in models)
    for m in models:
        m.set_state(**_state())
    before = models[1].leaf_totals()
    for _ in range(10):
        dt = models[0].stable_step()
        for m in models:
            m.step(dt)
    after = models[1].leaf_totals()
    for f in (0, 1, 3, 4, 5):
        assert abs(after[f] - before[f]) < 1e-13 * max(1.0, abs(before[f])), f
    assert models[1].composite_divergence() == 0.0
    assert models[1].divergence() < 1e-13
    a, b = models[0].dataset(0)["density"], models[1].dataset(0)["density"]
    assert float(np.mean(np.abs(a - b))) < 1e-5


@pytest.mark.gpu
def test_gpu_regrid_on_a_bounded_hierarchy_is_refused(gpu):
    from wellbalanced_amd.amr import AMRMHD2D

    model = AMRMHD2D(_hierarchy(), mean_field=B0, boundaries={"x": "outflow", "y": "wall"})
    model.set_state(**_state())
    with pytest.raises(NotImplementedError):
        model.regrid({0: np.zeros((32, 32), bool)})
