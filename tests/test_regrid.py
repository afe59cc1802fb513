"""Regridding a periodic 2D hierarchy on tags (MHD2D.regrid, mhd.py:1043-1056): the
clustering and nesting of the new patches, the move of the state (integrals kept,
div B zero, linear data exact) on the CPU, and the device transfer against the
oracle's restatement plus the reference's own regrid test (test_mhd_amr.py:69-90) on
the GPU."""

from __future__ import annotations

import numpy as np
import pytest

from oracle.amr2d import OracleAMR2D, regrid_state, unpack
from tests.helpers import amr_oracle, composite_mismatch
from wellbalanced_amd import amr_maps, initial
from wellbalanced_amd import regrid as rg
from wellbalanced_amd.hierarchy import AMRHierarchy


def _boxes_cover(boxes, flags):
    cov = np.zeros(flags.shape, int)
    for (i0, i1), (j0, j1) in boxes:
        cov[i0:i1, j0:j1] += 1
    return cov


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_clusters_cover_every_tag_once_within_the_size_limit(seed):
    rng = np.random.default_rng(seed)
    flags = np.zeros((48, 40), bool)
    for _ in range(5):
        i, j = rng.integers(0, 40, 2)
        flags[i : i + rng.integers(1, 9), j : j + rng.integers(1, 12)] = True
    flags |= rng.random(flags.shape) < 0.01
    boxes = rg.cluster(flags, efficiency=0.7, max_size=8)
    cov = _boxes_cover(boxes, flags)
    assert np.all(cov[flags] == 1) and cov.max() == 1
    assert all(i1 - i0 <= 8 and j1 - j0 <= 8 for (i0, i1), (j0, j1) in boxes)


def test_a_full_block_is_one_box_and_two_blobs_are_two():
    flags = np.zeros((32, 32), bool)
    flags[4:10, 6:12] = True
    assert rg.cluster(flags) == [((4, 10), (6, 12))]
    flags[20:24, 20:30] = True
    assert rg.cluster(flags) == [((4, 10), (6, 12)), ((20, 24), (20, 30))]


def _nested(h, nest=rg.NEST):
    """Every patch of level L + 1, its footprint widened by ``nest`` cells of level L,
    lies on level L."""
    for L in range(1, h.max_level + 1):
        shape = h.level_shape(L - 1)
        cover = np.zeros(shape, bool)
        for p in h.patches_at_level(L - 1):
            (i0, i1), (j0, j1) = p.box
            cover[i0:i1, j0:j1] = True
        for p in h.patches_at_level(L):
            (i0, i1), (j0, j1) = p.box
            foot = np.zeros(shape, bool)
            foot[i0 // 2 : i1 // 2, j0 // 2 : j1 // 2] = True
            if L - 1 > 0 and not cover[rg.dilate(foot, nest)].all():
                return False
    return True


def _two_level_regrid():
    h = AMRHierarchy((16, 16), periodic=(True, True))
    tags = {0: np.zeros((16, 16), bool)}
    tags[0][6:9, 5:12] = True
    h1 = rg.regrid(h, tags, max_patch_size=16)
    tags1 = {0: tags[0], 1: np.zeros((32, 32), bool)}
    tags1[1][14:16, 12:18] = True
    tags1[1][20, 3] = True  # outside level 1: the nesting grows level 1 around it
    h2 = rg.regrid(h1, tags1, max_patch_size=16)
    return h, h1, h2, tags1


def test_regridded_levels_nest_and_build_tables():
    _, h1, h2, tags1 = _two_level_regrid()
    assert h1.max_level == 1 and h2.max_level == 2
    assert _nested(h2)
    amr_maps.build(h2)
    cover = np.zeros(h2.level_shape(2), bool)
    for p in h2.patches_at_level(2):
        (i0, i1), (j0, j1) = p.box
        cover[i0:i1, j0:j1] = True
    fine_tags = np.repeat(np.repeat(tags1[1], 2, axis=0), 2, axis=1)
    assert cover[fine_tags].all()
    assert all(p.shape[0] <= 16 and p.shape[1] <= 16 for p in h2.patches if p.level > 0)


def _state(kind):
    if kind == "linear":  # reproduced exactly by the limited linear prolongation
        return dict(density=lambda x, y: 2.0 + 0.25 * x - 0.5 * y, velocity=(0.0, 0.0, 0.0),
                    pressure=1.0, vector_potential=lambda x, y: 0.3 * x - 0.2 * y + 0.1 * x * y,
                    bz=lambda x, y: 0.1 * x)
    return dict(density=lambda x, y: 1.0 + 0.3 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y),
                velocity=lambda x, y: (np.sin(2 * np.pi * y), 0.5 * np.cos(2 * np.pi * x), 0.1),
                pressure=lambda x, y: 1.0 + 0.2 * np.cos(2 * np.pi * (x + y)),
                vector_potential=lambda x, y: 0.05 * np.cos(2 * np.pi * x) * np.sin(4 * np.pi * y),
                bz=lambda x, y: 0.2 * np.sin(2 * np.pi * x))


def _oracle_on(h, state):
    return amr_oracle(h, state)


def _patches(h):
    return [(p.level, p.box) for p in h.patches]


def test_the_move_keeps_the_integrals_and_div_b():
    _, h1, h2, _ = _two_level_regrid()
    before = _oracle_on(h1, _state("smooth"))
    before.step()
    after = regrid_state(before, _patches(h2))
    np.testing.assert_allclose(after.leaf_totals(), before.leaf_totals(), rtol=0, atol=1e-15)
    # div B on every patch and coarse nodes equal to the fine nodes on them
    assert after.divergence() < 1e-14
    assert composite_mismatch(after) == 0.0


def test_linear_data_moves_exactly_onto_new_patches():
    """A linear density and a bilinear potential are reproduced on the new level 2 to
    round-off: the prolongation and the node interpolation are exact for them."""
    _, h1, h2, _ = _two_level_regrid()
    before = _oracle_on(h1, _state("linear"))
    moved = regrid_state(before, _patches(h2))
    exact = _oracle_on(h2, _state("linear"))
    for (L, _), (c1, a1), (c2, a2) in zip(moved.patches, moved.state, exact.state):
        if L != 2:
            continue
        np.testing.assert_allclose(c1[0, 2:-2, 2:-2], c2[0, 2:-2, 2:-2], rtol=0, atol=1e-14)
        np.testing.assert_allclose(a1[2:-2, 2:-2], a2[2:-2, 2:-2], rtol=0, atol=1e-14)


def test_a_planted_prolongation_slope_error_breaks_the_linear_move(monkeypatch):
    """The linear-data check pins the prolongation: a child on the wrong side of its
    coarse cell (the slope's sign flipped) is caught."""
    import oracle.amr2d as OA

    _, h1, h2, _ = _two_level_regrid()
    before = _oracle_on(h1, _state("linear"))
    monkeypatch.setattr(OA, "_minmod", lambda a, b: -np.where(a * b > 0.0, np.where(
        np.abs(a) < np.abs(b), a, b), 0.0))
    moved = regrid_state(before, _patches(h2))
    exact = _oracle_on(h2, _state("linear"))
    worst = max(np.abs(c1[0, 2:-2, 2:-2] - c2[0, 2:-2, 2:-2]).max()
                for (L, _), (c1, _a1), (c2, _a2) in zip(moved.patches, moved.state, exact.state)
                if L == 2)
    assert worst > 1e-3


# -- GPU ----------------------------------------------------------------------------------
KW = 2.0 * np.pi


def _orszag_tang_state():
    """The Orszag-Tang state of test_mhd_amr.py:52-66."""
    return dict(
        density=25.0 / (36.0 * np.pi),
        velocity=(lambda x, y: -np.sin(KW * y), lambda x, y: np.sin(KW * x), 0.0),
        pressure=5.0 / (12.0 * np.pi),
        vector_potential=lambda x, y: (np.cos(2 * KW * x) / (2 * KW) + np.cos(KW * y) / KW)
        / np.sqrt(4 * np.pi),
        bz=lambda x, y: 0.1 * np.cos(KW * x),
    )


def _amr_orszag_tang(n, **options):
    from wellbalanced_amd.amr import AMRMHD2D

    h = AMRHierarchy((n, n), ((0.0, 1.0), (0.0, 1.0)), periodic=(True, True))
    model = AMRMHD2D(h, **{"limiter": "vanleer", **options})
    model.set_state(**_orszag_tang_state())
    return model


@pytest.mark.gpu
def test_gpu_move_is_bitwise_the_oracle(gpu):
    """Two regrids of a stepped state (new level 1, then level 2 and a changed level 1):
    the device copies, prolongations, ghost fills and coarsening give the oracle's bits."""
    model = _amr_orszag_tang(16, limiter="minmod")
    for tags in ({0: np.pad(np.ones((4, 7), bool), ((6, 6), (5, 4)))},
                 {0: np.pad(np.ones((3, 6), bool), ((7, 6), (4, 6))),
                  1: np.pad(np.ones((2, 6), bool), ((14, 16), (12, 14)))}):
        model.step(model.stable_step())
        old_h, old_buf = model.hierarchy, model.buffer()
        old_oracle = OracleAMR2D(_patches(old_h), old_h.base_shape, old_h.bounds, old_buf,
                                 limiter="minmod")
        tags = {L: t for L, t in tags.items() if L <= old_h.max_level}
        model.regrid(tags, max_patch_size=16)
        want = regrid_state(old_oracle, _patches(model.hierarchy))
        got = unpack(model.buffer(), want.patches)
        for (cells, az), (oc, oa) in zip(got, want.state):  # the patches' own cells and nodes
            np.testing.assert_array_equal(cells[:, 2:-2, 2:-2], oc[:, 2:-2, 2:-2])
            np.testing.assert_array_equal(az[2:-2, 2:-2], oa[2:-2, 2:-2])
    assert model.hierarchy.max_level == 2


@pytest.mark.gpu
def test_gpu_regrids_conserve_and_keep_the_field_divergence_free(gpu):
    """test_mhd_amr.py:69-90: three regrids on the current, four steps after each."""
    model = _amr_orszag_tang(16, resistivity=1e-3)
    start = model.totals()
    for _ in range(3):
        before = model.totals()
        model.regrid(model.current_tags(0.1, levels=2), max_patch_size=16)
        after = model.totals()
        for name in before:
            assert abs(after[name] - before[name]) < 1e-14
        assert model.composite_divergence() < 1e-12
        for _ in range(4):
            model.step()
    assert model.hierarchy.max_level == 2
    end = model.totals()
    for name in start:
        assert abs(end[name] - start[name]) < 1e-13
    assert model.divergence() < 1e-12
    assert model.composite_divergence() < 1e-12
    leaf = model.leaf_values()
    assert np.all(leaf["pressure"] > 0) and np.all(leaf["density"] > 0)
    assert np.unique(leaf["level"]).max() == 2


@pytest.mark.gpu
def test_gpu_leaf_values_of_one_level_are_the_uniform_model(gpu):
    """test_mhd_amr.py:93-105 against the uniform model's dataset."""
    from wellbalanced_amd import MHD2D

    amr = _amr_orszag_tang(16)
    uni = MHD2D(AMRHierarchy((16, 16), ((0.0, 1.0), (0.0, 1.0)), periodic=(True, True)))
    uni.set_state(**_orszag_tang_state())
    leaf = amr.leaf_values()
    data = uni.dataset()
    order = np.lexsort((leaf["y"], leaf["x"]))
    np.testing.assert_allclose(leaf["density"][order], data["density"].ravel(), rtol=1e-14)
    np.testing.assert_allclose(leaf["magnetic_field_y"][order], data["magnetic_field_y"].ravel(),
                               rtol=1e-14, atol=1e-15)
    assert np.all(leaf["level"] == 0)


@pytest.mark.gpu
def test_gpu_diagnostics_match_the_host_sums(gpu):
    """The device totals are the leaf sums to round-off; the leaf cells tile the domain."""
    model = _amr_orszag_tang(16, limiter="minmod")
    model.regrid({0: np.pad(np.ones((4, 7), bool), ((6, 6), (5, 4)))}, max_patch_size=16)
    model.step(model.stable_step())
    leaf = model.leaf_values()
    h = model.hierarchy
    area = np.array([np.prod(h.spacing(L)) for L in leaf["level"]])
    assert abs(area.sum() - 1.0) < 1e-14
    totals = model.totals()
    assert abs(totals["density"] - np.sum(leaf["density"] * area)) < 1e-14


# -- reconnection diagnostics (mhd.py:1231-1329) --------------------------------------------
def test_harris_sheets_balance_pressure_and_reverse_the_field():
    h = AMRHierarchy((32, 16), bounds=((0, 4), (0, 2)), periodic=(True, True))
    state = initial.harris_sheets(h, width=0.1, perturbation=0.0)
    y = np.linspace(0.0, 2.0, 401)
    eps = 1e-6
    bx = (state["vector_potential"](0.3, y + eps) - state["vector_potential"](0.3, y - eps)) / (2 * eps)
    want = np.tanh((y - 0.5) / 0.1) - np.tanh((y - 1.5) / 0.1) - 1.0
    np.testing.assert_allclose(bx, want, atol=1e-7)
    np.testing.assert_allclose(state["pressure"](0.3, y) + 0.5 * want**2, 0.6, atol=1e-12)
    assert abs(bx[0] + 1.0) < 2e-4 and abs(bx[200] - 1.0) < 2e-4  # -1 outside, +1 between


class _Leafy:
    """A stand-in model: a hierarchy and the time (what the diagnostics read)."""

    def __init__(self, h):
        self.hierarchy = h
        self.time = 0.0


def test_sheet_diagnostics_count_the_turns_of_b_y_and_its_flux():
    h = AMRHierarchy((64, 32), bounds=((0, 4), (0, 2)), periodic=(True, True))
    x, y = h.cell_centers(h.patches[0])
    x, y = x.ravel(), y.ravel()
    from wellbalanced_amd.diagnostics import sheet_diagnostics

    leaf = {"x": x, "y": y, "level": np.zeros(x.size, np.int64),
            "magnetic_field_y": np.sin(2 * np.pi * x / 4.0), "current_z": np.cos(x),
            "density": np.ones(x.size)}
    d = sheet_diagnostics(_Leafy(h), leaf)
    assert (d["o_points"], d["x_points"]) == (1, 1)
    dx = 4.0 / 64
    xs = (np.arange(64) + 0.5) * dx
    assert d["psi"] == pytest.approx(0.5 * np.sum(np.abs(np.sin(2 * np.pi * xs / 4.0))) * dx, rel=1e-14)
    assert d["leaf_cells"] == 64 * 32


@pytest.mark.gpu
def test_gpu_reconnection_on_patches_follows_the_uniform_mesh(gpu):
    """test_mhd_amr.py:150-178: the double Harris sheet on 32 x 16 cells refined once
    around the sheets (regridded every 5 steps) against the uniform 64 x 32 mesh."""
    from wellbalanced_amd import MHD2D
    from wellbalanced_amd.amr import AMRMHD2D
    from wellbalanced_amd.diagnostics import sheet_diagnostics

    def run(shape, levels):
        mesh = AMRHierarchy(shape, bounds=((0, 4), (0, 2)), periodic=(True, True))
        state = initial.harris_sheets(mesh, width=0.1, perturbation=0.05)
        if levels:
            model = AMRMHD2D(mesh, resistivity=4e-3, limiter="vanleer")
        else:
            model = MHD2D(mesh, resistivity=4e-3, kernels="fused")
        model.set_state(**state)
        for _ in range(levels):
            model.regrid(model.current_tags(0.5, levels=levels), buffer=1)
            model.set_state(**state)
        first = sheet_diagnostics(model)
        while model.time < 0.2:
            if levels and model.steps % 5 == 0:
                model.regrid(model.current_tags(0.5, levels=levels), buffer=1)
            model.step(min(model.stable_step(), 0.2 - model.time))
        return first, sheet_diagnostics(model), model

    first, refined, model = run((32, 16), 1)
    start, uniform, _ = run((64, 32), 0)
    assert model.hierarchy.max_level == 1
    assert refined["leaf_cells"] < 0.75 * uniform["leaf_cells"]
    assert (first["o_points"], first["x_points"]) == (1, 1)
    assert abs(first["psi"] - start["psi"]) < 1e-3 * start["psi"]
    change = abs(uniform["psi"] - start["psi"])
    assert change > 1e-2 * start["psi"]
    assert abs(refined["psi"] - uniform["psi"]) < 0.05 * change


@pytest.mark.gpu
def test_gpu_uniform_diagnostics_are_the_dataset(gpu):
    """MHD2D's device leaf values, totals and divergence against its host dataset."""
    from wellbalanced_amd import MHD2D

    model = MHD2D.from_problem(initial.orszag_tang(24), kernels="fused")
    model.step(model.stable_step())
    leaf, data = model.leaf_values(), model.dataset()
    np.testing.assert_array_equal(leaf["density"], data["density"].ravel())
    np.testing.assert_allclose(leaf["pressure"], data["pressure"].ravel(), rtol=1e-13)
    cons, _ = model.arrays()
    inner = cons[:, 2:-2, 2:-2]
    area = model.dx * model.dy
    for k, (name, v) in enumerate(model.totals().items()):
        assert abs(v - inner[k].sum() * area) < 1e-14 * max(1.0, abs(v))
    assert model.divergence() < 1e-14


@pytest.mark.gpu
def test_gpu_mhd2d_on_refined_patches_is_the_hierarchy_model(gpu):
    """MHD2D(hierarchy with patches) gives the hierarchy model, with the reference's
    defaults (van Leer), run() to an end time and the per-level dataset."""
    from wellbalanced_amd import MHD2D
    from wellbalanced_amd.amr import AMRMHD2D

    h = AMRHierarchy((16, 16), ((0.0, 1.0), (0.0, 1.0)), periodic=(True, True))
    h.add_patch(1, ((8, 24), (8, 24)))
    model = MHD2D(h, resistivity=1e-3)
    assert isinstance(model, AMRMHD2D) and model.limiter == "vanleer"
    model.set_state(**_orszag_tang_state())
    before = model.totals()
    assert model.run(0.02) > 0 and abs(model.time - 0.02) < 1e-15
    after = model.totals()
    for name in before:
        assert abs(after[name] - before[name]) < 1e-13
    fine = model.dataset(1)
    assert np.isfinite(fine["density"][8:24, 8:24]).all() and np.isnan(fine["density"][0, 0])
    assert np.isfinite(model.dataset(0)["pressure"]).all()


@pytest.mark.gpu
def test_gpu_hierarchy_and_3d_bad_cells_raise_the_reference_error(gpu):
    """A non-positive pressure in a stage raises FloatingPointError (mhd.py:852-857) on
    the refined 2D hierarchy and in 3D, as in 2D."""
    from wellbalanced_amd import MHD3D
    from wellbalanced_amd.amr import AMRMHD2D

    h = AMRHierarchy((16, 16), ((0.0, 1.0), (0.0, 1.0)), periodic=(True, True))
    h.add_patch(1, ((8, 24), (8, 24)))
    model = AMRMHD2D(h)
    model.set_state(**_orszag_tang_state())
    buf = model.buffer()
    lay = model.layouts[1]
    buf[lay.offset + 4 * lay.plane + 5 * lay.Y + 5] = -1.0  # energy of one fine cell
    model.load_buffer(buf)
    with pytest.raises(FloatingPointError, match="pressure"):
        model.step(1e-4)
    m3 = MHD3D.from_problem(initial.turbulence_3d(8), kernels="fused")
    cons, pots = m3.arrays()
    cons[4, 4, 4, 4] = -1.0
    m3.load_arrays(cons, pots)
    with pytest.raises(FloatingPointError, match="pressure"):
        m3.step(1e-4)
