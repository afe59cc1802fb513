"""The hierarchy's field interface (patchwork's subset that the reference models and
tests use, mhd.py:467-477,703-777,1149-1194): ``allocate_field``,
``patch.field(name).local`` / ``.local_with_halos``, ``gather``, ``integral``,
``exchange_halos`` and ``restrict``.

CPU: the declarations, shapes and leaf masks, and the reference's resistive
energy-flux test (test_mhd3d.py:191-232) on the oracle.  GPU: the reference's
tests that write a model's state through ``patch.field(...).local`` and read it
with ``gather`` (test_mhd.py:240-246, test_mhd3d.py:183-232), and the views,
integrals, halos and restriction of a refined hierarchy against the models'
own device diagnostics.
"""

from __future__ import annotations

import numpy as np
import pytest

from wellbalanced_amd.hierarchy import AMRHierarchy, Patch

BOX = ((0.0, 1.0),) * 3


# -- declarations, shapes, leaf masks (CPU) -------------------------------------------
def test_allocate_field_rejects_bad_declarations():
    h = AMRHierarchy((8, 8), periodic=(True, True))
    h.allocate_field("rho", halo=2, split=None, conservative=True)
    with pytest.raises(ValueError, match="already allocated"):
        h.allocate_field("rho")
    with pytest.raises(ValueError, match="centering"):
        h.allocate_field("q", centering="face")
    with pytest.raises(ValueError, match="no z-edge"):
        h.allocate_field("q", centering="z-edge")
    with pytest.raises(ValueError, match="non-negative"):
        h.allocate_field("q", halo=-1)
    with pytest.raises(KeyError, match="allocate_field"):
        h.field(h.patches[0], "missing")


def test_haloed_shapes_follow_the_centering():
    h = AMRHierarchy((4, 5, 6), BOX)
    for name, centering in (("c", "cell"), ("n", "node"), ("ex", "x-edge"), ("ey", "y-edge"),
                            ("ez", "z-edge")):
        h.allocate_field(name, halo=2, centering=centering)
    p = h.patches[0]
    assert h._haloed_shape(p, h._specs["c"]) == (8, 9, 10)
    assert h._haloed_shape(p, h._specs["n"]) == (9, 10, 11)
    # an edge along x is cell-centred along x and nodal across it (mhd3d.py:80)
    assert h._haloed_shape(p, h._specs["ex"]) == (8, 10, 11)
    assert h._haloed_shape(p, h._specs["ey"]) == (9, 9, 11)
    assert h._haloed_shape(p, h._specs["ez"]) == (9, 10, 10)


def test_leaf_masks_exclude_the_cells_under_finer_patches():
    h = AMRHierarchy((8, 8), periodic=(True, True))
    fine = h.add_patch(1, ((4, 10), (2, 8)))
    h.add_patch(2, ((10, 14), (6, 10)))
    base = h._leaf_mask(h.patches[0])
    want = np.ones((8, 8), bool)
    want[2:5, 1:4] = False
    np.testing.assert_array_equal(base, want)
    mask = h._leaf_mask(fine)
    want = np.ones((6, 6), bool)
    want[1:3, 1:3] = False
    np.testing.assert_array_equal(mask, want)
    # leaf cells tile the domain once: their volumes sum to the domain's
    volume = sum(h._leaf_mask(p).sum() * np.prod(h.spacing(p.level)) for p in h.patches)
    assert volume == pytest.approx(1.0, abs=1e-15)


def test_patches_know_their_hierarchy_also_after_a_regrid():
    h = AMRHierarchy((16, 16), periodic=(True, True))
    h.add_patch(1, ((8, 16), (8, 16)))
    tags = {0: np.zeros((16, 16), bool), 1: np.zeros((32, 32), bool)}
    tags[0][2:5, 2:5] = True
    new = h.regrid(tags)
    assert all(p._hierarchy is new for p in new.patches)
    assert all(p._hierarchy is h for p in h.patches)
    with pytest.raises(ValueError, match="no hierarchy"):
        Patch(0, ((0, 16), (0, 16))).field("rho")


# -- the reference's resistive energy flux (test_mhd3d.py:191-232) ---------------------
def _sheared_problem(axis: int):
    """``B_axis = sin(2 pi s)``, ``s`` the next axis, in pressure balance (32 cells along
    ``s``): the set-up of test_mhd3d.py:191-211."""
    shape = [4, 4, 4]
    across = (axis + 1) % 3
    shape[across] = 32
    mesh = AMRHierarchy(tuple(shape), BOX, periodic=(True,) * 3)
    k = 2.0 * np.pi
    potential = [0.0, 0.0, 0.0]
    potential[(axis + 2) % 3] = lambda x, y, z: -np.cos(k * (x, y, z)[across]) / k
    state = dict(density=1.0, velocity=(0.0, 0.0, 0.0),
                 pressure=lambda x, y, z: 10.0 - 0.5 * np.sin(k * (x, y, z)[across]) ** 2,
                 vector_potential=tuple(potential))
    return mesh, state, across


def _energy_rate_check(energies, across, eta, duration):
    k = 2.0 * np.pi
    s = (np.arange(32) + 0.5) / 32
    expected = eta * k**2 * np.cos(2.0 * k * s)
    rate = (energies[0] - energies[1]) / duration
    rate = rate.mean(axis=tuple(a for a in range(3) if a != across))
    np.testing.assert_allclose(rate, expected, rtol=0.0, atol=0.03 * eta * k**2)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_the_resistive_energy_flux_is_eta_j_cross_b(axis):
    """On the oracle: at rest the energy changes by ``-div(eta J x B) = eta k^2 cos(2ks)``
    beyond what the ideal run changes (test_mhd3d.py:214-232)."""
    from oracle.mhd3d import Oracle3D
    from wellbalanced_amd import initial

    eta, duration = 1e-3, 0.002
    energies = []
    for resistivity in (eta, 0.0):
        mesh, state, across = _sheared_problem(axis)
        cons, pots = initial.host_state_3d(mesh, mesh.patches[0], 5.0 / 3.0, (0.0, 0.0, 0.0),
                                           **state)
        model = Oracle3D(cons, pots, mesh.spacing(0), resistivity=resistivity)
        start = model.interior()[4].copy()
        steps = 0
        while model.time < duration * (1.0 - 1e-14) and steps < 100:
            model.step(min(model.stable_step(), duration - model.time))
            steps += 1
        energies.append(model.interior()[4] - start)
    _energy_rate_check(energies, across, eta, duration)


# -- GPU: the reference's tests through patch.field / gather ---------------------------
@pytest.mark.gpu
def test_reports_a_negative_pressure(gpu):
    """test_mhd.py:240-246 as written: the energy planted through the field view."""
    from wellbalanced_amd import MHD2D

    model = MHD2D(AMRHierarchy((8, 8), periodic=(True, True)))
    model.set_state(1.0, (0.0, 0.0, 0.0), 1e-3)
    for patch in model.hierarchy.patches:
        patch.field(model.names["energy"]).local[3, 3] = -1.0
    with pytest.raises(FloatingPointError, match="pressure"):
        model.step(1e-3)


@pytest.mark.gpu
def test_reports_a_negative_pressure_3d(gpu):
    """test_mhd3d.py:183-188 as written."""
    from wellbalanced_amd import MHD3D

    model = MHD3D(AMRHierarchy((4, 4, 4), BOX, periodic=(True,) * 3))
    model.set_state(1.0, (0.0, 0.0, 0.0), 1.0)
    patch = model.hierarchy.patches[0]
    patch.field(model.names["energy"]).local[1, 1, 1] = -1.0
    with pytest.raises(FloatingPointError, match="non-positive"):
        model.step(1e-3)


@pytest.mark.gpu
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_gpu_the_resistive_energy_flux_is_eta_j_cross_b(gpu, axis):
    """test_mhd3d.py:214-232 as written: energies read with ``hierarchy.gather``."""
    from wellbalanced_amd import MHD3D

    eta, duration = 1e-3, 0.002
    energies = []
    for resistivity in (eta, 0.0):
        mesh, state, across = _sheared_problem(axis)
        model = MHD3D(mesh, resistivity=resistivity)
        model.set_state(**state)
        (patch,) = model.hierarchy.patches
        start = model.hierarchy.gather(patch, model.names["energy"])
        model.run(duration, max_steps=100)
        energies.append(model.hierarchy.gather(patch, model.names["energy"]) - start)
    _energy_rate_check(energies, across, eta, duration)


@pytest.mark.gpu
def test_views_gather_and_integral_of_a_uniform_model(gpu):
    from wellbalanced_amd import MHD2D

    model = MHD2D.orszag_tang(32, limiter="minmod")
    h, patch = model.hierarchy, model.hierarchy.patches[0]
    cons, az = model.arrays()
    for k, name in enumerate(model.names.values()):
        field = patch.field(name)
        assert tuple(field.local.shape) == (32, 32)
        assert tuple(field.local_with_halos.shape) == (36, 36)
        np.testing.assert_array_equal(h.gather(patch, name), cons[k, 2:-2, 2:-2])
    # the potential is nodal: 33 x 33 nodes (mhd.py:1149,1194)
    np.testing.assert_array_equal(h.gather(patch, model.potential), az[2:-2, 2:-2])
    totals = model.totals()
    for key, name in model.names.items():
        assert h.integral(name) == totals[key]
    # a write through the view is the model's state
    patch.field(model.names["bz"]).local[5, 7] = 0.25
    assert model.arrays()[0][5, 7, 9] == 0.25


@pytest.mark.gpu
def test_exchange_halos_fills_the_periodic_images(gpu):
    from wellbalanced_amd import MHD2D

    model = MHD2D.orszag_tang(16, limiter="minmod")
    h, patch = model.hierarchy, model.hierarchy.patches[0]
    rho = patch.field(model.names["density"])
    rho.local[0, 3] = 7.0
    rho.local[15, 4] = 9.0
    h.exchange_halos(model.names["density"])
    halo = rho.local_with_halos.cpu().numpy()
    assert halo[18, 5] == 7.0  # the image of cell (0, 3) beyond the upper x face
    assert halo[1, 6] == 9.0  # the image of cell (15, 4) below the lower x face


@pytest.mark.gpu
def test_a_refined_hierarchy_through_the_field_interface(gpu):
    from wellbalanced_amd import MHD2D
    from wellbalanced_amd.amr import AMRMHD2D

    h = AMRHierarchy((16, 16), periodic=(True, True))
    h.add_patch(1, ((8, 20), (10, 24)))
    model = MHD2D(h, limiter="minmod")
    assert isinstance(model, AMRMHD2D)
    model.set_state(lambda x, y: 1.0 + 0.2 * np.sin(2 * np.pi * x), (0.1, 0.0, 0.0), 1.0,
                    vector_potential=lambda x, y: 0.05 * np.cos(2 * np.pi * y))
    fine = h.patches[1]
    data = model.dataset(level=1)
    np.testing.assert_allclose(h.gather(fine, model.names["density"]),
                               data["density"][8:20, 10:24], rtol=0, atol=0)
    totals = model.totals()
    for key, name in model.names.items():
        assert h.integral(name) == totals[key]
    # a write on the fine patch, then restrict: the coarse cells under it hold the mean
    rho = fine.field(model.names["density"])
    rho.local[0:2, 0:2] = torch_tensor([[1.0, 2.0], [3.0, 5.0]])
    h.restrict(model.names["density"])
    coarse = h.gather(h.patches[0], model.names["density"])
    assert coarse[4, 5] == pytest.approx(2.75, abs=1e-15)
    # the halo fill of the refined hierarchy writes the fine patch's ghost cells
    halo = fine.field(model.names["density"]).local_with_halos
    halo[:2].fill_(-5.0)
    halo[-2:].fill_(-5.0)
    h.exchange_halos(model.names["density"])
    assert float(halo[:2].min()) > 0.0 and float(halo[-2:].min()) > 0.0
    # a regrid moves the fields onto the new hierarchy's patches, integrals kept
    before = h.integral(model.names["density"])
    model.regrid({0: np.pad(np.ones((4, 4), bool), 6), 1: np.zeros((32, 32), bool)})
    new = model.hierarchy
    assert new is not h
    assert new.integral(model.names["density"]) == pytest.approx(before, abs=1e-14)
    for patch in new.patches:
        assert new.gather(patch, model.names["density"]).shape == patch.shape


def torch_tensor(rows):
    import torch

    return torch.tensor(rows, dtype=torch.float64)


@pytest.mark.gpu
def test_fields_no_model_owns(gpu):
    h = AMRHierarchy((8, 8), periodic=(True, True))
    h.add_patch(1, ((4, 8), (4, 8)))
    h.allocate_field("tracer", halo=1)
    for patch in h.patches:
        field = patch.field("tracer")
        assert tuple(field.local.shape) == patch.shape
        field.local.fill_(1.0)
    # ones over the leaf cells integrate to the domain's area
    assert h.integral("tracer") == pytest.approx(1.0, abs=1e-15)
    with pytest.raises(NotImplementedError):
        h.exchange_halos("tracer")
