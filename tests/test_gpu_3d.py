"""GPU parity of the 3D path against the CPU oracle."""

import numpy as np
import pytest

from tests.helpers import oracle3d
from wellbalanced_amd import MHD2D, MHD3D, initial

pytestmark = pytest.mark.gpu


def _compare(model, oracle, rtol=0.0):
    cons, pots = model.arrays()
    h = 2
    nx, ny, nz = model.n
    got = cons[:, h : h + nx, h : h + ny, h : h + nz]
    want = oracle.interior()
    if rtol == 0.0:
        np.testing.assert_array_equal(got, want)
        for c in range(3):
            idx = tuple(slice(h, h + model.n[k] + (k != c)) for k in range(3))
            np.testing.assert_array_equal(pots[c][idx], oracle.pot[c][idx])
    else:
        scale = np.abs(want).max(axis=(1, 2, 3), keepdims=True)
        assert np.all(np.abs(got - want) <= rtol * scale)


@pytest.mark.parametrize("limiter", ["minmod", "vanleer", "mc", "none"])
def test_turbulence_split_steps_are_bitwise_the_oracle(gpu, limiter):
    problem = initial.turbulence_3d(12, seed=5, limiter=limiter)
    model = MHD3D.from_problem(problem)
    oracle = oracle3d(problem)
    for _ in range(3):
        dt = oracle.stable_step()
        assert model.stable_step() == dt
        model.step(dt)
        oracle.step(dt)
    _compare(model, oracle)


def test_alfven_wave_matches_the_oracle(gpu):
    problem = initial.alfven_wave_3d(6)
    model = MHD3D.from_problem(problem)
    oracle = oracle3d(problem)
    for _ in range(4):
        dt = oracle.stable_step()
        model.step(dt)
        oracle.step(dt)
    _compare(model, oracle)


def test_outflow_faces_match_the_oracle(gpu):
    from wellbalanced_amd.hierarchy import AMRHierarchy

    h = AMRHierarchy((8, 6, 7), ((0.0, 1.0),) * 3, periodic=(False, True, False))
    problem = initial.Problem(h, {"mean_field": (0.3, -0.2, 0.5), "limiter": "mc"}, dict(
        density=lambda x, y, z: 1.0 + 0.2 * np.sin(3 * x + y) * np.cos(2 * z),
        velocity=(0.1, -0.2, 0.3), pressure=0.7,
        vector_potential=(0.0, lambda x, y, z: 0.02 * np.sin(4 * z), lambda x, y, z: 0.03 * np.cos(3 * x))))
    model = MHD3D.from_problem(problem)
    oracle = oracle3d(problem)
    for _ in range(3):
        dt = oracle.stable_step()
        model.step(dt)
        oracle.step(dt)
    _compare(model, oracle)


def test_a_state_uniform_along_z_is_the_2d_solution(gpu):
    flat = MHD3D.orszag_tang(16, layers=2)
    plane = MHD2D.orszag_tang(16)
    for _ in range(10):
        dt = plane.stable_step()
        plane.step(dt)
        flat.step(dt)
    solid, sheet = flat.dataset(), plane.dataset()
    for name in ("density", "pressure", "velocity_x", "magnetic_field_y"):
        np.testing.assert_allclose(solid[name], np.repeat(sheet[name][:, :, None], 2, axis=2), atol=1e-13)
    assert np.abs(solid["velocity_z"]).max() < 1e-15
    assert np.abs(solid["magnetic_field_z"]).max() < 1e-15


def test_turbulence_256_conserves_and_keeps_div_b_zero(gpu):
    """The 256^3 workload: size-independent properties after 3 steps."""
    model = MHD3D.turbulence(256)
    before = model.totals()
    for _ in range(3):
        model.step()
    after = model.totals()
    for name in before:
        assert abs(after[name] - before[name]) < 1e-12 * max(1.0, abs(before[name]))
    assert model.divergence() < 1e-13
