"""The reference's refinement estimator (current_tags): the oracle against closed
forms (CPU), and the GPU tags against the oracle, bit for bit."""

import numpy as np
import pytest

from oracle import tags as T
from tests.helpers import oracle2d, oracle3d
from wellbalanced_amd import initial
from wellbalanced_amd.hierarchy import AMRHierarchy


def _quadratic_2d(c, n=16, **options):
    """A_z = c (x^2 + y^2): a uniform J_z = -4c (outflow faces)."""
    hier = AMRHierarchy((n, n), ((0.0, 1.0), (0.0, 1.0)), periodic=(False, False))
    state = dict(density=1.0, velocity=(0.0, 0.0, 0.0), pressure=1.0,
                 vector_potential=lambda x, y: c * (x * x + y * y))
    return initial.Problem(hier, {"limiter": "minmod", **options}, state)


def _quadratic_3d(c, n=12):
    """A_z = c (x^2 + y^2), A_x = A_y = 0: |J| = 4c."""
    hier = AMRHierarchy((n, n, n), ((0.0, 1.0),) * 3, periodic=(False, False, False))
    zero = lambda x, y, z: 0.0 * x  # noqa: E731
    state = dict(density=1.0, velocity=(0.0, 0.0, 0.0), pressure=1.0,
                 vector_potential=(zero, zero, lambda x, y, z: c * (x * x + y * y)))
    return initial.Problem(hier, {"limiter": "minmod"}, state)


def test_oracle_tags_2d_closed_form():
    c, n = 0.7, 16
    h = 1.0 / n
    o = oracle2d(_quadratic_2d(c, n))
    inner = (slice(1, -1), slice(1, -1))  # the outflow-extrapolated halo nodes bend J at the edge
    assert T.current_tags2d(o, 4 * c * h * (1 - 1e-9))[inner].all()
    assert not T.current_tags2d(o, 4 * c * h * (1 + 1e-9))[inner].any()


def test_oracle_tags_3d_closed_form():
    c, n = 0.7, 12
    h = 1.0 / n
    o = oracle3d(_quadratic_3d(c, n))
    inner = (slice(1, -1),) * 3
    assert T.current_tags3d(o, 4 * c * h * (1 - 1e-9))[inner].all()
    assert not T.current_tags3d(o, 4 * c * h * (1 + 1e-9))[inner].any()


def test_oracle_tags_follow_the_current_sheet():
    """Orszag-Tang: the tags sit where |J| is large, and more cells pass a lower bar."""
    o = oracle2d(initial.orszag_tang(32))
    low, high = T.current_tags2d(o, 0.05), T.current_tags2d(o, 0.2)
    assert high.sum() < low.sum() and not (high & ~low).any()


@pytest.mark.gpu
@pytest.mark.parametrize("threshold", [0.02, 0.1])
def test_gpu_tags_2d_are_the_oracle(gpu, threshold):
    from wellbalanced_amd import MHD2D

    problem = initial.orszag_tang(48)
    model = MHD2D.from_problem(problem)
    oracle = oracle2d(problem)
    for _ in range(3):
        dt = oracle.stable_step()
        model.step(dt)
        oracle.step(dt)
    got = model.current_tags(threshold, levels=2)
    assert set(got) == {0}
    np.testing.assert_array_equal(got[0], T.current_tags2d(oracle, threshold))


@pytest.mark.gpu
def test_gpu_tags_3d_are_the_oracle(gpu):
    from wellbalanced_amd import MHD3D

    problem = initial.blast_turbulence_3d(16, seed=3)
    model = MHD3D.from_problem(problem)
    oracle = oracle3d(problem)
    for _ in range(2):
        dt = oracle.stable_step()
        model.step(dt)
        oracle.step(dt)
    ref = T.current_tags3d(oracle, 0.05)
    assert 0 < ref.sum() < ref.size
    np.testing.assert_array_equal(model.current_tags(0.05)[0], ref)


@pytest.mark.gpu
def test_gpu_tags_of_a_hierarchy(gpu):
    """A refined level covering the whole domain is tagged as the uniform fine grid."""
    from wellbalanced_amd import MHD2D
    from wellbalanced_amd.amr import AMRMHD2D

    h = AMRHierarchy((16, 16), ((0.0, 1.0), (0.0, 1.0)), periodic=(True, True))
    for box in [((0, 16), (0, 16)), ((16, 32), (0, 16)), ((0, 16), (16, 32)), ((16, 32), (16, 32))]:
        h.add_patch(1, box)
    problem = initial.orszag_tang(32, limiter="minmod")
    amr = AMRMHD2D(h, limiter="minmod")
    amr.set_state(**problem.state)
    uniform = MHD2D.from_problem(problem)
    got = amr.current_tags(0.05, levels=2)
    want = uniform.current_tags(0.05)[0]
    assert set(got) == {0, 1} and 0 < want.sum() < want.size
    np.testing.assert_array_equal(got[1], want)
