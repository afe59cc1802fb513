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


def _wavy_problem(shape, periodic, limiter="minmod"):
    from wellbalanced_amd.hierarchy import AMRHierarchy

    h = AMRHierarchy(shape, ((0.0, 1.0), (0.0, 0.7), (0.0, 1.3)), periodic=periodic)
    return initial.Problem(h, {"mean_field": (0.6, -0.3, 0.2), "limiter": limiter}, dict(
        density=lambda x, y, z: 1.0 + 0.3 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * z),
        velocity=lambda x, y, z: (0.2 * np.sin(2 * np.pi * y), 0.1 * np.cos(2 * np.pi * z + 1), 0.3 * np.sin(2 * np.pi * x)),
        pressure=lambda x, y, z: 0.8 + 0.1 * np.cos(2 * np.pi * y),
        vector_potential=(lambda x, y, z: 0.02 * np.sin(2 * np.pi * z), lambda x, y, z: 0.03 * np.cos(2 * np.pi * x),
                          lambda x, y, z: 0.01 * np.sin(2 * np.pi * (x + y)))))


@pytest.mark.parametrize("shape", [(20, 20, 20), (31, 17, 29), (3, 70, 5), (40, 2, 2)])
def test_stable_step_is_the_oracle(gpu, shape):
    """The FP64 rate (column march) gives the oracle's dt bit for bit; the mixed path's
    rate (reciprocals, one division per cell) agrees to 1e-14."""
    problem = _wavy_problem(shape, (True, True, True))
    oracle = oracle3d(problem)
    assert MHD3D.from_problem(problem, kernels="fused").stable_step() == oracle.stable_step()
    mixed = MHD3D.from_problem(problem, kernels="fused", precision="mixed").stable_step()
    assert mixed == pytest.approx(oracle.stable_step(), rel=1e-14)


@pytest.mark.parametrize("shape,nxs,periodic", [
    ((12, 12, 12), 32, (True, True, True)), ((31, 17, 29), 4, (True, True, True)),
    ((16, 30, 15), 7, (True, True, True)), ((12, 12, 12), 32, (False, True, False)),
    ((31, 17, 29), 4, (True, False, True)), ((16, 30, 15), 7, (False, False, False)),
    ((3, 2, 40), 5, (True, True, True))])
@pytest.mark.parametrize("limiter", ["minmod", "vanleer", "mc", "none"])
def test_fused_fp64_stage_is_bitwise_the_split_path(gpu, shape, nxs, periodic, limiter):
    """The one-launch FP64 stage gives the same bits as the split kernels (and so as
    the oracle), on uneven tiles and segments; along periodic axes it reads the
    periodic images itself (no halo fill), along outflow axes the filled halo."""
    from wellbalanced_amd.hierarchy import AMRHierarchy

    problem = _wavy_problem(shape, periodic, limiter)
    split = MHD3D.from_problem(problem, kernels="split")
    fused = MHD3D.from_problem(problem, kernels="fused")
    fused.nxs = nxs
    for _ in range(3):
        dt = split.stable_step()
        split.step(dt)
        fused.step(dt)
    a, pa = split.arrays()
    b, pb = fused.arrays()
    h2 = 2
    inner = (slice(None),) + tuple(slice(h2, h2 + n) for n in shape)
    np.testing.assert_array_equal(b[inner], a[inner])
    for c in range(3):
        idx = tuple(slice(h2, h2 + shape[k] + (k != c)) for k in range(3))
        np.testing.assert_array_equal(pb[c][idx], pa[c][idx])


def test_fused_turbulence_matches_the_oracle(gpu):
    problem = initial.turbulence_3d(20, seed=6)
    model = MHD3D.from_problem(problem, kernels="fused")
    oracle = oracle3d(problem)
    for _ in range(3):
        dt = oracle.stable_step()
        model.step(dt)
        oracle.step(dt)
    _compare(model, oracle)


def test_mixed_precision_is_close_and_keeps_div_b_and_conservation(gpu):
    """FP32 reconstruction/HLLD with FP64 CT and update: the density stays within
    1e-5 (L1) of the FP64 path, div B at FP64 round-off, totals conserved."""
    problem = initial.turbulence_3d(32, seed=3)
    fp64 = MHD3D.from_problem(problem, kernels="fused")
    mixed = MHD3D.from_problem(problem, kernels="fused", precision="mixed")
    before = mixed.totals()
    for _ in range(10):
        dt = fp64.stable_step()
        fp64.step(dt)
        mixed.step(dt)
    a = fp64.dataset()["density"]
    b = mixed.dataset()["density"]
    l1 = float(np.mean(np.abs(a - b)))
    assert 0.0 < l1 < 1e-5
    assert mixed.divergence() < 1e-13
    after = mixed.totals()
    for name in before:
        assert abs(after[name] - before[name]) < 1e-12 * max(1.0, abs(before[name]))


def test_advance_keeps_dt_on_the_device(gpu):
    """advance() computes dt on the device: the same steps as step() with the host dt."""
    problem = initial.turbulence_3d(16, seed=4)
    a = MHD3D.from_problem(problem, kernels="fused")
    b = MHD3D.from_problem(problem, kernels="fused")
    times = []
    for _ in range(3):
        times.append(a.step())
    b.advance(3)
    assert b.time == a.time == sum([0.0] + times) or abs(b.time - a.time) == 0.0
    np.testing.assert_array_equal(a.arrays()[0], b.arrays()[0])


@pytest.mark.parametrize("limiter", ["minmod", "mc"])
def test_mixed_stage_launch_shapes_keep_the_bits(gpu, limiter):
    """The mixed fused stage at 2 and 3 waves per SIMD (3: each Heun stage with its own
    set of early loads) and with segments of 8 and 32 planes leaves the same cells and
    potentials, bit for bit, over steps at one step size."""
    problem = initial.turbulence_3d(32, seed=4, limiter=limiter)
    results = {}
    dt = None
    for waves, nxs in [(3, 32), (3, 8), (2, 32), (2, 8)]:
        model = MHD3D.from_problem(problem, kernels="fused", precision="mixed")
        model.waves, model.nxs = waves, nxs
        dt = dt or 0.5 * model.stable_step()
        for _ in range(3):
            model.step(dt)
        model.check()
        results[waves, nxs] = model.arrays()
    cons0, pots0 = results[3, 32]
    for key, (cons, pots) in results.items():
        np.testing.assert_array_equal(cons, cons0, err_msg=str(key))
        for c in range(3):
            np.testing.assert_array_equal(pots[c], pots0[c], err_msg=str(key))
