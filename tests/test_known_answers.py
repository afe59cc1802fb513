"""The reference's whole-run checks on the uniform grid, on the oracle (CPU) and on the
GPU path (whose FP64 kernels must also equal the oracle bit for bit):

* ``tests/unit/test_mhd.py:78-86``: the oblique Alfvén wave converges at second order
  (log2 > 1.85) with div B < 1e-14;
* ``tests/unit/test_mhd.py:379-387``: the rotor keeps its point symmetry to 1e-10,
  positive pressure and div B < 1e-14;
* ``tests/unit/test_mhd.py:390-405``: the blast wave conserves to 1e-12, keeps its point
  symmetry to 1e-10 and runs further along the field than across it;
* ``tests/unit/test_mhd3d.py:13-21``: the 3D Alfvén wave converges at second order
  (log2 > 1.8) with div B < 1e-14.
"""

import numpy as np
import pytest

from oracle.mhd3d import cell_fields, face_fields
from tests.helpers import oracle2d, oracle3d
from wellbalanced_amd import initial


def _cell_fields_2d(o):
    a = o.nodes()
    bxf = (a[:, 1:] - a[:, :-1]) / o.dy + o.b0[0]
    byf = -(a[1:, :] - a[:-1, :]) / o.dx + o.b0[1]
    return 0.5 * (bxf[:-1] + bxf[1:]), 0.5 * (byf[:, :-1] + byf[:, 1:])


def _pressure_2d(o):
    c = o.interior()
    bx, by = _cell_fields_2d(o)
    rho = c[0]
    u, v, w = c[1] / rho, c[2] / rho, c[3] / rho
    return (o.gamma - 1.0) * (c[4] - 0.5 * rho * (u**2 + v**2 + w**2)
                              - 0.5 * (bx**2 + by**2 + c[5] ** 2))


def _cell_by_3d(o):
    h = 2
    local = [o.pot[c][tuple(slice(h, h + o.n[k] + (k != c)) for k in range(3))] for c in range(3)]
    return cell_fields(face_fields(local, o.spacing, o.b0))[1]


def _run3d(o, duration):
    end = o.time + duration
    while o.time < end * (1.0 - 1e-14):
        o.step(min(o.stable_step(), end - o.time))


# -- the oracle ----------------------------------------------------------------------------
def test_alfven_wave_converges_at_second_order():
    errors = []
    for cells in (8, 16):
        problem = initial.alfven_wave(cells, limiter="none")
        o = oracle2d(problem)
        exact = _cell_fields_2d(o)[1].copy()
        o.run(problem.period)
        errors.append(np.mean(np.abs(_cell_fields_2d(o)[1] - exact)))
        assert o.divergence() < 1e-14
    assert np.log2(errors[0] / errors[1]) > 1.85


def test_rotor_keeps_its_point_symmetry_and_div_b_zero():
    o = oracle2d(initial.rotor(32))
    o.run(0.05)
    density = o.interior()[0]
    np.testing.assert_allclose(density, density[::-1, ::-1], rtol=1e-10)
    assert _pressure_2d(o).min() > 0.0
    assert o.divergence() < 1e-14


def _blast_checks(totals_before, totals_after, pressure, x, y, divergence, rtol=1e-10):
    assert np.abs(totals_after - totals_before).max() < 1e-12
    np.testing.assert_allclose(pressure, pressure[::-1, ::-1], rtol=rtol)
    hot = pressure > 1.0
    X, Y = np.meshgrid(x, y, indexing="ij")
    assert np.abs(X + Y)[hot].max() > 1.5 * np.abs(X - Y)[hot].max()
    assert divergence < 1e-14


def test_blast_wave_conserves_and_runs_along_the_field():
    problem = initial.blast_wave(32)
    o = oracle2d(problem)
    before = o.totals()
    o.run(0.1)
    x, y = problem.hierarchy.cell_centers(problem.hierarchy.patches[0])
    _blast_checks(before, o.totals(), _pressure_2d(o), x[:, 0], y[0, :], o.divergence())


def test_3d_alfven_wave_converges_at_second_order():
    errors = []
    for cells in (8, 16):
        problem = initial.alfven_wave_3d(cells, limiter="none")
        o = oracle3d(problem)
        exact = _cell_by_3d(o).copy()
        _run3d(o, problem.period)
        errors.append(np.mean(np.abs(_cell_by_3d(o) - exact)))
        assert o.divergence() < 1e-14
    assert np.log2(errors[0] / errors[1]) > 1.8


# -- the GPU path --------------------------------------------------------------------------
def _same_2d(model, o):
    """The FP64 GPU state against the oracle: every conserved variable and A_z within
    1e-12 of its largest magnitude (the stated parity bar).  In practice bit for bit, but
    gfx950's FP64 division is not correctly rounded at exact near-ties (the rotor meets
    one: 1.3999999999999992 / 0.9999999999999994 gives 1.4, IEEE 1.4000000000000001), so
    a value that cancels to round-off can differ in its low digits."""
    cons, az = model.arrays()
    for got, want in [*zip(cons[:, 2:-2, 2:-2], o.interior()), (az[2:-2, 2:-2], o.nodes())]:
        scale = np.abs(want).max()
        if scale == 0.0:
            np.testing.assert_array_equal(got, want)
        else:
            assert np.abs(got - want).max() <= 1e-12 * scale


@pytest.mark.gpu
@pytest.mark.parametrize("kernels,precision", [("split", "fp64"), ("fused", "fp64"),
                                               ("fused", "mixed")])
def test_gpu_alfven_wave_converges_at_second_order(gpu, kernels, precision):
    from wellbalanced_amd import MHD2D

    errors = []
    for cells in (8, 16):
        model, exact = MHD2D.alfven_wave(cells, limiter="none", kernels=kernels,
                                         precision=precision)
        model.run(model.period)
        errors.append(np.mean(np.abs(model.dataset()["magnetic_field_y"]
                                     - exact["magnetic_field_y"])))
        assert model.divergence() < 1e-14
        if precision == "fp64" and cells == 8:
            problem = initial.alfven_wave(cells, limiter="none")
            o = oracle2d(problem)
            o.run(problem.period)
            assert o.time == model.time
            _same_2d(model, o)
    assert np.log2(errors[0] / errors[1]) > 1.85


@pytest.mark.gpu
@pytest.mark.parametrize("kernels", ["split", "fused"])
def test_gpu_rotor_keeps_its_point_symmetry_and_div_b_zero(gpu, kernels):
    from wellbalanced_amd import MHD2D

    model = MHD2D.rotor(32, kernels=kernels)
    model.run(0.05)
    data = model.dataset()
    density = data["density"]
    np.testing.assert_allclose(density, density[::-1, ::-1], rtol=1e-10)
    assert data["pressure"].min() > 0.0
    assert model.divergence() < 1e-14
    o = oracle2d(initial.rotor(32))
    o.run(0.05)
    _same_2d(model, o)


@pytest.mark.gpu
def test_gpu_mixed_rotor_stays_near_fp64_and_conserves(gpu):
    """The mixed path on the rotor, advanced at the FP64 path's step: FP32 rounding in
    the fluxes is not mirror-symmetric, and the limiters' switches at the rotor's sharp
    edge amplify it, so the point symmetry holds to about 6e-3 (not 1e-10) and the
    density differs from FP64 by L1 1.0e-4, L-inf 3.5e-3 at t = 0.05; div B stays zero
    and the pressure positive."""
    from wellbalanced_amd import MHD2D

    fp64 = MHD2D.rotor(32, kernels="fused")
    mixed = MHD2D.rotor(32, kernels="fused", precision="mixed")
    while fp64.time < 0.05 * (1.0 - 1e-14):
        dt = min(fp64.stable_step(), 0.05 - fp64.time)
        fp64.step(dt)
        mixed.step(dt)
    a, b = fp64.dataset()["density"], mixed.dataset()["density"]
    assert np.mean(np.abs(a - b)) < 2e-4
    assert np.abs(a - b).max() < 1e-2
    assert mixed.divergence() < 1e-14
    assert mixed.dataset()["pressure"].min() > 0.0


@pytest.mark.gpu
@pytest.mark.parametrize("kernels,precision", [("split", "fp64"), ("fused", "fp64"),
                                               ("fused", "mixed")])
def test_gpu_blast_wave_conserves_and_runs_along_the_field(gpu, kernels, precision):
    from wellbalanced_amd import MHD2D

    model = MHD2D.blast_wave(32, kernels=kernels, precision=precision)
    before = np.array(list(model.totals().values()))
    model.run(0.1)
    data = model.dataset()
    # FP32 rounding in the mixed fluxes is not mirror-symmetric: the point symmetry holds
    # to its level (9.5e-8 of the largest pressure measured), FP64 to 1e-10
    _blast_checks(before, np.array(list(model.totals().values())), data["pressure"], data["x"],
                  data["y"], model.divergence(), rtol=1e-6 if precision == "mixed" else 1e-10)


@pytest.mark.gpu
@pytest.mark.parametrize("kernels,precision", [("split", "fp64"), ("fused", "fp64"),
                                               ("fused", "mixed")])
def test_gpu_3d_alfven_wave_converges_at_second_order(gpu, kernels, precision):
    from wellbalanced_amd import MHD3D

    errors = []
    for cells in (8, 16):
        model, exact = MHD3D.alfven_wave(cells, limiter="none", kernels=kernels,
                                         precision=precision)
        model.run(model.period)
        errors.append(np.mean(np.abs(model.dataset()["magnetic_field_y"]
                                     - exact["magnetic_field_y"])))
        assert model.divergence() < 1e-14
    assert np.log2(errors[0] / errors[1]) > 1.8
