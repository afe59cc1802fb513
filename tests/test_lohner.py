"""The Löhner estimator: closed forms of the oracle (CPU) and the GPU kernel
against the oracle (GPU, bit for bit)."""

import numpy as np
import pytest

from oracle import lohner as L
from tests.helpers import oracle3d
from wellbalanced_amd import initial


def test_linear_field_has_no_second_derivative():
    x, y, z = np.meshgrid(*(np.arange(12.0),) * 3, indexing="ij")
    u = 2.0 + 0.3 * x - 0.1 * y + 0.05 * z
    assert np.abs(L.estimate(u)).max() < 1e-15


def test_a_jump_is_flagged_and_smooth_is_not():
    x = np.arange(20.0)
    u = np.where(x < 10, 1.0, 2.0)[:, None, None] * np.ones((20, 8, 8))
    e = L.estimate(u)
    assert e.max() > 0.8 and e[:, 2, 2].argmax() in (6, 7, 8)
    smooth = 1.0 + 0.01 * np.sin(0.3 * x)[:, None, None] * np.ones((20, 8, 8))
    assert L.estimate(smooth).max() < 0.1


def test_a_pure_second_derivative_in_1d():
    """u = x^2 along x, constant along y and z, at x = 4: only (a, b) = (x, x) has a
    numerator, (8)^2; the denominator sums all nine (a, b) pairs of FLASH's form."""
    x = np.arange(9.0)
    u = (x * x)[:, None, None] * np.ones((9, 5, 5))
    e = L.estimate(u)
    eps = 0.01
    u2, u3, u4, u5, u6 = 4.0, 9.0, 16.0, 25.0, 36.0
    xx = (abs(u6 - u4) + abs(u4 - u2) + eps * ((u6 + u4) + (u4 + u2))) ** 2
    xy = (2 * abs(u5 - u3) + eps * 2 * (u5 + u3)) ** 2          # b = y or z
    yx = (eps * ((u5 + u5) + (u3 + u3))) ** 2                   # a = y or z, b = x
    yy = (eps * (2 * u4 + 2 * u4)) ** 2                         # a, b in {y, z}
    den = xx + 2 * xy + 2 * yx + 4 * yy
    want = np.sqrt((u6 - 2 * u4 + u2) ** 2 / den)
    np.testing.assert_allclose(e[2, 0, 0], want, rtol=1e-14)


def test_block_flags():
    f = np.zeros((32, 32, 32), bool)
    f[17, 3, 30] = True
    b = L.block_flags(f, 16)
    assert b.shape == (2, 2, 2) and b.sum() == 1 and b[1, 0, 1]


@pytest.mark.gpu
def test_gpu_lohner_is_bitwise_the_oracle(gpu):
    from wellbalanced_amd import MHD3D
    from wellbalanced_amd.indicator import lohner

    problem = initial.turbulence_3d(20, seed=3)
    model = MHD3D.from_problem(problem, kernels="fused")
    oracle = oracle3d(problem)
    for _ in range(2):
        dt = oracle.stable_step()
        model.step(dt)
        oracle.step(dt)
    got = lohner(model).cpu().numpy()
    oracle.fill_halos()
    prim, _ = oracle.primitive()
    er, _ = L.flags(prim[0], prim[4])
    np.testing.assert_array_equal(got, er)
