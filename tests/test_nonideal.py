"""The non-ideal face terms of MHD2D: the resistive terms (``mhd_resistive``) and the
anisotropic heat flux (``mhd_conductive``), added to the ideal fluxes before the walls
(mhd.py:868-871), and the explicit diffusion limit of the stable step (mhd.py:954-961).

CPU: the oracle against the reference kernels' golden vectors, closed forms that a
transposed field or a dropped term fails, and the reference's own diffusion tests
(test_mhd_amr.py:109-141).  GPU: the per-kernel twins and the fused stages against the
oracle.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from oracle import kernels as K
from tests.helpers import oracle2d
from wellbalanced_amd import initial
from wellbalanced_amd.hierarchy import AMRHierarchy

GOLDEN = Path(__file__).resolve().parent / "golden" / "golden.npz"
#: dx, dy, eta, parallel, perpendicular of the golden vectors (tests/golden/make_golden.py)
NONIDEAL = (0.1, 0.07, 0.37, 0.8, 0.05)
KW = 2.0 * np.pi


# -- golden vectors of the reference's kernels ------------------------------------------------
@pytest.mark.parametrize("seed", [1, 7])
def test_resistive_equals_the_reference_kernel(seed):
    g = np.load(GOLDEN)
    dx, dy, eta, _, _ = NONIDEAL
    fx, fy, ez = g["ni_fx"].copy(), g["ni_fy"].copy(), g["ni_ez"].copy()
    K.resistive(g[f"prim_{seed}"], g[f"bxf_{seed}"], g[f"byf_{seed}"], dx, dy, eta, fx, fy, ez)
    np.testing.assert_array_equal(fx, g[f"res_fx_{seed}"])
    np.testing.assert_array_equal(fy, g[f"res_fy_{seed}"])
    np.testing.assert_array_equal(ez, g[f"res_ez_{seed}"])


@pytest.mark.parametrize("seed", [1, 7])
def test_conductive_equals_the_reference_kernel(seed):
    g = np.load(GOLDEN)
    dx, dy, _, par, perp = NONIDEAL
    fx, fy = g["ni_fx"].copy(), g["ni_fy"].copy()
    K.conductive(g[f"con_prim_{seed}"], g[f"bxf_{seed}"], g[f"byf_{seed}"], dx, dy, par, perp, fx, fy)
    np.testing.assert_array_equal(fx, g[f"con_fx_{seed}"])
    np.testing.assert_array_equal(fy, g[f"con_fy_{seed}"])


def test_the_golden_conductive_block_has_a_field_free_face():
    """The unit field of a face with no field is 0 / TINY = 0, not NaN."""
    g = np.load(GOLDEN)
    prim = g["con_prim_1"]
    assert g["bxf_1"][4, 2] == 0.0 and np.all(prim[5:8, 3:5, 2] == 0.0)
    assert np.all(np.isfinite(g["con_fx_1"]))


# -- closed forms -------------------------------------------------------------------------------
def _quadratic_block(a=0.3, b=-0.7, c=0.45, X=12, Y=10, dx=0.1, dy=0.07):
    """A_z = a x^2 + b y^2 on the nodes and B_z = c x in the cells: the node current is
    -2 (a + b) everywhere, B_x = 2 b y and B_y = -2 a x in the cells."""
    xn = np.arange(X + 1)[:, None] * dx
    yn = np.arange(Y + 1)[None, :] * dy
    az = a * xn**2 + b * yn**2
    bxf = (az[:, 1:] - az[:, :-1]) / dy
    byf = -(az[1:, :] - az[:-1, :]) / dx
    xc = (np.arange(X)[:, None] + 0.5) * dx + 0 * yn[:, :Y]
    prim = np.zeros((8, X, Y))
    prim[0] = prim[4] = 1.0
    prim[5] = 0.5 * (bxf[:-1] + bxf[1:])
    prim[6] = 0.5 * (byf[:, :-1] + byf[:, 1:])
    prim[7] = c * xc
    return prim, bxf, byf, (a, b, c, dx, dy)


def _resistive_gain(prim, bxf, byf, coef, eta=0.37):
    X, Y = prim.shape[1:]
    fx, fy = np.zeros((7, X - 3, Y)), np.zeros((7, X, Y - 3))
    ez = np.zeros((X - 3, Y - 3))
    K.resistive(prim, bxf, byf, coef[3], coef[4], eta, fx, fy, ez)
    return fx, fy, ez


def _resistive_closed_form_holds(prim, bxf, byf, coef, eta=0.37) -> bool:
    a, b, c, dx, dy = coef
    fx, fy, ez = _resistive_gain(prim, bxf, byf, coef, eta)
    X, Y = prim.shape[1:]
    h, nx, ny = 2, X - 4, Y - 4
    xf = (h + np.arange(nx + 1))[:, None] * dx  # x-faces of the block
    yf = (h + np.arange(ny + 1))[None, :] * dy
    jz = -2.0 * (a + b)
    ok = np.allclose(ez, eta * jz, rtol=0, atol=1e-12)
    ok &= np.allclose(fx[6, :, h : h + ny], -eta * c, rtol=0, atol=1e-12)
    want = -eta * xf * (c * c + 4.0 * a * (a + b)) + 0 * yf[:, :ny]
    ok &= np.allclose(fx[4, :, h : h + ny], want, rtol=0, atol=1e-12)
    ok &= np.allclose(fy[6, h : h + nx, :], 0.0, rtol=0, atol=1e-12)
    want = -4.0 * eta * b * (a + b) * yf + 0 * xf[:nx]
    ok &= np.allclose(fy[4, h : h + nx, :], want, rtol=0, atol=1e-12)
    return bool(ok)


def test_resistive_closed_form_of_a_quadratic_potential():
    assert _resistive_closed_form_holds(*_quadratic_block())


def test_a_planted_transposed_field_fails_the_resistive_closed_form():
    prim, bxf, byf, coef = _quadratic_block()
    swapped = prim[[0, 1, 2, 3, 4, 6, 5, 7]]
    assert not _resistive_closed_form_holds(swapped, bxf, byf, coef)


def _uniform_field_block(theta, alpha, beta, X=10, Y=9, dx=0.1, dy=0.07):
    """rho = 1, p = T = alpha x + beta y, the field (cos theta, sin theta, 0)."""
    xc = (np.arange(X)[:, None] + 0.5) * dx
    yc = (np.arange(Y)[None, :] + 0.5) * dy
    prim = np.zeros((8, X, Y))
    prim[0] = 1.0
    prim[4] = 5.0 + alpha * xc + beta * yc
    prim[5] = np.cos(theta)
    prim[6] = np.sin(theta)
    bxf = np.full((X + 1, Y), np.cos(theta))
    byf = np.full((X, Y + 1), np.sin(theta))
    return prim, bxf, byf, dx, dy


@pytest.mark.parametrize("theta", [0.0, 0.6, np.pi / 2])
def test_conductive_closed_forms(theta):
    """q_n = parallel b_n (b_n dT/dn + b_t dT/dt) + perpendicular dT/dn for a linear T:
    along x-faces dT/dn = alpha, dT/dt = beta (zero on the block's first and last row,
    where no limited gradient is taken); along y-faces the other way round."""
    par, perp, alpha, beta = 0.8, 0.05, 0.3, -1.1
    prim, bxf, byf, dx, dy = _uniform_field_block(theta, alpha, beta)
    X, Y = prim.shape[1:]
    fx, fy = np.zeros((7, X - 3, Y)), np.zeros((7, X, Y - 3))
    K.conductive(prim, bxf, byf, dx, dy, par, perp, fx, fy)
    c, s = np.cos(theta), np.sin(theta)
    inner = par * c * (c * alpha + s * beta) + perp * alpha
    edge = par * c * (c * alpha) + perp * alpha
    np.testing.assert_allclose(fx[4][:, 1:-1], -inner, rtol=0, atol=1e-11)
    np.testing.assert_allclose(fx[4][:, [0, -1]], -edge, rtol=0, atol=1e-11)
    inner = par * s * (s * beta + c * alpha) + perp * beta
    edge = par * s * (s * beta) + perp * beta
    np.testing.assert_allclose(fy[4][1:-1, :], -inner, rtol=0, atol=1e-11)
    np.testing.assert_allclose(fy[4][[0, -1], :], -edge, rtol=0, atol=1e-11)
    assert not np.any(fx[[0, 1, 2, 3, 5, 6]]) and not np.any(fy[[0, 1, 2, 3, 5, 6]])


def test_mc_limiter_closed_forms():
    a = np.array([1.0, 1.0, -1.0, 0.0, 2.0])
    b = np.array([3.0, -1.0, -4.0, 5.0, 2.0])
    np.testing.assert_array_equal(K.mc(a, b), [2.0, 0.0, -2.0, 0.0, 2.0])


# -- the reference's diffusion tests (test_mhd_amr.py:109-141), on one level -------------------
def _decay_problem(n=16, eta=0.02, amp=1e-3):
    h = AMRHierarchy((n, n), ((0.0, 1.0), (0.0, 1.0)), periodic=(True, True))
    state = dict(density=1.0, velocity=(0.0, 0.0, 0.0), pressure=1.0,
                 vector_potential=lambda x, y: amp * np.cos(KW * x) / KW,
                 bz=lambda x, y: amp * np.cos(KW * y))
    return initial.Problem(h, {"resistivity": eta, "limiter": "none"}, state)


def test_resistivity_damps_the_field_at_the_diffusion_rate():
    n, eta, amp = 16, 0.02, 1e-3
    model = oracle2d(_decay_problem(n, eta, amp))
    start = model.totals()
    model.run(0.25)
    h = 2
    prim, *_ = K.primitive(*model.cons, model.az, model.dx, model.dy, *model.b0, model.gamma)
    xc = (np.arange(n)[:, None] + 0.5) / n + np.zeros((1, n))
    yc = (np.arange(n)[None, :] + 0.5) / n + np.zeros((n, 1))
    decay = np.exp(-eta * KW * KW * model.time)
    by = np.mean(np.abs(prim[6, h:-h, h:-h] - amp * np.sin(KW * xc) * decay))
    bz = np.mean(np.abs(prim[7, h:-h, h:-h] - amp * np.cos(KW * yc) * decay))
    assert by < 3e-3 * amp * decay
    assert bz < 4e-3 * amp * decay
    # Ohmic heating stays in the energy
    assert abs(model.totals()[4] - start[4]) < 1e-12


def test_the_step_keeps_within_the_diffusion_limit():
    h = AMRHierarchy((16, 16), ((0.0, 1.0), (0.0, 1.0)), periodic=(True, True))
    problem = initial.Problem(h, {"resistivity": 1.0},
                              dict(density=1.0, velocity=(0.0, 0.0, 0.0), pressure=1e-3))
    model = oracle2d(problem)
    dx = 1.0 / 16
    assert model.stable_step() == pytest.approx(model.cfl * 0.5 / (2.0 / dx**2))


def test_conduction_limits_the_step_by_the_smallest_density():
    """D = (gamma - 1) kappa / rho_min when it exceeds eta."""
    h = AMRHierarchy((16, 8), ((0.0, 1.0), (0.0, 0.5)), periodic=(True, True))
    problem = initial.Problem(h, {"conduction": (3.0, 0.5), "resistivity": 0.1},
                              dict(density=lambda x, y: 1.0 + 0.5 * np.sin(KW * x),
                                   velocity=(0.0, 0.0, 0.0), pressure=1e-4))
    model = oracle2d(problem)
    rho_min = float(model.interior()[0].min())
    d = (2.0 / 3.0) * 3.0 / rho_min
    assert model.stable_step() == pytest.approx(model.cfl * 0.5 / (d * 2 * 16.0**2), rel=1e-14)


def test_conduction_conserves_and_relaxes_the_temperature_along_the_field():
    """An entropy mode in pressure balance (static without conduction), with the field
    along x: a temperature varying in x relaxes, one varying in y stays as it is (no
    perpendicular conduction); the total energy is kept."""
    def run(direction):
        h = AMRHierarchy((16, 16), ((0.0, 1.0), (0.0, 1.0)), periodic=(True, True))
        rho = (lambda x, y: 1.0 / (1.0 + 0.1 * np.cos(KW * x))) if direction == 0 else \
            (lambda x, y: 1.0 / (1.0 + 0.1 * np.cos(KW * y)))
        problem = initial.Problem(h, {"conduction": (0.5, 0.0), "mean_field": (1.0, 0.0),
                                      "limiter": "minmod"},
                                  dict(density=rho, velocity=(0.0, 0.0, 0.0), pressure=1.0))
        model = oracle2d(problem)
        start = model.totals()
        model.run(0.05)
        assert abs(model.totals()[4] - start[4]) < 1e-13
        prim, *_ = K.primitive(*model.cons, model.az, model.dx, model.dy, *model.b0, model.gamma)
        t = prim[4, 2:-2, 2:-2] / prim[0, 2:-2, 2:-2]
        return np.ptp(t)

    along, across = run(0), run(1)
    assert across > 0.19
    assert along < 0.7 * across


# -- GPU ----------------------------------------------------------------------------------------
def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def _call(name, *args):
    import torch

    from wellbalanced_amd import _lib

    _lib.call(name, *args, _lib.stream_handle())
    torch.cuda.synchronize()


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [1, 7])
def test_resistive_twin_is_the_reference_kernel(gpu, seed):
    g = np.load(GOLDEN)
    dx, dy, eta, _, _ = NONIDEAL
    prim, bxf, byf = (_dev(g[k]) for k in (f"prim_{seed}", f"bxf_{seed}", f"byf_{seed}"))
    fx, fy, ez = _dev(g["ni_fx"]), _dev(g["ni_fy"]), _dev(g["ni_ez"])
    X, Y = g[f"prim_{seed}"].shape[1:]
    _call("wb_resistive2d", prim.data_ptr(), bxf.data_ptr(), byf.data_ptr(), X, Y, dx, dy, eta,
          fx.data_ptr(), fy.data_ptr(), ez.data_ptr())
    np.testing.assert_array_equal(fx.cpu().numpy(), g[f"res_fx_{seed}"])
    np.testing.assert_array_equal(fy.cpu().numpy(), g[f"res_fy_{seed}"])
    np.testing.assert_array_equal(ez.cpu().numpy(), g[f"res_ez_{seed}"])


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [1, 7])
def test_conductive_twin_is_the_reference_kernel(gpu, seed):
    g = np.load(GOLDEN)
    dx, dy, _, par, perp = NONIDEAL
    prim, bxf, byf = (_dev(g[k]) for k in (f"con_prim_{seed}", f"bxf_{seed}", f"byf_{seed}"))
    fx, fy = _dev(g["ni_fx"]), _dev(g["ni_fy"])
    X, Y = g[f"prim_{seed}"].shape[1:]
    _call("wb_conductive2d", prim.data_ptr(), bxf.data_ptr(), byf.data_ptr(), X, Y, dx, dy, par,
          perp, fx.data_ptr(), fy.data_ptr())
    np.testing.assert_array_equal(fx.cpu().numpy(), g[f"con_fx_{seed}"])
    np.testing.assert_array_equal(fy.cpu().numpy(), g[f"con_fy_{seed}"])


def _nonideal_problems():
    """Orszag-Tang (periodic) and a Harris sheet between walls, both with resistivity
    and anisotropic conduction."""
    ot = initial.orszag_tang(40, limiter="minmod", resistivity=1e-3, conduction=(0.02, 0.002))
    harris = initial.harris_sheet(70, 36, length=3.2, noise=1e-2, seed=1, limiter="mc",
                                  resistivity=4e-3, conduction=(0.01, 0.0))
    return ot, harris


def _bitwise(model, oracle):
    cons, az = model.arrays()
    h, nx, ny = 2, model.nx, model.ny
    np.testing.assert_array_equal(cons[:, h : h + nx, h : h + ny], oracle.interior())
    np.testing.assert_array_equal(az[h : h + nx + 1, h : h + ny + 1], oracle.nodes())


@pytest.mark.gpu
@pytest.mark.parametrize("kernels,kernel,tile,nxs", [
    ("split", "march", 126, 32), ("fused", "march", 126, 32), ("fused", "march", 62, 5),
    ("fused", "march", 254, 64), ("fused", "march", 30, 4), ("fused", "tiles", 126, 32)])
def test_nonideal_steps_are_bitwise_the_oracle(gpu, kernels, kernel, tile, nxs):
    from wellbalanced_amd import MHD2D

    for problem in _nonideal_problems():
        model = MHD2D.from_problem(problem, kernels=kernels)
        model.kernel, model.tile, model.nxs = kernel, tile, nxs
        oracle = oracle2d(problem)
        for _ in range(4):
            dt = oracle.stable_step()
            assert model.stable_step() == dt
            model.step(dt)
            oracle.step(dt)
        _bitwise(model, oracle)


@pytest.mark.gpu
def test_nonideal_mixed_is_close_to_fp64_and_conserves(gpu):
    """Mixed precision with the diffusive terms in FP64 beside the FP32 HLLD flux:
    L1(rho) < 1e-5 of FP64 after 10 steps, conservation to 1e-13, div B at round-off."""
    from wellbalanced_amd import MHD2D

    problem = initial.orszag_tang(128, limiter="minmod", resistivity=1e-3,
                                  conduction=(0.05, 0.005))
    fp64 = MHD2D.from_problem(problem, kernels="fused")
    mixed = MHD2D.from_problem(problem, kernels="fused", precision="mixed")
    before = mixed.totals()
    for _ in range(10):
        dt = fp64.stable_step()
        fp64.step(dt)
        mixed.step(dt)
    l1 = float(np.mean(np.abs(fp64.dataset()["density"] - mixed.dataset()["density"])))
    assert 0.0 < l1 < 1e-5
    assert mixed.divergence() < 1e-14
    after = mixed.totals()
    for name in before:
        assert abs(after[name] - before[name]) < 1e-13


@pytest.mark.gpu
def test_gpu_resistivity_damps_the_field_at_the_diffusion_rate(gpu):
    """test_mhd_amr.py:109-134 on one level, through the fused stage."""
    from wellbalanced_amd import MHD2D

    n, eta, amp = 16, 0.02, 1e-3
    model = MHD2D.from_problem(_decay_problem(n, eta, amp), kernels="fused")
    start = model.totals()
    model.run(0.25)
    d = model.dataset()
    decay = np.exp(-eta * KW * KW * model.time)
    by = np.mean(np.abs(d["magnetic_field_y"] - amp * np.sin(KW * d["x"])[:, None] * decay))
    bz = np.mean(np.abs(d["magnetic_field_z"] - amp * np.cos(KW * d["y"])[None, :] * decay))
    assert by < 3e-3 * amp * decay
    assert bz < 4e-3 * amp * decay
    assert abs(model.totals()["energy"] - start["energy"]) < 1e-12


@pytest.mark.gpu
def test_gpu_the_step_keeps_within_the_diffusion_limit(gpu):
    """test_mhd_amr.py:136-141."""
    from wellbalanced_amd import MHD2D

    model = MHD2D(AMRHierarchy((16, 16), ((0.0, 1.0), (0.0, 1.0)), periodic=(True, True)),
                  resistivity=1.0)
    model.set_state(1.0, (0.0, 0.0, 0.0), 1e-3)
    dx = 1.0 / 16
    assert model.stable_step() == pytest.approx(model.cfl * 0.5 / (2.0 / dx**2))


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", ["tiles", "march"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gpu_hierarchy_with_diffusion_matches_the_oracle_and_conserves(gpu, precision, kernel):
    """Three levels with resistivity and conduction: the refluxing carries the diffusive
    fluxes, so the leaf totals are kept (to 1e-13 of the largest, mixed as FP64: the
    refluxing's recomputed FP32 face flux has the stage's bits);
    FP64 is bit for bit the oracle, including the step limited by each patch's diffusion
    number."""
    from tests.helpers import amr_oracle
    from tests.test_amr import _assert_interiors_equal, _refined, _smooth_state
    from wellbalanced_amd.amr import AMRMHD2D

    h = _refined(3)
    diff = dict(resistivity=2e-3, conduction=(0.03, 0.003))
    oracle = amr_oracle(h, _smooth_state(), limiter="minmod", **diff)
    model = AMRMHD2D(h, limiter="minmod", precision=precision, kernel=kernel, **diff)
    model.set_state(**_smooth_state())
    before = model.leaf_totals()
    for _ in range(3):
        dt = oracle.stable_step()
        if precision == "fp64":
            assert model.stable_step() == dt
        model.step(dt)
        oracle.step(dt)
    after = model.leaf_totals()
    assert np.abs(after - before).max() < 1e-13 * np.abs(before).max()
    assert model.composite_divergence() == 0.0
    if precision == "fp64":
        _assert_interiors_equal(model.buffer(), oracle)


# -- 3D resistivity (mhd3d.py:561-602, 674-724) ---------------------------------------------
def _ot3d_in_plane(normal, cells=16, layers=2, **options):
    """The Orszag-Tang vortex in the plane of the two axes after ``normal`` (cyclic
    order: 2D x -> axis a, 2D y -> axis b), uniform along ``normal``."""
    from wellbalanced_amd.hierarchy import AMRHierarchy as H

    a, b = (normal + 1) % 3, (normal + 2) % 3
    shape = [cells] * 3
    shape[normal] = layers
    bounds = [(0.0, 1.0)] * 3
    bounds[normal] = (0.0, layers / cells)
    h = H(tuple(shape), tuple(bounds), periodic=(True, True, True))
    b0 = 1.0 / np.sqrt(4.0 * np.pi)

    def velocity(*x):
        v = [0.0 * x[0]] * 3
        v[a] = -np.sin(KW * x[b])
        v[b] = np.sin(KW * x[a])
        return tuple(v)

    pot = [0.0, 0.0, 0.0]
    pot[normal] = lambda *x: b0 * (np.cos(2 * KW * x[a]) / (2 * KW) + np.cos(KW * x[b]) / KW)
    state = dict(density=25.0 / (36.0 * np.pi), velocity=velocity, pressure=5.0 / (12.0 * np.pi),
                 vector_potential=tuple(pot))
    return initial.Problem(h, options, state)


@pytest.mark.parametrize("normal", [0, 1, 2])
def test_3d_resistive_run_uniform_along_an_axis_is_the_2d_run(normal):
    """The resistive 3D oracle on a state uniform along ``normal`` equals the resistive
    2D oracle (whose terms the reference kernel's golden vectors pin), for every
    orientation: J_x, J_y and J_z and the face means of B all take part."""
    from tests.helpers import oracle3d

    eta = 5e-3
    flat = oracle3d(_ot3d_in_plane(normal, limiter="minmod", resistivity=eta))
    plane = oracle2d(initial.orszag_tang(16, limiter="minmod", resistivity=eta))
    a, b = (normal + 1) % 3, (normal + 2) % 3
    for _ in range(4):
        dt = plane.stable_step()
        plane.step(dt)
        flat.step(dt)
    c3 = np.moveaxis(flat.interior(), (1 + a, 1 + b, 1 + normal), (1, 2, 3))
    c2 = plane.interior()
    pairs = [(0, 0), (1 + a, 1), (1 + b, 2), (1 + normal, 3), (4, 4)]
    for r3, r2 in pairs:
        np.testing.assert_allclose(c3[r3], np.repeat(c2[r2][:, :, None], 2, axis=2), atol=1e-13)
    assert flat.divergence() < 1e-14


def test_3d_resistive_step_keeps_within_the_diffusion_limit():
    """mhd3d.py:717-721: cfl * 0.5 / (eta * sum(d^-2)) when it binds."""
    from tests.helpers import oracle3d

    problem = _ot3d_in_plane(2, cells=8, layers=4, resistivity=1.0)
    model = oracle3d(problem)
    dx = 1.0 / 8
    assert model.stable_step() == pytest.approx(model.cfl * 0.5 / (3.0 / dx**2))


def test_3d_resistivity_conserves_energy_and_div_b():
    from tests.helpers import oracle3d

    model = oracle3d(initial.turbulence_3d(8, limiter="minmod", resistivity=0.01))
    before = model.totals()
    for _ in range(3):
        model.step()
    np.testing.assert_allclose(model.totals(), before, rtol=0, atol=1e-13)
    assert model.divergence() < 1e-14


def _compare3d(model, oracle, rtol):
    cons = model.arrays()[0]
    h = 2
    got = cons[(slice(None),) + tuple(slice(h, h + n) for n in model.n)]
    want = oracle.interior()
    scale = np.abs(want).max(axis=(1, 2, 3), keepdims=True)
    assert np.all(np.abs(got - want) <= rtol * scale)


@pytest.mark.gpu
@pytest.mark.parametrize("kernels", ["split", "fused"])
def test_gpu_3d_resistive_steps_match_the_oracle(gpu, kernels):
    """MHD3D with resistivity against the oracle: the split stage adds the resistive
    terms into its fluxes and edges before the update, as the reference does, and is bit
    for bit the oracle; the fused stage adds them after the stage (to 1e-12 of each
    field's largest value).  The diffusion-limited step exactly."""
    from tests.helpers import oracle3d
    from wellbalanced_amd import MHD3D

    problem = initial.turbulence_3d(16, limiter="minmod", resistivity=0.02)
    model = MHD3D.from_problem(problem, kernels=kernels)
    oracle = oracle3d(problem)
    for _ in range(3):
        dt = oracle.stable_step()
        assert model.stable_step() == dt
        model.step(dt)
        oracle.step(dt)
    _compare3d(model, oracle, 0.0 if kernels == "split" else 1e-12)
    assert model.divergence() < 1e-13


@pytest.mark.gpu
def test_gpu_3d_resistive_mixed_is_close_and_conserves(gpu):
    from wellbalanced_amd import MHD3D

    problem = initial.turbulence_3d(32, limiter="minmod", resistivity=0.01)
    fp64 = MHD3D.from_problem(problem, kernels="fused")
    mixed = MHD3D.from_problem(problem, kernels="fused", precision="mixed")
    before = mixed.totals()
    for _ in range(5):
        dt = fp64.stable_step()
        fp64.step(dt)
        mixed.step(dt)
    a, b = fp64.arrays()[0][0, 2:-2, 2:-2, 2:-2], mixed.arrays()[0][0, 2:-2, 2:-2, 2:-2]
    assert 0.0 < float(np.mean(np.abs(a - b))) < 1e-5
    after = mixed.totals()
    for name in before:
        assert abs(after[name] - before[name]) < 1e-12
    assert mixed.divergence() < 1e-13


def _beltrami(eta=0.01):
    """test_mhd3d.py:95-123: A = (cos kz, sin kz, 0) / k on 4 x 4 x 16 periodic cells gives
    B = (cos kz, sin kz, 0), J = k B: no force, the field decays as exp(-eta k^2 t)."""
    from wellbalanced_amd.hierarchy import AMRHierarchy as H

    h = H((4, 4, 16), ((0.0, 1.0),) * 3, periodic=(True,) * 3)
    k = KW
    state = dict(density=1.0, velocity=(0.0, 0.0, 0.0), pressure=1.0,
                 vector_potential=(lambda x, y, z: np.cos(k * z) / k,
                                   lambda x, y, z: np.sin(k * z) / k, 0.0))
    return initial.Problem(h, {"resistivity": eta}, state)


def _beltrami_ratio(pb_before, pb_after, eta, duration):
    h = 1.0 / 16
    k_discrete = 2.0 * np.sin(KW * h / 2) / h
    return pb_after / pb_before, np.exp(-2 * eta * k_discrete**2 * duration)


def test_3d_resistivity_damps_a_force_free_field_at_the_diffusion_rate():
    """The oracle on test_mhd3d.py:95-123 (ratio to 2e-3 of the discrete rate)."""
    from oracle.mhd3d import cell_fields, face_fields
    from tests.helpers import oracle3d

    model = oracle3d(_beltrami())

    def magnetic_pressure():
        h = 2
        local = [model.pot[c][tuple(slice(h, h + model.n[k] + (k != c)) for k in range(3))]
                 for c in range(3)]
        bx, by, bz = cell_fields(face_fields(local, model.spacing, model.b0))
        return float(np.mean(0.5 * (bx * bx + by * by + bz * bz)))

    before = magnetic_pressure()
    while model.time < 0.5 * (1 - 1e-14):
        model.step(min(model.stable_step(), 0.5 - model.time))
    ratio, want = _beltrami_ratio(before, magnetic_pressure(), 0.01, model.time)
    assert ratio == pytest.approx(want, rel=2e-3)
    assert np.abs(model.interior()[1]).max() < 1e-12


@pytest.mark.gpu
def test_gpu_3d_resistivity_damps_a_force_free_field_at_the_diffusion_rate(gpu):
    """test_mhd3d.py:95-123 on the device, through the fused stage."""
    from wellbalanced_amd import MHD3D

    model = MHD3D.from_problem(_beltrami(), kernels="fused")
    before = float(model.dataset()["magnetic_pressure"].mean())
    model.run(0.5)
    after = model.dataset()
    ratio, want = _beltrami_ratio(before, float(after["magnetic_pressure"].mean()), 0.01, 0.5)
    assert ratio == pytest.approx(want, rel=2e-3)
    assert np.abs(after["velocity_x"]).max() < 1e-12


@pytest.mark.gpu
def test_gpu_tiled_nonideal_mixed_stage_is_bitwise_the_marching_one(gpu):
    """With resistivity and conduction the tiled mixed stage keeps the marching one's
    bits (both add the FP64 face terms to the same FP32 fluxes), walls included."""
    from wellbalanced_amd import MHD2D

    for problem in _nonideal_problems():
        models = []
        for kernel in ("march", "tiles"):
            m = MHD2D.from_problem(problem, kernels="fused", precision="mixed")
            m.kernel = kernel
            models.append(m)
        for _ in range(4):
            dt = models[0].stable_step()
            for m in models:
                m.step(dt)
        a, aa = models[0].arrays()
        b, bb = models[1].arrays()
        np.testing.assert_array_equal(b, a)
        np.testing.assert_array_equal(bb, aa)
