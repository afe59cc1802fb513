"""Pins of the CPU oracle (no GPU): golden vectors written by the reference's own
kernels, the closed forms of the reference's tests, and whole-run invariants."""

from pathlib import Path

import numpy as np
import pytest

from oracle import kernels as K
from oracle.mhd2d import Oracle2D
from tests.helpers import block, conserved, oracle2d, oracle3d
from wellbalanced_amd import initial
from wellbalanced_amd.hierarchy import AMRHierarchy

GAMMA = 5.0 / 3.0
GOLDEN = Path(__file__).resolve().parent / "golden" / "golden.npz"


def _state(rho, vn, vt1, vt2, p, bt1, bt2, size=1):
    return np.array([[rho], [vn], [vt1], [vt2], [p], [bt1], [bt2]], float) * np.ones(size)


def _physical_flux(state, bn, gamma=GAMMA):
    rho, vn, vt1, vt2, p, bt1, bt2 = state
    b2 = bn**2 + bt1**2 + bt2**2
    pt = p + 0.5 * b2
    energy = p / (gamma - 1) + 0.5 * rho * (vn**2 + vt1**2 + vt2**2) + 0.5 * b2
    vb = vn * bn + vt1 * bt1 + vt2 * bt2
    return np.array([
        rho * vn, rho * vn**2 + pt - bn**2, rho * vn * vt1 - bn * bt1, rho * vn * vt2 - bn * bt2,
        (energy + pt) * vn - bn * vb, bt1 * vn - bn * vt1, bt2 * vn - bn * vt2,
    ])


# -- golden vectors of the reference's scalar kernels (tests/golden/make_golden.py) --------
@pytest.mark.parametrize("seed", [1, 7])
@pytest.mark.parametrize("limiter", range(4))
@pytest.mark.parametrize("axis", [0, 1])
def test_sweep_equals_the_reference_kernel(seed, limiter, axis):
    g = np.load(GOLDEN)
    bn = (g[f"bxf_{seed}"], g[f"byf_{seed}"])[axis]
    got = K.sweep(g[f"prim_{seed}"], bn, axis, limiter, GAMMA)
    np.testing.assert_array_equal(got, g[f"flux_{seed}_{axis}_{limiter}"])


def test_ct_contact_equals_the_reference_kernel():
    g = np.load(GOLDEN)
    fx, fy, ecc = g["ct_fx"], g["ct_fy"], g["ct_ecc"]
    np.testing.assert_array_equal(K.ct_contact(fx[5], fy[5], fx[0], fy[0], ecc), g["ct_ez"])


def test_golden_inputs_are_the_reference_test_blocks():
    g = np.load(GOLDEN)
    prim, bxf, byf = block(1)
    np.testing.assert_array_equal(g["prim_1"], prim)
    np.testing.assert_array_equal(g["bxf_1"], bxf)


def test_a_planted_sign_error_in_the_double_star_state_fails_the_golden(monkeypatch):
    """The golden vectors catch a sign error in the HLLD double-star energy (the
    rotational-discontinuity term of the left and right states swapped in sign)."""
    g = np.load(GOLDEN)
    assert np.array_equal(K.sweep(g["prim_1"], g["bxf_1"], 0, 0, GAMMA), g["flux_1_0_0"])
    monkeypatch.setattr(K, "DOUBLE_STAR_SENSE", (1.0, -1.0))
    got = K.sweep(g["prim_1"], g["bxf_1"], 0, 0, GAMMA)
    assert not np.array_equal(got, g["flux_1_0_0"])


# -- closed forms of the Riemann solver (test_mhd.py:38-74) ---------------------------------
@pytest.mark.parametrize("bn", [0.0, 0.3, -1.2])
def test_hlld_is_consistent(bn):
    state = _state(1.3, 0.4, -0.2, 0.7, 0.9, 0.5, -0.8)
    flux = K.hlld_flux(state, state, np.array([bn]), GAMMA)
    np.testing.assert_allclose(flux, _physical_flux(state, bn), atol=1e-13)


def test_hlld_upwinds_supersonic_flow():
    left = _state(1.0, 20.0, 0.1, 0.0, 1.0, 0.2, 0.1)
    right = _state(0.5, 19.0, -0.3, 0.2, 0.4, -0.1, 0.0)
    flux = K.hlld_flux(left, right, np.array([0.5]), GAMMA)
    np.testing.assert_allclose(flux, _physical_flux(left, 0.5), atol=1e-12)
    left[1], right[1] = -19.0, -20.0
    flux = K.hlld_flux(left, right, np.array([0.5]), GAMMA)
    np.testing.assert_allclose(flux, _physical_flux(right, 0.5), atol=1e-12)


def test_hlld_mirror_symmetry():
    left = _state(1.0, 0.3, 0.1, -0.2, 1.0, 0.6, 0.2)
    right = _state(0.3, -0.1, 0.4, 0.1, 0.2, -0.5, 0.3)
    bn = np.array([0.7])
    flux = K.hlld_flux(left, right, bn, GAMMA)
    mirror = np.array([[1], [-1], [1], [1], [1], [1], [1]])
    mirrored = K.hlld_flux(right * mirror, left * mirror, -bn, GAMMA)
    expected = flux * np.array([[-1], [1], [-1], [-1], [-1], [-1], [-1]])
    np.testing.assert_allclose(mirrored, expected, atol=1e-13)


def test_hlld_resolves_an_isolated_contact():
    left = _state(1.0, 0.0, 0.0, 0.0, 1.0, 0.5, 0.0)
    right = _state(0.2, 0.0, 0.0, 0.0, 1.0, 0.5, 0.0)
    flux = K.hlld_flux(left, right, np.array([0.8]), GAMMA)
    assert abs(flux[0, 0]) < 1e-14
    np.testing.assert_allclose(flux[1, 0], 1.0 + 0.5 * 0.25 - 0.5 * 0.64)


def test_hlld_without_field_is_hllc():
    """B = 0: HLLD reduces to HLLC (Toro), whose star flux is written out here."""
    left = _state(1.0, 0.5, 0.2, -0.1, 1.0, 0.0, 0.0)
    right = _state(0.4, -0.3, -0.4, 0.3, 0.3, 0.0, 0.0)
    flux = K.hlld_flux(left, right, np.array([0.0]), GAMMA)
    rl, ul, pl = 1.0, 0.5, 1.0
    rr, ur, pr = 0.4, -0.3, 0.3
    cl, cr = np.sqrt(GAMMA * pl / rl), np.sqrt(GAMMA * pr / rr)
    sl, sr = min(ul, ur) - max(cl, cr), max(ul, ur) + max(cl, cr)
    sm = (pr - pl + rl * ul * (sl - ul) - rr * ur * (sr - ur)) / (rl * (sl - ul) - rr * (sr - ur))
    assert sl < 0 < sm  # the left star region
    el = pl / (GAMMA - 1) + 0.5 * rl * (ul**2 + 0.2**2 + 0.1**2)
    ustar = rl * (sl - ul) / (sl - sm) * np.array([
        1.0, sm, 0.2, -0.1,
        el / rl + (sm - ul) * (sm + pl / (rl * (sl - ul))),
    ])
    u = np.array([rl, rl * ul, rl * 0.2, -rl * 0.1, el])
    f = _physical_flux(left, 0.0)[:5, 0]
    np.testing.assert_allclose(flux[:5, 0], f + sl * (ustar - u), rtol=1e-12, atol=1e-14)
    assert flux[5, 0] == 0.0 and flux[6, 0] == 0.0


# -- limiters (limiters.py:18-46) -------------------------------------------------------------
def test_limiters_closed_forms():
    a = np.array([1.0, -2.0, 3.0, 0.0, 1.0])
    b = np.array([2.0, -1.0, -1.0, 1.0, 4.0])
    np.testing.assert_array_equal(K.limited_slope(a, b, 0), [1.0, -1.0, 0.0, 0.0, 1.0])
    np.testing.assert_allclose(K.limited_slope(a, b, 1), [4 / 3, -4 / 3, 0.0, 0.0, 1.6])
    np.testing.assert_allclose(K.limited_slope(a, b, 2), [1.5, -1.5, 0.0, 0.0, 2.0])
    np.testing.assert_array_equal(K.limited_slope(a, b, 3), 0.5 * (a + b))


# -- primitive and rate -----------------------------------------------------------------------
def test_primitive_round_trip():
    rng = np.random.default_rng(11)
    shape = (9, 7)
    rho = rng.uniform(0.5, 2.0, shape)
    u, v, w, bz = (rng.normal(size=shape) for _ in range(4))
    p = rng.uniform(0.1, 1.0, shape)
    az = 0.1 * rng.normal(size=(shape[0] + 1, shape[1] + 1))
    dx, dy, b0 = 0.1, 0.2, (0.3, -0.4)
    bxf = (az[:, 1:] - az[:, :-1]) / dy + b0[0]
    byf = -(az[1:, :] - az[:-1, :]) / dx + b0[1]
    bx, by = 0.5 * (bxf[:-1] + bxf[1:]), 0.5 * (byf[:, :-1] + byf[:, 1:])
    energy = p / (GAMMA - 1) + 0.5 * rho * (u**2 + v**2 + w**2) + 0.5 * (bx**2 + by**2 + bz**2)
    prim, bxf2, byf2, ecc, bad = K.primitive(rho, rho * u, rho * v, rho * w, energy, bz, az,
                                              dx, dy, *b0, GAMMA)
    assert bad == 0
    np.testing.assert_allclose(prim, np.stack([rho, u, v, w, p, bx, by, bz]), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(ecc, bx * v - by * u, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize(("row", "value"),
                         [(0, -1.0), (4, -1.0e3), (1, np.inf), (2, np.nan), (5, -np.inf), (0, np.nan)])
def test_primitive_counts_the_bad_cells(row, value):
    fields, az = conserved()
    fields[row][3, 4] = value
    fields[row][7, 1] = value
    assert K.primitive(*fields, az, 0.1, 0.2, 0.3, -0.4, GAMMA)[-1] == 2


def test_rate_closed_forms():
    """Without field the fast speed is the sound speed; with only B_x along x it is
    max(a, |B_x|/sqrt(rho)) along x and sqrt(a^2 + B_x^2/rho) along y."""
    shape = (8, 6)
    rho = np.full(shape, 2.0)
    u, v = 0.3, -0.7
    p = 0.5
    fields = [rho, rho * u, rho * v, 0 * rho]
    a2 = GAMMA * p / 2.0
    energy = np.full(shape, p / (GAMMA - 1) + 0.5 * 2.0 * (u * u + v * v))
    az = np.zeros((shape[0] + 1, shape[1] + 1))
    largest, smallest = K.rate(*fields, energy, 0 * rho, az, 2, 0.1, 0.2, 0.0, 0.0, GAMMA)
    np.testing.assert_allclose(largest, (abs(u) + np.sqrt(a2)) / 0.1 + (abs(v) + np.sqrt(a2)) / 0.2)
    assert smallest == 2.0
    bx = 1.5
    energy = energy + 0.5 * bx * bx
    largest, _ = K.rate(*fields, energy, 0 * rho, az, 2, 0.1, 0.2, bx, 0.0, GAMMA)
    cx = max(np.sqrt(a2), bx / np.sqrt(2.0))
    cy = np.sqrt(a2 + bx * bx / 2.0)
    np.testing.assert_allclose(largest, (abs(u) + cx) / 0.1 + (abs(v) + cy) / 0.2, rtol=1e-14)


def test_ct_contact_of_a_uniform_state_is_the_cell_field():
    shape = (10, 8)
    ecc = np.full(shape, 0.37)
    fx = np.full((7, shape[0] - 3, shape[1]), 0.0)
    fy = np.full((7, shape[0], shape[1] - 3), 0.0)
    fx[5] = -0.37  # x-flux of B_y is -E_z
    fy[5] = 0.37
    fx[0] = 1.0
    fy[0] = -1.0
    np.testing.assert_allclose(K.ct_contact(fx[5], fy[5], fx[0], fy[0], ecc), 0.37, atol=1e-16)


# -- whole runs (test_mhd.py:89-151, test_mhd3d.py:24-38) -----------------------------------
def test_orszag_tang_keeps_div_b_zero_and_conserves():
    model = oracle2d(initial.orszag_tang(32))
    before = model.totals()
    model.run(0.1)
    after = model.totals()
    assert model.divergence() < 1e-14
    np.testing.assert_allclose(after, before, rtol=0, atol=1e-13)


def test_uniform_state_stays_uniform_with_outflow_faces():
    h = AMRHierarchy((8, 6), ((0.0, 1.0), (0.0, 1.0)))
    problem = initial.Problem(h, {"mean_field": (0.3, -0.2), "limiter": "mc"},
                              dict(density=1.0, velocity=(0.2, -0.1, 0.05), pressure=0.5, bz=0.4))
    model = oracle2d(problem)
    model.run(0.2)
    c = model.interior()
    np.testing.assert_allclose(c[0], 1.0, atol=1e-13)
    np.testing.assert_allclose(c[5], 0.4, atol=1e-13)


def _brio_wu(axis, cells=100):
    if axis == 0:
        h = AMRHierarchy((cells, 2), ((0.0, 1.0), (0.0, 2.0 / cells)), periodic=(False, True))
    else:
        h = AMRHierarchy((2, cells), ((0.0, 2.0 / cells), (0.0, 1.0)), periodic=(True, False))
    mean = (0.75, 0.0) if axis == 0 else (0.0, 0.75)

    def s(x, y):
        return x if axis == 0 else y

    sign = 1.0 if axis == 0 else -1.0
    problem = initial.Problem(h, {"gamma": 2.0, "mean_field": mean}, dict(
        density=lambda x, y: np.where(s(x, y) < 0.5, 1.0, 0.125), velocity=(0.0, 0.0, 0.0),
        pressure=lambda x, y: np.where(s(x, y) < 0.5, 1.0, 0.1),
        vector_potential=lambda x, y: sign * np.abs(s(x, y) - 0.5)))
    model = oracle2d(problem)
    model.run(0.1)
    return model


def test_brio_wu_along_x_and_y_agree():
    rx = _brio_wu(0).interior()[0][:, 0]
    ry = _brio_wu(1).interior()[0][0, :]
    np.testing.assert_allclose(rx, ry, atol=1e-12)
    assert 0.11 < rx.min() and rx.max() <= 1.0 + 1e-12
    assert 0.6 < rx[50] < 0.75


def test_a_state_uniform_along_z_is_the_2d_solution():
    flat = oracle3d(initial.orszag_tang_3d(16, 2))
    plane = oracle2d(initial.orszag_tang(16))
    for _ in range(4):
        dt = plane.stable_step()
        plane.step(dt)
        flat.step(dt)
    c3 = flat.interior()
    c2 = plane.interior()
    for row in range(4):  # rho, mx, my and mz
        np.testing.assert_allclose(c3[row], np.repeat(c2[row][:, :, None], 2, axis=2), atol=1e-13)
    assert flat.divergence() < 1e-14


def test_3d_conserves_and_keeps_div_b_zero():
    model = oracle3d(initial.turbulence_3d(8, limiter="minmod"))
    before = model.totals()
    for _ in range(3):
        model.step()
    np.testing.assert_allclose(model.totals(), before, rtol=0, atol=1e-13)
    assert model.divergence() < 1e-14


def test_turbulence_fft_evaluation_is_exact():
    """The FFT evaluation equals the direct sum of sines at the staggered points."""
    problem = initial.turbulence_3d(8, seed=3)
    h = problem.hierarchy
    az = problem.state["vector_potential"][2]
    x, y, z = h.edge_coordinates(h.patches[0], 2)
    # direct sum with the same generator draws
    rng = np.random.default_rng(3)
    modes = np.array([(a, b, c) for a in range(-4, 5) for b in range(-4, 5) for c in range(-4, 5)
                      if 1 <= a * a + b * b + c * c <= 16], dtype=float)
    k = 2 * np.pi * modes
    k2 = np.sum(k * k, axis=1)

    def draw():
        amp = rng.standard_normal((len(modes), 3))
        amp -= (np.sum(amp * k, axis=1) / k2)[:, None] * k
        return amp, rng.uniform(0, 2 * np.pi, len(modes))

    draw()
    amp, ph = draw()
    curl = np.cross(k, amp)
    amp = amp * (0.1 / np.sqrt(np.sum(curl**2) / 2.0))
    direct = np.zeros(x.shape)
    for m in range(len(modes)):
        direct += amp[m, 2] * np.sin(k[m, 0] * x + k[m, 1] * y + k[m, 2] * z + ph[m])
    np.testing.assert_allclose(az, direct, atol=1e-12)


def test_oracle_step_is_second_order_in_time():
    """On a fixed grid the smooth Alfvén wave run with dt and dt/2 differs from a run
    with dt/16 by a ratio of about 4: the step is second order in time (Heun), where
    forward Euler or a wrong stage average would give about 2."""
    def run(dt, steps):
        o = oracle2d(initial.alfven_wave(8, limiter="none"))
        for _ in range(steps):
            o.step(dt)
        return o.interior().copy()

    base = 0.01
    ref = run(base / 16, 160)
    errors = [np.abs(run(base / m, 10 * m) - ref).max() for m in (1, 2)]
    assert 3.5 < errors[0] / errors[1] < 4.5, errors


def test_a_first_order_stage_combination_fails_the_time_order(monkeypatch):
    """The time-order check catches a wrong combination of the stages (the second
    stage's result kept instead of the Heun average)."""
    from oracle import mhd2d

    def euler_euler(self, dt=None):
        self.advance(dt)
        self.time += dt
        self.steps += 1
        return dt

    monkeypatch.setattr(mhd2d.Oracle2D, "step", euler_euler)

    def run(dt, steps):
        o = oracle2d(initial.alfven_wave(8, limiter="none"))
        for _ in range(steps):
            o.step(dt)
        return o.interior().copy()

    ref = run(0.01 / 16, 160)
    errors = [np.abs(run(0.01 / m, 10 * m) - ref).max() for m in (1, 2)]
    assert errors[0] / errors[1] < 3.0, errors


@pytest.mark.parametrize("limiter", range(4))
def test_the_sweep_on_a_face_at_rest_is_the_golden(limiter):
    """The rotor columns at rest (B_n at round-off): the oracle is the reference's own
    sweep bit for bit."""
    g = np.load(GOLDEN)
    got = K.sweep(g["rotor_prim"], g["rotor_byf"], 1, limiter, float(g["rotor_gamma"]))
    np.testing.assert_array_equal(got, g[f"rotor_flux_{limiter}"])
