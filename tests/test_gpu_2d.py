"""GPU parity of the 2D path against the CPU oracle (FP64: bit for bit)."""

from pathlib import Path

import numpy as np
import pytest

from oracle import kernels as K
from tests.helpers import block, conserved, oracle2d
from wellbalanced_amd import MHD2D, initial
from wellbalanced_amd.hierarchy import AMRHierarchy

pytestmark = pytest.mark.gpu
GAMMA = 5.0 / 3.0
GOLDEN = Path(__file__).resolve().parent / "golden" / "golden.npz"


def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def _empty(shape, dtype=None):
    import torch

    return torch.full(shape, float("nan"), dtype=torch.float64, device="cuda")


def _call(name, *args):
    from wellbalanced_amd import _lib

    import torch

    _lib.call(name, *args, _lib.stream_handle())
    torch.cuda.synchronize()


def _p(t):
    return t.data_ptr()


@pytest.mark.parametrize("seed", [1, 7])
@pytest.mark.parametrize("limiter", range(4))
@pytest.mark.parametrize("axis", [0, 1])
def test_sweep_twin_is_bitwise_the_oracle(gpu, seed, limiter, axis):
    prim, bxf, byf = block(seed)
    X, Y = prim.shape[1:]
    shape = (7, X - 3, Y) if axis == 0 else (7, X, Y - 3)
    out = _empty(shape)
    dp, db = _dev(prim), _dev((bxf, byf)[axis])
    _call("wb_sweep2d", _p(dp), _p(db), X, Y, axis, limiter, GAMMA, _p(out))
    expected = K.sweep(prim, (bxf, byf)[axis], axis, limiter, GAMMA)
    np.testing.assert_array_equal(out.cpu().numpy(), expected)


def test_ct_twin_is_bitwise_the_oracle(gpu):
    rng = np.random.default_rng(2)
    X, Y = 13, 9
    fx = rng.normal(size=(7, X - 3, Y))
    fy = rng.normal(size=(7, X, Y - 3))
    fx[0, 3, 4] = fy[0, 2, 5] = 0.0
    ecc = rng.normal(size=(X, Y))
    d = [_dev(a) for a in (fx[5], fy[5], fx[0], fy[0], ecc)]
    out = _empty((X - 3, Y - 3))
    _call("wb_ct_contact", *map(_p, d), X, Y, _p(out))
    np.testing.assert_array_equal(out.cpu().numpy(), K.ct_contact(fx[5], fy[5], fx[0], fy[0], ecc))


def _primitive_gpu(fields, az):
    import torch

    X, Y = fields[0].shape
    d = [_dev(f) for f in fields] + [_dev(az)]
    prim, bxf, byf, ecc = _empty((8, X, Y)), _empty((X + 1, Y)), _empty((X, Y + 1)), _empty((X, Y))
    bad = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    _call("wb_primitive2d", *map(_p, d), X, Y, 0.1, 0.2, 0.3, -0.4, GAMMA, _p(prim), _p(bxf),
          _p(byf), _p(ecc), _p(bad))
    return prim.cpu().numpy(), bxf.cpu().numpy(), byf.cpu().numpy(), ecc.cpu().numpy(), int(bad.item())


def test_primitive_twin_is_bitwise_the_oracle(gpu):
    fields, az = conserved()
    got = _primitive_gpu(fields, az)
    want = K.primitive(*fields, az, 0.1, 0.2, 0.3, -0.4, GAMMA)
    assert got[-1] == want[-1] == 0
    for a, b in zip(got[:-1], want[:-1]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize(("row", "value"),
                         [(0, -1.0), (4, -1.0e3), (1, np.inf), (2, np.nan), (5, -np.inf), (0, np.nan)])
def test_primitive_twin_counts_the_bad_cells(gpu, row, value):
    fields, az = conserved()
    fields[row][3, 4] = value
    fields[row][7, 1] = value
    assert _primitive_gpu(fields, az)[-1] == 2


def test_rate_twin_is_bitwise_the_oracle(gpu):
    fields, az = conserved()
    X, Y = fields[0].shape
    d = [_dev(f) for f in fields] + [_dev(az)]
    out = _empty((2,))
    _call("wb_rate2d", *map(_p, d), X, Y, 2, 0.1, 0.2, 0.3, -0.4, GAMMA, _p(out))
    largest, smallest = K.rate(*fields, az, 2, 0.1, 0.2, 0.3, -0.4, GAMMA)
    got = out.cpu().numpy()
    assert got[0] == largest
    assert got[1] == smallest == np.min(fields[0][2:-2, 2:-2])


def _compare(model, oracle, atol=0.0, rtol=0.0):
    cons, az = model.arrays()
    h = 2
    nx, ny = model.nx, model.ny
    got = cons[:, h : h + nx, h : h + ny]
    want = oracle.interior()
    if atol == 0.0 and rtol == 0.0:
        np.testing.assert_array_equal(got, want)
        np.testing.assert_array_equal(az[h : h + nx + 1, h : h + ny + 1], oracle.nodes())
    else:
        scale = np.abs(want).max(axis=(1, 2), keepdims=True)
        assert np.all(np.abs(got - want) <= atol + rtol * scale)


@pytest.mark.parametrize("limiter", ["minmod", "vanleer", "mc", "none"])
def test_orszag_tang_steps_are_bitwise_the_oracle(gpu, limiter):
    problem = initial.orszag_tang(32, limiter=limiter)
    model = MHD2D.from_problem(problem)
    oracle = oracle2d(problem)
    for _ in range(5):
        dt = oracle.stable_step()
        assert model.stable_step() == dt
        model.step(dt)
        oracle.step(dt)
    _compare(model, oracle)


def test_orszag_tang_256_matches_to_1e12(gpu):
    """Config 1's grid: 256^2 OT, minmod, 20 steps; every conserved variable within
    1e-12 of its largest magnitude (in practice bit for bit)."""
    problem = initial.orszag_tang(256, limiter="minmod")
    model = MHD2D.from_problem(problem)
    oracle = oracle2d(problem)
    for _ in range(20):
        dt = oracle.stable_step()
        model.step(dt)
        oracle.step(dt)
    _compare(model, oracle, rtol=1e-12)


def test_walls_and_outflow_match_the_oracle(gpu):
    problem = initial.harris_sheet(64, 32, length=3.2, noise=1e-2, seed=4)
    model = MHD2D.from_problem(problem)
    oracle = oracle2d(problem)
    for _ in range(8):
        dt = oracle.stable_step()
        model.step(dt)
        oracle.step(dt)
    _compare(model, oracle)
    h = AMRHierarchy((24, 20), ((0.0, 1.0), (0.0, 1.0)))
    problem = initial.Problem(h, {"limiter": "mc", "mean_field": (0.2, 0.1)}, dict(
        density=lambda x, y: 1.0 + 0.3 * np.exp(-((x - 0.4) ** 2 + (y - 0.6) ** 2) / 0.01),
        velocity=(0.3, -0.2, 0.1), pressure=1.0,
        vector_potential=lambda x, y: 0.05 * np.sin(3 * x) * np.cos(2 * y), bz=0.2))
    model = MHD2D.from_problem(problem)
    oracle = oracle2d(problem)
    for _ in range(6):
        dt = oracle.stable_step()
        model.step(dt)
        oracle.step(dt)
    _compare(model, oracle)


def test_orszag_tang_keeps_div_b_zero_and_conserves(gpu):
    model = MHD2D.orszag_tang(32)
    before = model.totals()
    model.run(0.1)
    after = model.totals()
    assert model.divergence() < 1e-14
    for name in before:
        assert abs(after[name] - before[name]) < 1e-13
    assert float(model.dataset()["pressure"].min()) > 0.0


def test_walls_conserve_and_keep_the_normal_field_zero(gpu):
    h = AMRHierarchy((24, 24))
    model = MHD2D(h, boundaries={"x": "wall", "y": "wall"})
    model.set_state(
        1.0, (0.3, -0.2, 0.1),
        lambda x, y: 1.0 + 2.0 * np.exp(-((x - 0.3) ** 2 + (y - 0.6) ** 2) / 0.005),
        vector_potential=lambda x, y: 0.1 * np.sin(np.pi * x) * np.sin(np.pi * y), bz=0.2,
    )
    before = model.totals()
    model.run(0.3)
    after = model.totals()
    for name in ("density", "energy", "momentum_z", "bz"):
        assert abs(after[name] - before[name]) < 1e-13
    _, az = model.arrays()
    a = az[2:-2, 2:-2]
    for edge in (a[0], a[-1], a[:, 0], a[:, -1]):
        assert np.abs(edge).max() < 1e-15


def test_reports_a_negative_pressure(gpu):
    model = MHD2D(AMRHierarchy((8, 8), periodic=(True, True)))
    model.set_state(1.0, (0.0, 0.0, 0.0), 1e-3)
    cons, az = model.arrays()
    cons[4, 5, 5] = -1.0
    model.load_arrays(cons, az)
    with pytest.raises(FloatingPointError, match="pressure"):
        model.step(1e-3)


def test_uniform_state_stays_uniform(gpu):
    model = MHD2D(AMRHierarchy((8, 6), ((0.0, 1.0), (0.0, 1.0))), mean_field=(0.3, -0.2), limiter="mc")
    model.set_state(1.0, (0.2, -0.1, 0.05), 0.5, bz=0.4)
    model.run(0.2)
    data = model.dataset()
    np.testing.assert_allclose(data["density"], 1.0, atol=1e-13)
    np.testing.assert_allclose(data["magnetic_field_x"], 0.3, atol=1e-13)
    np.testing.assert_allclose(data["pressure"], 0.5, atol=1e-12)


def test_needs_a_state_first(gpu):
    model = MHD2D(AMRHierarchy((4, 4), periodic=(True, True)))
    with pytest.raises(RuntimeError, match="set_state"):
        model.step()
    with pytest.raises(ValueError, match="positive"):
        model.set_state(1.0, (0.0, 0.0, 0.0), -1.0)
    with pytest.raises(ValueError, match="three"):
        model.set_state(1.0, (0.0, 0.0), 1.0)


@pytest.mark.parametrize("tile,nxs", [(126, 32), (62, 5), (254, 64), (30, 4)])
@pytest.mark.parametrize("limiter", ["minmod", "vanleer", "mc", "none"])
def test_fused_fp64_stage_is_bitwise_the_split_path(gpu, tile, nxs, limiter):
    for problem in (initial.orszag_tang(40, limiter=limiter),
                    initial.harris_sheet(70, 36, length=3.2, noise=1e-2, seed=1, limiter=limiter)):
        split = MHD2D.from_problem(problem)
        fused = MHD2D.from_problem(problem, kernels="fused")
        fused.kernel, fused.tile, fused.nxs = "march", tile, nxs
        for _ in range(4):
            dt = split.stable_step()
            split.step(dt)
            fused.step(dt)
        a, aa = split.arrays()
        b, bb = fused.arrays()
        np.testing.assert_array_equal(b[:, 2:-2, 2:-2], a[:, 2:-2, 2:-2])
        np.testing.assert_array_equal(bb[2:-2, 2:-2], aa[2:-2, 2:-2])


@pytest.mark.parametrize("limiter", ["minmod", "vanleer", "mc", "none"])
def test_tiled_fp64_stage_is_bitwise_the_split_path(gpu, limiter):
    """The tiled stage (a workgroup per 16 x 16 tile) on grids that are not multiples of
    the tile: the same bits as the split kernels, walls included."""
    for problem in (initial.orszag_tang(40, limiter=limiter),
                    initial.harris_sheet(70, 36, length=3.2, noise=1e-2, seed=1, limiter=limiter)):
        split = MHD2D.from_problem(problem)
        tiled = MHD2D.from_problem(problem, kernels="fused")
        tiled.kernel = "tiles"
        for _ in range(4):
            dt = split.stable_step()
            split.step(dt)
            tiled.step(dt)
        a, aa = split.arrays()
        b, bb = tiled.arrays()
        np.testing.assert_array_equal(b[:, 2:-2, 2:-2], a[:, 2:-2, 2:-2])
        np.testing.assert_array_equal(bb[2:-2, 2:-2], aa[2:-2, 2:-2])


@pytest.mark.parametrize("limiter", ["minmod", "vanleer", "mc", "none"])
def test_tiled_mixed_stage_is_bitwise_the_marching_one(gpu, limiter):
    for problem in (initial.orszag_tang(40, limiter=limiter),
                    initial.harris_sheet(70, 36, length=3.2, noise=1e-2, seed=1, limiter=limiter)):
        march = MHD2D.from_problem(problem, kernels="fused", precision="mixed")
        march.kernel = "march"
        tiled = MHD2D.from_problem(problem, kernels="fused", precision="mixed")
        tiled.kernel = "tiles"
        for _ in range(4):
            dt = march.stable_step()
            march.step(dt)
            tiled.step(dt)
        a, aa = march.arrays()
        b, bb = tiled.arrays()
        np.testing.assert_array_equal(b, a)
        np.testing.assert_array_equal(bb, aa)


def test_fused_outflow_and_walls_match_the_oracle(gpu):
    h = AMRHierarchy((37, 29), ((0.0, 1.0), (0.0, 1.0)))
    problem = initial.Problem(h, {"limiter": "mc", "boundaries": {"x": "wall"}, "mean_field": (0.0, 0.1)}, dict(
        density=lambda x, y: 1.0 + 0.3 * np.exp(-((x - 0.4) ** 2 + (y - 0.6) ** 2) / 0.01),
        velocity=(0.3, -0.2, 0.1), pressure=1.0,
        vector_potential=lambda x, y: 0.05 * np.sin(3 * x) * np.cos(2 * y), bz=0.2))
    for kernel in ("march", "tiles"):
        model = MHD2D.from_problem(problem, kernels="fused")
        model.nxs = 8
        model.kernel = kernel
        oracle = oracle2d(problem)
        for _ in range(6):
            dt = oracle.stable_step()
            model.step(dt)
            oracle.step(dt)
        _compare(model, oracle)


def test_mixed_2d_is_close_to_fp64(gpu):
    problem = initial.orszag_tang(128, limiter="minmod")
    fp64 = MHD2D.from_problem(problem, kernels="fused")
    mixed = MHD2D.from_problem(problem, kernels="fused", precision="mixed")
    before = mixed.totals()
    for _ in range(20):
        dt = fp64.stable_step()
        fp64.step(dt)
        mixed.step(dt)
    l1 = float(np.mean(np.abs(fp64.dataset()["density"] - mixed.dataset()["density"])))
    assert 0.0 < l1 < 1e-5
    assert mixed.divergence() < 1e-14
    after = mixed.totals()
    for name in before:
        assert abs(after[name] - before[name]) < 1e-13


def test_mixed_harris_wings_stay_close_to_fp64(gpu):
    """The Harris sheet's uniform wings (|B_x| = 1, no tangential field, c_A > c_s, white
    velocity noise) put the FP32 HLLD star denominator at its rounding noise: the
    mixed path must treat it as degenerate there.  Tolerances after 20 steps: max |dp|
    < 1e-4 (before the FP32 guard, single A_z nodes jumped and p fell to 0.003), L1 of
    rho < 1e-6 (measured 1.9e-7)."""
    from wellbalanced_amd import MHD2D

    problem = initial.harris_sheet(2048, 1024)
    a = MHD2D.from_problem(problem, kernels="fused", precision="fp64")
    b = MHD2D.from_problem(problem, kernels="fused", precision="mixed")
    for _ in range(20):
        dt = a.stable_step()
        a.step(dt)
        b.step(dt)
    da, db = a.dataset(), b.dataset()
    assert np.abs(da["pressure"] - db["pressure"]).max() < 1e-4
    assert np.abs(da["density"] - db["density"]).mean() < 1e-6
    assert b.divergence() < 1e-14


@pytest.mark.parametrize("limiter", range(4))
def test_sweep_twin_on_a_face_at_rest_is_the_golden_to_round_off(gpu, limiter):
    """The reference's own sweep on four columns of the rotor at rest (B_n at round-off,
    tangential fluxes cancelling to 1e-15): the GPU's fluxes equal the golden vector to
    1e-15 of the largest flux.  Not always bit for bit: gfx950's FP64 division misrounds
    the exact near-tie 1.3999999999999992 / 0.9999999999999994 (1.4, IEEE
    1.4000000000000001) that gamma p / rho meets there, and the fast speed's last bit
    moves the tangential fluxes by about 4e-16."""
    g = np.load(GOLDEN)
    prim, byf, gamma = g["rotor_prim"], g["rotor_byf"], float(g["rotor_gamma"])
    X, Y = prim.shape[1:]
    out = _empty((7, X, Y - 3))
    dp, db = _dev(prim), _dev(byf)  # held: their memory must outlive the call
    _call("wb_sweep2d", _p(dp), _p(db), X, Y, 1, limiter, gamma, _p(out))
    want = g[f"rotor_flux_{limiter}"]
    assert np.abs(out.cpu().numpy() - want).max() <= 1e-15 * np.abs(want).max()
