"""MHD3D on a refined 3D hierarchy of patches of any size (the reference's MHD3D on
patchwork patches, mhd3d.py:448-473,726-821): the oracle's invariants on the CPU and the
GPU path against the oracle (FP64), plus the reference's refined and regridding checks
(tests/unit/test_mhd3d.py:41-92)."""

import numpy as np
import pytest

from oracle.mhd3d import Oracle3D
from oracle.patches3d import OraclePatches3D, pack, regrid_state, unpack
from wellbalanced_amd import initial
from wellbalanced_amd.hierarchy import AMRHierarchy

CONSERVED = ("density", "momentum_x", "momentum_y", "momentum_z", "energy")


def _alfven_mesh():
    """test_mhd3d.py:41-48: the Alfvén wave's 8-cell mesh doubled along x with a refined
    patch of 12 x 8 x 8 fine cells."""
    mesh = AMRHierarchy((16, 8, 8), ((0.0, 2.0), (0.0, 1.0), (0.0, 1.0)), periodic=(True,) * 3)
    mesh.add_patch(1, ((8, 20), (4, 12), (2, 10)))
    return mesh


def _alfven_state():
    """test_mhd3d.py:49-70."""
    wavevector = 2 * np.pi * np.array([0.5, 1.0, 1.0])
    k = np.linalg.norm(wavevector)
    n = wavevector / k
    e1 = np.cross(n, [0.0, 0.0, 1.0])
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(n, e1)

    def perpendicular(x, y, z):
        phase = wavevector[0] * x + wavevector[1] * y + wavevector[2] * z
        return [0.1 * (np.sin(phase) * e1[i] + np.cos(phase) * e2[i]) for i in range(3)]

    def potential(i):
        return lambda x, y, z: perpendicular(x, y, z)[i] / k

    return tuple(n), dict(density=1.0, velocity=lambda x, y, z: [-b for b in perpendicular(x, y, z)],
                          pressure=0.1, vector_potential=(potential(0), potential(1), potential(2)))


def _oracle(mesh, state, mean_field, **options):
    patches = [(p.level, p.box) for p in mesh.patches]
    per = []
    for p in mesh.patches:
        cons, pots = initial.host_state_3d(mesh, p, 5.0 / 3.0, mean_field, **state)
        per.append([cons, *pots])
    o = OraclePatches3D(patches, mesh.base_shape, mesh.bounds, pack(per, patches),
                        mean_field=mean_field, **options)
    o.coarsen(o.state, cells=True)
    return o


def _composite(o):
    """The largest mismatch of a restricted coarse edge and the mean of its fine edges."""
    worst = 0.0
    for L in range(o.nlev - 1):
        fshape = o.shape(L + 1)
        for name in ("ax", "ay", "az"):
            fine = o._level_values(o.state, L + 1, name)[0]
            c = "xyz".index(name[1])
            for k, (lev, box) in enumerate(o.patches):
                if lev != L:
                    continue
                arr = o.state[k][1 + c]
                nd = [a != c for a in range(3)]
                for C in np.ndindex(*[box[a][1] - box[a][0] + nd[a] for a in range(3)]):
                    G = [box[a][0] + C[a] for a in range(3)]
                    t = [[(2 * G[a] + (s if a == c else 0)) % fshape[a] for a in range(3)] for s in (0, 1)]
                    v = [fine[tuple(x)] for x in t]
                    if np.isnan(v).any():
                        continue
                    worst = max(worst, abs(arr[tuple(C[a] + 2 for a in range(3))] - 0.5 * (v[0] + v[1])))
    return worst


def _div(o):
    from oracle.mhd3d import face_fields

    largest = 0.0
    for k, (L, box) in enumerate(o.patches):
        n = [hi - lo for lo, hi in box]
        sp = o.spacing(L)
        local = [o.state[k][1 + c][tuple(slice(2, 2 + n[a] + (a != c)) for a in range(3))]
                 for c in range(3)]
        faces = face_fields(local, sp, o.b0)
        div = sum(np.diff(f, axis=a) / sp[a] for a, f in enumerate(faces))
        largest = max(largest, float(np.abs(div).max()) * min(sp))
    return largest


# -- the oracle ----------------------------------------------------------------------------
def test_the_refined_alfven_mesh_conserves_and_keeps_div_b_zero():
    b0, state = _alfven_state()
    o = _oracle(_alfven_mesh(), state, b0, limiter="minmod")
    before = o.leaf_totals()
    for _ in range(2):
        o.step()
    np.testing.assert_allclose(o.leaf_totals(), before, rtol=0, atol=1e-14)
    assert _div(o) < 1e-14
    assert _composite(o) == 0.0


def test_a_refined_level_covering_the_domain_is_the_uniform_run():
    """Eight level-1 patches over the whole periodic box evolve bit for bit as one
    uniform grid of their resolution: same-level copies and shared edges are exact."""
    mesh = AMRHierarchy((4, 4, 4), ((0.0, 1.0),) * 3, periodic=(True,) * 3)
    for d in np.ndindex(2, 2, 2):
        mesh.add_patch(1, tuple((4 * d[a], 4 * d[a] + 4) for a in range(3)))
    two_pi = 2.0 * np.pi
    state = dict(
        density=lambda x, y, z: 1.0 + 0.2 * np.sin(two_pi * x) * np.cos(two_pi * y),
        velocity=lambda x, y, z: (0.3 * np.sin(two_pi * z), -0.2 * np.cos(two_pi * x), 0.1 + 0 * y),
        pressure=lambda x, y, z: 1.0 + 0.1 * np.cos(two_pi * (y + z)),
        # cosines: the same value at both ends of a periodic axis, so that the patches'
        # nodes on the seam start equal to the uniform grid's
        vector_potential=(lambda x, y, z: 0.05 * np.cos(two_pi * y),
                          lambda x, y, z: 0.05 * np.cos(two_pi * z),
                          lambda x, y, z: 0.05 * np.cos(two_pi * x)))
    o = _oracle(mesh, state, (1.0, 0.0, 0.0), limiter="minmod")
    h = AMRHierarchy((8, 8, 8), ((0.0, 1.0),) * 3, periodic=(True,) * 3)
    cons, pots = initial.host_state_3d(h, h.patches[0], 5.0 / 3.0, (1.0, 0.0, 0.0), **state)
    uniform = Oracle3D(cons, pots, h.spacing(0), mean_field=(1.0, 0.0, 0.0), limiter="minmod")
    for _ in range(2):
        dt = uniform.stable_step()
        uniform.step(dt)
        o.step(dt)
    for k, (L, box) in enumerate(o.patches):
        if L != 1:
            continue
        sl = tuple(slice(2 + lo, 2 + hi) for lo, hi in box)
        np.testing.assert_array_equal(o.state[k][0][:, 2:-2, 2:-2, 2:-2], uniform.cons[(slice(None), *sl)])


def test_the_regrid_move_keeps_the_integrals_and_div_b():
    b0, state = _alfven_state()
    o = _oracle(_alfven_mesh(), state, b0, limiter="minmod")
    o.step()
    new = [(0, ((0, 16), (0, 8), (0, 8))), (1, ((4, 12), (2, 10), (0, 8))),
           (1, ((12, 20), (2, 10), (0, 8)))]
    moved = regrid_state(o, new)
    np.testing.assert_allclose(moved.leaf_totals(), o.leaf_totals(), rtol=0, atol=1e-14)
    assert _div(moved) < 1e-14
    assert _composite(moved) == 0.0


# -- the GPU path --------------------------------------------------------------------------
def _model(mesh, state, mean_field, **options):
    from wellbalanced_amd import MHD3D

    model = MHD3D(mesh, mean_field=mean_field, **options)
    model.set_state(state["density"], state["velocity"], state["pressure"],
                    vector_potential=state["vector_potential"])
    return model


def _assert_same(model, o, exact=True):
    got = unpack(model.buffer(), o.patches)
    for arrays, want in zip(got, o.state):
        for a, b in zip(arrays, want):
            inner = tuple(slice(2, s - 2) for s in a.shape[-3:])
            x, y = a[(...,) + inner], b[(...,) + inner]
            if exact:
                np.testing.assert_array_equal(x, y)
            else:
                assert np.abs(x - y).max() <= 1e-12 * max(np.abs(y).max(), 1e-300)


@pytest.mark.gpu
@pytest.mark.parametrize("resistivity", [0.0, 2e-3])
def test_gpu_refined_patches_are_the_oracle(gpu, resistivity):
    b0, state = _alfven_state()
    mesh = _alfven_mesh()
    model = _model(mesh, state, b0, limiter="minmod", resistivity=resistivity)
    assert type(model).__name__ == "RefinedMHD3D"
    o = _oracle(mesh, state, b0, limiter="minmod", resistivity=resistivity)
    np.testing.assert_array_equal(model.buffer(), o.buf)
    for _ in range(2):
        dt = o.stable_step()
        assert model.stable_step() == dt
        model.step(dt)
        o.step(dt)
    _assert_same(model, o)


@pytest.mark.gpu
def test_gpu_a_refined_patch_conserves_and_keeps_div_b_zero(gpu):
    """test_mhd3d.py:72-81."""
    b0, state = _alfven_state()
    model = _model(_alfven_mesh(), state, b0)
    start = model.totals()
    model.run(0.1)
    end = model.totals()
    for name in CONSERVED:
        assert abs(end[name] - start[name]) < 1e-14
    assert model.divergence() < 1e-14
    assert model.composite_divergence() < 1e-14


@pytest.mark.gpu
def test_gpu_regridding_on_the_current_keeps_div_b_zero(gpu):
    """test_mhd3d.py:84-92: |J| of the circularly polarised wave is uniform, so the whole
    base level is tagged; the regrid into patches of at most 8 cells keeps div B zero and
    the energy."""
    b0, state = _alfven_state()
    model = _model(_alfven_mesh(), state, b0)
    model.regrid(model.current_tags(0.1, levels=1), max_patch_size=8)
    assert model.hierarchy.max_level == 1
    assert len(model.hierarchy.patches_at_level(1)) > 1
    start = model.totals()
    model.run(0.05)
    assert model.composite_divergence() < 1e-14
    assert abs(model.totals()["energy"] - start["energy"]) < 1e-13


@pytest.mark.gpu
def test_gpu_regrid_move_is_the_oracle(gpu):
    b0, state = _alfven_state()
    mesh = _alfven_mesh()
    model = _model(mesh, state, b0, limiter="minmod")
    model.step(model.stable_step())
    old = OraclePatches3D.from_hierarchy(model.hierarchy, model.buffer(), mean_field=b0,
                                         limiter="minmod")
    tags = {0: np.zeros((16, 8, 8), bool)}
    tags[0][5:11, 2:6, 1:7] = True
    model.regrid(tags, max_patch_size=8)
    want = regrid_state(old, [(p.level, p.box) for p in model.hierarchy.patches])
    _assert_same(model, want)


@pytest.mark.gpu
def test_gpu_resistive_refined_run_conserves(gpu):
    """Resistivity on the refined mesh: the Ohmic heating stays in the energy, the
    resistive energy fluxes are refluxed (test_mhd_amr.py:109-134 in 3D)."""
    b0, state = _alfven_state()
    model = _model(_alfven_mesh(), state, b0, resistivity=5e-3)
    start = model.totals()
    model.run(0.05)
    end = model.totals()
    for name in CONSERVED:
        assert abs(end[name] - start[name]) < 1e-14
    assert model.composite_divergence() < 1e-14
