"""The configured workloads at their full sizes, through size-independent properties
(the oracle is too slow there): the leaf integrals of the conserved fields, div B on
every patch and across levels, and mixed precision within the stated L1 of FP64.
C1 (256^2) and C2 (256^3) are covered in test_gpu_2d / test_gpu_3d."""

from __future__ import annotations

import numpy as np
import pytest

from wellbalanced_amd import initial

pytestmark = pytest.mark.gpu


def test_c3_blast_on_three_levels_of_512_squared(gpu):
    """C3: 512^2 base, 3 levels of patches of 16..128 cells; FP64 and mixed."""
    from wellbalanced_amd.amr import AMRMHD2D, blast_hierarchy, blast_state

    h = blast_hierarchy(512, 3)
    b0 = (2**-0.5, 2**-0.5)
    models = {p: AMRMHD2D(h, limiter="minmod", mean_field=b0, precision=p) for p in ("fp64", "mixed")}
    before = {}
    for p, m in models.items():
        m.set_state(**blast_state())
        before[p] = m.leaf_totals()
    for _ in range(3):
        dt = models["fp64"].stable_step()
        for m in models.values():
            m.step(dt)
    for p, m in models.items():
        after = m.leaf_totals()
        assert np.abs(after - before[p]).max() < 1e-13 * np.abs(before[p]).max()
        assert m.divergence() < 1e-13
        assert m.composite_divergence() < 1e-14
    a = models["fp64"].dataset(0)["density"]
    b = models["mixed"].dataset(0)["density"]
    assert float(np.mean(np.abs(a - b))) < 1e-5


def test_c4_harris_sheet_2048_by_1024(gpu):
    """C4: walls in y; mass, energy, z-momentum and B_z kept, div B at round-off."""
    from wellbalanced_amd import MHD2D

    problem = initial.harris_sheet(2048, 1024)
    for precision in ("fp64", "mixed"):
        model = MHD2D.from_problem(problem, kernels="fused", precision=precision)
        before = model.totals()
        model.advance(10)
        model.check()
        after = model.totals()
        for name in ("density", "energy", "momentum_z", "bz"):
            assert abs(after[name] - before[name]) < 1e-13 * max(1.0, abs(before[name]))
        assert model.divergence() < 1e-13


def test_c5_three_level_blocks_on_a_128_cubed_base(gpu):
    """C5's shape: a 128^3 blast in turbulence on 3 levels of 16^3 blocks (started from a
    uniform run by prolongation), FP64 and mixed for 3 steps."""
    from wellbalanced_amd import MHD3D
    from wellbalanced_amd.amr3d import AMRMHD3D
    from wellbalanced_amd.amr3d_maps import Hierarchy3D

    problem = initial.blast_turbulence_3d(128, limiter="minmod", seed=0)
    uni = MHD3D.from_problem(problem, kernels="fused")
    uni.advance(2)
    n = 128
    x = (np.arange(n) + 0.5) / n
    r = np.sqrt((x[:, None, None] - 0.5) ** 2 + (x[None, :, None] - 0.5) ** 2
                + (x[None, None, :] - 0.5) ** 2)
    h = Hierarchy3D(n)
    h.refine_flagged(0, r < 0.25)
    cells = np.argwhere(r < 0.15)
    h.refine(1, sorted({tuple(int(v) for v in c // 8) for c in cells} & h.levels[1]))
    assert len(h.levels) == 3
    models = {p: AMRMHD3D.from_uniform(uni, h, precision=p) for p in ("fp64", "mixed")}
    before = {p: m.leaf_totals() for p, m in models.items()}
    for _ in range(3):
        dt = models["fp64"].stable_step()
        for m in models.values():
            m.step(dt)
    for p, m in models.items():
        m.check()
        after = m.leaf_totals()
        assert np.abs(after - before[p]).max() < 1e-13 * np.abs(before[p]).max()
        assert m.divergence() < 1e-13
        assert m.composite_divergence() < 1e-12
    a, _ = models["fp64"].level_state(0)
    b, _ = models["mixed"].level_state(0)
    assert float(np.mean(np.abs(a[0] - b[0]))) < 1e-5
