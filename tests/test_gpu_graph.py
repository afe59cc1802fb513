"""HIP-graph captured steps equal eager steps, replay after replay."""

import numpy as np
import pytest

from wellbalanced_amd import MHD2D, MHD3D, initial
from wellbalanced_amd.graph import StepGraph

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kernels,precision", [("fused", "fp64"), ("fused", "mixed"), ("split", "fp64")])
def test_captured_2d_steps_equal_eager_steps(gpu, kernels, precision):
    import torch

    problem = initial.harris_sheet(96, 48, length=3.2, noise=1e-2, seed=2)
    eager = MHD2D.from_problem(problem, kernels=kernels, precision=precision)
    graphed = MHD2D.from_problem(problem, kernels=kernels, precision=precision)
    graph = StepGraph(graphed, steps=2)
    for replay in range(4):
        graph.replay()
        eager.advance(2)
        # eager work of the program between replays
        _ = eager.totals()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(graphed.arrays()[0], eager.arrays()[0])
        np.testing.assert_array_equal(graphed.arrays()[1], eager.arrays()[1])
        assert graphed.time == eager.time
    assert graphed.steps == eager.steps == 8


def test_captured_3d_steps_equal_eager_steps(gpu):
    import torch

    problem = initial.turbulence_3d(24, seed=7)
    eager = MHD3D.from_problem(problem, kernels="fused", precision="mixed")
    graphed = MHD3D.from_problem(problem, kernels="fused", precision="mixed")
    graph = StepGraph(graphed, steps=3)
    for replay in range(3):
        graph.replay()
        eager.advance(3)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(graphed.arrays()[0], eager.arrays()[0])
        assert graphed.time == eager.time


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_captured_refined_2d_steps_equal_eager_steps(gpu, precision):
    """A 3-level hierarchy of uneven patches: the captured steps (fills, the persistent
    stage with its work counter cleared by a kernel, refluxing, coarsening, the CFL
    reduction and the device dt) replay bit for bit as eager steps."""
    import torch

    from wellbalanced_amd import amr
    from wellbalanced_amd.amr import AMRMHD2D

    h = amr.blast_hierarchy(base=64, levels=3, sizes=(8, 24), seed=3)
    eager = AMRMHD2D(h, mean_field=(0.5**0.5, 0.5**0.5), precision=precision)
    graphed = AMRMHD2D(h, mean_field=(0.5**0.5, 0.5**0.5), precision=precision)
    for m in (eager, graphed):
        m.set_state(**amr.blast_state())
        m.check_interval = 0
    graph = StepGraph(graphed, steps=2)
    for replay in range(3):
        graph.replay()
        eager.advance(2)
        _ = eager.totals()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(graphed.buffer(), eager.buffer())
        assert graphed.time == eager.time
    graphed.check()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_captured_refined_3d_steps_equal_eager_steps(gpu, precision):
    import torch

    from wellbalanced_amd.amr3d import AMRMHD3D
    from wellbalanced_amd.amr3d_maps import Hierarchy3D

    h = Hierarchy3D(32)
    h.refine(0, [(0, 0, 0)])
    h.refine(1, [(1, 1, 1)])
    state = dict(density=1.0, velocity=(0.1, -0.05, 0.02),
                 pressure=lambda x, y, z: 1.0 + 4.0 * np.exp(
                     -((x - 0.3) ** 2 + (y - 0.35) ** 2 + (z - 0.3) ** 2) / 0.01),
                 vector_potential=(lambda x, y, z: 0.0 * x,
                                   lambda x, y, z: 0.1 * np.sin(2 * np.pi * z),
                                   lambda x, y, z: 0.1 * np.sin(2 * np.pi * x)))
    eager = AMRMHD3D(h, mean_field=(0.3, 0.2, 0.1), limiter="minmod", precision=precision)
    graphed = AMRMHD3D(h, mean_field=(0.3, 0.2, 0.1), limiter="minmod", precision=precision)
    for m in (eager, graphed):
        m.set_state(**state)
        m.check_interval = 0
    graph = StepGraph(graphed, steps=2)
    for replay in range(3):
        graph.replay()
        eager.advance(2)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(graphed.buffer(), eager.buffer())
        assert graphed.time == eager.time
    graphed.check()


def test_a_graph_refuses_to_replay_after_a_regrid_or_a_new_tiling(gpu):
    """The graph holds the addresses of the buffers and tables at capture: after a
    regrid (new buffers and tables) or a new tiling (new work list) it raises before
    any launch, and the model's state is untouched."""
    from tests.test_regrid import _amr_orszag_tang

    model = _amr_orszag_tang(16, limiter="minmod")
    tags = {0: np.pad(np.ones((4, 7), bool), ((6, 6), (5, 4)))}
    model.regrid(tags, max_patch_size=16)
    graph = StepGraph(model, steps=1)
    graph.replay()
    model.set_tiling(tile=30)
    before = model.buffer()
    with pytest.raises(RuntimeError, match="capture a new StepGraph"):
        graph.replay()
    np.testing.assert_array_equal(model.buffer(), before)
    graph = StepGraph(model, steps=1)
    graph.replay()
    model.regrid(tags, max_patch_size=8)
    with pytest.raises(RuntimeError, match="capture a new StepGraph"):
        graph.replay()


@pytest.mark.parametrize("dims", [2, 3])
def test_lockstep_graph_equals_eager_steps_at_the_first_models_dt(gpu, dims):
    """FP64 and mixed models captured in lockstep: each replay equals eager steps in
    which the mixed model takes the FP64 model's stable step, and the follower's time
    is the leader's."""
    import torch

    from wellbalanced_amd.graph import LockstepGraph

    if dims == 2:
        problem = initial.harris_sheet(96, 48, length=3.2, noise=1e-2, seed=2)
        make = lambda p: MHD2D.from_problem(problem, kernels="fused", precision=p)
    else:
        problem = initial.turbulence_3d(24, seed=7)
        make = lambda p: MHD3D.from_problem(problem, kernels="fused", precision=p)
    eager = [make("fp64"), make("mixed")]
    graphed = [make("fp64"), make("mixed")]
    graph = LockstepGraph(graphed, steps=2)
    for replay in range(3):
        graph.replay()
        for _ in range(2):
            dt = eager[0].stable_step()
            for m in eager:
                m.step(dt)
        _ = eager[1].totals()
        torch.cuda.synchronize()
        for g, e in zip(graphed, eager):
            np.testing.assert_array_equal(g.arrays()[0], e.arrays()[0])
            for a, b in zip(g.arrays()[1], e.arrays()[1]):
                np.testing.assert_array_equal(a, b)
        assert graphed[1].time == graphed[0].time == eager[0].time
    assert [m.steps for m in graphed] == [6, 6]
    # the follower keeps its own dt: an eager step of it after the replays is its own
    assert graphed[1].stable_step() == eager[1].stable_step()
