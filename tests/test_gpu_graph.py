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
