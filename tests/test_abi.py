"""The C ABI: the library loads and exports every entry point the header declares
(no compute calls, so this runs without a GPU)."""

import ctypes
import re
from pathlib import Path

import pytest

from wellbalanced_amd import _lib

HEADER = Path(__file__).resolve().parent.parent / "include" / "wbmhd.h"


def _declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t)\s+(wb_\w+)\s*\(", text, re.M)))


def test_header_declares_the_entry_points():
    names = _declared()
    assert "wb_sweep2d" in names and "wb_primitive3d" in names
    assert set(names) == set(_lib.SIGNATURES), "the ctypes table and the header disagree"


def test_library_exports_every_declared_symbol():
    if not _lib.LIBRARY.exists():
        pytest.skip("libwbmhd.so is not built")
    lib = ctypes.CDLL(str(_lib.LIBRARY))
    for name in _declared():
        assert hasattr(lib, name), name


def test_invalid_arguments_are_rejected_without_a_launch():
    if not _lib.LIBRARY.exists():
        pytest.skip("libwbmhd.so is not built")
    lib = _lib.load()
    assert lib.wb_mhd3d_state_size(1, 4, 4) == _lib.WB_EINVAL
    assert lib.wb_mhd3d_state_size(4, 4, 4) == 5 * 8**3 + 3 * 8 * 9 * 9


def test_the_product_fails_loudly_without_a_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from wellbalanced_amd import MHD2D
    from wellbalanced_amd.hierarchy import AMRHierarchy

    with pytest.raises(RuntimeError, match="GPU"):
        MHD2D(AMRHierarchy((8, 8), periodic=(True, True)))


def test_validation_messages_follow_the_reference():
    from wellbalanced_amd import MHD2D
    from wellbalanced_amd.hierarchy import AMRHierarchy

    with pytest.raises(ValueError, match="2D"):
        MHD2D(AMRHierarchy((8,)))
    with pytest.raises(NotImplementedError, match="Cartesian"):
        MHD2D(AMRHierarchy((4, 4), ((0.0, 1.0), (0.0, 1.0)), geometry="cylindrical"))
    with pytest.raises(ValueError, match="limiter"):
        MHD2D(AMRHierarchy((4, 4)), limiter="superbee")
    with pytest.raises(ValueError, match="cfl"):
        MHD2D(AMRHierarchy((4, 4)), cfl=1.5)
    with pytest.raises(ValueError, match="gamma"):
        MHD2D(AMRHierarchy((4, 4)), gamma=1.0)
    periodic = AMRHierarchy((4, 4), periodic=(True, False))
    with pytest.raises(ValueError, match="periodic"):
        MHD2D(periodic, boundaries={"x": "wall"})
    with pytest.raises(ValueError, match="boundaries"):
        MHD2D(AMRHierarchy((4, 4)), boundaries={"z": "wall"})
    with pytest.raises(ValueError, match="mean field"):
        MHD2D(AMRHierarchy((4, 4)), boundaries={"y": "wall"}, mean_field=(0.0, 1.0))


def test_new_entry_points_reject_bad_arguments_before_any_launch():
    """Every check happens on the host, before a kernel is enqueued (NULL pointers are
    never touched), so these calls run without a GPU."""
    if not _lib.LIBRARY.exists():
        pytest.skip("libwbmhd.so is not built")
    lib = _lib.load()
    E = _lib.WB_EINVAL
    assert lib.wb_resistive2d(None, None, None, 2, 9, 0.1, 0.1, 0.1, None, None, None, None) == E
    assert lib.wb_resistive2d(None, None, None, 9, 9, 0.1, 0.1, -1.0, None, None, None, None) == E
    assert lib.wb_conductive2d(None, None, None, 9, 9, 0.1, 0.1, -1.0, 0.0, None, None, None) == E
    assert lib.wb_stable_dt_diffusive(None, 0.4, 0.0, -1.0, 0.0, 1.0, None, None, None) == E
    assert lib.wb_mhd2d_stage(None, None, 8, 8, 0.1, 0.1, 0.0, 0.0, 5 / 3, -1.0, 0.0, 0.0, 0,
                              None, 0, 0, None, 8, 126, None) == E
    assert lib.wb_amr2d_reflux(None, None, None, None, None, 1, 0.0, 0.0, 5 / 3, 0.0, -1.0,
                               0.0, 0, None, 0, None, None) == E
    # a register is needed when there is a coarse cell to reflux
    assert lib.wb_amr2d_reflux(None, None, None, None, None, 1, 0.0, 0.0, 5 / 3, 0.0, 0.0,
                               0.0, 0, None, 1, None, None) == E
    assert lib.wb_amr2d_reflux_mixed(None, None, None, None, None, 1, 0.0, 0.0, 5 / 3, 0.0,
                                     0.0, 0.0, 0, None, 0, None, None) == E
    assert lib.wb_mhd3d_resistive_workspace(0, 4, 4) == E
    assert lib.wb_mhd3d_resistive(None, None, 0, 4, 4, 0.1, 0.1, 0.1, 0.0, 0.0, 0.0, 0.1, None,
                                  0, None, None) == E
    assert lib.wb_amr2d_regrid_copy(None, None, None, None, None, None, -1, None, None, 0,
                                    None) == E
    assert lib.wb_amr2d_divergence(None, None, None, 0, 0.0, 0.0, None, 0, None, 0, None,
                                   None) == E
    assert lib.wb_amr2d_leaf_values(None, None, None, -1, 0.0, 0.0, 5 / 3, None, None) == E
    assert lib.wb_amr2d_rate(None, None, None, 0, 0.0, 0.0, 5 / 3, 0.0, 0.0, None, None,
                             None) == E
    # the hierarchies' fused glue: negative counts are refused, empty work enqueues nothing
    fill = lambda nc, nn, ni, nj: lib.wb_amr2d_fill_level(
        None, None, None, None, None, nc, None, None, nn, None, None, None, None, None, None,
        None, ni, None, None, None, None, None, None, nj, None)
    assert fill(-1, 0, 0, 0) == E and fill(0, 0, 0, -1) == E and fill(0, 0, 0, 0) == 0
    assert lib.wb_amr2d_coarsen(None, None, -1, None, 0, None) == E
    assert lib.wb_amr2d_coarsen(None, None, 0, None, 0, None) == 0
    assert lib.wb_amr3d_fill_regions(None, None, None, None, -1, None) == E
    assert lib.wb_amr3d_fill_regions(None, None, None, None, 0, None) == 0
    for name in ("wb_amr2d_reflux_both", "wb_amr2d_reflux_both_mixed"):
        both2 = getattr(lib, name)
        rec2d = ctypes.c_double(0.0)
        assert both2(None, None, None, None, -1, None, None, None, None) == E
        assert both2(None, None, None, None, 0, None, None, None, None) == 0
        # rows, groups and both stages' records are needed when there is a group
        assert both2(None, None, None, None, 1, ctypes.byref(rec2d), None, None, None) == E
    both = lib.wb_amr3d_reflux_both
    rec0 = ctypes.c_double(0.0)
    assert both(None, None, None, None, -1, None, None, None, None, None) == E
    assert both(None, None, None, None, 0, None, None, None, None, None) == 0
    # the slots and both stages' face records are needed when there is a cell to reflux
    assert both(None, None, None, None, 1, None, None, ctypes.byref(rec0), None, None) == E
    # the stage takes the face record's slots and the record together or neither
    rec = ctypes.c_double(0.0)
    assert lib.wb_amr3d_stage(None, None, None, None, 1, 0, None, None, 0.0, 0.0, 0.0, 5 / 3,
                              0, None, 0, None, None, ctypes.byref(rec), None) == E
    for name in ("wb_amr3d_rate", "wb_amr3d_rate_mixed"):
        assert getattr(lib, name)(None, None, 0, 0.0, 0.0, 0.0, 5 / 3, None, None) == E


def test_negative_diffusivities_are_refused():
    """test_mhd_amr.py:144-147 (and the conductivities, mhd.py:437-440)."""
    from wellbalanced_amd import MHD2D, MHD3D
    from wellbalanced_amd.amr import AMRMHD2D
    from wellbalanced_amd.hierarchy import AMRHierarchy

    with pytest.raises(ValueError, match="resistivity"):
        MHD2D(AMRHierarchy((8, 8), periodic=(True, True)), resistivity=-1.0)
    with pytest.raises(ValueError, match="conductivities"):
        MHD2D(AMRHierarchy((8, 8), periodic=(True, True)), conduction=(1.0, -1.0))
    with pytest.raises(ValueError, match="resistivity"):
        AMRMHD2D(AMRHierarchy((8, 8), periodic=(True, True)), resistivity=-1.0)
    with pytest.raises(ValueError, match="resistivity"):
        MHD3D(AMRHierarchy((4, 4, 4), periodic=(True, True, True)), resistivity=-1.0)


def test_a_step_graph_checks_the_model_generation_before_launching():
    """Host logic of StepGraph.replay (no GPU): a model whose buffers or tables were
    replaced since the capture is refused before the library is called."""
    from types import SimpleNamespace

    from wellbalanced_amd import autotune
    from wellbalanced_amd.graph import StepGraph

    model = SimpleNamespace(generation=3, device=None, steps=0)
    graph = object.__new__(StepGraph)
    graph.model, graph.steps, graph._graph, graph._generation = model, 1, None, 2
    with pytest.raises(RuntimeError, match="capture a new StepGraph"):
        graph.replay()
    # the autotuner's choice bumps the generation of a model without set_tiling
    autotune._apply(model, {"nxs": 8})
    assert model.generation == 4 and model.nxs == 8
