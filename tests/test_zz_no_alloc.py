"""The allocation-free entry point ``wb_mhd3d_resistive_noalloc`` (no counterpart in
the reference: for callers that cannot allocate) against ``wb_mhd3d_resistive`` with
a caller-owned workspace, through the C ABI: the same bits on two calls in a row, the
second on a smaller grid (the static workspace keeps the first call's contents)."""

import numpy as np
import pytest
import torch

from wellbalanced_amd import _lib


def _pair(nx, ny, nz, seed, heun):
    lib = _lib.load()
    size = lib.wb_mhd3d_state_size(nx, ny, nz)
    rng = np.random.default_rng(seed)
    dev = torch.device("cuda")
    src = torch.as_tensor(rng.normal(size=size), device=dev)
    start = torch.as_tensor(rng.normal(size=size), device=dev)
    dt = torch.tensor([1e-3], dtype=torch.float64, device=dev)
    work = torch.empty(lib.wb_mhd3d_resistive_workspace(nx, ny, nz), dtype=torch.float64,
                       device=dev)
    outs = []
    stream = _lib.stream_handle(dev)
    args = (nx, ny, nz, 1.0 / nx, 1.0 / ny, 1.0 / nz, 0.3, -0.2, 0.1, 0.02, _lib.ptr(dt), heun)
    want = start.clone()
    _lib.call("wb_mhd3d_resistive", _lib.ptr(src), _lib.ptr(want), *args, _lib.ptr(work), stream)
    got = start.clone()
    _lib.call("wb_mhd3d_resistive_noalloc", _lib.ptr(src), _lib.ptr(got), *args, stream)
    torch.cuda.synchronize()
    outs = want.cpu().numpy(), got.cpu().numpy(), start.cpu().numpy()
    return outs


@pytest.mark.gpu
def test_no_alloc_resistive_terms_equal_the_workspace_entry_point(gpu):
    for (nx, ny, nz), seed, heun in (((32, 24, 20), 1, 0), ((12, 10, 8), 2, 1)):
        want, got, start = _pair(nx, ny, nz, seed, heun)
        assert not np.array_equal(want, start)  # the terms changed the state
        np.testing.assert_array_equal(got, want)


@pytest.mark.gpu
def test_no_alloc_refuses_a_grid_beyond_its_workspace(gpu):
    lib = _lib.load()
    limit = lib.wb_mhd3d_resistive_noalloc_limit()
    assert lib.wb_mhd3d_resistive_workspace(64, 64, 64) <= limit
    assert lib.wb_mhd3d_resistive_workspace(80, 80, 80) > limit
    rc = lib.wb_mhd3d_resistive_noalloc(None, None, 80, 80, 80, 0.1, 0.1, 0.1, 0.0, 0.0, 0.0,
                                        0.02, None, 0, None)
    assert rc == _lib.WB_ESCRATCH
    with pytest.raises(RuntimeError, match="static scratch"):
        _lib.call("wb_mhd3d_resistive_noalloc", None, None, 80, 80, 80, 0.1, 0.1, 0.1, 0.0, 0.0,
                  0.0, 0.02, None, 0, None)
