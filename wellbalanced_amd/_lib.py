"""ctypes binding of ``libwbmhd.so`` (the C ABI of ``include/wbmhd.h``).

The library is loaded from the package folder.  There is no CPU fallback: if it
is missing or a GPU is not available, the product raises.
"""

from __future__ import annotations

import ctypes
from ctypes import c_double, c_float, c_int, c_int64, c_void_p
from pathlib import Path

LIBRARY = Path(__file__).resolve().parent / "libwbmhd.so"
WB_EINVAL = -100000
WB_ESCRATCH = -100001

P = c_void_p
I64 = c_int64
D = c_double
I = c_int
F = c_float

#: Argument types of every exported entry point (``include/wbmhd.h``).
SIGNATURES: dict[str, tuple[type, list]] = {
    "wb_primitive2d": (I, [P] * 7 + [I64, I64] + [D] * 5 + [P] * 5 + [P]),
    "wb_sweep2d": (I, [P, P, I64, I64, I, I, D, P, P]),
    "wb_ct_contact": (I, [P] * 5 + [I64, I64, P, P]),
    "wb_rate2d": (I, [P] * 7 + [I64, I64, I] + [D] * 5 + [P, P]),
    "wb_mhd2d_fill_halos": (I, [P, I64, I64, I, I, P]),
    "wb_mhd2d_walls": (I, [P, P, P, I64, I64, I, P]),
    "wb_mhd2d_update": (I, [P, P, P, P, P, I64, I64, P, D, D, I, P]),
    "wb_resistive2d": (I, [P, P, P, I64, I64, D, D, D, P, P, P, P]),
    "wb_conductive2d": (I, [P, P, P, I64, I64, D, D, D, D, P, P, P]),
    "wb_stable_dt": (I, [P, D, D, P, P, P]),
    "wb_stable_dt_diffusive": (I, [P, D, D, D, D, D, P, P, P]),
    "wb_mhd2d_stage": (I, [P, P, I64, I64] + [D] * 8 + [I, P, I, I, P, I, I, P]),
    "wb_mhd2d_stage_mixed": (I, [P, P, I64, I64] + [D] * 8 + [I, P, I, I, P, I, I, P]),
    "wb_mhd3d_state_size": (I64, [I64, I64, I64]),
    "wb_mhd3d_resistive_workspace": (I64, [I64, I64, I64]),
    "wb_mhd3d_resistive": (I, [P, P, I64, I64, I64] + [D] * 7 + [P, I, P, P]),
    "wb_mhd3d_resistive_noalloc_limit": (I64, []),
    "wb_mhd3d_resistive_noalloc": (I, [P, P, I64, I64, I64] + [D] * 7 + [P, I, P]),
    "wb_mhd3d_fill_halos": (I, [P, I64, I64, I64, I, P]),
    "wb_mhd3d_fill_outflow": (I, [P, I64, I64, I64, I, P]),
    "wb_primitive3d": (I, [P, I64, I64, I64] + [D] * 7 + [P] * 5 + [P]),
    "wb_sweep3d": (I, [P, P, I64, I64, I64, I, I, D, P, P]),
    "wb_edge3d": (I, [P, P, P, I64, I64, I64, I, P, P]),
    "wb_mhd3d_update": (I, [P, P] + [P] * 6 + [I64] * 3 + [P] + [D] * 3 + [I, P]),
    "wb_mhd3d_stage": (I, [P, P, I64, I64, I64] + [D] * 7 + [I, P, I, P, I, I, P]),
    "wb_mhd3d_stage_mixed": (I, [P, P, I64, I64, I64] + [D] * 7 + [I, P, I, P, I, I, I, P]),
    "wb_rate3d": (I, [P, I64, I64, I64] + [D] * 7 + [P, P]),
    "wb_rate3d_mixed": (I, [P, I64, I64, I64] + [D] * 7 + [P, P]),
    "wb_lohner3d": (I, [P, I64, I64, I64] + [D] * 8 + [P, P, P]),
    "wb_graph_capture_begin": (I, [P]),
    "wb_graph_capture_end": (I, [P, P]),
    "wb_graph_launch": (I, [P, I, P]),
    "wb_graph_destroy": (I, [P]),
    "wb_current_tags2d": (I, [P, P, I64, I64, D, D, D, P, P]),
    "wb_current_tags3d": (I, [P, I64, I64, I64, D, D, D, D, P, P]),
    "wb_amr3d_fill": (I, [P, P, I, P, P, P]),
    "wb_amr3d_fill_list": (I, [P, P, P, I, I, P, P, P]),
    "wb_amr3d_fill_regions": (I, [P, P, P, P, I, P]),
    "wb_amr3d_coarsen": (I, [P, P, P, I, I, P, P, P]),
    "wb_patches3d_gather": (I, [P, P, P, P, P, P, P, I64, P]),
    "wb_patches3d_prolong": (I, [P, P, I64, P]),
    "wb_patches3d_reflux": (I, [P, P, P, P, P, I64, P, I, P, P]),
    "wb_resistive3d": (I, [P, P, P, I64, I64, I64, D, D, D, D, P, P, P, P, P, P, P]),
    "wb_amr3d_prolong": (I, [P, P, P, I, P, P, P]),
    "wb_amr3d_rate": (I, [P, P, I, D, D, D, D, P, P]),
    "wb_amr3d_rate_mixed": (I, [P, P, I, D, D, D, D, P, P]),
    "wb_amr3d_stage": (I, [P, P, P, P, I, I, P, P, D, D, D, D, I, P, I, P, P, P, P]),
    "wb_amr3d_stage_mixed": (I, [P, P, P, P, I, I, P, P, D, D, D, D, I, P, I, P, I, P, P, P]),
    "wb_amr3d_reflux_both": (I, [P, P, P, P, I, P, P, P, P, P]),
    "wb_cnn_prep": (I, [P, P, P]),
    "wb_cnn_moments": (I, [P, I64, I64, I64] + [D] * 7 + [P, I, P, P]),
    "wb_cnn_input": (I, [P, I64, I64, I64] + [D] * 7 + [P, P, P]),
    "wb_cnn_forward": (I, [P, P, P, I64, I64, I64, P, P, P, I, P]),
    "wb_cnn_backward": (I, [P] * 6 + [I64, I64, I64] + [P, P, P, P, I, P, P, P]),
    "wb_cnn_adam": (I, [P, P, P, P, I, F, F, F, F, I, P]),
    "wb_cnn_conv": (I, [P, I, P, P, I64, I64, I64, I, P, P, P, I, P]),
    "wb_cnn_wgrad": (I, [P, I, P, I64, I64, I64, P, I, P, P]),
    "wb_amr2d_fill_level": (I, [P] * 5 + [I, P, P, I] + [P] * 7 + [I] + [P] * 6 + [I, P]),
    "wb_amr2d_coarsen": (I, [P, P, I, P, I, P]),
    "wb_amr2d_domain_faces": (I, [P, P, P, I, P]),
    "wb_amr2d_regrid_copy": (I, [P] * 6 + [I, P, P, I, P]),
    "wb_amr2d_prolong": (I, [P] * 8 + [I] + [P] * 5 + [I, P]),
    "wb_amr2d_leaf_values": (I, [P, P, P, I, D, D, D, P, P]),
    "wb_amr2d_totals_workspace": (I, []),
    "wb_amr2d_totals": (I, [P, P, P, I, P, P, P]),
    "wb_amr2d_divergence": (I, [P, P, P, I, D, D, P, I, P, I, P, P]),
    "wb_amr2d_rate": (I, [P, P, P, I, D, D, D, D, D, P, P, P]),
    "wb_amr2d_stage": (I, [P, P, P, P, I, P] + [D] * 6 + [I, P, I, P, I, I, I, P]),
    "wb_amr2d_stage_mixed": (I, [P, P, P, P, I, P] + [D] * 6 + [I, P, I, P, I, I, I, P]),
    "wb_mhd2d_stage_tiles": (I, [P, P, I64, I64] + [D] * 8 + [I, P, I, I, P, P]),
    "wb_mhd2d_stage_tiles_mixed": (I, [P, P, I64, I64] + [D] * 8 + [I, P, I, I, P, P]),
    "wb_amr2d_stage_tiles": (I, [P, P, P, P, I] + [D] * 6 + [I, P, I, P, P, P, P, P]),
    "wb_amr2d_stage_tiles_mixed": (I, [P, P, P, P, I] + [D] * 6 + [I, P, I, P, P, P, P, P]),
    "wb_amr2d_reflux_both": (I, [P, P, P, P, I, P, P, P, P]),
    "wb_amr2d_reflux_both_mixed": (I, [P, P, P, P, I, P, P, P, P]),
    "wb_amr2d_reflux": (I, [P, P, P, P, P, I] + [D] * 6 + [I, P, I, P, P]),
    "wb_amr2d_reflux_mixed": (I, [P, P, P, P, P, I] + [D] * 6 + [I, P, I, P, P]),
}

_lib = None


def load() -> ctypes.CDLL:
    """Load ``libwbmhd.so`` once and set the argument types; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIBRARY.exists():
        raise RuntimeError(
            f"{LIBRARY} is not built: run `python -m wellbalanced_amd.build` "
            "(the HIP path has no CPU fallback)"
        )
    lib = ctypes.CDLL(str(LIBRARY))
    for name, (restype, argtypes) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    _lib = lib
    return lib


def call(name: str, *args) -> int:
    """Call entry point ``name`` and raise ``RuntimeError`` on a negative return."""
    rc = getattr(load(), name)(*args)
    if rc < 0 and SIGNATURES[name][0] is I:
        what = ("invalid arguments" if rc == WB_EINVAL else
                "the input needs more than the static scratch buffer" if rc == WB_ESCRATCH
                else f"HIP error {-rc}")
        raise RuntimeError(f"{name} failed: {what}")
    return rc


def ptr(tensor) -> int:
    """Device address of a contiguous torch tensor."""
    if not tensor.is_contiguous():
        raise ValueError("wbmhd needs contiguous tensors")
    return tensor.data_ptr()


def stream_handle(device=None) -> int:
    """The raw HIP stream of torch's current stream on ``device``."""
    import torch

    return torch.cuda.current_stream(device).cuda_stream
