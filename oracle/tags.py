"""The reference's refinement estimator ``current_tags`` in NumPy (oracle only).

2D: ``mhd.py:1011-1041`` with ``_current`` (``mhd.py:826-831``): |J_z| at the nodes
from the face fields, averaged over a cell's four nodes, times max(dx, dy).
3D: ``mhd3d.py:771-801``: |J| of the cell field's curl by ``np.gradient``, times
max(dx, dy, dz).  A cell is tagged where the product exceeds the threshold.
"""

from __future__ import annotations

import numpy as np

from .kernels import primitive

HALO = 2


def current_tags2d(oracle, threshold: float) -> np.ndarray:
    """Tags of the interior cells of an `Oracle2D` (its halos are filled first)."""
    oracle.fill_halos()
    _, bxf, byf, _, _ = primitive(*oracle.cons, oracle.az, oracle.dx, oracle.dy,
                                  *oracle.b0, oracle.gamma)
    dx, dy = oracle.dx, oracle.dy
    jz = np.abs((byf[1:, 1:-1] - byf[:-1, 1:-1]) / dx - (bxf[1:-1, 1:] - bxf[1:-1, :-1]) / dy)
    h, nx, ny = HALO, oracle.nx, oracle.ny
    cell = 0.25 * sum(
        jz[h - 1 + a : h - 1 + a + nx, h - 1 + b : h - 1 + b + ny] for a in (0, 1) for b in (0, 1)
    )
    return cell * max(dx, dy) > threshold


def current_tags3d(oracle, threshold: float) -> np.ndarray:
    """Tags of the interior cells of an `Oracle3D` (its halos are filled first)."""
    oracle.fill_halos()
    prim, _ = oracle.primitive()
    cells = prim[5:8]
    spacing = oracle.spacing
    square = 0.0
    for c in range(3):
        a, b = (c + 1) % 3, (c + 2) % 3
        jc = np.gradient(cells[b], spacing[a], axis=a) - np.gradient(cells[a], spacing[b], axis=b)
        square = square + jc**2
    size = np.sqrt(square)[(slice(HALO, -HALO),) * 3]
    return size * max(spacing) > threshold
