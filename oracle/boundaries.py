"""Halo filling of one patch that covers its whole level (oracle only).

The reference fills halos with patchwork's ``exchange_halos`` (periodic images,
coarse-fine interpolation) and then its own domain-face rules
(``mhd.py:298-353,744-777``; 3D ``mhd3d.py:448-473``).  For a single patch the
exchange is the periodic copy, done axis by axis over the full extent of the other
axes so that corner halos end up as images of images.  On a periodic axis the
node at index ``n`` is the image of node 0 and is refreshed from it (the patch
stores ``n+1`` nodes along a node-centred axis).

``centred`` gives, per axis, whether the array is cell-centred (True) or
node-centred (False) along it.  Arrays are haloed with ``HALO`` on every side.
"""

from __future__ import annotations

import numpy as np

HALO = 2


def periodic(array: np.ndarray, axis: int, n: int, centred: bool) -> None:
    """Fill the halos along ``axis`` with the periodic images (``n`` interior cells)."""
    h = HALO
    m = np.moveaxis(array, axis, 0)
    m[:h] = m[n : n + h]
    if centred:
        m[n + h :] = m[h : 2 * h]
    else:
        m[n + h :] = m[h : 2 * h + 1]


def outflow_cells(array: np.ndarray, axis: int, lower: bool, upper: bool) -> None:
    """Zero-gradient halo cells (``mhd.py:298-305``)."""
    h = HALO
    m = np.moveaxis(array, axis, 0)
    if lower:
        m[:h] = m[h]
    if upper:
        m[-h:] = m[-h - 1]


def outflow_nodes(array: np.ndarray, axis: int, lower: bool, upper: bool) -> None:
    """Linearly extrapolated halo nodes (``mhd.py:308-325``)."""
    h = HALO
    m = np.moveaxis(array, axis, 0)
    if lower:
        edge, inner = m[h].copy(), m[h + 1].copy()
        for k in range(1, h + 1):
            m[h - k] = edge + k * (edge - inner)
    if upper:
        edge, inner = m[-h - 1].copy(), m[-h - 2].copy()
        for k in range(1, h + 1):
            m[-h - 1 + k] = edge + k * (edge - inner)


def wall_cells(array: np.ndarray, axis: int, lower: bool, upper: bool, sign: float) -> None:
    """Mirrored halo cells at a wall (``mhd.py:328-339``)."""
    h = HALO
    m = np.moveaxis(array, axis, 0)
    if lower:
        for k in range(h):
            m[h - 1 - k] = sign * m[h + k]
    if upper:
        n = m.shape[0]
        for k in range(h):
            m[n - h + k] = sign * m[n - h - 1 - k]


def wall_nodes(array: np.ndarray, axis: int, lower: bool, upper: bool) -> None:
    """Halo nodes odd about the wall node (``mhd.py:342-353``)."""
    h = HALO
    m = np.moveaxis(array, axis, 0)
    if lower:
        wall = m[h].copy()
        for k in range(1, h + 1):
            m[h - k] = 2.0 * wall - m[h + k]
    if upper:
        n = m.shape[0]
        wall = m[n - h - 1].copy()
        for k in range(1, h + 1):
            m[n - h - 1 + k] = 2.0 * wall - m[n - h - 1 - k]


def fill(cells, potentials, shape, periodic_axes, walls, normal_rows):
    """Fill all halos of a single-patch level.

    ``cells``: list of cell-centred haloed arrays; ``potentials``: list of
    ``(array, centred_per_axis)``; ``walls[axis]``: wall (True) or outflow;
    ``normal_rows[axis]``: index into ``cells`` of the momentum normal to ``axis``.
    Periodic axes first, then each other axis in order (``mhd.py:744-777``).
    """
    ndim = len(shape)
    for axis in range(ndim):
        if periodic_axes[axis]:
            for a in cells:
                periodic(a, axis, shape[axis], True)
            for a, centred in potentials:
                periodic(a, axis, shape[axis], centred[axis])
    for axis in range(ndim):
        if periodic_axes[axis]:
            continue
        if walls[axis]:
            for k, a in enumerate(cells):
                wall_cells(a, axis, True, True, -1.0 if k == normal_rows[axis] else 1.0)
            for a, _ in potentials:
                wall_nodes(a, axis, True, True)
        else:
            for a in cells:
                outflow_cells(a, axis, True, True)
            for a, _ in potentials:
                outflow_nodes(a, axis, True, True)
