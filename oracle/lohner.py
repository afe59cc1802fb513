"""The second-derivative refinement estimator of Löhner (1987), oracle only.

The reference's only estimator is ``current_tags`` (|J| times the cell size,
``mhd.py:1011-1041``, ``mhd3d.py:771-801``); the gradient/Löhner estimator that
labels the learned indicator is this package's, in the normalised form of FLASH
(Fryxell et al. 2000, ApJS 131, 273, sec. 6):

for a variable u at cell i, with the central differences ``d_a(i) = u(i+e_a) -
u(i-e_a)`` and ``s_a(i) = |u(i+e_a)| + |u(i-e_a)|`` at the neighbours,

    num = sum_{a,b} ( d_a(i+e_b) - d_a(i-e_b) )^2
    den = sum_{a,b} ( |d_a(i+e_b)| + |d_a(i-e_b)| + eps * (s_a(i+e_b) + s_a(i-e_b)) )^2
    E   = sqrt(num / den)            (0 where den = 0)

with eps = 0.01.  The estimate of a cell is the largest over density and
pressure; a cell is flagged where E > 0.8.  Needs a halo of 2.
"""

from __future__ import annotations

import numpy as np

EPS = 0.01
CUTOFF = 0.8


def estimate(u: np.ndarray, eps: float = EPS) -> np.ndarray:
    """E of every interior cell of a haloed (halo 2) 2D or 3D array."""
    nd = u.ndim
    h = 2
    inner = tuple(slice(h, s - h) for s in u.shape)

    def shift(arr, axis, k):
        idx = [slice(None)] * nd
        n = arr.shape[axis]
        idx[axis] = slice(1 + k, n - 1 + k)
        return arr[tuple(idx)]

    # first differences on the cells 1 .. n-2 of every axis (one ring inside)
    ring = tuple(slice(1, s - 1) for s in u.shape)
    d, s = [], []
    for a in range(nd):
        lo = shift(u, a, -1)
        hi = shift(u, a, 1)
        # trim the other axes to the ring
        idx = tuple(slice(None) if b == a else slice(1, u.shape[b] - 1) for b in range(nd))
        d.append((hi - lo)[idx])
        s.append((np.abs(hi) + np.abs(lo))[idx])
    num = np.zeros(u[inner].shape)
    den = np.zeros(u[inner].shape)
    for a in range(nd):
        for b in range(nd):
            da_hi = _take(d[a], b, 2, nd)
            da_lo = _take(d[a], b, 0, nd)
            sa_hi = _take(s[a], b, 2, nd)
            sa_lo = _take(s[a], b, 0, nd)
            num = num + (da_hi - da_lo) * (da_hi - da_lo)
            t = np.abs(da_hi) + np.abs(da_lo) + eps * (sa_hi + sa_lo)
            den = den + t * t
    del ring
    with np.errstate(all="ignore"):
        e = np.where(den > 0.0, np.sqrt(num / den), 0.0)
    return e


def _take(arr, axis, start, nd):
    """The interior window of a ring-sized array shifted by start-1 along ``axis``."""
    idx = []
    for b in range(nd):
        n = arr.shape[b]
        if b == axis:
            idx.append(slice(start, n - 2 + start))
        else:
            idx.append(slice(1, n - 1))
    return arr[tuple(idx)]


def flags(rho: np.ndarray, p: np.ndarray, cutoff: float = CUTOFF) -> tuple[np.ndarray, np.ndarray]:
    """The estimate max(E(rho), E(p)) and the flags E > cutoff."""
    e = np.maximum(estimate(rho), estimate(p))
    return e, e > cutoff


def block_flags(cell_flags: np.ndarray, block: int) -> np.ndarray:
    """A block is flagged if any of its cells is (blocks of ``block`` cells per side)."""
    shape = tuple(s // block for s in cell_flags.shape)
    f = cell_flags[tuple(slice(0, s * block) for s in shape)]
    nd = f.ndim
    f = f.reshape(sum(((s, block) for s in shape), ()))
    return f.any(axis=tuple(range(1, 2 * nd, 2)))
