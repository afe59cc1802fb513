"""A minimal block-structured mesh with the interface of ``patchwork.AMRHierarchy``
that ``MHD2D``/``MHD3D`` use (carafe ``mhd.py:467-477,703-734,748-777``).

The reference allocates its fields on the external ``patchwork`` package (not
available here).  This stand-in holds the geometry only -- levels, patches and
their boxes; the models keep the field data on the device themselves.  Level 0
covers the domain with one patch; finer patches are added with `add_patch`.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class Patch:
    """A box of cells ``((i0, i1), (j0, j1)[, (k0, k1)])`` on one level."""

    level: int
    box: tuple

    @property
    def shape(self) -> tuple[int, ...]:
        return tuple(hi - lo for lo, hi in self.box)


class AMRHierarchy:
    """Levels of uniform Cartesian patches over a box, refinement ratio 2."""

    def __init__(self, shape, bounds=None, periodic=None, *, geometry: str = "cartesian",
                 refinement_ratio: int = 2):
        shape = tuple(int(n) for n in shape)
        self.ndim = len(shape)
        if bounds is None:
            bounds = ((0.0, 1.0),) * self.ndim
        if periodic is None:
            periodic = (False,) * self.ndim
        if len(bounds) != self.ndim or len(periodic) != self.ndim:
            raise ValueError("bounds and periodic need one entry per axis")
        if any(n < 1 for n in shape):
            raise ValueError("every axis needs at least one cell")
        self.base_shape = shape
        self.bounds = tuple((float(lo), float(hi)) for lo, hi in bounds)
        self.periodic = tuple(bool(p) for p in periodic)
        self.cylindrical = geometry == "cylindrical"
        self.graded = False
        self.refinement_ratio = int(refinement_ratio)
        self.patches: list[Patch] = [Patch(0, tuple((0, n) for n in shape))]

    @property
    def lower_bound(self) -> tuple[float, ...]:
        return tuple(lo for lo, _ in self.bounds)

    @property
    def max_level(self) -> int:
        return max(p.level for p in self.patches)

    @property
    def local_patches(self) -> list[Patch]:
        return self.patches

    def level_shape(self, level: int) -> tuple[int, ...]:
        r = self.refinement_ratio**level
        return tuple(n * r for n in self.base_shape)

    def spacing(self, level: int) -> tuple[float, ...]:
        shape = self.level_shape(level)
        return tuple((hi - lo) / n for (lo, hi), n in zip(self.bounds, shape))

    def patches_at_level(self, level: int) -> list[Patch]:
        return [p for p in self.patches if p.level == level]

    def add_patch(self, level: int, box) -> Patch:
        """Add a patch on ``level`` (>= 1) covering ``box`` in that level's cells."""
        box = tuple((int(lo), int(hi)) for lo, hi in box)
        shape = self.level_shape(level)
        if level < 1 or len(box) != self.ndim:
            raise ValueError("patches are added on levels >= 1 with one range per axis")
        if any(lo < 0 or hi > n or hi <= lo for (lo, hi), n in zip(box, shape)):
            raise ValueError(f"box {box} is outside level {level} of shape {shape}")
        patch = Patch(level, box)
        self.patches.append(patch)
        return patch

    def regrid(self, tags: dict, **options) -> "AMRHierarchy":
        """A new hierarchy whose level L + 1 covers the tagged cells of level L
        (``tags``: level -> bool array of the level's shape), as patchwork's regrid is
        called by MHD2D.regrid (mhd.py:1052); ``options``: ``max_patch_size``,
        ``buffer``, ``efficiency`` (`wellbalanced_amd.regrid.regrid`)."""
        from .regrid import regrid

        return regrid(self, tags, **options)

    # coordinates ----------------------------------------------------------
    def cell_centers(self, patch: Patch) -> list[np.ndarray]:
        """The cell-centre coordinates of a patch, one array per axis (``indexing='ij'``)."""
        return self._grid(patch, (True,) * self.ndim)

    def node_coordinates(self, patch: Patch) -> list[np.ndarray]:
        return self._grid(patch, (False,) * self.ndim)

    def edge_coordinates(self, patch: Patch, axis: int) -> list[np.ndarray]:
        """Midpoints of the edges along ``axis`` (cell-centred along it, nodes across)."""
        return self._grid(patch, tuple(a == axis for a in range(self.ndim)))

    def _grid(self, patch: Patch, centred) -> list[np.ndarray]:
        spacing = self.spacing(patch.level)
        axes = []
        for a, ((lo, hi), c) in enumerate(zip(patch.box, centred)):
            index = np.arange(lo, hi if c else hi + 1, dtype=np.float64)
            offset = 0.5 if c else 0.0
            axes.append(self.bounds[a][0] + (index + offset) * spacing[a])
        return list(np.meshgrid(*axes, indexing="ij"))
