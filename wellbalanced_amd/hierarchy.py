"""A minimal block-structured mesh with the interface of ``patchwork.AMRHierarchy``
that ``MHD2D``/``MHD3D`` use (carafe ``mhd.py:467-477,703-734,748-777``).

The reference allocates its fields on the external ``patchwork`` package (not
available here).  This stand-in holds the levels, patches and boxes, and the
subset of patchwork's field interface the reference models and tests touch:
``allocate_field``, ``patch.field(name).local`` / ``.local_with_halos``,
``gather``, ``integral``, ``exchange_halos`` and ``restrict``.  A model's fields
are views into the model's own state buffer in HBM (``torch`` tensors on the HIP
device): writing into ``patch.field(name).local`` changes the model's state, and
``exchange_halos``/``restrict``/``integral`` of those fields run the model's HIP
kernels.  Fields no model owns are plain device arrays per patch.  Level 0 covers
the domain with one patch; finer patches are added with `add_patch`.
"""

from __future__ import annotations

import weakref
from dataclasses import dataclass

import numpy as np

CENTERINGS = ("cell", "node", "x-edge", "y-edge", "z-edge")


@dataclass(frozen=True)
class Patch:
    """A box of cells ``((i0, i1), (j0, j1)[, (k0, k1)])`` on one level."""

    level: int
    box: tuple

    @property
    def shape(self) -> tuple[int, ...]:
        return tuple(hi - lo for lo, hi in self.box)

    def field(self, name: str) -> "PatchField":
        """The field ``name`` on this patch (views into device memory)."""
        hierarchy = getattr(self, "_hierarchy", None)
        if hierarchy is None:
            raise ValueError("the patch belongs to no hierarchy")
        return hierarchy.field(self, name)

    def local_cell_centers(self, hierarchy: "AMRHierarchy", name: str | None = None):
        """The cell-centre coordinates of the patch's cells, one array per axis."""
        return hierarchy.cell_centers(self)

    def local_node_coordinates(self, hierarchy: "AMRHierarchy", name: str | None = None):
        """The coordinates of the patch's nodes, one array per axis."""
        return hierarchy.node_coordinates(self)


class PatchField:
    """One field on one patch: ``local_with_halos`` is the haloed device array,
    ``local`` the view of its interior (the patch's cells, or its nodes/edges)."""

    def __init__(self, with_halos, halo: int) -> None:
        self.local_with_halos = with_halos
        self.halo = halo
        self.local = with_halos[tuple(slice(halo, n - halo) for n in with_halos.shape)]


@dataclass
class FieldSpec:
    halo: int
    centering: str
    conservative: bool
    owner: object = None  # a weak reference to the model whose state holds the field

    def model(self):
        """The owning model, None for a field no model owns."""
        if self.owner is None:
            return None
        model = self.owner()
        if model is None:
            raise RuntimeError("the model that held this field no longer exists")
        return model


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
        self.patches: list[Patch] = []
        self._specs: dict[str, FieldSpec] = {}
        self._store: dict[tuple[str, Patch], object] = {}
        self._adopt(Patch(0, tuple((0, n) for n in shape)))

    def _adopt(self, patch: Patch) -> Patch:
        object.__setattr__(patch, "_hierarchy", self)
        self.patches.append(patch)
        return patch

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
        return self._adopt(Patch(level, box))

    def regrid(self, tags: dict, **options) -> "AMRHierarchy":
        """A new hierarchy whose level L + 1 covers the tagged cells of level L
        (``tags``: level -> bool array of the level's shape), as patchwork's regrid is
        called by MHD2D.regrid (mhd.py:1052); ``options``: ``max_patch_size``,
        ``buffer``, ``efficiency`` (`wellbalanced_amd.regrid.regrid`)."""
        from .regrid import regrid

        return regrid(self, tags, **options)

    # fields (patchwork's field interface, as mhd.py:467-477,703-777,1149-1194 use it)
    def allocate_field(self, name: str, *, halo: int = 2, split=None, conservative: bool = False,
                       centering: str = "cell") -> None:
        """Declare the field ``name`` with ``halo`` ghost layers on every patch.
        ``centering``: "cell", "node" or an edge ("x-edge", ...).  Storage of a field no
        model owns is allocated on the current HIP device at its first use."""
        if centering not in CENTERINGS:
            raise ValueError(f"centering is one of {CENTERINGS}")
        if centering.endswith("-edge") and "xyz".index(centering[0]) >= self.ndim:
            raise ValueError(f"a {self.ndim}D hierarchy has no {centering} fields")
        if halo < 0:
            raise ValueError("halo must be non-negative")
        if name in self._specs:
            raise ValueError(f"the field {name!r} is already allocated")
        self._specs[name] = FieldSpec(int(halo), centering, bool(conservative))

    def _bind(self, owner, fields: dict[str, str], halo: int) -> None:
        """Make ``owner`` (a model) the holder of ``fields`` (name -> centering): their
        patch arrays become views into its state buffer.  A later model with the same
        field names takes them over."""
        for name, centering in fields.items():
            self._specs[name] = FieldSpec(halo, centering, centering == "cell",
                                          weakref.ref(owner))
            for key in [k for k in self._store if k[0] == name]:
                del self._store[key]

    def _spec(self, name: str) -> FieldSpec:
        spec = self._specs.get(name)
        if spec is None:
            raise KeyError(f"no field {name!r} on the hierarchy (allocate_field first)")
        return spec

    def _index(self, patch: Patch) -> int:
        for k, p in enumerate(self.patches):
            if p == patch:
                return k
        raise ValueError(f"{patch} is not a patch of this hierarchy")

    def _haloed_shape(self, patch: Patch, spec: FieldSpec) -> tuple[int, ...]:
        nodes = {"cell": (), "node": tuple(range(self.ndim))}.get(spec.centering)
        if nodes is None:  # an edge along axis a: nodal across it
            axis = "xyz".index(spec.centering[0])
            nodes = tuple(b for b in range(self.ndim) if b != axis)
        return tuple(n + 2 * spec.halo + (1 if a in nodes else 0)
                     for a, n in enumerate(patch.shape))

    def field(self, patch: Patch, name: str) -> PatchField:
        """``patch.field(name)``: the device views of the field on the patch."""
        spec = self._spec(name)
        k = self._index(patch)
        model = spec.model()
        if model is not None:
            return PatchField(model._field_view(k, name), spec.halo)
        key = (name, self.patches[k])
        array = self._store.get(key)
        if array is None:
            import torch

            if not torch.cuda.is_available():
                raise RuntimeError("field storage lives on a HIP device; there is none")
            array = torch.zeros(self._haloed_shape(patch, spec), dtype=torch.float64,
                                device="cuda")
            self._store[key] = array
        return PatchField(array, spec.halo)

    def gather(self, patch: Patch, name: str) -> np.ndarray:
        """The field's values on the patch (without halos) as a host array."""
        return self.field(patch, name).local.detach().cpu().numpy().copy()

    def _leaf_mask(self, patch: Patch) -> np.ndarray:
        """True on the patch's cells that no patch of the next level covers."""
        mask = np.ones(patch.shape, bool)
        r = self.refinement_ratio
        for fine in self.patches_at_level(patch.level + 1):
            cut = []
            for (lo, hi), (flo, fhi) in zip(patch.box, fine.box):
                a, b = max(lo, flo // r), min(hi, -(-fhi // r))
                if a >= b:
                    break
                cut.append(slice(a - lo, b - lo))
            else:
                mask[tuple(cut)] = False
        return mask

    def integral(self, name: str) -> float:
        """The integral of a cell field over the domain, from the leaf cells.  A model's
        conserved field is summed on the device by the model (a fixed-order FP64 tree)."""
        spec = self._spec(name)
        if spec.centering != "cell":
            raise ValueError("integral is defined for cell fields")
        model = spec.model()
        if model is not None:
            return float(model._integral(name))
        import torch

        total = 0.0
        for patch in self.patches:
            values = self.field(patch, name).local
            mask = torch.as_tensor(self._leaf_mask(patch), device=values.device)
            total += float((values * mask).sum()) * float(np.prod(self.spacing(patch.level)))
        return total

    def exchange_halos(self, name: str, interpolation: str = "linear") -> None:
        """Fill the ghost layers of ``name`` on every patch.  A model's fields are
        filled together by the model's halo kernels (periodic copies, coarse-fine
        interpolation, outflow and wall rows); ``interpolation`` is then fixed by the
        model (bilinear cells, quadratic nodes)."""
        model = self._spec(name).model()
        if model is None:
            raise NotImplementedError(
                "halo exchange is provided for the fields of a model (its HIP kernels)")
        model._exchange_halos()

    def restrict(self, name: str) -> None:
        """Average the field from each level onto the level below (a model's conserved
        cells and its potential on shared nodes, finest level first)."""
        model = self._spec(name).model()
        if model is not None:
            model._restrict_fields()
        elif self.max_level > 0:
            raise NotImplementedError(
                "restriction is provided for the fields of a model (its HIP kernels)")

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
