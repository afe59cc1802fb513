"""``AMRMHD2D``: the 2D MHD step on a refined patch hierarchy.

The reference runs ``MHD2D`` on patchwork patches of several levels that all take
the same time step (``mhd.py:40-44``): per stage the halos from the hierarchy, the
fluxes of every patch and the register update with scale ``0.5 dt``; A_z restricted
between the stages; after the Heun average one reflux and the restriction
(``mhd.py:889-994``).  Here every patch's state sits in one FP64 buffer in HBM (two,
ping-ponged as in `MHD2D`); a stage is

1. the halo fill, level by level: same-level copies, then bilinear cells and
   quadratic A_z nodes from the level below (gather tables of `amr_maps`), then on a
   non-periodic axis the outflow or wall ghosts of the level's patches on the domain
   faces (``mhd.py:748-777``), which the next level's interpolation may read;
2. ONE launch over every tile of every patch, which updates every patch cell, covered
   coarse cells included: a workgroup per 16 x 16 tile (`k_stage2d_tiles`, the default)
   or the persistent marching stage (`k_stage2d_batched`: workgroups pulling (patch,
   tile, segment) items from a counter; ``kernel="march"``);
3. the flux register of the coarse cells next to finer patches (one launch for every
   level pair): stage 1 records half its correction, stage 2 (whose output is the Heun
   average) adds the recorded half and its own;
4. after stage 1 the restriction of A_z (coarse nodes on fine nodes take the fine
   value); after stage 2 the average of covered coarse cells over their children and
   the restriction, finest level pair first.

dt comes from the largest fast-wave rate over every patch cell (the finest cells
bind it) and, with resistivity or conduction, the explicit diffusion limit of each
patch.  ``precision="mixed"`` runs the FP32 reconstruction/HLLD stage.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib, amr_maps
from .hierarchy import AMRHierarchy
from .mhd import BOUNDARIES, CONSERVED, HALO, LIMITER_CODES, PRECISIONS, _device

_PATCH = np.dtype([("off", np.int64), ("nx", np.int32), ("ny", np.int32),
                   ("dx", np.float64), ("dy", np.float64), ("walls", np.int32),
                   ("pad", np.int32)])
_ITEM = np.dtype([("patch", np.int32), ("tile", np.int32), ("seg", np.int32)])
KERNELS = ("tiles", "march")
TILE = 16  # cells per side of the tiled stage's tiles


class AMRMHD2D:
    """MHD on every patch of a 2D `AMRHierarchy` (see the module).  ``boundaries`` maps
    a non-periodic axis ('x' or 'y') to 'outflow' (the default) or 'wall'
    (``mhd.py:410-455``)."""

    def __init__(self, hierarchy: AMRHierarchy, *, gamma: float = 5.0 / 3.0,
                 mean_field=(0.0, 0.0), limiter: str = "minmod", cfl: float = 0.4,
                 resistivity: float = 0.0, conduction: tuple[float, float] = (0.0, 0.0),
                 boundaries: dict[str, str] | None = None, prefix: str = "mhd_",
                 precision: str = "fp64", tile: int = 62, nxs: int = 16, kernel: str = "tiles",
                 device=None) -> None:
        if hierarchy.ndim != 2:
            raise ValueError("AMRMHD2D needs a 2D hierarchy")
        if limiter not in LIMITER_CODES:
            raise ValueError(f"limiter is one of {sorted(LIMITER_CODES)}")
        if not 0.0 < cfl <= 1.0:
            raise ValueError("cfl must be in (0, 1]")
        if gamma <= 1.0:
            raise ValueError("gamma must be larger than 1")
        if not resistivity >= 0.0:
            raise ValueError("resistivity must be non-negative")
        if not min(conduction) >= 0.0:
            raise ValueError("the conductivities must be non-negative")
        if precision not in PRECISIONS:
            raise ValueError(f"precision is one of {PRECISIONS}")
        if tile not in (30, 62, 126):
            raise ValueError("tile is 30, 62 or 126 cells")
        if kernel not in KERNELS:
            raise ValueError(f"kernel is one of {KERNELS}")
        kinds = {"x": "outflow", "y": "outflow"}
        for axis_name, kind in (boundaries or {}).items():
            if axis_name not in kinds or kind not in BOUNDARIES:
                raise ValueError(
                    f"boundaries maps 'x' or 'y' to one of {BOUNDARIES}, "
                    f"got {axis_name!r}: {kind!r}"
                )
            axis = "xy".index(axis_name)
            if hierarchy.periodic[axis]:
                raise ValueError(f"the hierarchy is periodic along {axis_name}")
            if kind == "wall" and mean_field[axis] != 0.0:
                raise ValueError(f"a wall normal to {axis_name} needs a zero mean field along it")
            kinds[axis_name] = kind
        self.walls = tuple(kinds[name] == "wall" and not hierarchy.periodic[a]
                           for a, name in enumerate("xy"))
        self.periodic = tuple(bool(p) for p in hierarchy.periodic)
        self.hierarchy = hierarchy
        self.gamma = float(gamma)
        self.mean_field = (float(mean_field[0]), float(mean_field[1]))
        self.limiter = limiter
        self.cfl = float(cfl)
        self.resistivity = float(resistivity)
        self.conduction = (float(conduction[0]), float(conduction[1]))
        self.precision = precision
        self.tile = tile
        self.nxs = nxs
        self.kernel = kernel
        self.names = {name: prefix + name for name in CONSERVED}
        self.potential = prefix + "az"
        self.device = _device(device)
        self._use(hierarchy, amr_maps.build(hierarchy), upload=False)
        opts = dict(dtype=torch.float64, device=self.device)
        self._dt = torch.zeros(1, **opts)
        self._t = torch.zeros(1, **opts)
        self._rate = torch.zeros(3, **opts)
        self._bad = torch.zeros(2, dtype=torch.int32, device=self.device)
        # the bad-cell count of kernels outside the step (tagging): never checked
        self._scratch_bad = torch.zeros(1, dtype=torch.int32, device=self.device)
        self._counter = torch.zeros(1, dtype=torch.int32, device=self.device)
        self._upload_tables()
        self.workgroups = 2048
        self.check_interval = 1
        self.steps = 0
        self._ready = False

    def _use(self, hierarchy: AMRHierarchy, maps, upload: bool = True) -> None:
        """Take ``hierarchy`` (its tables ``maps``) with new, zeroed state buffers."""
        self.hierarchy = hierarchy
        hierarchy._bind(self, {**{n: "cell" for n in self.names.values()},
                               self.potential: "node"}, HALO)
        self.maps = maps
        self.layouts = maps.layouts
        self.nlev = hierarchy.max_level + 1
        opts = dict(dtype=torch.float64, device=self.device)
        self._buf = [torch.zeros(maps.total, **opts), torch.zeros(maps.total, **opts)]
        if upload:
            self._upload_tables()

    def set_tiling(self, tile: int | None = None, nxs: int | None = None,
                   workgroups: int | None = None, kernel: str | None = None) -> None:
        """Change the launch shape (no effect on the results) and rebuild the work list:
        ``kernel`` "tiles" (a workgroup per 16 x 16 tile of a patch) or "march" (the
        persistent marching stage with ``tile``, ``nxs`` and ``workgroups``)."""
        if kernel is not None:
            if kernel not in KERNELS:
                raise ValueError(f"kernel is one of {KERNELS}")
            self.kernel = kernel
        if tile is not None:
            if tile not in (30, 62, 126):
                raise ValueError("tile is 30, 62 or 126 cells")
            self.tile = tile
        if nxs is not None:
            self.nxs = int(nxs)
        if workgroups is not None:
            self.workgroups = int(workgroups)
        self._upload_tables()

    # -- tables on the device ------------------------------------------------------------
    def _dev(self, a, dtype=torch.int64):
        return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=self.device)

    def _upload_tables(self) -> None:
        # new device tables (and, after _use, new state buffers): a captured graph holds
        # the old addresses (wellbalanced_amd.graph)
        self.generation = getattr(self, "generation", 0) + 1
        m = self.maps
        self._copy, self._interp = {}, {}
        for L in range(self.nlev):
            cd, cs, cpd, cps = m.copy_cells[L]
            nd, ns = m.copy_nodes[L]
            self._copy[L] = [self._dev(x) for x in (cd, cs, cpd, cps, nd, ns)]
            dst, pd, c00, cy, pc, wx, wy = m.cell_interp[L]
            nid, nbase, nsx, nsy, nkx, nky = m.node_interp[L]
            self._interp[L] = ([self._dev(x) for x in (dst, pd, c00, cy, pc)]
                               + [self._dev(wx, torch.float64), self._dev(wy, torch.float64)]
                               + [self._dev(x) for x in (nid, nbase, nsx, nsy, nkx, nky)])
        self._average, self._restrict = {}, {}
        # the refluxing of every level pair in one table: each level pair writes only its
        # coarse level's cells and reads only the stage input, so one launch does them all
        # before the coarsening; a group is one coarse cell's rows
        rows_all, groups = [], []
        base = 0
        for L in range(self.nlev - 1):
            rows = m.reflux[L]
            i = 0
            while i < len(rows):
                j = i
                while j < len(rows) and tuple(rows[j, :3]) == tuple(rows[i, :3]):
                    j += 1
                groups.append((base + i, j - i))
                i = j
            rows_all.append(rows.reshape(-1, 13))
            base += len(rows)
        rows_all = np.concatenate(rows_all) if rows_all else np.zeros((0, 13), np.int64)
        self._reflux = (self._dev(rows_all.reshape(-1)),
                        self._dev(np.array(groups, np.int32).reshape(-1), torch.int32), len(groups))
        # the flux register: half the first stage's correction of each refluxed cell
        self._reg = torch.zeros(6 * max(len(groups), 1), dtype=torch.float64, device=self.device)
        for L in range(self.nlev - 1):
            self._average[L] = (self._dev(m.average[L].reshape(-1)), len(m.average[L]))
            self._restrict[L] = (self._dev(m.restrict[L].reshape(-1)), len(m.restrict[L]))
        patches = np.zeros(len(self.layouts), _PATCH)
        spatch = np.zeros(len(self.layouts))
        items = []
        # the patches on a non-periodic domain face, per level: rows {patch, faces, walls}
        wall_bits = int(self.walls[0]) | (int(self.walls[1]) << 1)
        faces = np.zeros(len(self.layouts), np.int64)
        self._faces = {}
        for L in range(self.nlev):
            rows = m.domain_faces.get(L, np.zeros((0, 2), np.int64))
            faces[rows[:, 0]] = rows[:, 1]
            table = np.zeros((len(rows), 3), np.int32)
            table[:, :2] = rows
            table[:, 2] = wall_bits
            self._faces[L] = (self._dev(table.reshape(-1), torch.int32), len(rows))
        # a face is a wall face if its axis has walls: x faces bits 0-1, y faces bits 2-3
        wall_faces = faces & ((3 if self.walls[0] else 0) | (12 if self.walls[1] else 0))
        for k, lay in enumerate(self.layouts):
            dx, dy = self.hierarchy.spacing(lay.level)
            patches[k] = (lay.offset, lay.nx, lay.ny, dx, dy, int(wall_faces[k]), 0)
            spatch[k] = dx**-2 + dy**-2  # the diffusion limit's sum (mhd.py:961)
            for t in range((lay.ny + self.tile - 1) // self.tile):
                for s in range((lay.nx + self.nxs - 1) // self.nxs):
                    items.append((k, t, s))
        # the largest patches first, so the persistent workgroups end together
        sizes = np.array([self.layouts[i[0]].nx * self.layouts[i[0]].ny for i in items])
        order = np.argsort(-sizes, kind="stable")
        items = np.array([items[i] for i in order], _ITEM)
        # the tiled stage's items: (patch, tile along y, tile along x), a patch's tiles
        # together and y fastest, so neighbouring tiles run on one XCD
        tiles = [(k, ty, tx) for k, lay in enumerate(self.layouts)
                 for tx in range(-(-lay.nx // TILE)) for ty in range(-(-lay.ny // TILE))]
        self._tiles = torch.from_numpy(np.array(tiles, _ITEM).view(np.uint8).copy()).to(self.device)
        self.ntiles = len(tiles)
        self._record_tables(rows_all)
        self._patches = torch.from_numpy(patches.view(np.uint8).copy()).to(self.device)
        self._items = torch.from_numpy(items.view(np.uint8).copy()).to(self.device)
        self._spatch = self._dev(spatch, torch.float64)
        self.nitems = len(items)
        # the CFL reduction walks every patch's cells in chunks of 256
        chunks = [(k, c) for k, lay in enumerate(self.layouts)
                  for c in range(0, lay.nx * lay.ny, 256)]
        self._chunks = self._dev(np.array(chunks, np.int32).reshape(-1), torch.int32)
        self.nchunks = len(chunks)
        fine = [c for c in chunks if self.layouts[c[0]].level > 0]
        self._fine_chunks = self._dev(np.array(fine, np.int32).reshape(-1), torch.int32)
        self.nfine = len(fine)
        # the leaf cells (not covered by a finer patch): patch, haloed local cell
        covered = np.zeros(self.maps.total, bool)
        for L in range(self.nlev - 1):
            rows = self.maps.average[L]
            if len(rows):
                covered[rows[:, 0]] = True
        leaf, self._leaf_xy = [], []
        for k, lay in enumerate(self.layouts):
            a, b = np.meshgrid(np.arange(lay.nx) + HALO, np.arange(lay.ny) + HALO, indexing="ij")
            keep = ~covered[lay.cell(a, b)]
            a, b = a[keep], b[keep]
            leaf.append(np.stack([np.full(a.size, k), a * lay.Y + b], axis=1))
            dx, dy = self.hierarchy.spacing(lay.level)
            lo = self.hierarchy.lower_bound
            (i0, _), (j0, _) = lay.box
            self._leaf_xy.append(np.stack([lo[0] + (i0 + a - HALO + 0.5) * dx,
                                           lo[1] + (j0 + b - HALO + 0.5) * dy,
                                           np.full(a.size, lay.level)], axis=1))
        leaf = np.concatenate(leaf).astype(np.int32)
        self._leaf_xy = np.concatenate(self._leaf_xy)
        self._leaf = self._dev(leaf.reshape(-1), torch.int32)
        self.nleaf = len(leaf)
        restrict = [self.maps.restrict[L] for L in range(self.nlev - 1) if len(self.maps.restrict[L])]
        restrict = np.concatenate(restrict) if restrict else np.zeros((0, 2), np.int64)
        self._restrict_all = self._dev(restrict.reshape(-1))
        self.nrestrict = len(restrict)
        self._work = torch.zeros(_lib.call("wb_amr2d_totals_workspace"), dtype=torch.float64,
                                 device=self.device)

    def _record_tables(self, rows: np.ndarray) -> None:
        """The tiled stage records the fluxes of the refluxing's faces and the refluxing
        reads them: a slot per distinct face (patch, axis, face a, face b) of the reflux
        rows, each face listed by the tile that holds it (the tile of the cell on its
        high side, the patch's last tile for a face on its high boundary)."""
        keys = np.concatenate([rows[:, [0, 3, 5, 6]], rows[:, [7, 3, 8, 9]],
                               rows[:, [10, 3, 11, 12]]]) if len(rows) else np.zeros((0, 4), np.int64)
        # one integer per face, ordered as the rows (patch, axis, face a, face b) are
        packed = ((keys[:, 0] << 41) | (keys[:, 1] << 40) | (keys[:, 2] << 20) | keys[:, 3])
        unique, inverse = np.unique(packed, return_inverse=True)
        inverse = inverse.reshape(-1)
        faces = np.stack([unique >> 41, (unique >> 40) & 1, (unique >> 20) & 0xFFFFF,
                          unique & 0xFFFFF], axis=1)
        n = len(rows)
        rows8 = np.zeros((n, 8), np.int32)
        if n:
            rows8[:, :5] = rows[:, :5]
            rows8[:, 5], rows8[:, 6], rows8[:, 7] = inverse[:n], inverse[n:2 * n], inverse[2 * n:]
        nx = np.array([lay.nx for lay in self.layouts], np.int64)
        ny = np.array([lay.ny for lay in self.layouts], np.int64)
        ntx, nty = -(-nx // TILE), -(-ny // TILE)
        first_item = np.concatenate([[0], np.cumsum(ntx * nty)])
        p, axis, fa, fb = (faces[:, k] for k in range(4))
        tx, ty = (fa - HALO) // TILE, (fb - HALO) // TILE
        tx = np.where(axis == 0, np.minimum(tx, ntx[p] - 1), tx)
        ty = np.where(axis == 1, np.minimum(ty, nty[p] - 1), ty)
        i, j = fa - (HALO + TILE * tx), fb - (HALO + TILE * ty)
        index = np.where(axis == 0, i * (TILE + 2) + j + 1, (i + 1) * (TILE + 1) + j)
        item = first_item[p] + tx * nty[p] + ty
        # per item: its count, then (code, slot) pairs
        counts = np.bincount(item, minlength=self.ntiles)
        first = np.concatenate([[0], np.cumsum(1 + 2 * counts)[:-1]])
        flat = np.zeros(int(np.sum(1 + 2 * counts)), np.int32)
        flat[first] = counts
        order = np.argsort(item, kind="stable")
        starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
        rank = np.arange(len(order)) - starts[item[order]]
        at = first[item[order]] + 1 + 2 * rank
        flat[at] = (axis[order] | (index[order] << 1)).astype(np.int32)
        flat[at + 1] = order.astype(np.int32)
        self._rows8 = self._dev(rows8.reshape(-1), torch.int32)
        self._rec_first = self._dev(first, torch.int32)
        self._rec_lists = self._dev(flat, torch.int32)
        # the recorded fluxes of stage 1 and stage 2
        self._rec1 = torch.zeros(6 * max(len(faces), 1), dtype=torch.float64, device=self.device)
        self._rec = torch.zeros(6 * max(len(faces), 1), dtype=torch.float64, device=self.device)

    # -- state ---------------------------------------------------------------------------
    def set_state(self, density, velocity, pressure, *, vector_potential=0.0, bz=0.0) -> None:
        """Set the state on every patch, then average and restrict (mhd.py:683-737)."""
        buf = amr_maps.pack_state(self.hierarchy, self.layouts, self.maps.total, self.gamma,
                                  self.mean_field, dict(density=density, velocity=velocity,
                                                        pressure=pressure,
                                                        vector_potential=vector_potential, bz=bz))
        self.load_buffer(buf)
        stream = _lib.stream_handle(self.device)
        for L in range(self.nlev - 2, -1, -1):
            self._coarsen(self._buf[0], L, stream)

    def load_buffer(self, buf: np.ndarray) -> None:
        self._buf[0].copy_(torch.from_numpy(np.asarray(buf, np.float64)))
        self._t.zero_()
        self.steps = 0
        self._ready = True

    def buffer(self) -> np.ndarray:
        return self._buf[0].cpu().numpy()

    @property
    def time(self) -> float:
        return float(self._t.item())

    @property
    def cells(self) -> int:
        """Cells on every level (the zone-updates of one step)."""
        return int(sum(lay.nx * lay.ny for lay in self.layouts))

    # -- the hierarchy's field interface (patch.field(name), gather, integral, ...) -------
    def _field_view(self, index: int, name: str) -> torch.Tensor:
        """The haloed device view of field ``name`` on patch ``index`` (the hierarchy's
        order) in the current state."""
        lay = self.layouts[index]
        buf = self._buf[0]
        if name == self.potential:
            start = lay.offset + 6 * lay.plane
            return buf[start : lay.offset + lay.size].view(lay.X + 1, lay.Y + 1)
        start = lay.offset + list(self.names.values()).index(name) * lay.plane
        return buf[start : start + lay.plane].view(lay.X, lay.Y)

    def _exchange_halos(self) -> None:
        self._fill(self._buf[0], _lib.stream_handle(self.device))

    def _restrict_fields(self) -> None:
        stream = _lib.stream_handle(self.device)
        for L in range(self.nlev - 2, -1, -1):
            self._coarsen(self._buf[0], L, stream)

    def _integral(self, name: str) -> float:
        key = next(k for k, v in self.names.items() if v == name)
        return self.totals()[key]

    # -- the step ------------------------------------------------------------------------
    def _fill_gathers(self, buf: torch.Tensor, L: int, stream: int) -> None:
        cd, cs, cpd, cps, nd, ns = self._copy[L]
        t = self._interp[L]
        _lib.call("wb_amr2d_fill_level", _lib.ptr(buf), _lib.ptr(cd), _lib.ptr(cs),
                  _lib.ptr(cpd), _lib.ptr(cps), len(cd), _lib.ptr(nd), _lib.ptr(ns), len(nd),
                  *[_lib.ptr(x) for x in t[:7]], len(t[0]), *[_lib.ptr(x) for x in t[7:]],
                  len(t[7]), stream)

    def _fill(self, buf: torch.Tensor, stream: int) -> None:
        for L in range(self.nlev):
            self._fill_level(buf, L, stream)

    def _fill_level(self, buf: torch.Tensor, L: int, stream: int) -> None:
        self._fill_gathers(buf, L, stream)
        rows, n = self._faces[L]
        if n:  # the domain faces of the level's patches (mhd.py:748-777)
            _lib.call("wb_amr2d_domain_faces", _lib.ptr(buf), _lib.ptr(self._patches),
                      _lib.ptr(rows), n, stream)

    def _coarsen(self, buf: torch.Tensor, L: int, stream: int) -> None:
        arows, na = self._average[L]
        rrows, nr = self._restrict[L]
        _lib.call("wb_amr2d_coarsen", _lib.ptr(buf), _lib.ptr(arows), na, _lib.ptr(rrows), nr,
                  stream)

    def _restrict_nodes(self, buf: torch.Tensor, L: int, stream: int) -> None:
        rrows, nr = self._restrict[L]
        _lib.call("wb_amr2d_coarsen", _lib.ptr(buf), None, 0, _lib.ptr(rrows), nr, stream)

    def _stage(self, src, dst, heun: bool, slot: int, stream: int) -> None:
        self._fill(src, stream)
        suffix = "" if self.precision == "fp64" else "_mixed"
        code = LIMITER_CODES.index(self.limiter)
        diff = (self.resistivity, *self.conduction)
        rows, groups, n = self._reflux
        if self.kernel == "tiles":
            # the stage records the fluxes of the refluxing's faces (each stage in its
            # own record), which the refluxing reads
            record = self._rec if heun else self._rec1
            rec = ((_lib.ptr(self._rec_first), _lib.ptr(self._rec_lists), _lib.ptr(record))
                   if n else (None, None, None))
            _lib.call("wb_amr2d_stage_tiles" + suffix, _lib.ptr(src), _lib.ptr(dst),
                      _lib.ptr(self._patches), _lib.ptr(self._tiles), self.ntiles,
                      *self.mean_field, self.gamma, *diff, code, _lib.ptr(self._dt), int(heun),
                      self._bad.data_ptr() + 4 * slot, *rec, stream)
            if n and heun:  # both stages' register and the reflux (mhd.py:911-919,989)
                _lib.call("wb_amr2d_reflux_both" + suffix, _lib.ptr(dst),
                          _lib.ptr(self._patches), _lib.ptr(self._rows8), _lib.ptr(groups), n,
                          _lib.ptr(self._rec1), _lib.ptr(self._rec), _lib.ptr(self._dt),
                          stream)
        else:
            _lib.call("wb_amr2d_stage" + suffix, _lib.ptr(src), _lib.ptr(dst),
                      _lib.ptr(self._patches), _lib.ptr(self._items), self.nitems,
                      _lib.ptr(self._counter), *self.mean_field, self.gamma, *diff, code,
                      _lib.ptr(self._dt), int(heun), self._bad.data_ptr() + 4 * slot,
                      int(self.nxs), int(self.tile), int(self.workgroups), stream)
            if n:  # stage 1 records into the register, stage 2 refluxes (mhd.py:911-919,989)
                _lib.call("wb_amr2d_reflux" + suffix, _lib.ptr(src), _lib.ptr(dst),
                          _lib.ptr(self._patches), _lib.ptr(rows), _lib.ptr(groups), n,
                          *self.mean_field, self.gamma, *diff, code, _lib.ptr(self._dt),
                          int(not heun), _lib.ptr(self._reg), stream)
        for L in range(self.nlev - 2, -1, -1):
            if heun:  # the synchronisation after the Heun average (mhd.py:989-991)
                self._coarsen(dst, L, stream)
            else:  # A_z between the stages (mhd.py:981)
                self._restrict_nodes(dst, L, stream)

    @property
    def _diffusive(self) -> bool:
        return bool(self.resistivity) or max(self.conduction) > 0.0

    def _rate_kernel(self, stream: int) -> None:
        spatch = _lib.ptr(self._spatch) if self._diffusive else None
        _lib.call("wb_amr2d_rate", _lib.ptr(self._buf[0]), _lib.ptr(self._patches),
                  _lib.ptr(self._chunks), self.nchunks, *self.mean_field, self.gamma,
                  self.resistivity, (self.gamma - 1.0) * max(self.conduction), spatch,
                  _lib.ptr(self._rate), stream)

    def _dt_kernel(self, time_ptr, stream: int) -> None:
        if self._diffusive:  # the largest D s of the patches is rate[2]
            _lib.call("wb_stable_dt_diffusive", _lib.ptr(self._rate), self.cfl, 0.0,
                      self.resistivity, (self.gamma - 1.0) * max(self.conduction), 0.0,
                      _lib.ptr(self._dt), time_ptr, stream)
        else:
            _lib.call("wb_stable_dt", _lib.ptr(self._rate), self.cfl, 0.0, _lib.ptr(self._dt),
                      time_ptr, stream)

    def stable_step(self) -> float:
        stream = _lib.stream_handle(self.device)
        self._rate_kernel(stream)
        self._dt_kernel(None, stream)
        return float(self._dt[0].item())

    def _heun(self, stream: int, events=None) -> None:
        a, b = self._buf
        if events is None:
            self._stage(a, b, False, 0, stream)
            self._stage(b, a, True, 1, stream)
            return
        marks = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        marks[0].record()
        self._stage(a, b, False, 0, stream)
        marks[1].record()
        marks[2].record()
        self._stage(b, a, True, 1, stream)
        marks[3].record()
        events.extend([(marks[0], marks[1]), (marks[2], marks[3])])

    def step(self, dt: float | None = None) -> float:
        if not self._ready:
            raise RuntimeError("set the state first (set_state)")
        if dt is None:
            dt = self.stable_step()
        dt = float(dt)
        stream = _lib.stream_handle(self.device)
        self._dt.fill_(dt)
        self._heun(stream)
        self._t.add_(dt)
        self.steps += 1
        if self.check_interval and self.steps % self.check_interval == 0:
            self.check()
        return dt

    def advance(self, steps: int = 1, events: bool = False):
        stream = _lib.stream_handle(self.device)
        marks = [] if events else None
        for _ in range(steps):
            self._rate_kernel(stream)
            self._dt_kernel(_lib.ptr(self._t), stream)
            self._heun(stream, marks)
            self.steps += 1
        return marks

    def run(self, duration: float, *, max_steps: int = 1_000_000) -> int:
        """Advance by ``duration``, the last step shortened (mhd.py:996-1008)."""
        if not np.isfinite(duration) or duration < 0.0:
            raise ValueError("duration must be a finite, non-negative time")
        end = self.time + duration
        taken = 0
        while self.time < end * (1.0 - 1e-14) and taken < max_steps:
            self.step(min(self.stable_step(), end - self.time))
            taken += 1
        return taken

    def current_tags(self, threshold: float, levels: int) -> dict[int, np.ndarray]:
        """Tag the cells where |J_z| times the cell size exceeds ``threshold`` on levels
        ``0 .. levels - 1`` that exist, patch by patch (``mhd.py:1011-1041``)."""
        stream = _lib.stream_handle(self.device)
        buf = self._buf[0]
        self._fill(buf, stream)
        tags = {L: np.zeros(self.hierarchy.level_shape(L), dtype=bool)
                for L in range(min(levels, self.nlev))}
        big = max(lay.X * lay.Y for lay in self.layouts)
        opts = dict(dtype=torch.float64, device=self.device)
        prim = torch.empty(8 * big, **opts)
        bxf = torch.empty(big + max(lay.Y for lay in self.layouts), **opts)
        byf = torch.empty(big + max(lay.X for lay in self.layouts), **opts)
        ecc = torch.empty(big, **opts)
        out = torch.empty(big, dtype=torch.uint8, device=self.device)
        for lay in self.layouts:
            if lay.level not in tags:
                continue
            dx, dy = self.hierarchy.spacing(lay.level)
            base = buf.data_ptr() + 8 * lay.offset
            plane = 8 * lay.plane
            _lib.call("wb_primitive2d", *[base + k * plane for k in range(6)], base + 6 * plane,
                      lay.X, lay.Y, dx, dy, *self.mean_field, self.gamma, _lib.ptr(prim),
                      _lib.ptr(bxf), _lib.ptr(byf), _lib.ptr(ecc), _lib.ptr(self._scratch_bad),
                      stream)
            _lib.call("wb_current_tags2d", _lib.ptr(bxf), _lib.ptr(byf), lay.nx, lay.ny, dx, dy,
                      float(threshold), _lib.ptr(out), stream)
            cell = out[: lay.nx * lay.ny].view(lay.nx, lay.ny).cpu().numpy().astype(bool)
            (i0, i1), (j0, j1) = lay.box
            tags[lay.level][i0:i1, j0:j1] |= cell
        return tags

    def regrid(self, tags: dict, **options) -> None:
        """Regrid on ``tags`` (level -> bool cells, as `current_tags` gives them) and move
        the state onto the new patches (mhd.py:1043-1056): same-level cells and nodes
        are copied, new cells take the conservative prolongation of their coarse cell
        (integrals kept), new A_z nodes the bilinear interpolation (div B stays zero).
        ``options``: ``max_patch_size``, ``buffer``, ``efficiency`` of
        `wellbalanced_amd.regrid.regrid`."""
        if not self._ready:
            raise RuntimeError("set the state first (set_state)")
        if not all(self.periodic):
            raise NotImplementedError("regridding runs on periodic hierarchies")
        from . import regrid as rg

        old_h, old = self.hierarchy, self._buf[0]
        new_h = rg.regrid(old_h, tags, **options)
        maps = amr_maps.build(new_h)
        tables = rg.transfer_tables(old_h, new_h, maps)
        self._use(new_h, maps)
        new = self._buf[0]
        stream = _lib.stream_handle(self.device)
        ptr = _lib.ptr
        for L in range(self.nlev):
            t = tables[L]
            cd, cs, cpd, cps = (self._dev(x) for x in t["copy_cells"])
            nd, ns = (self._dev(x) for x in t["copy_nodes"])
            _lib.call("wb_amr2d_regrid_copy", ptr(old), ptr(new), ptr(cd), ptr(cs), ptr(cpd),
                      ptr(cps), len(cd), ptr(nd), ptr(ns), len(nd), stream)
            cells = [self._dev(x) for x in t["cells"]]
            nodes = [self._dev(x) for x in t["nodes"]]
            _lib.call("wb_amr2d_prolong", ptr(new), *(ptr(x) for x in cells), len(cells[0]),
                      *(ptr(x) for x in nodes), len(nodes[0]), stream)
            self._fill_level(new, L, stream)
        for L in range(self.nlev - 2, -1, -1):
            self._coarsen(new, L, stream)
        torch.cuda.synchronize(self.device)

    def check(self) -> None:
        if int(self._bad.cpu().sum()):
            raise FloatingPointError(f"non-positive density or pressure at t = {self.time}")

    # -- diagnostics (on the device; only the results come back) ----------------------
    def leaf_totals(self) -> np.ndarray:
        """Integrals of the 6 conserved fields over the leaf cells (a fixed-order FP64
        tree on the device)."""
        out = torch.zeros(6, dtype=torch.float64, device=self.device)
        _lib.call("wb_amr2d_totals", _lib.ptr(self._buf[0]), _lib.ptr(self._patches),
                  _lib.ptr(self._leaf), self.nleaf, _lib.ptr(self._work), _lib.ptr(out),
                  _lib.stream_handle(self.device))
        return out.cpu().numpy()

    def totals(self) -> dict[str, float]:
        """Integrals of the conserved fields over the domain, from the leaf cells
        (mhd.py:1172-1175)."""
        return {name: float(v) for name, v in zip(CONSERVED, self.leaf_totals())}

    def _divergences(self) -> np.ndarray:
        out = torch.zeros(3, dtype=torch.float64, device=self.device)
        _lib.call("wb_amr2d_divergence", _lib.ptr(self._buf[0]), _lib.ptr(self._patches),
                  _lib.ptr(self._chunks), self.nchunks, *self.mean_field,
                  _lib.ptr(self._restrict_all), self.nrestrict, _lib.ptr(self._fine_chunks),
                  self.nfine, _lib.ptr(out), _lib.stream_handle(self.device))
        return out.cpu().numpy()

    def composite_divergence(self) -> float:
        """The largest mismatch of a coarse node and the fine node on it, relative to the
        largest A_z difference of neighbouring fine nodes (mhd.py:1133-1170): the
        coarse faces' flux equals the sum of the fine faces' over them."""
        _, worst, scale = self._divergences()
        return float(worst / scale) if scale else 0.0

    def divergence(self) -> float:
        """Largest |div B| times the cell size over the cells of every patch
        (mhd.py:1114-1131)."""
        return float(self._divergences()[0])

    def dataset(self, level: int = 0) -> dict[str, np.ndarray]:
        """The primitive variables of ``level`` over its cell centres, the level's patches
        gathered into arrays of the level's shape with NaN where it has no patch
        (mhd.py:1177-1226; plain arrays instead of xarray)."""
        if not 0 <= level < self.nlev:
            raise ValueError(f"the hierarchy has levels 0 .. {self.nlev - 1}")
        buf = self.buffer()
        nx, ny = self.hierarchy.level_shape(level)
        dx, dy = self.hierarchy.spacing(level)
        names = ("density", "velocity_x", "velocity_y", "velocity_z", "pressure",
                 "magnetic_field_x", "magnetic_field_y", "magnetic_field_z", "magnetic_pressure")
        out = {name: np.full((nx, ny), np.nan) for name in names}
        for lay in self.layouts:
            if lay.level != level:
                continue
            c = buf[lay.offset : lay.offset + 6 * lay.plane].reshape(6, lay.X, lay.Y)
            c = c[:, HALO:-HALO, HALO:-HALO]
            a = buf[lay.offset + 6 * lay.plane : lay.offset + lay.size].reshape(lay.X + 1, lay.Y + 1)
            a = a[HALO:-HALO, HALO:-HALO]
            bxf = (a[:, 1:] - a[:, :-1]) / dy + self.mean_field[0]
            byf = -(a[1:, :] - a[:-1, :]) / dx + self.mean_field[1]
            bx, by = 0.5 * (bxf[:-1] + bxf[1:]), 0.5 * (byf[:, :-1] + byf[:, 1:])
            rho = c[0]
            u, v, w = c[1] / rho, c[2] / rho, c[3] / rho
            b2 = bx**2 + by**2 + c[5] ** 2
            p = (self.gamma - 1.0) * (c[4] - 0.5 * rho * (u**2 + v**2 + w**2) - 0.5 * b2)
            (i0, i1), (j0, j1) = lay.box
            for name, value in zip(names, (rho, u, v, w, p, bx, by, c[5], 0.5 * b2)):
                out[name][i0:i1, j0:j1] = value
        lo = self.hierarchy.lower_bound
        return {"x": lo[0] + (np.arange(nx) + 0.5) * dx, "y": lo[1] + (np.arange(ny) + 0.5) * dy,
                **out, "time": self.time}

    def leaf_values(self) -> dict[str, np.ndarray]:
        """The primitive variables on the leaf cells (mhd.py:1059-1112): ``x``, ``y``,
        ``level``, the variables of `MHD2D.dataset` and ``current_z`` (the mean of the
        cell's corners), patch by patch in row-major order."""
        stream = _lib.stream_handle(self.device)
        self._fill(self._buf[0], stream)
        out = torch.empty(9 * max(self.nleaf, 1), dtype=torch.float64, device=self.device)
        _lib.call("wb_amr2d_leaf_values", _lib.ptr(self._buf[0]), _lib.ptr(self._patches),
                  _lib.ptr(self._leaf), self.nleaf, *self.mean_field, self.gamma, _lib.ptr(out),
                  stream)
        v = out[: 9 * self.nleaf].view(9, self.nleaf).cpu().numpy()
        names = ("density", "velocity_x", "velocity_y", "velocity_z", "pressure",
                 "magnetic_field_x", "magnetic_field_y", "magnetic_field_z", "current_z")
        return {"x": self._leaf_xy[:, 0].copy(), "y": self._leaf_xy[:, 1].copy(),
                "level": self._leaf_xy[:, 2].astype(np.int64),
                **{name: v[k].copy() for k, name in enumerate(names)}}


def blast_hierarchy(base: int = 512, levels: int = 3, sizes=(16, 128), seed: int = 0) -> AMRHierarchy:
    """The refined blast workload: a periodic unit square of ``base``^2 cells, and on
    each finer level a square around the centre (half the extent of the level below)
    tiled with patches whose sides are drawn uniformly in ``sizes``."""
    h = AMRHierarchy((base, base), ((0.0, 1.0), (0.0, 1.0)), periodic=(True, True))
    for level in range(1, levels):
        n = base * 2**level
        half = n // (2 ** (level + 1))
        lo = (n // 2 - half) // 2 * 2
        hi = (n // 2 + half) // 2 * 2
        amr_maps.random_patches(h, level, ((lo, hi), (lo, hi)), sizes=sizes, seed=seed + level)
    return h


def blast_state():
    """Pressure 100 inside r < 0.1 of the centre, 1 outside; rho = 1, v = 0; the
    uniform field B = (1, 1)/sqrt(2) is the mean field (A_z = 0)."""
    return dict(density=1.0, velocity=(0.0, 0.0, 0.0),
                pressure=lambda x, y: np.where(np.hypot(x - 0.5, y - 0.5) < 0.1, 100.0, 1.0))
