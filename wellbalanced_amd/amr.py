"""``AMRMHD2D``: the 2D MHD step on a refined, periodic patch hierarchy.

The reference runs ``MHD2D`` on patchwork patches of several levels that all take
the same time step (``mhd.py:40-44``): halos from the hierarchy, fluxes per patch,
the register update, A_z restriction after each stage, refluxing at the end
(``mhd.py:889-994``).  Here every patch's state sits in one FP64 buffer in HBM (two,
ping-ponged as in `MHD2D`); a stage is

1. the halo fill, level by level: same-level copies, then bilinear cells and
   4-point nodes from the level below (gather tables of `amr_maps`);
2. ONE persistent launch over every tile of every patch (`k_stage2d_batched`: the
   fused 2D stage, workgroups pulling (patch, tile, segment) items from a counter);
3. per coarse/fine level pair, finest first: refluxing of the uncovered coarse cells
   next to finer patches, the average of covered coarse cells, the restriction of
   A_z.

dt comes from the largest fast-wave rate over every patch cell (the finest cells
bind it) and, with resistivity or conduction, the explicit diffusion limit of each
patch.  ``precision="mixed"`` runs the FP32 reconstruction/HLLD stage.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib, amr_maps
from .hierarchy import AMRHierarchy
from .mhd import HALO, LIMITER_CODES, PRECISIONS, _device

_PATCH = np.dtype([("off", np.int64), ("nx", np.int32), ("ny", np.int32),
                   ("dx", np.float64), ("dy", np.float64)])
_ITEM = np.dtype([("patch", np.int32), ("tile", np.int32), ("seg", np.int32)])


class AMRMHD2D:
    """Ideal MHD on every patch of a periodic 2D `AMRHierarchy` (see the module)."""

    def __init__(self, hierarchy: AMRHierarchy, *, gamma: float = 5.0 / 3.0,
                 mean_field=(0.0, 0.0), limiter: str = "minmod", cfl: float = 0.4,
                 resistivity: float = 0.0, conduction: tuple[float, float] = (0.0, 0.0),
                 precision: str = "fp64", tile: int = 62, nxs: int = 16, device=None) -> None:
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
        if tile not in (62, 126):
            raise ValueError("tile is 62 or 126 cells")
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
        self.device = _device(device)
        self.maps = amr_maps.build(hierarchy)
        self.layouts = self.maps.layouts
        self.nlev = hierarchy.max_level + 1
        opts = dict(dtype=torch.float64, device=self.device)
        self._buf = [torch.zeros(self.maps.total, **opts), torch.zeros(self.maps.total, **opts)]
        self._dt = torch.zeros(1, **opts)
        self._t = torch.zeros(1, **opts)
        self._rate = torch.zeros(3, **opts)
        self._bad = torch.zeros(2, dtype=torch.int32, device=self.device)
        self._counter = torch.zeros(1, dtype=torch.int32, device=self.device)
        self._upload_tables()
        self.workgroups = 2048
        self.check_interval = 1
        self.steps = 0
        self._ready = False

    def set_tiling(self, tile: int | None = None, nxs: int | None = None,
                   workgroups: int | None = None) -> None:
        """Change the launch shape (no effect on the results) and rebuild the work list."""
        if tile is not None:
            if tile not in (62, 126):
                raise ValueError("tile is 62 or 126 cells")
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
        self._reflux, self._average, self._restrict = {}, {}, {}
        for L in range(self.nlev - 1):
            rows = m.reflux[L]
            groups = []
            i = 0
            while i < len(rows):
                j = i
                while j < len(rows) and tuple(rows[j, :3]) == tuple(rows[i, :3]):
                    j += 1
                groups.append((i, j - i))
                i = j
            self._reflux[L] = (self._dev(rows.reshape(-1)),
                               self._dev(np.array(groups, np.int32).reshape(-1), torch.int32),
                               len(groups))
            self._average[L] = (self._dev(m.average[L].reshape(-1)), len(m.average[L]))
            self._restrict[L] = (self._dev(m.restrict[L].reshape(-1)), len(m.restrict[L]))
        patches = np.zeros(len(self.layouts), _PATCH)
        spatch = np.zeros(len(self.layouts))
        items = []
        for k, lay in enumerate(self.layouts):
            dx, dy = self.hierarchy.spacing(lay.level)
            patches[k] = (lay.offset, lay.nx, lay.ny, dx, dy)
            spatch[k] = dx**-2 + dy**-2  # the diffusion limit's sum (mhd.py:961)
            for t in range((lay.ny + self.tile - 1) // self.tile):
                for s in range((lay.nx + self.nxs - 1) // self.nxs):
                    items.append((k, t, s))
        # the largest patches first, so the persistent workgroups end together
        sizes = np.array([self.layouts[i[0]].nx * self.layouts[i[0]].ny for i in items])
        order = np.argsort(-sizes, kind="stable")
        items = np.array([items[i] for i in order], _ITEM)
        self._patches = torch.from_numpy(patches.view(np.uint8).copy()).to(self.device)
        self._items = torch.from_numpy(items.view(np.uint8).copy()).to(self.device)
        self._spatch = self._dev(spatch, torch.float64)
        self.nitems = len(items)
        # the CFL reduction walks every patch's cells in chunks of 256
        chunks = [(k, c) for k, lay in enumerate(self.layouts)
                  for c in range(0, lay.nx * lay.ny, 256)]
        self._chunks = self._dev(np.array(chunks, np.int32).reshape(-1), torch.int32)
        self.nchunks = len(chunks)

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

    # -- the step ------------------------------------------------------------------------
    def _fill(self, buf: torch.Tensor, stream: int) -> None:
        p = _lib.ptr(buf)
        for L in range(self.nlev):
            cd, cs, cpd, cps, nd, ns = self._copy[L]
            _lib.call("wb_amr2d_copy", p, _lib.ptr(cd), _lib.ptr(cs), _lib.ptr(cpd), _lib.ptr(cps),
                      len(cd), _lib.ptr(nd), _lib.ptr(ns), len(nd), stream)
            t = self._interp[L]
            _lib.call("wb_amr2d_interp", p, *[_lib.ptr(x) for x in t[:7]], len(t[0]),
                      *[_lib.ptr(x) for x in t[7:]], len(t[7]), stream)

    def _coarsen(self, buf: torch.Tensor, L: int, stream: int) -> None:
        rows, n = self._average[L]
        _lib.call("wb_amr2d_average", _lib.ptr(buf), _lib.ptr(rows), n, stream)
        rows, n = self._restrict[L]
        _lib.call("wb_amr2d_restrict", _lib.ptr(buf), _lib.ptr(rows), n, stream)

    def _stage(self, src, dst, heun: bool, slot: int, stream: int) -> None:
        self._fill(src, stream)
        suffix = "" if self.precision == "fp64" else "_mixed"
        code = LIMITER_CODES.index(self.limiter)
        diff = (self.resistivity, *self.conduction)
        _lib.call("wb_amr2d_stage" + suffix, _lib.ptr(src), _lib.ptr(dst), _lib.ptr(self._patches),
                  _lib.ptr(self._items), self.nitems, _lib.ptr(self._counter), *self.mean_field,
                  self.gamma, *diff, code, _lib.ptr(self._dt), int(heun),
                  self._bad.data_ptr() + 4 * slot, int(self.nxs), int(self.tile),
                  int(self.workgroups), stream)
        for L in range(self.nlev - 2, -1, -1):
            rows, groups, n = self._reflux[L]
            _lib.call("wb_amr2d_reflux" + suffix, _lib.ptr(src), _lib.ptr(dst),
                      _lib.ptr(self._patches), _lib.ptr(rows), _lib.ptr(groups), n,
                      *self.mean_field, self.gamma, *diff, code, _lib.ptr(self._dt), int(heun),
                      stream)
            self._coarsen(dst, L, stream)

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
                      _lib.ptr(bxf), _lib.ptr(byf), _lib.ptr(ecc), self._bad.data_ptr() + 4,
                      stream)
            _lib.call("wb_current_tags2d", _lib.ptr(bxf), _lib.ptr(byf), lay.nx, lay.ny, dx, dy,
                      float(threshold), _lib.ptr(out), stream)
            cell = out[: lay.nx * lay.ny].view(lay.nx, lay.ny).cpu().numpy().astype(bool)
            (i0, i1), (j0, j1) = lay.box
            tags[lay.level][i0:i1, j0:j1] |= cell
        return tags

    def check(self) -> None:
        if int(self._bad.cpu().sum()):
            raise FloatingPointError(f"non-positive density or pressure at t = {self.time}")

    # -- diagnostics ---------------------------------------------------------------------
    def leaf_totals(self) -> np.ndarray:
        """Integrals of the 6 conserved fields over the leaf cells."""
        buf = self.buffer()
        covered = np.zeros(self.maps.total, bool)
        for L in range(self.nlev - 1):
            rows = self.maps.average[L]
            if len(rows):
                covered[rows[:, 0]] = True
        totals = np.zeros(6)
        for lay in self.layouts:
            dx, dy = self.hierarchy.spacing(lay.level)
            a, b = np.meshgrid(np.arange(lay.nx) + HALO, np.arange(lay.ny) + HALO, indexing="ij")
            offs = lay.cell(a, b)
            keep = ~covered[offs]
            for f in range(6):
                totals[f] += float(np.sum(buf[offs + f * lay.plane][keep])) * (dx * dy)
        return totals

    def composite_divergence(self) -> float:
        """Largest mismatch of a coarse node and the fine node on it, relative to the
        largest fine A_z difference (mhd.py:1133-1170): zero after every stage."""
        buf = self.buffer()
        worst, scale = 0.0, 0.0
        for L in range(self.nlev - 1):
            rows = self.maps.restrict[L]
            if len(rows):
                worst = max(worst, float(np.abs(buf[rows[:, 0]] - buf[rows[:, 1]]).max()))
        for lay in self.layouts:
            if lay.level == 0:
                continue
            az = buf[lay.offset + 6 * lay.plane : lay.offset + lay.size].reshape(lay.X + 1, lay.Y + 1)
            a = az[HALO:-HALO, HALO:-HALO]
            scale = max(scale, float(np.abs(np.diff(a, axis=0)).max()), float(np.abs(np.diff(a, axis=1)).max()))
        return worst / scale if scale else 0.0

    def divergence(self) -> float:
        """Largest |div B| times the cell size over the cells of every patch."""
        buf = self.buffer()
        largest = 0.0
        for lay in self.layouts:
            dx, dy = self.hierarchy.spacing(lay.level)
            az = buf[lay.offset + 6 * lay.plane : lay.offset + lay.size].reshape(lay.X + 1, lay.Y + 1)
            a = az[HALO:-HALO, HALO:-HALO]
            bxf = (a[:, 1:] - a[:, :-1]) / dy + self.mean_field[0]
            byf = -(a[1:, :] - a[:-1, :]) / dx + self.mean_field[1]
            div = (bxf[1:] - bxf[:-1]) / dx + (byf[:, 1:] - byf[:, :-1]) / dy
            largest = max(largest, float(np.abs(div).max()) * min(dx, dy))
        return largest


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
