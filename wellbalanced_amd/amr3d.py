"""The ideal-MHD step on a refined 3D hierarchy of equal blocks (periodic).

`AMRMHD3D` runs ``MHD3D.step`` (``mhd3d.py:726-754``) on every block of every level
of a `wellbalanced_amd.amr3d_maps.Hierarchy3D`: one time step for all levels, the
stage of every block in one launch, the flux register of the coarse-fine faces over
both stages, the potential restricted between the stages (the mean of the fine edges
along a coarse edge, ``mhd3d.py:31-36``), and after the Heun average one reflux, the
covered cells averaged down and the potential restricted, so that div B is zero on
the composite mesh.
The stage and the refluxing read the cells around a block in place from the blocks
of its level (``direct``); the ghosts its level has no block for are interpolated from
the level below (``wb_amr3d_fill_regions``; without ``direct``, ``wb_amr3d_fill``
fills every ghost).  ``precision="mixed"`` runs the FP32 reconstruction/HLLD stage and
the mixed CFL rate.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib, amr3d_maps
from .amr3d_maps import HALO, Hierarchy3D, block_views
from .mhd import LIMITER_CODES

PRECISIONS = ("fp64", "mixed")
_LEVELS = np.dtype([("nlev", np.int32), ("nb", np.int32, 6), ("base", np.int32, 6)])


class AMRMHD3D:
    def __init__(self, hierarchy: Hierarchy3D, *, gamma: float = 5.0 / 3.0,
                 mean_field=(0.0, 0.0, 0.0), limiter: str = "minmod", cfl: float = 0.3,
                 precision: str = "fp64", device=None) -> None:
        if limiter not in LIMITER_CODES:
            raise ValueError(f"limiter is one of {LIMITER_CODES}")
        if precision not in PRECISIONS:
            raise ValueError(f"precision is one of {PRECISIONS}")
        if len(hierarchy.levels) > 6:
            raise ValueError("at most 6 levels")
        _lib.load()
        self.hierarchy = hierarchy
        self.gamma = float(gamma)
        self.mean_field = tuple(float(b) for b in mean_field)
        self.limiter = limiter
        self.cfl = float(cfl)
        self.precision = precision
        self.device = torch.device(device or "cuda")
        self.tables = amr3d_maps.build(hierarchy)
        self.nlev = len(hierarchy.levels)
        self.nblocks = len(self.tables.blocks)
        opts = dict(dtype=torch.float64, device=self.device)
        self._buf = [torch.zeros(self.tables.total, **opts), torch.zeros(self.tables.total, **opts)]
        self._rate = torch.zeros(2, **opts)
        self._dt = torch.zeros(1, **opts)
        self._t = torch.zeros(1, **opts)
        self._bad = torch.zeros(2, dtype=torch.int32, device=self.device)
        self.steps = 0
        self.check_interval = 1
        # the mixed block stage's launch shape: 3 = half a block per workgroup at 3 waves
        # per SIMD, 2 = a block per workgroup at 2 (the autotuner picks)
        self.waves = 3
        # the stage and the refluxing read the cells around a block in place from the
        # blocks of its level, so that only the ghosts the level has no block for are
        # filled (a launch-shape choice; the autotuner picks)
        self.direct = True
        self._upload_tables()

    # -- tables --------------------------------------------------------------------------
    def _bytes(self, a: np.ndarray) -> torch.Tensor:
        return torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).copy()).to(self.device)

    def _upload_tables(self) -> None:
        # new device tables: a captured graph holds the old addresses
        self.generation = getattr(self, "generation", 0) + 1
        t = self.tables
        self._blocks = self._bytes(t.blocks)
        self._lookup = torch.as_tensor(t.lookup, dtype=torch.int32, device=self.device)
        lv = np.zeros(1, _LEVELS)
        lv["nlev"] = self.nlev
        lv["nb"][0, : self.nlev] = t.level_nb
        lv["base"][0, : self.nlev] = t.level_base
        self._levels = self._bytes(lv)
        # the refined blocks of each level (coarsened finest level first)
        level_of = t.blocks["level"]
        self._children = []
        for L in range(self.nlev - 1):
            rows = t.children[level_of[t.children[:, 0]] == L] if len(t.children) else t.children
            self._children.append((torch.as_tensor(rows.reshape(-1), dtype=torch.int32,
                                                   device=self.device), len(rows)))
        self._reflux_items = self._bytes(t.reflux_items) if len(t.reflux_items) else self._blocks
        self._reflux_cells = self._bytes(t.reflux_cells) if len(t.reflux_cells) else self._blocks
        self.nreflux = len(t.reflux_cells)
        # the flux register: half the first stage's correction of each refluxed cell
        # the coarse-fine block faces whose fluxes the stage records for the refluxing
        slots, nslots = amr3d_maps.face_slots(self.nblocks, t.reflux_items)
        self._rec_slots = torch.as_tensor(slots.reshape(-1), device=self.device)
        # the fluxes of the coarse-fine block faces each stage solves (stage 1, stage 2)
        self._rec1 = torch.zeros(max(nslots, 1) * 256 * 5, dtype=torch.float64, device=self.device)
        self._rec = torch.zeros(max(nslots, 1) * 256 * 5, dtype=torch.float64, device=self.device)
        self._hoods = torch.as_tensor(t.hoods, device=self.device)
        # the blocks with a missing neighbour on their level: with direct reads, the only
        # ones whose ghosts (those the level has no block for) are filled
        ids = np.arange(self.nblocks, dtype=np.int32)
        self._missing = torch.as_tensor(ids[t.missing], device=self.device)
        self._regions = torch.as_tensor(t.fill_regions.reshape(-1), device=self.device)
        self.nregions = len(t.fill_regions)

    # -- state ---------------------------------------------------------------------------
    def set_state(self, density, velocity, pressure, *, vector_potential=(0.0, 0.0, 0.0)) -> None:
        """Every block from the state functions on its own level, then the covered
        cells and edges coarsened and the ghosts filled."""
        state = dict(density=density, velocity=velocity, pressure=pressure,
                     vector_potential=vector_potential)
        buf = amr3d_maps.pack_state(self.hierarchy, self.tables, self.gamma, self.mean_field, state)
        self.load_buffer(buf)

    def load_buffer(self, buf: np.ndarray) -> None:
        self._buf[0].copy_(torch.from_numpy(np.asarray(buf, np.float64)))
        stream = _lib.stream_handle(self.device)
        for L in range(self.nlev - 2, -1, -1):
            self._coarsen(self._buf[0], L, stream)
        self._fill(self._buf[0], stream)
        self._t.zero_()
        self.steps = 0

    def buffer(self) -> np.ndarray:
        return self._buf[0].cpu().numpy()

    @classmethod
    def from_uniform(cls, model, hierarchy: Hierarchy3D, **options) -> "AMRMHD3D":
        """A hierarchy started from the state of a uniform `MHD3D` model of the base
        resolution: the level-0 blocks copy it, every finer block is interpolated from
        the level below (as a regrid fills new patches), then the covered cells and
        edges are coarsened and the ghosts filled."""
        opts = dict(gamma=model.gamma, mean_field=model.mean_field, limiter=model.limiter,
                    cfl=model.cfl, device=model.device)
        opts.update(options)
        amr = cls(hierarchy, **opts)
        if tuple(model.n) != hierarchy.level_shape(0):
            raise ValueError("the uniform model must have the base resolution")
        cons, pots = model.arrays()
        buf = np.zeros(amr.tables.total)
        b, h = hierarchy.block, HALO
        for row in amr.tables.blocks:
            if row["level"] != 0:
                continue
            cells, views = block_views(buf, int(row["off"]), b)
            o = [int(row[a]) * b for a in ("bx", "by", "bz")]
            cells[:, h:-h, h:-h, h:-h] = cons[:, h + o[0] : h + o[0] + b, h + o[1] : h + o[1] + b,
                                             h + o[2] : h + o[2] + b]
            for c in range(3):
                sl = tuple(slice(h, h + b + (a != c)) for a in range(3))
                src = tuple(slice(h + o[a], h + o[a] + b + (a != c)) for a in range(3))
                views[c][sl] = pots[c][src]
        amr._buf[0].copy_(torch.from_numpy(buf))
        stream = _lib.stream_handle(amr.device)
        amr._fill(amr._buf[0], stream)
        levels = amr.tables.blocks["level"]
        for L in range(1, amr.nlev):
            ids = torch.as_tensor(np.flatnonzero(levels == L).astype(np.int32), device=amr.device)
            _lib.call("wb_amr3d_prolong", _lib.ptr(amr._buf[0]), _lib.ptr(amr._blocks),
                      _lib.ptr(ids), len(ids), _lib.ptr(amr._lookup), _lib.ptr(amr._levels), stream)
        for L in range(amr.nlev - 2, -1, -1):
            amr._coarsen(amr._buf[0], L, stream)
        amr._fill(amr._buf[0], stream)
        amr._t.fill_(model.time)
        return amr

    def level_state(self, level: int = 0):
        """The state of a level that blocks cover entirely (level 0) as uniform haloed
        arrays ``cons (5, n+4, n+4, n+4)`` and the three potentials (halos zero)."""
        n = self.hierarchy.level_shape(level)[0]
        b, h = self.hierarchy.block, HALO
        buf = self.buffer()
        X = n + 4
        cons = np.zeros((5, X, X, X))
        pots = [np.zeros((X, X + 1, X + 1)), np.zeros((X + 1, X, X + 1)), np.zeros((X + 1, X + 1, X))]
        for row in self.tables.blocks:
            if row["level"] != level:
                continue
            cells, views = block_views(buf, int(row["off"]), b)
            o = [int(row[a]) * b for a in ("bx", "by", "bz")]
            cons[:, h + o[0] : h + o[0] + b, h + o[1] : h + o[1] + b, h + o[2] : h + o[2] + b] = \
                cells[:, h:-h, h:-h, h:-h]
            for c in range(3):
                sl = tuple(slice(h, h + b + (a != c)) for a in range(3))
                dst = tuple(slice(h + o[a], h + o[a] + b + (a != c)) for a in range(3))
                pots[c][dst] = views[c][sl]
        return cons, pots

    @property
    def time(self) -> float:
        return float(self._t.item())

    @property
    def cells(self) -> int:
        """Cells of every block of every level (the zone-updates of one step)."""
        return self.nblocks * self.hierarchy.block**3

    # -- the step ------------------------------------------------------------------------
    def _fill(self, buf: torch.Tensor, stream: int) -> None:
        _lib.call("wb_amr3d_fill", _lib.ptr(buf), _lib.ptr(self._blocks), self.nblocks,
                  _lib.ptr(self._lookup), _lib.ptr(self._levels), stream)

    def _coarsen(self, buf: torch.Tensor, L: int, stream: int, cells: bool = True) -> None:
        """The covered cells averaged (``cells``) and the edges restricted."""
        rows, n = self._children[L]
        if n:
            _lib.call("wb_amr3d_coarsen", _lib.ptr(buf), _lib.ptr(self._blocks), _lib.ptr(rows), n,
                      int(cells), _lib.ptr(self._lookup), _lib.ptr(self._levels), stream)

    def _stage(self, src, dst, heun: bool, slot: int, stream: int, marks=None) -> None:
        # direct reads: the stage and the refluxing read the cells around a block in
        # place from the blocks of its level, so only the ghosts the level has no block
        # for are filled
        if not self.direct:
            self._fill(src, stream)
        elif self.nregions:
            _lib.call("wb_amr3d_fill_regions", _lib.ptr(src), _lib.ptr(self._blocks),
                      _lib.ptr(self._hoods), _lib.ptr(self._regions), self.nregions, stream)
        suffix = "" if self.precision == "fp64" else "_mixed"
        code = LIMITER_CODES.index(self.limiter)
        if marks is not None:
            marks[0].record()
        extra = () if self.precision == "fp64" else (int(self.waves),)
        # the stage records the fluxes of the coarse-fine block faces for the refluxing,
        # each stage in its own record
        record = self._rec if heun else self._rec1
        rec = (_lib.ptr(self._rec_slots), _lib.ptr(record)) if self.nreflux else (None, None)
        _lib.call("wb_amr3d_stage" + suffix, _lib.ptr(src), _lib.ptr(dst), _lib.ptr(self._blocks),
                  None, self.nblocks, int(self.direct), _lib.ptr(self._lookup),
                  _lib.ptr(self._levels), *self.mean_field, self.gamma, code, _lib.ptr(self._dt),
                  int(heun), self._bad.data_ptr() + 4 * slot, *extra, *rec, stream)
        if marks is not None:
            marks[1].record()
        if self.nreflux and heun:  # both stages' refluxing after the Heun average
            _lib.call("wb_amr3d_reflux_both", _lib.ptr(dst), _lib.ptr(self._blocks),
                      _lib.ptr(self._reflux_items), _lib.ptr(self._reflux_cells), self.nreflux,
                      _lib.ptr(self._rec_slots), _lib.ptr(self._rec1), _lib.ptr(self._rec),
                      _lib.ptr(self._dt), stream)
        # after stage 1 the potential is restricted (mhd3d.py:740-741), after the Heun
        # average the covered cells are averaged down and the potential restricted
        for L in range(self.nlev - 2, -1, -1):
            self._coarsen(dst, L, stream, cells=heun)

    def _heun(self, stream: int, events=None):
        marks = None
        if events is not None:
            marks = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        self._stage(self._buf[0], self._buf[1], False, 0, stream, marks[:2] if marks else None)
        self._stage(self._buf[1], self._buf[0], True, 1, stream, marks[2:] if marks else None)
        if events is not None:
            events.append((marks[0], marks[1]))
            events.append((marks[2], marks[3]))

    def _rate_kernel(self, stream: int) -> None:
        suffix = "" if self.precision == "fp64" else "_mixed"
        _lib.call("wb_amr3d_rate" + suffix, _lib.ptr(self._buf[0]), _lib.ptr(self._blocks),
                  self.nblocks, *self.mean_field, self.gamma, _lib.ptr(self._rate), stream)

    def stable_step(self) -> float:
        """``cfl`` over the largest fast-wave rate of any block cell (mhd3d.py:674-724)."""
        self._rate_kernel(_lib.stream_handle(self.device))
        return self.cfl * (1.0 / float(self._rate[0].item()))

    def step(self, dt: float | None = None) -> float:
        if dt is None:
            dt = self.stable_step()
        dt = float(dt)
        self._dt.fill_(dt)
        self._heun(_lib.stream_handle(self.device))
        self._t.add_(dt)
        self.steps += 1
        if self.check_interval and self.steps % self.check_interval == 0:
            self.check()
        return dt

    def advance(self, steps: int = 1, events: bool = False):
        """``steps`` steps at the stable step with dt and the time on the device."""
        stream = _lib.stream_handle(self.device)
        marks = [] if events else None
        for _ in range(steps):
            self._rate_kernel(stream)
            _lib.call("wb_stable_dt", _lib.ptr(self._rate), self.cfl, 0.0, _lib.ptr(self._dt),
                      _lib.ptr(self._t), stream)
            self._heun(stream, marks)
            self.steps += 1
        return marks

    def check(self) -> None:
        if int(self._bad.cpu().sum()):
            raise FloatingPointError(f"non-positive density or pressure at t = {self.time}")

    # -- diagnostics ---------------------------------------------------------------------
    def _covered(self) -> np.ndarray:
        """Per block: True where a finer block covers it."""
        t = self.tables
        cov = np.zeros(self.nblocks, bool)
        if len(t.children):
            cov[t.children[:, 0]] = True
        return cov

    def leaf_totals(self) -> np.ndarray:
        """Integrals of the 5 conserved fields over the leaf cells."""
        buf = self.buffer()
        totals = np.zeros(5)
        cov = self._covered()
        h = HALO
        for k, row in enumerate(self.tables.blocks):
            if cov[k]:
                continue
            cells, _ = block_views(buf, int(row["off"]), self.hierarchy.block)
            vol = float(row["dx"] * row["dy"] * row["dz"])
            totals += cells[:, h:-h, h:-h, h:-h].sum(axis=(1, 2, 3)) * vol
        return totals

    def divergence(self) -> float:
        """Largest |div B| times the cell size over the cells of every block."""
        from .mhd3d import face_fields

        buf = self.buffer()
        largest = 0.0
        h = HALO
        for row in self.tables.blocks:
            _, pots = block_views(buf, int(row["off"]), self.hierarchy.block)
            spacing = (float(row["dx"]), float(row["dy"]), float(row["dz"]))
            local = [p[tuple(slice(h, p.shape[a] - h) for a in range(3))] for p in pots]
            faces = face_fields(local, spacing, self.mean_field)
            div = sum(np.diff(f, axis=a) / spacing[a] for a, f in enumerate(faces))
            largest = max(largest, float(np.abs(div).max()) * min(spacing))
        return largest

    def composite_divergence(self) -> float:
        """Largest mismatch of a coarse edge of a refined block and the mean of the fine
        edges along it, relative to the largest fine potential difference: zero after
        every stage (the flux through a coarse face is the sum of the fine fluxes)."""
        buf = self.buffer()
        t = self.tables
        b = self.hierarchy.block
        worst, scale = 0.0, 0.0
        for row in t.children:
            _, cp = block_views(buf, int(t.blocks[row[0]]["off"]), b)
            kids = [block_views(buf, int(t.blocks[k]["off"]), b)[1] for k in row[1:]]
            for c in range(3):
                ext = [b if a == c else b + 1 for a in range(3)]
                for q in np.ndindex(*ext):
                    vals = []
                    for s in (0, 1):
                        f = [2 * q[a] + (s if a == c else 0) for a in range(3)]
                        oct_ = [1 if f[a] >= 2 * b else f[a] // b for a in range(3)]
                        loc = tuple(f[a] - b * oct_[a] + 2 for a in range(3))
                        vals.append(kids[(oct_[0] * 2 + oct_[1]) * 2 + oct_[2]][c][loc])
                    coarse = cp[c][tuple(v + 2 for v in q)]
                    worst = max(worst, abs(coarse - 0.5 * (vals[0] + vals[1])))
                    scale = max(scale, abs(vals[0] - vals[1]))
        return worst / scale if scale else worst
