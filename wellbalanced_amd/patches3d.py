"""``RefinedMHD3D``: ``MHD3D`` on a refined 3D patch hierarchy (patches of any size).

The reference's ``MHD3D`` runs on patchwork hierarchies (``mhd3d.py:163-965``): per
stage the halos of every level, the fluxes of every patch (primitive, three sweeps,
CT-contact edge field, the resistive terms), the register update with scale ``0.5
dt`` and ``A -= dt E``; the potential restricted between the stages; after the Heun
average one reflux and the restriction (``mhd3d.py:726-754``); ``current_tags`` and
``regrid`` (``mhd3d.py:771-821``).  Here every patch's state sits in one FP64 buffer in
HBM (two, ping-ponged), and a stage is

1. the ghost fill, one gather launch per level (tables of `patches3d_maps`);
2. per patch the FP64 split kernels of the uniform grid (``wb_primitive3d``,
   ``wb_sweep3d`` x 3, ``wb_edge3d`` x 3, ``wb_resistive3d``, ``wb_mhd3d_update``), each
   patch's sweeps kept for
3. the flux register of every coarse-fine face (``wb_patches3d_reflux``): stage 1
   records half its correction, stage 2 adds it and its own half;
4. after stage 1 the restriction of the edges, after stage 2 the average of covered
   cells and the restriction, finest level pair first (one gather launch each).

FP64 only; periodic domains.  The block hierarchy of `wellbalanced_amd.amr3d` is the
throughput path for refined 3D runs; this is the drop-in for arbitrary patches.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib, patches3d_maps
from .hierarchy import AMRHierarchy
from .mhd import LIMITER_CODES, _device
from .patches3d_maps import HALO, NCELL

CONSERVED_3D = ("density", "momentum_x", "momentum_y", "momentum_z", "energy")


class RefinedMHD3D:
    """3D MHD on every patch of a periodic refined `AMRHierarchy` (see the module)."""

    def __init__(self, hierarchy: AMRHierarchy, *, gamma: float = 5.0 / 3.0,
                 mean_field=(0.0, 0.0, 0.0), resistivity: float = 0.0, limiter: str = "vanleer",
                 cfl: float = 0.3, prefix: str = "mhd_", device=None) -> None:
        if hierarchy.ndim != 3:
            raise ValueError("MHD3D needs a 3D hierarchy")
        if limiter not in LIMITER_CODES:
            raise ValueError(f"limiter is one of {sorted(LIMITER_CODES)}")
        if not 0.0 < cfl <= 1.0:
            raise ValueError("cfl must be in (0, 1]")
        if gamma <= 1.0:
            raise ValueError("gamma must be larger than 1")
        if not resistivity >= 0.0:
            raise ValueError("resistivity must be non-negative")
        if not all(hierarchy.periodic):
            raise NotImplementedError("refined 3D hierarchies run on periodic domains")
        self.gamma = float(gamma)
        self.mean_field = tuple(float(b) for b in mean_field)
        self.resistivity = float(resistivity)
        self.limiter = limiter
        self.cfl = float(cfl)
        self.precision = "fp64"
        self.names = {name: prefix + name for name in CONSERVED_3D}
        self.potential = tuple(prefix + n for n in ("ax", "ay", "az"))
        self.device = _device(device)
        self._dt = torch.zeros(1, dtype=torch.float64, device=self.device)
        self._t = torch.zeros(1, dtype=torch.float64, device=self.device)
        self.check_interval = 1
        self.steps = 0
        self.period = float("nan")
        self._use(hierarchy, patches3d_maps.build(hierarchy))
        self._ready = False

    # -- buffers and tables -------------------------------------------------------------
    def _use(self, hierarchy: AMRHierarchy, maps) -> None:
        """Take ``hierarchy`` (its tables ``maps``) with new, zeroed state buffers."""
        self.hierarchy = hierarchy
        hierarchy._bind(self, {**{n: "cell" for n in self.names.values()},
                               **dict(zip(self.potential, ("x-edge", "y-edge", "z-edge")))},
                        HALO)
        self.maps = maps
        self.layouts = maps.layouts
        self.nlev = hierarchy.max_level + 1
        self.generation = getattr(self, "generation", 0) + 1
        opts = dict(dtype=torch.float64, device=self.device)
        self._buf = [torch.zeros(maps.total, **opts), torch.zeros(maps.total, **opts)]
        self._fill_t = {L: self._stencils(s) for L, s in maps.fill.items()}
        self._coarsen_t = {L: self._stencils(s) for L, s in maps.coarsen.items()}
        self._restrict_t = {L: self._stencils(s) for L, s in maps.restrict.items()}
        # per patch its sweeps (kept until the register has read them), shared work arrays
        big = max(int(np.prod(lay.dims)) for lay in self.layouts)
        nodes = max(int(np.prod([d + 1 for d in lay.dims])) for lay in self.layouts)
        self._prim = torch.empty(8 * big, **opts)
        self._faces = [torch.empty(nodes, **opts) for _ in range(3)]
        self._edges = [torch.empty(nodes, **opts) for _ in range(3)]
        self._sweep_off = []
        total = 0
        for lay in self.layouts:
            X, Y, Z = lay.dims
            sizes = (7 * (X - 3) * Y * Z, 7 * X * (Y - 3) * Z, 7 * X * Y * (Z - 3))
            offs = [total, total + sizes[0], total + sizes[0] + sizes[1]]
            self._sweep_off.append(offs)
            total += sum(sizes)
        self._sweeps = torch.zeros(max(total, 1), **opts)
        self._reflux_tables()
        self._bad = torch.zeros(2 * len(self.layouts), dtype=torch.int32, device=self.device)
        self._rates = torch.zeros(2 * len(self.layouts), **opts)

    def _dev(self, a, dtype=torch.int64):
        return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=self.device)

    def _stencils(self, s):
        return (self._dev(s.dst), self._dev(s.start), self._dev(s.count, torch.int32),
                self._dev(s.src), self._dev(s.w, torch.float64), len(s))

    def _reflux_tables(self) -> None:
        cells, faces, spacing = [], [], []
        for k, a, b, c, L, fl in self.maps.reflux:
            lay = self.layouts[k]
            cells.append((lay.index("cells", a, b, c), int(np.prod(lay.extent("cells"))),
                          len(faces), len(fl)))
            h = self.hierarchy.spacing(L)
            for axis, side, coarse, fines in fl:
                row = [axis, side]
                for p, fa, fb, fc in [coarse, *fines]:
                    off, stride = self._sweep_index(p, axis, fa, fb, fc)
                    row += [off, stride]
                faces.append(row)
                spacing.append(h[axis])
        self.nreflux = len(cells)
        self._rcells = self._dev(np.array(cells, np.int64).reshape(-1))
        self._rfaces = self._dev(np.array(faces, np.int64).reshape(-1))
        self._rspacing = self._dev(np.array(spacing, np.float64), torch.float64)
        self._reg = torch.zeros(5 * max(self.nreflux, 1), dtype=torch.float64, device=self.device)

    def _sweep_index(self, k, axis, a, b, c):
        """Offset (row 0) and row stride in the sweeps buffer of the flux through the low
        face of haloed local cell (a, b, c) along ``axis`` of patch k (the sweep holds
        face f between cells f + 1 and f + 2)."""
        lay = self.layouts[k]
        dims = list(lay.dims)
        dims[axis] -= 3
        idx = [a, b, c]
        idx[axis] -= 2
        plane = dims[0] * dims[1] * dims[2]
        return self._sweep_off[k][axis] + (idx[0] * dims[1] + idx[1]) * dims[2] + idx[2], plane

    # -- state ---------------------------------------------------------------------------
    def set_state(self, density, velocity, pressure, *, vector_potential=(0.0, 0.0, 0.0)) -> None:
        """Set the state on every patch, then average the covered cells and restrict the
        potential (mhd3d.py:388-440)."""
        buf = patches3d_maps.pack_state(self.hierarchy, self.layouts, self.maps.total,
                                        self.gamma, self.mean_field,
                                        dict(density=density, velocity=velocity,
                                             pressure=pressure, vector_potential=vector_potential))
        self.load_buffer(buf)
        self._coarsen(self._buf[0], cells=True)

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
        return int(sum(np.prod(lay.n) for lay in self.layouts))

    def _check_ready(self) -> None:
        if not self._ready:
            raise RuntimeError("set the state first (set_state)")

    # -- the hierarchy's field interface --------------------------------------------------
    def _field_view(self, index: int, name: str) -> torch.Tensor:
        lay = self.layouts[index]
        buf = self._buf[0]
        if name in self.potential:
            f = ("ax", "ay", "az")[self.potential.index(name)]
            start = lay.base(f)
            return buf[start : start + int(np.prod(lay.extent(f)))].view(*lay.extent(f))
        comp = list(self.names.values()).index(name)
        plane = int(np.prod(lay.dims))
        start = lay.offset + comp * plane
        return buf[start : start + plane].view(*lay.dims)

    def _exchange_halos(self) -> None:
        self._fill(self._buf[0], _lib.stream_handle(self.device))

    def _restrict_fields(self) -> None:
        self._coarsen(self._buf[0], cells=True)

    def _integral(self, name: str) -> float:
        key = next(k for k, v in self.names.items() if v == name)
        return self.totals()[key]

    # -- the step ------------------------------------------------------------------------
    def _gather(self, dst, src, table, stream) -> None:
        d, st, cnt, s, w, n = table
        if n:
            _lib.call("wb_patches3d_gather", _lib.ptr(dst), _lib.ptr(src), _lib.ptr(d),
                      _lib.ptr(st), _lib.ptr(cnt), _lib.ptr(s), _lib.ptr(w), n, stream)

    def _fill(self, buf, stream) -> None:
        for L in range(self.nlev):
            self._gather(buf, buf, self._fill_t[L], stream)

    def _coarsen(self, buf, cells: bool, stream=None) -> None:
        stream = _lib.stream_handle(self.device) if stream is None else stream
        tables = self._coarsen_t if cells else self._restrict_t
        for L in range(self.nlev - 2, -1, -1):
            self._gather(buf, buf, tables[L], stream)

    def _patch_stage(self, src, dst, k, heun: bool, slot: int, stream: int) -> None:
        lay = self.layouts[k]
        nx, ny, nz = lay.n
        dx, dy, dz = self.hierarchy.spacing(lay.level)
        call, ptr = _lib.call, _lib.ptr
        base = 8 * lay.offset
        s_in, s_out = src.data_ptr() + base, dst.data_ptr() + base
        bad = self._bad.data_ptr() + 4 * (2 * k + slot)
        f = [ptr(x) for x in self._faces]
        call("wb_primitive3d", s_in, nx, ny, nz, dx, dy, dz, *self.mean_field, self.gamma,
             ptr(self._prim), *f, bad, stream)
        code = LIMITER_CODES.index(self.limiter)
        sw = [self._sweeps.data_ptr() + 8 * off for off in self._sweep_off[k]]
        for s in range(3):
            call("wb_sweep3d", ptr(self._prim), f[s], nx, ny, nz, s, code, self.gamma, sw[s],
                 stream)
        e = [ptr(x) for x in self._edges]
        for c in range(3):
            a, b = (c + 1) % 3, (c + 2) % 3
            call("wb_edge3d", ptr(self._prim), sw[a], sw[b], nx, ny, nz, c, e[c], stream)
        if self.resistivity:  # mhd3d.py:631-632
            call("wb_resistive3d", *f, nx, ny, nz, dx, dy, dz, self.resistivity, *sw, *e, stream)
        call("wb_mhd3d_update", s_in, s_out, *sw, *e, nx, ny, nz, ptr(self._dt), dx, dy, dz,
             int(heun), stream)

    def _stage(self, src, dst, heun: bool, slot: int, stream: int) -> None:
        self._fill(src, stream)
        for k in range(len(self.layouts)):
            self._patch_stage(src, dst, k, heun, slot, stream)
        if self.nreflux:  # stage 1 records into the register, stage 2 refluxes
            _lib.call("wb_patches3d_reflux", _lib.ptr(dst), _lib.ptr(self._sweeps),
                      _lib.ptr(self._rcells), _lib.ptr(self._rfaces), _lib.ptr(self._rspacing),
                      self.nreflux, _lib.ptr(self._dt), int(not heun), _lib.ptr(self._reg), stream)
        self._coarsen(dst, cells=heun, stream=stream)

    def stable_step(self) -> float:
        """cfl times the smallest of 1 / (largest summed fast-wave rate) over the patches'
        cells and, with a resistivity, the diffusion limit of the finest cells
        (mhd3d.py:674-724); synchronises."""
        self._check_ready()
        stream = _lib.stream_handle(self.device)
        buf = self._buf[0]
        for k, lay in enumerate(self.layouts):
            _lib.call("wb_rate3d", buf.data_ptr() + 8 * lay.offset, *lay.n,
                      *self.hierarchy.spacing(lay.level), *self.mean_field, self.gamma,
                      self._rates.data_ptr() + 16 * k, stream)
        rates = self._rates.view(-1, 2)[:, 0].cpu().numpy()
        smallest = np.inf
        for k, lay in enumerate(self.layouts):
            smallest = min(smallest, 1.0 / float(rates[k]))
            if self.resistivity:
                spacing = self.hierarchy.spacing(lay.level)
                smallest = min(smallest, 0.5 / (self.resistivity * sum(d**-2 for d in spacing)))
        return self.cfl * float(smallest)

    def step(self, dt: float | None = None) -> float:
        self._check_ready()
        if dt is None:
            dt = self.stable_step()
        dt = float(dt)
        stream = _lib.stream_handle(self.device)
        self._dt.fill_(dt)
        a, b = self._buf
        self._stage(a, b, False, 0, stream)
        self._stage(b, a, True, 1, stream)
        self._t.add_(dt)
        self.steps += 1
        if self.check_interval and self.steps % self.check_interval == 0:
            self.check()
        return dt

    def run(self, duration: float, *, max_steps: int = 1_000_000) -> int:
        """mhd3d.py:756-768: the last step shortened to land on the end."""
        self._check_ready()
        if not np.isfinite(duration) or duration < 0.0:
            raise ValueError("duration must be a finite, non-negative time")
        end = self.time + duration
        taken = 0
        while self.time < end * (1.0 - 1e-14) and taken < max_steps:
            self.step(min(self.stable_step(), end - self.time))
            taken += 1
        return taken

    def check(self) -> None:
        if int(self._bad.cpu().sum()):
            raise FloatingPointError(f"non-positive density or pressure at t = {self.time}")

    # -- refinement ----------------------------------------------------------------------
    def current_tags(self, threshold: float, levels: int) -> dict[int, np.ndarray]:
        """Tag the cells where |J| (the curl of the cell field by central differences)
        times the largest spacing exceeds ``threshold`` on levels 0 .. levels - 1 that
        exist (mhd3d.py:771-801)."""
        self._check_ready()
        stream = _lib.stream_handle(self.device)
        buf = self._buf[0]
        self._fill(buf, stream)
        tags = {L: np.zeros(self.hierarchy.level_shape(L), bool)
                for L in range(min(levels, self.nlev))}
        scratch = torch.zeros(1, dtype=torch.int32, device=self.device)
        f = [_lib.ptr(x) for x in self._faces]
        for lay in self.layouts:
            if lay.level not in tags:
                continue
            nx, ny, nz = lay.n
            spacing = self.hierarchy.spacing(lay.level)
            _lib.call("wb_primitive3d", buf.data_ptr() + 8 * lay.offset, nx, ny, nz, *spacing,
                      *self.mean_field, self.gamma, _lib.ptr(self._prim), *f, _lib.ptr(scratch),
                      stream)
            out = torch.empty(nx * ny * nz, dtype=torch.uint8, device=self.device)
            _lib.call("wb_current_tags3d", _lib.ptr(self._prim), nx, ny, nz, *spacing,
                      float(threshold), _lib.ptr(out), stream)
            cell = out.view(nx, ny, nz).cpu().numpy().astype(bool)
            tags[lay.level][tuple(slice(lo, hi) for lo, hi in lay.box)] |= cell
        return tags

    def regrid(self, tags: dict, **options) -> None:
        """Regrid on ``tags`` and move the state onto the new patches (mhd3d.py:803-821):
        same-level cells and edges copied, new cells the conservative minmod prolongation
        of their coarse cell (integrals kept), new edges interpolated (div B stays zero),
        then the covered cells averaged and the potential restricted.  ``options``:
        ``max_patch_size``, ``buffer``, ``efficiency`` of `wellbalanced_amd.regrid`."""
        self._check_ready()
        from . import regrid as rg

        old_h, old_maps, old = self.hierarchy, self.maps, self._buf[0]
        new_h = rg.regrid(old_h, tags, **options)
        maps = patches3d_maps.build(new_h)
        tables = patches3d_maps.transfer_tables(old_h, old_maps, new_h, maps)
        time = self._t.clone()
        self._use(new_h, maps)
        self._t.copy_(time)
        new = self._buf[0]
        stream = _lib.stream_handle(self.device)
        for L in range(self.nlev):
            t = tables[L]
            self._gather(new, old, self._stencils(t["copies"]), stream)
            rows = self._dev(t["prolong"].reshape(-1))
            if len(t["prolong"]):
                _lib.call("wb_patches3d_prolong", _lib.ptr(new), _lib.ptr(rows),
                          len(t["prolong"]), stream)
            self._gather(new, new, self._stencils(t["interp"]), stream)
            self._gather(new, new, self._fill_t[L], stream)
        self._coarsen(new, cells=True, stream=stream)
        torch.cuda.synchronize(self.device)

    # -- diagnostics ---------------------------------------------------------------------
    def _views(self, buf: np.ndarray, k: int):
        lay = self.layouts[k]
        plane = int(np.prod(lay.dims))
        cells = buf[lay.offset : lay.offset + NCELL * plane].reshape(NCELL, *lay.dims)
        pots = [buf[lay.base(f) : lay.base(f) + int(np.prod(lay.extent(f)))].reshape(lay.extent(f))
                for f in ("ax", "ay", "az")]
        return cells, pots

    def totals(self) -> dict[str, float]:
        """Integrals of the conserved fields over the leaf cells (mhd3d.py:903-907)."""
        self._check_ready()
        buf = self.buffer()
        tot = np.zeros(NCELL)
        for k, lay in enumerate(self.layouts):
            cells, _ = self._views(buf, k)
            inner = cells[:, HALO:-HALO, HALO:-HALO, HALO:-HALO]
            vol = float(np.prod(self.hierarchy.spacing(lay.level)))
            tot += (inner * ~self.maps.covered[k]).sum(axis=(1, 2, 3)) * vol
        return {name: float(v) for name, v in zip(CONSERVED_3D, tot)}

    def _local_faces(self, buf, k):
        from .mhd3d import face_fields

        lay = self.layouts[k]
        _, pots = self._views(buf, k)
        local = [p[tuple(slice(HALO, HALO + lay.n[a] + (a != c)) for a in range(3))]
                 for c, p in enumerate(pots)]
        return local, face_fields(local, self.hierarchy.spacing(lay.level), self.mean_field)

    def divergence(self) -> float:
        """The largest |div B| times the cell size over the cells of every patch
        (mhd3d.py:823-841)."""
        self._check_ready()
        buf = self.buffer()
        largest = 0.0
        for k, lay in enumerate(self.layouts):
            spacing = self.hierarchy.spacing(lay.level)
            _, faces = self._local_faces(buf, k)
            div = sum(np.diff(f, axis=a) / spacing[a] for a, f in enumerate(faces))
            largest = max(largest, float(np.abs(div).max()) * min(spacing))
        return largest

    def composite_divergence(self) -> float:
        """The largest mismatch of a coarse edge over a finer patch and the mean of the
        two fine edges along it, relative to the largest fine potential difference
        (mhd3d.py:843-907): zero after every step."""
        self._check_ready()
        buf = self.buffer()
        worst = 0.0
        for s in self.maps.restrict.values():
            for i in range(len(s)):
                t = s.start[i]
                want = s.w[t] * buf[s.src[t]] + s.w[t + 1] * buf[s.src[t + 1]]
                worst = max(worst, abs(buf[s.dst[i]] - want))
        scale = 0.0
        for k, lay in enumerate(self.layouts):
            if lay.level == 0:
                continue
            local, _ = self._local_faces(buf, k)
            for c, p in enumerate(local):
                for a in range(3):
                    if p.shape[a] > 1:
                        scale = max(scale, float(np.abs(np.diff(p, axis=a)).max()))
        return worst / scale if scale else worst

    def dataset(self, level: int = 0) -> dict[str, np.ndarray]:
        """The primitive variables of ``level`` over its cell centres, NaN where it has
        no patch (mhd3d.py:909-965; plain arrays instead of xarray)."""
        from .mhd3d import cell_fields

        if not 0 <= level < self.nlev:
            raise ValueError(f"the hierarchy has levels 0 .. {self.nlev - 1}")
        buf = self.buffer()
        shape = self.hierarchy.level_shape(level)
        names = ("density", "velocity_x", "velocity_y", "velocity_z", "pressure",
                 "magnetic_field_x", "magnetic_field_y", "magnetic_field_z", "magnetic_pressure")
        out = {name: np.full(shape, np.nan) for name in names}
        for k, lay in enumerate(self.layouts):
            if lay.level != level:
                continue
            cells, _ = self._views(buf, k)
            c = cells[:, HALO:-HALO, HALO:-HALO, HALO:-HALO]
            _, faces = self._local_faces(buf, k)
            bx, by, bz = cell_fields(faces)
            rho = c[0]
            u, v, w = c[1] / rho, c[2] / rho, c[3] / rho
            b2 = bx**2 + by**2 + bz**2
            p = (self.gamma - 1.0) * (c[4] - 0.5 * rho * (u**2 + v**2 + w**2) - 0.5 * b2)
            sl = tuple(slice(lo, hi) for lo, hi in lay.box)
            for name, value in zip(names, (rho, u, v, w, p, bx, by, bz, 0.5 * b2)):
                out[name][sl] = value
        out["time"] = self.time
        return out
