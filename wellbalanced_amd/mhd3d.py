"""``MHD3D``: the reference's 3D ideal-MHD model with its step on the GPU.

Same constructor, set-ups, ``set_state``, ``stable_step``, ``step``, ``run`` and
diagnostics as carafe ``mhd3d.py:163-965``.  The state of the patch is one FP64
buffer in HBM: 5 haloed cell planes (rho, m_x, m_y, m_z, E), then the edge vector
potential A_x, A_y, A_z (haloed).  Each Heun stage runs as HIP kernels through
``libwbmhd.so``; two buffers ping-pong as in `MHD2D`.

``kernels="split"`` runs one kernel per piece (halos, primitive, three sweeps,
three edge fields, update) in FP64: the parity path.  ``kernels="fused"`` runs the
whole stage in one LDS-tiled kernel, in ``precision="fp64"`` or ``"mixed"`` (FP32
reconstruction and HLLD, FP64 primitive conversion, CT, potential, update).
"""

from __future__ import annotations

from typing import Any

import numpy as np
import torch

from . import _lib, initial
from .hierarchy import AMRHierarchy, Patch
from .mhd import LIMITER_CODES, PRECISIONS, Value, _device

HALO = 2
CONSERVED_3D = ("density", "momentum_x", "momentum_y", "momentum_z", "energy")
KERNELS = ("split", "fused")


def face_fields(potential, spacing, b0):
    """``B_a = B0_a + d_b A_c - d_c A_b`` on the faces (mhd3d.py:361-378); NumPy or torch."""
    faces = []
    for a in range(3):
        b, c = (a + 1) % 3, (a + 2) % 3
        faces.append(_diff(potential[c], b) / spacing[b] - _diff(potential[b], c) / spacing[c] + b0[a])
    return faces


def cell_fields(faces):
    """The mean of the two faces of each cell along its axis (mhd3d.py:380-386)."""
    out = []
    for a, f in enumerate(faces):
        n = f.shape[a]
        out.append(0.5 * (_take(f, a, 0, n - 1) + _take(f, a, 1, n)))
    return out


def _take(x, axis, lo, hi):
    index = [slice(None)] * x.ndim
    index[axis] = slice(lo, hi)
    return x[tuple(index)]


def _diff(x, axis):
    n = x.shape[axis]
    return _take(x, axis, 1, n) - _take(x, axis, 0, n - 1)


class MHD3D:
    """3D ideal MHD on one uniform patch of an `AMRHierarchy` (see the module).  On a
    hierarchy with refined patches the constructor returns a `wellbalanced_amd.
    patches3d.RefinedMHD3D` with the same API (periodic domains, FP64)."""

    def __new__(cls, hierarchy: AMRHierarchy, **options):
        if cls is MHD3D and len(getattr(hierarchy, "patches", ())) > 1:
            from .patches3d import RefinedMHD3D

            if any(options.get("gravity", (0.0, 0.0, 0.0))):
                raise NotImplementedError("the gravity source is outside this MHD path")
            if options.get("precision", "fp64") != "fp64":
                raise NotImplementedError("refined 3D patches run the FP64 split kernels; the "
                                          "block hierarchy (wellbalanced_amd.amr3d) runs mixed")
            keep = ("gamma", "mean_field", "resistivity", "limiter", "cfl", "prefix", "device")
            return RefinedMHD3D(hierarchy, **{k: v for k, v in options.items() if k in keep})
        return super().__new__(cls)

    def __init__(self, hierarchy: AMRHierarchy, *, gamma: float = 5.0 / 3.0,
                 mean_field=(0.0, 0.0, 0.0), resistivity: float = 0.0, limiter: str = "vanleer",
                 cfl: float = 0.3, gravity=(0.0, 0.0, 0.0), prefix: str = "mhd_",
                 precision: str = "fp64", kernels: str = "split", device=None) -> None:
        if hierarchy.ndim != 3:
            raise ValueError("MHD3D needs a 3D hierarchy")
        if hierarchy.cylindrical or hierarchy.graded:
            raise NotImplementedError("MHD3D is Cartesian on uniform levels: no graded axes")
        if limiter not in LIMITER_CODES:
            raise ValueError(f"limiter is one of {sorted(LIMITER_CODES)}")
        if not 0.0 < cfl <= 1.0:
            raise ValueError("cfl must be in (0, 1]")
        if gamma <= 1.0:
            raise ValueError("gamma must be larger than 1")
        if not resistivity >= 0.0:
            raise ValueError("resistivity must be non-negative")
        if len(mean_field) != 3 or len(gravity) != 3:
            raise ValueError("mean_field and gravity have three components")
        if precision not in PRECISIONS:
            raise ValueError(f"precision is one of {PRECISIONS}")
        if kernels not in KERNELS:
            raise ValueError(f"kernels is one of {KERNELS}")
        if kernels == "split" and precision != "fp64":
            raise ValueError("the split kernels are the FP64 parity path; mixed needs kernels='fused'")
        if any(gravity):
            raise NotImplementedError("the gravity source is outside this MHD path")
        self.hierarchy = hierarchy
        self.gamma = float(gamma)
        self.mean_field = tuple(float(b) for b in mean_field)
        self.resistivity = float(resistivity)
        self._resistive_work = None
        self.limiter = limiter
        self.cfl = float(cfl)
        self.gravity = (0.0, 0.0, 0.0)
        self.periodic = tuple(hierarchy.periodic)
        self.names = {name: prefix + name for name in CONSERVED_3D}
        self.potential = tuple(prefix + n for n in ("ax", "ay", "az"))
        hierarchy._bind(self, {**{n: "cell" for n in self.names.values()},
                               **dict(zip(self.potential, ("x-edge", "y-edge", "z-edge")))},
                        HALO)
        self.precision = precision
        self.kernels = kernels
        self.device = _device(device)
        self.patch: Patch = hierarchy.patches[0]
        self.n = self.patch.shape
        self.spacing = hierarchy.spacing(0)
        nx, ny, nz = self.n
        self.dims = (nx + 4, ny + 4, nz + 4)
        X, Y, Z = self.dims
        self._cells = 5 * X * Y * Z
        self._pshape = [(X, Y + 1, Z + 1), (X + 1, Y, Z + 1), (X + 1, Y + 1, Z)]
        self._poff = [self._cells]
        for s in self._pshape[:2]:
            self._poff.append(self._poff[-1] + int(np.prod(s)))
        size = self._poff[-1] + int(np.prod(self._pshape[2]))
        opts = dict(dtype=torch.float64, device=self.device)
        self._buf = [torch.zeros(size, **opts), torch.zeros(size, **opts)]
        self._work: dict[str, torch.Tensor] = {}
        self._bad = torch.zeros(2, dtype=torch.int32, device=self.device)
        # the bad-cell count of kernels outside the step (tagging): never checked
        self._scratch_bad = torch.zeros(1, dtype=torch.int32, device=self.device)
        self._rate = torch.zeros(2, **opts)
        self._dt = torch.zeros(1, **opts)
        self._t = torch.zeros(1, **opts)
        self.check_interval = 1
        self.nxs = 32
        self.waves = 2
        self.steps = 0
        self.period = float("nan")
        self._ready = False

    # -- set-ups (mhd3d.py:236-351) ------------------------------------------------
    @classmethod
    def from_problem(cls, problem, **options: Any) -> MHD3D:
        """A model of a `wellbalanced_amd.initial.Problem` with its state set."""
        model = cls(problem.hierarchy, **{**problem.options, **options})
        model.set_state(**problem.state)
        model.period = problem.period
        return model

    @classmethod
    def alfven_wave(cls, cells: int, *, amplitude: float = 0.1, **options: Any):
        """The oblique circularly polarised Alfvén wave (mhd3d.py:236-286)."""
        model = cls.from_problem(initial.alfven_wave_3d(cells, amplitude=amplitude, **options))
        return model, model.dataset()

    @classmethod
    def orszag_tang(cls, cells: int, layers: int = 4, **options: Any) -> MHD3D:
        """The Orszag–Tang vortex, uniform along z (mhd3d.py:288-325)."""
        return cls.from_problem(initial.orszag_tang_3d(cells, layers, **options))

    @classmethod
    def turbulence(cls, cells: int = 256, **options: Any) -> MHD3D:
        """Seeded Alfvénic turbulence on the periodic unit cube (`initial.turbulence_3d`)."""
        seed = options.pop("seed", 0)
        return cls.from_problem(initial.turbulence_3d(cells, seed=seed, **options))

    @classmethod
    def blast_wave(cls, cells: int, **options: Any) -> MHD3D:
        """The magnetised 3D blast wave (mhd3d.py:327-351)."""
        hierarchy = AMRHierarchy(
            (cells, cells, (3 * cells) // 2), ((-0.5, 0.5), (-0.5, 0.5), (-0.75, 0.75)),
            periodic=(True, True, True),
        )
        field = 1.0 / np.sqrt(3.0)
        options = {"limiter": "minmod", **options}
        model = cls(hierarchy, mean_field=(field, field, field), **options)
        model.set_state(1.0, (0.0, 0.0, 0.0),
                        lambda x, y, z: np.where(x**2 + y**2 + z**2 < 0.01, 10.0, 0.1))
        return model

    # -- state ------------------------------------------------------------------------
    def _views(self, buf: torch.Tensor):
        X, Y, Z = self.dims
        cells = buf[: self._cells].view(5, X, Y, Z)
        pots = [buf[o : o + int(np.prod(s))].view(*s) for o, s in zip(self._poff, self._pshape)]
        return cells, pots

    @property
    def state(self) -> torch.Tensor:
        return self._buf[0]

    def set_state(self, density: Value, velocity, pressure: Value, *,
                  vector_potential=(0.0, 0.0, 0.0)) -> None:
        """Set the state on the patch (mhd3d.py:388-440)."""
        cons, pots = initial.host_state_3d(
            self.hierarchy, self.patch, self.gamma, self.mean_field, density, velocity, pressure,
            vector_potential=vector_potential,
        )
        self.load_arrays(cons, pots)

    def load_arrays(self, cons: np.ndarray, pots) -> None:
        """Upload haloed host arrays ``cons (5, X, Y, Z)`` and the three potentials."""
        flat = np.concatenate([np.asarray(cons, np.float64).ravel()] +
                              [np.asarray(p, np.float64).ravel() for p in pots])
        self._buf[0].copy_(torch.from_numpy(flat))
        self.time = 0.0
        self.steps = 0
        self._ready = True

    @property
    def time(self) -> float:
        """The model time (kept on the device; reading it synchronises)."""
        return float(self._t.item())

    @time.setter
    def time(self, value: float) -> None:
        self._t.fill_(float(value))

    def arrays(self):
        cells, pots = self._views(self._buf[0])
        return cells.cpu().numpy(), [p.cpu().numpy() for p in pots]

    def _check_ready(self) -> None:
        if not self._ready:
            raise RuntimeError("set the state first (set_state)")

    # -- the step -------------------------------------------------------------------------
    def _workspace(self) -> dict[str, torch.Tensor]:
        if self._work:
            return self._work
        X, Y, Z = self.dims
        opts = dict(dtype=torch.float64, device=self.device)
        nx, ny, nz = self.n
        w = {
            "prim": torch.empty(8 * X * Y * Z, **opts),
            "f0": torch.empty((X + 1) * Y * Z, **opts),
            "f1": torch.empty(X * (Y + 1) * Z, **opts),
            "f2": torch.empty(X * Y * (Z + 1), **opts),
            "s0": torch.empty(7 * (X - 3) * Y * Z, **opts),
            "s1": torch.empty(7 * X * (Y - 3) * Z, **opts),
            "s2": torch.empty(7 * X * Y * (Z - 3), **opts),
            "e0": torch.empty(nx * (ny + 1) * (nz + 1), **opts),
            "e1": torch.empty((nx + 1) * ny * (nz + 1), **opts),
            "e2": torch.empty((nx + 1) * (ny + 1) * nz, **opts),
        }
        self._work = w
        return w

    def _stage_split(self, src, dst, heun, slot, stream):
        nx, ny, nz = self.n
        call, ptr = _lib.call, _lib.ptr
        w = self._workspace()
        pmask = sum(int(p) << a for a, p in enumerate(self.periodic))
        call("wb_mhd3d_fill_halos", ptr(src), nx, ny, nz, pmask, stream)
        bad = self._bad.data_ptr() + 4 * slot
        call("wb_primitive3d", ptr(src), nx, ny, nz, *self.spacing, *self.mean_field, self.gamma,
             ptr(w["prim"]), ptr(w["f0"]), ptr(w["f1"]), ptr(w["f2"]), bad, stream)
        code = LIMITER_CODES.index(self.limiter)
        for s in range(3):
            call("wb_sweep3d", ptr(w["prim"]), ptr(w[f"f{s}"]), nx, ny, nz, s, code, self.gamma,
                 ptr(w[f"s{s}"]), stream)
        for c in range(3):
            a, b = (c + 1) % 3, (c + 2) % 3
            call("wb_edge3d", ptr(w["prim"]), ptr(w[f"s{a}"]), ptr(w[f"s{b}"]), nx, ny, nz, c,
                 ptr(w[f"e{c}"]), stream)
        if self.resistivity:  # into the fluxes and edges before the update (mhd3d.py:631-632)
            call("wb_resistive3d", ptr(w["f0"]), ptr(w["f1"]), ptr(w["f2"]), nx, ny, nz,
                 *self.spacing, self.resistivity, ptr(w["s0"]), ptr(w["s1"]), ptr(w["s2"]),
                 ptr(w["e0"]), ptr(w["e1"]), ptr(w["e2"]), stream)
        call("wb_mhd3d_update", ptr(src), ptr(dst), ptr(w["s0"]), ptr(w["s1"]), ptr(w["s2"]),
             ptr(w["e0"]), ptr(w["e1"]), ptr(w["e2"]), nx, ny, nz, ptr(self._dt), *self.spacing,
             int(heun), stream)

    def _resistive(self, src, dst, heun, stream) -> None:
        """The resistive terms (mhd3d.py:561-602) on top of the fused stage just run (the
        split stage adds them into its fluxes and edges before the update)."""
        if not self.resistivity:
            return
        import torch

        nx, ny, nz = self.n
        if self._resistive_work is None:
            size = _lib.call("wb_mhd3d_resistive_workspace", nx, ny, nz)
            self._resistive_work = torch.empty(size, dtype=torch.float64, device=self.device)
        _lib.call("wb_mhd3d_resistive", _lib.ptr(src), _lib.ptr(dst), nx, ny, nz, *self.spacing,
                  *self.mean_field, self.resistivity, _lib.ptr(self._dt), int(heun),
                  _lib.ptr(self._resistive_work), stream)

    def _stage_fused(self, src, dst, heun, slot, stream, marks=None):
        nx, ny, nz = self.n
        pmask = sum(int(p) << a for a, p in enumerate(self.periodic))
        # the stage reads the periodic images itself: only the outflow axes are filled
        # (all axes with resistivity, whose terms read the halo of the stage input)
        if self.resistivity:
            _lib.call("wb_mhd3d_fill_halos", _lib.ptr(src), nx, ny, nz, pmask, stream)
        elif pmask != 7:
            _lib.call("wb_mhd3d_fill_outflow", _lib.ptr(src), nx, ny, nz, pmask, stream)
        if marks is not None:
            marks[0].record()
        args = (_lib.ptr(src), _lib.ptr(dst), nx, ny, nz, *self.spacing, *self.mean_field,
                self.gamma, LIMITER_CODES.index(self.limiter), _lib.ptr(self._dt), int(heun),
                self._bad.data_ptr() + 4 * slot, int(self.nxs))
        if self.precision == "fp64":
            _lib.call("wb_mhd3d_stage", *args, pmask, stream)
        else:
            _lib.call("wb_mhd3d_stage_mixed", *args, int(self.waves), pmask, stream)
        self._resistive(src, dst, heun, stream)
        if marks is not None:
            marks[1].record()

    def _rate_kernel(self, stream: int) -> None:
        nx, ny, nz = self.n
        name = "wb_rate3d" if self.precision == "fp64" else "wb_rate3d_mixed"
        _lib.call(name, _lib.ptr(self._buf[0]), nx, ny, nz, *self.spacing,
                  *self.mean_field, self.gamma, _lib.ptr(self._rate), stream)

    def _dt_kernel(self, time_ptr, stream: int) -> None:
        if self.resistivity:  # mhd3d.py:717-721
            _lib.call("wb_stable_dt_diffusive", _lib.ptr(self._rate), self.cfl, 0.0,
                      self.resistivity, 0.0, sum(d**-2 for d in self.spacing),
                      _lib.ptr(self._dt), time_ptr, stream)
        else:
            _lib.call("wb_stable_dt", _lib.ptr(self._rate), self.cfl, 0.0, _lib.ptr(self._dt),
                      time_ptr, stream)

    def stable_step(self) -> float:
        """cfl times the smallest of 1 / (largest summed fast-wave rate) and, with a
        resistivity, the diffusion limit (mhd3d.py:674-724); synchronises."""
        self._check_ready()
        stream = _lib.stream_handle(self.device)
        self._rate_kernel(stream)
        self._dt_kernel(None, stream)
        return float(self._dt[0].item())

    def _heun(self, stream: int, events=None) -> None:
        import torch

        a, b = self._buf
        stage = self._stage_split if self.kernels == "split" else self._stage_fused
        if events is None:
            stage(a, b, False, 0, stream)
            stage(b, a, True, 1, stream)
            return
        marks = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        if self.kernels == "fused":
            # the events bracket the stage kernel alone (the halo fill runs before)
            self._stage_fused(a, b, False, 0, stream, marks[0:2])
            self._stage_fused(b, a, True, 1, stream, marks[2:4])
        else:
            marks[0].record()
            stage(a, b, False, 0, stream)
            marks[1].record()
            marks[2].record()
            stage(b, a, True, 1, stream)
            marks[3].record()
        events.extend([(marks[0], marks[1]), (marks[2], marks[3])])

    def step(self, dt: float | None = None) -> float:
        """Advance by one Heun (SSP-RK2) step (mhd3d.py:726-754); return the step taken."""
        self._check_ready()
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
        """``steps`` steps at the stable step with dt and the time kept on the device (no
        host synchronisation; `check` tests for bad cells).  With ``events`` returns the
        HIP event pairs bracketing each stage launch (halo fill included)."""
        self._check_ready()
        stream = _lib.stream_handle(self.device)
        marks = [] if events else None
        for _ in range(steps):
            self._rate_kernel(stream)
            self._dt_kernel(_lib.ptr(self._t), stream)
            self._heun(stream, marks)
            self.steps += 1
        return marks

    def current_tags(self, threshold: float, levels: int = 1) -> dict[int, np.ndarray]:
        """Tag the cells where |J| (the cell field's curl by central differences) times
        the cell size exceeds ``threshold`` (``mhd3d.py:771-801``).  The uniform grid
        is level 0."""
        self._check_ready()
        if levels < 1:
            return {}
        stream = _lib.stream_handle(self.device)
        nx, ny, nz = self.n
        w = self._workspace()
        src = self._buf[0]
        pmask = sum(int(p) << a for a, p in enumerate(self.periodic))
        _lib.call("wb_mhd3d_fill_halos", _lib.ptr(src), nx, ny, nz, pmask, stream)
        _lib.call("wb_primitive3d", _lib.ptr(src), nx, ny, nz, *self.spacing, *self.mean_field,
                  self.gamma, _lib.ptr(w["prim"]), _lib.ptr(w["f0"]), _lib.ptr(w["f1"]),
                  _lib.ptr(w["f2"]), _lib.ptr(self._scratch_bad), stream)
        out = torch.empty(nx * ny * nz, dtype=torch.uint8, device=self.device)
        _lib.call("wb_current_tags3d", _lib.ptr(w["prim"]), nx, ny, nz, *self.spacing,
                  float(threshold), _lib.ptr(out), stream)
        return {0: out.view(nx, ny, nz).cpu().numpy().astype(bool)}

    def check(self) -> None:
        bad = self._bad.cpu()
        if int(bad.sum()):
            raise FloatingPointError(
                f"non-positive density or pressure on the level-0 patch {self.patch.box} "
                f"at t = {self.time}"
            )

    def run(self, duration: float, *, max_steps: int = 1_000_000) -> int:
        """mhd3d.py:756-768."""
        self._check_ready()
        if not np.isfinite(duration) or duration < 0.0:
            raise ValueError("duration must be a finite, non-negative time")
        end = self.time + duration
        taken = 0
        while self.time < end * (1.0 - 1e-14) and taken < max_steps:
            dt = min(self.stable_step(), end - self.time)
            self.step(dt)
            taken += 1
        return taken

    # -- diagnostics -----------------------------------------------------------------------
    def _local_potentials(self, buf=None):
        _, pots = self._views(self._buf[0] if buf is None else buf)
        nx, ny, nz = self.n
        out = []
        for c, p in enumerate(pots):
            out.append(p[tuple(slice(HALO, HALO + self.n[k] + (k != c)) for k in range(3))])
        return out

    # -- the hierarchy's field interface (patch.field(name), gather, integral, ...) -------
    def _field_view(self, index: int, name: str) -> torch.Tensor:
        """The haloed device view of field ``name`` in the current state."""
        cells, pots = self._views(self._buf[0])
        if name in self.potential:
            return pots[self.potential.index(name)]
        return cells[list(self.names.values()).index(name)]

    def _exchange_halos(self) -> None:
        nx, ny, nz = self.n
        pmask = sum(int(p) << a for a, p in enumerate(self.periodic))
        _lib.call("wb_mhd3d_fill_halos", _lib.ptr(self._buf[0]), nx, ny, nz, pmask,
                  _lib.stream_handle(self.device))

    def _restrict_fields(self) -> None:
        """One level: nothing to restrict."""

    def _integral(self, name: str) -> float:
        key = next(k for k, v in self.names.items() if v == name)
        return self._totals()[key]

    def totals(self) -> dict[str, float]:
        self._check_ready()
        return self._totals()

    def _totals(self) -> dict[str, float]:
        cells, _ = self._views(self._buf[0])
        nx, ny, nz = self.n
        inner = cells[:, HALO : HALO + nx, HALO : HALO + ny, HALO : HALO + nz]
        vol = self.spacing[0] * self.spacing[1] * self.spacing[2]
        sums = inner.sum(dim=(1, 2, 3)) * vol
        return {name: float(v) for name, v in zip(CONSERVED_3D, sums.cpu().tolist())}

    def divergence(self) -> float:
        """mhd3d.py:823-841."""
        self._check_ready()
        faces = face_fields(self._local_potentials(), self.spacing, self.mean_field)
        div = sum(_diff(f, a) / self.spacing[a] for a, f in enumerate(faces))
        return float(div.abs().max()) * min(self.spacing)

    def dataset(self, level: int = 0) -> dict[str, np.ndarray]:
        """Primitive variables over the cell centres (mhd3d.py:909-965), as plain arrays."""
        self._check_ready()
        if level != 0:
            raise ValueError("a uniform model has level 0 only")
        cells, _ = self._views(self._buf[0])
        nx, ny, nz = self.n
        c = cells[:, HALO : HALO + nx, HALO : HALO + ny, HALO : HALO + nz].cpu().numpy()
        local = [p.cpu().numpy() for p in self._local_potentials()]
        bx, by, bz = cell_fields(face_fields(local, self.spacing, self.mean_field))
        rho = c[0]
        u, v, w = c[1] / rho, c[2] / rho, c[3] / rho
        b2 = bx**2 + by**2 + bz**2
        p = (self.gamma - 1.0) * (c[4] - 0.5 * rho * (u**2 + v**2 + w**2) - 0.5 * b2)
        return {
            "density": rho, "velocity_x": u, "velocity_y": v, "velocity_z": w, "pressure": p,
            "magnetic_field_x": bx, "magnetic_field_y": by, "magnetic_field_z": bz,
            "magnetic_pressure": 0.5 * b2, "time": self.time,
        }
