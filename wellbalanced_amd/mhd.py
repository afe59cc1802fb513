"""``MHD2D``: the reference's 2.5D ideal-MHD model with its step on the GPU.

Same constructor, problem set-ups, ``set_state``, ``stable_step``, ``step``,
``run`` and diagnostics as carafe ``mhd.py:386-1226``.  The state of the patch
lives in HBM as one FP64 buffer (6 haloed cell planes, then the haloed nodal
``A_z``); each stage of Heun's method runs as HIP kernels through
``libwbmhd.so`` (``include/wbmhd.h``).  Two state buffers ping-pong: stage 1 reads
``U^n`` and writes ``U^1``, stage 2 reads ``U^1`` and writes the Heun average over
``U^n`` in place, so no start copy is made (``mhd.py:972-988``).

``precision="fp64"`` is the parity path; ``"mixed"`` computes the reconstruction
and the HLLD flux in FP32 and keeps the primitive conversion, CT, face field,
update and sums in FP64.
"""

from __future__ import annotations

from collections.abc import Callable
from typing import Any

import numpy as np
import torch

from . import _lib, initial
from .hierarchy import AMRHierarchy, Patch

HALO = 2
CONSERVED = ("density", "momentum_x", "momentum_y", "momentum_z", "energy", "bz")
BOUNDARIES = ("outflow", "wall")
LIMITER_CODES = ("minmod", "vanleer", "mc", "none")
PRECISIONS = ("fp64", "mixed")
KERNELS = ("split", "fused")

Value = float | np.ndarray | Callable[..., Any]


def _device(device) -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("wellbalanced_amd needs a GPU (HIP device); there is no CPU fallback")
    dev = torch.device("cuda") if device is None else torch.device(device)
    if dev.type != "cuda":
        raise ValueError("the state lives on a HIP device ('cuda' in PyTorch-ROCm)")
    if dev.index is None:
        dev = torch.device("cuda", torch.cuda.current_device())
    return dev


class MHD2D:
    """2.5D MHD on one uniform patch of an `AMRHierarchy` (see the module).  On a
    hierarchy with refined patches the constructor returns a `wellbalanced_amd.amr.
    AMRMHD2D` with the same API (no gravity)."""

    def __new__(cls, hierarchy: AMRHierarchy, **options):
        if cls is MHD2D and len(getattr(hierarchy, "patches", ())) > 1:
            from .amr import AMRMHD2D

            if any(options.get("gravity", (0.0, 0.0))):
                raise NotImplementedError("the gravity source is outside this MHD path")
            keep = ("gamma", "mean_field", "limiter", "cfl", "resistivity", "conduction",
                    "boundaries", "prefix", "precision", "device")
            return AMRMHD2D(hierarchy, **{"limiter": "vanleer",
                                          **{k: v for k, v in options.items() if k in keep}})
        return super().__new__(cls)

    def __init__(self, hierarchy: AMRHierarchy, *, gamma: float = 5.0 / 3.0,
                 mean_field: tuple[float, float] = (0.0, 0.0), resistivity: float = 0.0,
                 limiter: str = "vanleer", cfl: float = 0.4,
                 conduction: tuple[float, float] = (0.0, 0.0),
                 gravity: tuple[float, float] = (0.0, 0.0),
                 boundaries: dict[str, str] | None = None, prefix: str = "mhd_",
                 precision: str = "fp64", kernels: str = "split", device=None) -> None:
        # validation as mhd.py:425-455
        if hierarchy.ndim != 2:
            raise ValueError("MHD2D needs a 2D hierarchy")
        if hierarchy.cylindrical or hierarchy.graded:
            raise NotImplementedError(
                "MHD2D is Cartesian on uniform levels: no r–z metric, no graded axes"
            )
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
        if kernels not in KERNELS:
            raise ValueError(f"kernels is one of {KERNELS}")
        if kernels == "split" and precision != "fp64":
            raise ValueError("the split kernels are the FP64 parity path; mixed needs kernels='fused'")
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
        if any(gravity):
            raise NotImplementedError("the gravity source is outside this MHD path")
        self.hierarchy = hierarchy
        self.gamma = float(gamma)
        self.mean_field = (float(mean_field[0]), float(mean_field[1]))
        self.resistivity = float(resistivity)
        self.limiter = limiter
        self.cfl = float(cfl)
        self.conduction = (float(conduction[0]), float(conduction[1]))
        self.gravity = (0.0, 0.0)
        self.walls = tuple(kinds[name] == "wall" for name in "xy")
        self.periodic = tuple(hierarchy.periodic)
        self.names = {name: prefix + name for name in CONSERVED}
        self.potential = prefix + "az"
        hierarchy._bind(self, {**{n: "cell" for n in self.names.values()},
                               self.potential: "node"}, HALO)
        self.precision = precision
        self.kernels = kernels
        self.nxs = 32
        self.tile = 126
        # the fused stage's kernel: "tiles" (a workgroup per 16 x 16 tile) or "march"
        # (strips of columns marched along x, the launch shape `tile` x `nxs`); the
        # results are the same
        self.kernel = "tiles"
        self.device = _device(device)
        self.patch: Patch = hierarchy.patches[0]
        self.nx, self.ny = self.patch.shape
        self.dx, self.dy = hierarchy.spacing(0)
        X, Y = self.nx + 2 * HALO, self.ny + 2 * HALO
        self.X, self.Y = X, Y
        self._cells = 6 * X * Y
        size = self._cells + (X + 1) * (Y + 1)
        opts = dict(dtype=torch.float64, device=self.device)
        self._buf = [torch.zeros(size, **opts), torch.zeros(size, **opts)]
        self._prim = torch.empty(8 * X * Y, **opts)
        self._bxf = torch.empty((X + 1) * Y, **opts)
        self._byf = torch.empty(X * (Y + 1), **opts)
        self._ecc = torch.empty(X * Y, **opts)
        self._fx = torch.empty(7 * (X - 3) * Y, **opts)
        self._fy = torch.empty(7 * X * (Y - 3), **opts)
        self._ez = torch.empty((self.nx + 1) * (self.ny + 1), **opts)
        self._bad = torch.zeros(2, dtype=torch.int32, device=self.device)
        # the bad-cell count of kernels outside the step (tagging): never checked
        self._scratch_bad = torch.zeros(1, dtype=torch.int32, device=self.device)
        self._rate = torch.zeros(2, **opts)
        self._dt = torch.zeros(1, **opts)
        self._t = torch.zeros(1, **opts)
        self.check_interval = 1
        self.steps = 0
        self.period = float("nan")
        self._ready = False

    # -- construction of common cases (mhd.py:486-665) --------------------------
    @classmethod
    def from_problem(cls, problem, **options: Any) -> MHD2D:
        """A model of a `wellbalanced_amd.initial.Problem` with its state set."""
        model = cls(problem.hierarchy, **{**problem.options, **options})
        model.set_state(**problem.state)
        model.period = problem.period
        return model

    @classmethod
    def alfven_wave(cls, cells: int, *, amplitude: float = 0.1, **options: Any):
        """The oblique circularly polarised Alfvén wave (mhd.py:486-530) and its exact state."""
        model = cls.from_problem(initial.alfven_wave(cells, amplitude=amplitude, **options))
        return model, model.dataset()

    @classmethod
    def orszag_tang(cls, cells: int, **options: Any) -> MHD2D:
        """The Orszag–Tang vortex on the periodic unit square (mhd.py:532-559)."""
        return cls.from_problem(initial.orszag_tang(cells, **options))

    @classmethod
    def rotor(cls, cells: int, **options: Any) -> MHD2D:
        """The MHD rotor (mhd.py:561-592)."""
        return cls.from_problem(initial.rotor(cells, **options))

    @classmethod
    def blast_wave(cls, cells: int, **options: Any) -> MHD2D:
        """The magnetised blast wave (mhd.py:594-620)."""
        return cls.from_problem(initial.blast_wave(cells, **options))

    # -- state ---------------------------------------------------------------------
    def _views(self, buf: torch.Tensor):
        cells = buf[: self._cells].view(6, self.X, self.Y)
        az = buf[self._cells :].view(self.X + 1, self.Y + 1)
        return cells, az

    @property
    def state(self) -> torch.Tensor:
        """The current FP64 state buffer (6 haloed cell planes then haloed A_z)."""
        return self._buf[0]

    def set_state(self, density: Value, velocity, pressure: Value, *,
                  vector_potential: Value = 0.0, bz: Value = 0.0) -> None:
        """Set the state (mhd.py:683-737): functions of the cell centres / nodes or values."""
        cons, az = initial.host_state_2d(
            self.hierarchy, self.patch, self.gamma, self.mean_field, density, velocity, pressure,
            vector_potential=vector_potential, bz=bz,
        )
        self.load_arrays(cons, az)

    def load_arrays(self, cons: np.ndarray, az: np.ndarray) -> None:
        """Upload haloed host arrays ``cons (6, X, Y)`` and ``az (X+1, Y+1)`` as the state."""
        flat = np.concatenate([np.asarray(cons, dtype=np.float64).ravel(),
                               np.asarray(az, dtype=np.float64).ravel()])
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

    def arrays(self) -> tuple[np.ndarray, np.ndarray]:
        """The haloed state on the host: ``cons (6, X, Y)``, ``az (X+1, Y+1)``."""
        cells, az = self._views(self._buf[0])
        return cells.cpu().numpy(), az.cpu().numpy()

    def _check_ready(self) -> None:
        if not self._ready:
            raise RuntimeError("set the state first (set_state)")

    # -- the hierarchy's field interface (patch.field(name), gather, integral, ...) ------
    def _field_view(self, index: int, name: str) -> torch.Tensor:
        """The haloed device view of field ``name`` in the current state."""
        cells, az = self._views(self._buf[0])
        if name == self.potential:
            return az
        return cells[list(self.names.values()).index(name)]

    def _exchange_halos(self) -> None:
        pmask = int(self.periodic[0]) | (int(self.periodic[1]) << 1)
        wmask = int(self.walls[0]) | (int(self.walls[1]) << 1)
        _lib.call("wb_mhd2d_fill_halos", _lib.ptr(self._buf[0]), self.nx, self.ny, pmask, wmask,
                  _lib.stream_handle(self.device))

    def _restrict_fields(self) -> None:
        """One level: nothing to restrict."""

    def _integral(self, name: str) -> float:
        key = next(k for k, v in self.names.items() if v == name)
        return self._totals()[key]

    # -- the step --------------------------------------------------------------------
    def _wall_mask(self) -> int:
        mask = 0
        if self.walls[0]:
            mask |= 3
        if self.walls[1]:
            mask |= 12
        return mask

    def _stage(self, src: torch.Tensor, dst: torch.Tensor, heun: bool, slot: int,
               stream: int) -> None:
        X, Y, nx, ny = self.X, self.Y, self.nx, self.ny
        call, ptr = _lib.call, _lib.ptr
        pmask = int(self.periodic[0]) | (int(self.periodic[1]) << 1)
        wmask = int(self.walls[0]) | (int(self.walls[1]) << 1)
        call("wb_mhd2d_fill_halos", ptr(src), nx, ny, pmask, wmask, stream)
        base = src.data_ptr()
        plane = X * Y * 8
        planes = [base + k * plane for k in range(6)]
        az = base + self._cells * 8
        bad = self._bad.data_ptr() + 4 * slot
        call("wb_primitive2d", *planes, az, X, Y, self.dx, self.dy, *self.mean_field,
             self.gamma, ptr(self._prim), ptr(self._bxf), ptr(self._byf), ptr(self._ecc),
             bad, stream)
        code = LIMITER_CODES.index(self.limiter)
        call("wb_sweep2d", ptr(self._prim), ptr(self._bxf), X, Y, 0, code, self.gamma,
             ptr(self._fx), stream)
        call("wb_sweep2d", ptr(self._prim), ptr(self._byf), X, Y, 1, code, self.gamma,
             ptr(self._fy), stream)
        fxp, fyp = self._fx.data_ptr(), self._fy.data_ptr()
        sx, sy = (X - 3) * Y * 8, X * (Y - 3) * 8
        call("wb_ct_contact", fxp + 5 * sx, fyp + 5 * sy, fxp, fyp, ptr(self._ecc), X, Y,
             ptr(self._ez), stream)
        # mhd.py:868-871: the heat flux, then the resistive terms, before the walls
        if max(self.conduction) > 0.0:
            call("wb_conductive2d", ptr(self._prim), ptr(self._bxf), ptr(self._byf), X, Y,
                 self.dx, self.dy, *self.conduction, fxp, fyp, stream)
        if self.resistivity:
            call("wb_resistive2d", ptr(self._prim), ptr(self._bxf), ptr(self._byf), X, Y,
                 self.dx, self.dy, self.resistivity, fxp, fyp, ptr(self._ez), stream)
        if self._wall_mask():
            call("wb_mhd2d_walls", fxp, fyp, ptr(self._ez), nx, ny, self._wall_mask(), stream)
        call("wb_mhd2d_update", ptr(src), ptr(dst), fxp, fyp, ptr(self._ez), nx, ny,
             ptr(self._dt), self.dx, self.dy, int(heun), stream)

    def _rate_kernel(self, stream: int) -> None:
        base = self._buf[0].data_ptr()
        plane = self.X * self.Y * 8
        _lib.call("wb_rate2d", *[base + k * plane for k in range(6)], base + self._cells * 8,
                  self.X, self.Y, HALO, self.dx, self.dy, *self.mean_field, self.gamma,
                  _lib.ptr(self._rate), stream)

    @property
    def _diffusive(self) -> bool:
        return bool(self.resistivity) or max(self.conduction) > 0.0

    def _dt_kernel(self, remaining: float, time_ptr, stream: int) -> None:
        """dt on the device from the rate: cfl / (largest fast-wave rate), and with a
        diffusivity the explicit diffusion limit (mhd.py:931-963)."""
        if self._diffusive:
            _lib.call("wb_stable_dt_diffusive", _lib.ptr(self._rate), self.cfl, remaining,
                      self.resistivity, (self.gamma - 1.0) * max(self.conduction),
                      self.dx**-2 + self.dy**-2, _lib.ptr(self._dt), time_ptr, stream)
        else:
            _lib.call("wb_stable_dt", _lib.ptr(self._rate), self.cfl, remaining,
                      _lib.ptr(self._dt), time_ptr, stream)

    def stable_step(self) -> float:
        """cfl times the smallest of the fast-wave crossing time of a cell and, with
        resistivity or conduction, the explicit diffusion limit (mhd.py:930-963);
        synchronises."""
        self._check_ready()
        stream = _lib.stream_handle(self.device)
        self._rate_kernel(stream)
        self._dt_kernel(0.0, None, stream)
        return float(self._dt[0].item())

    def _stage_fused(self, src: torch.Tensor, dst: torch.Tensor, heun: bool, slot: int,
                     stream: int) -> None:
        pmask = int(self.periodic[0]) | (int(self.periodic[1]) << 1)
        wmask = int(self.walls[0]) | (int(self.walls[1]) << 1)
        _lib.call("wb_mhd2d_fill_halos", _lib.ptr(src), self.nx, self.ny, pmask, wmask, stream)
        suffix = "" if self.precision == "fp64" else "_mixed"
        if self.kernel == "tiles":
            _lib.call("wb_mhd2d_stage_tiles" + suffix, _lib.ptr(src), _lib.ptr(dst), self.nx,
                      self.ny, self.dx, self.dy, *self.mean_field, self.gamma, self.resistivity,
                      *self.conduction, LIMITER_CODES.index(self.limiter), _lib.ptr(self._dt),
                      int(heun), wmask, self._bad.data_ptr() + 4 * slot, stream)
            return
        name = "wb_mhd2d_stage" + suffix
        _lib.call(name, _lib.ptr(src), _lib.ptr(dst), self.nx, self.ny, self.dx, self.dy,
                  *self.mean_field, self.gamma, self.resistivity, *self.conduction,
                  LIMITER_CODES.index(self.limiter),
                  _lib.ptr(self._dt), int(heun), wmask, self._bad.data_ptr() + 4 * slot,
                  int(self.nxs), int(self.tile), stream)

    def _heun(self, stream: int, events=None) -> None:
        a, b = self._buf
        stage = self._stage if self.kernels == "split" else self._stage_fused
        if events is None:
            stage(a, b, False, 0, stream)
            stage(b, a, True, 1, stream)
            return
        marks = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        marks[0].record()
        stage(a, b, False, 0, stream)
        marks[1].record()
        marks[2].record()
        stage(b, a, True, 1, stream)
        marks[3].record()
        events.extend([(marks[0], marks[1]), (marks[2], marks[3])])

    def step(self, dt: float | None = None) -> float:
        """Advance by one Heun (SSP-RK2) step (mhd.py:965-994); return the step taken."""
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
        """``steps`` steps at the stable step, with dt and the time kept on the device:
        no host synchronisation (call `check` to test for bad cells).  With ``events``
        returns the HIP event pairs bracketing each stage launch."""
        self._check_ready()
        stream = _lib.stream_handle(self.device)
        marks = [] if events else None
        for _ in range(steps):
            self._rate_kernel(stream)
            self._dt_kernel(0.0, _lib.ptr(self._t), stream)
            self._heun(stream, marks)
            self.steps += 1
        return marks

    def current_tags(self, threshold: float, levels: int = 1) -> dict[int, np.ndarray]:
        """Tag the cells where |J_z| times the cell size exceeds ``threshold``
        (``mhd.py:1011-1041``): the node current from the face fields, averaged over a
        cell's four nodes.  The uniform grid is level 0."""
        self._check_ready()
        if levels < 1:
            return {}
        stream = _lib.stream_handle(self.device)
        X, Y, nx, ny = self.X, self.Y, self.nx, self.ny
        src = self._buf[0]
        pmask = int(self.periodic[0]) | (int(self.periodic[1]) << 1)
        wmask = int(self.walls[0]) | (int(self.walls[1]) << 1)
        _lib.call("wb_mhd2d_fill_halos", _lib.ptr(src), nx, ny, pmask, wmask, stream)
        base, plane = src.data_ptr(), X * Y * 8
        _lib.call("wb_primitive2d", *[base + k * plane for k in range(6)], base + self._cells * 8,
                  X, Y, self.dx, self.dy, *self.mean_field, self.gamma, _lib.ptr(self._prim),
                  _lib.ptr(self._bxf), _lib.ptr(self._byf), _lib.ptr(self._ecc),
                  _lib.ptr(self._scratch_bad), stream)
        out = torch.empty(nx * ny, dtype=torch.uint8, device=self.device)
        _lib.call("wb_current_tags2d", _lib.ptr(self._bxf), _lib.ptr(self._byf), nx, ny, self.dx,
                  self.dy, float(threshold), _lib.ptr(out), stream)
        return {0: out.view(nx, ny).cpu().numpy().astype(bool)}

    def check(self) -> None:
        """Raise FloatingPointError if a stage met a bad cell (mhd.py:852-857); synchronises."""
        bad = self._bad.cpu()
        if int(bad.sum()):
            raise FloatingPointError(
                f"non-positive density or pressure on the level-0 patch {self.patch.box} "
                f"at t = {self.time}"
            )

    def run(self, duration: float, *, max_steps: int = 1_000_000) -> int:
        """Advance by ``duration``, the last step shortened (mhd.py:996-1008)."""
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

    # -- diagnostics (mhd.py:1059-1226), on the device -----------------------------------
    def _diag_tables(self):
        """The patch and leaf tables of the device diagnostics: the grid as one patch."""
        if getattr(self, "_diag", None) is None:
            from .amr import _PATCH

            patch = np.zeros(1, _PATCH)
            patch[0] = (0, self.nx, self.ny, self.dx, self.dy, 0, 0)
            a, b = np.meshgrid(np.arange(self.nx) + HALO, np.arange(self.ny) + HALO, indexing="ij")
            leaf = np.stack([np.zeros(a.size, np.int64), (a * self.Y + b).ravel()], axis=1)
            chunks = [(0, c) for c in range(0, self.nx * self.ny, 256)]
            dev = dict(device=self.device)
            self._diag = (
                torch.from_numpy(patch.view(np.uint8).copy()).to(self.device),
                torch.as_tensor(leaf.astype(np.int32).reshape(-1), **dev),
                torch.as_tensor(np.array(chunks, np.int32).reshape(-1), **dev),
                len(chunks),
                torch.zeros(_lib.call("wb_amr2d_totals_workspace"), dtype=torch.float64, **dev),
            )
        return self._diag

    def totals(self) -> dict[str, float]:
        """Integrals of the conserved fields (a fixed-order FP64 tree on the device)."""
        self._check_ready()
        return self._totals()

    def _totals(self) -> dict[str, float]:
        patch, leaf, _, _, work = self._diag_tables()
        out = torch.zeros(6, dtype=torch.float64, device=self.device)
        _lib.call("wb_amr2d_totals", _lib.ptr(self._buf[0]), _lib.ptr(patch), _lib.ptr(leaf),
                  self.nx * self.ny, _lib.ptr(work), _lib.ptr(out), _lib.stream_handle(self.device))
        return {name: float(v) for name, v in zip(CONSERVED, out.cpu().tolist())}

    def divergence(self) -> float:
        """Largest |div B| times the cell size (mhd.py:1114-1131), on the device."""
        self._check_ready()
        patch, _, chunks, nchunks, _ = self._diag_tables()
        out = torch.zeros(3, dtype=torch.float64, device=self.device)
        _lib.call("wb_amr2d_divergence", _lib.ptr(self._buf[0]), _lib.ptr(patch), _lib.ptr(chunks),
                  nchunks, *self.mean_field, None, 0, None, 0, _lib.ptr(out),
                  _lib.stream_handle(self.device))
        return float(out[0].item())

    def leaf_values(self) -> dict[str, np.ndarray]:
        """The primitive variables of every cell (one level: the leaf cells), with ``x``,
        ``y``, ``level`` and ``current_z`` (the mean of the cell's corners), as
        mhd.py:1059-1112 gives them; computed on the device."""
        self._check_ready()
        stream = _lib.stream_handle(self.device)
        pmask = int(self.periodic[0]) | (int(self.periodic[1]) << 1)
        wmask = int(self.walls[0]) | (int(self.walls[1]) << 1)
        _lib.call("wb_mhd2d_fill_halos", _lib.ptr(self._buf[0]), self.nx, self.ny, pmask, wmask,
                  stream)
        patch, leaf, _, _, _ = self._diag_tables()
        n = self.nx * self.ny
        out = torch.empty(9 * n, dtype=torch.float64, device=self.device)
        _lib.call("wb_amr2d_leaf_values", _lib.ptr(self._buf[0]), _lib.ptr(patch), _lib.ptr(leaf),
                  n, *self.mean_field, self.gamma, _lib.ptr(out), stream)
        v = out.view(9, n).cpu().numpy()
        x, y = self.hierarchy.cell_centers(self.patch)
        names = ("density", "velocity_x", "velocity_y", "velocity_z", "pressure",
                 "magnetic_field_x", "magnetic_field_y", "magnetic_field_z", "current_z")
        return {"x": x.ravel(), "y": y.ravel(), "level": np.zeros(n, np.int64),
                **{name: v[k].copy() for k, name in enumerate(names)}}

    def dataset(self, level: int = 0) -> dict[str, np.ndarray]:
        """Primitive variables over the cell centres (mhd.py:1177-1226), as plain arrays."""
        self._check_ready()
        if level != 0:
            raise ValueError("a uniform model has level 0 only")
        cons, az = self.arrays()
        h = HALO
        c = cons[:, h : h + self.nx, h : h + self.ny]
        a = az[h : h + self.nx + 1, h : h + self.ny + 1]
        bxf = (a[:, 1:] - a[:, :-1]) / self.dy + self.mean_field[0]
        byf = -(a[1:, :] - a[:-1, :]) / self.dx + self.mean_field[1]
        bx, by = 0.5 * (bxf[:-1] + bxf[1:]), 0.5 * (byf[:, :-1] + byf[:, 1:])
        rho = c[0]
        u, v, w = c[1] / rho, c[2] / rho, c[3] / rho
        b2 = bx**2 + by**2 + c[5] ** 2
        p = (self.gamma - 1.0) * (c[4] - 0.5 * rho * (u**2 + v**2 + w**2) - 0.5 * b2)
        x, y = self.hierarchy.cell_centers(self.patch)
        return {
            "x": x[:, 0], "y": y[0, :],
            "density": rho, "velocity_x": u, "velocity_y": v, "velocity_z": w,
            "pressure": p, "magnetic_field_x": bx, "magnetic_field_y": by,
            "magnetic_field_z": c[5], "magnetic_pressure": 0.5 * b2,
            "time": self.time,
        }
