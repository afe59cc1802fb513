"""The learned refinement indicator: a 3D CNN trained on the GPU by this package's kernels.

Three 3x3x3 periodic convolutions, 8 -> 16 -> 16 -> 1 channels with ReLU, over the
8 primitives of a uniform 3D state (``mhd3d.py:475-495``), normalised per channel;
the output is one logit per cell.  Operands are BF16 on the matrix cores with FP32
accumulation (``wb_cnn_forward`` / ``wb_cnn_backward``), the masters and Adam are
FP32 (``wb_cnn_adam``).  The reference has no learned indicator (its estimator is
``current_tags``, ``mhd3d.py:771-801``): the labels come from the Löhner estimator
(`wellbalanced_amd.indicator.lohner`), which stays the fallback.

Parameter layout (FP32, on the device): per layer ``W[27][16][16]`` (tap, input
channel, output channel; the channels a layer lacks are zero) and a bias of 16.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib

TAPS = 27
LAYER = TAPS * 256 + 16
NPARAMS = 3 * LAYER
KW_SIZE = TAPS * 16 * 8 + 4 * TAPS * 256
CHANNELS = ((8, 16), (16, 16), (16, 1))


def he_init(seed: int = 0) -> list[np.ndarray]:
    """He-initialised weights ``(27, cin, cout)`` and zero biases, FP32."""
    rng = np.random.default_rng(seed)
    out = []
    for cin, cout in CHANNELS:
        w = rng.standard_normal((TAPS, cin, cout)).astype(np.float32)
        out += [w * np.float32(np.sqrt(2.0 / (TAPS * cin))), np.zeros(cout, np.float32)]
    return out


def pack(params: list[np.ndarray]) -> np.ndarray:
    """Six arrays (w1, b1, w2, b2, w3, b3) into the flat padded layout."""
    flat = np.zeros(NPARAMS, np.float32)
    for layer, (cin, cout) in enumerate(CHANNELS):
        w = np.zeros((TAPS, 16, 16), np.float32)
        w[:, :cin, :cout] = params[2 * layer]
        b = np.zeros(16, np.float32)
        b[:cout] = params[2 * layer + 1]
        base = layer * LAYER
        flat[base : base + TAPS * 256] = w.ravel()
        flat[base + TAPS * 256 : base + LAYER] = b
    return flat


def unpack(flat: np.ndarray) -> list[np.ndarray]:
    """The flat layout back into six arrays of the true shapes."""
    out = []
    for layer, (cin, cout) in enumerate(CHANNELS):
        base = layer * LAYER
        w = flat[base : base + TAPS * 256].reshape(TAPS, 16, 16)
        out += [w[:, :cin, :cout].copy(), flat[base + TAPS * 256 : base + TAPS * 256 + cout].copy()]
    return out


class Indicator:
    """The network and its training state on one device, for grids of ``shape``.

    ``shape[2]`` must be a multiple of 16 (an MFMA tile takes 16 cells along z).
    """

    def __init__(self, shape, *, device="cuda", seed: int = 0, lr: float = 2e-3,
                 betas=(0.9, 0.999), eps: float = 1e-8, workgroups: int = 512):
        nx, ny, nz = (int(s) for s in shape)
        if nz % 16 or nz < 16 or min(nx, ny) < 1:
            raise ValueError("the grid's z extent must be a positive multiple of 16")
        self.shape = (nx, ny, nz)
        self.n = nx * ny * nz
        self.device = torch.device(device)
        self.lr, self.betas, self.eps = lr, betas, eps
        self.workgroups = workgroups
        f32 = dict(dtype=torch.float32, device=self.device)
        i16 = dict(dtype=torch.int16, device=self.device)
        self.params = torch.zeros(NPARAMS, **f32)
        self.grad = torch.zeros(NPARAMS, **f32)
        self.m = torch.zeros(NPARAMS, **f32)
        self.v = torch.zeros(NPARAMS, **f32)
        self.kw = torch.zeros(KW_SIZE, **i16)
        self.norm = torch.tensor([0.0] * 8 + [1.0] * 8, **f32)
        self.a1 = torch.empty(self.n * 16, **i16)
        self.a2 = torch.empty(self.n * 16, **i16)
        self.logit = torch.empty(self.n, **f32)
        self.g3 = torch.empty(self.n, **i16)  # the logits' gradient: one channel
        self.g2 = torch.empty(self.n * 16, **i16)
        self.g1 = torch.empty(self.n * 16, **i16)
        self.partial = torch.empty(workgroups * LAYER, **f32)
        self.loss = torch.zeros(1, **f32)
        self.steps = 0
        self.set_parameters(he_init(seed))

    # -- parameters -------------------------------------------------------------------
    def _stream(self) -> int:
        return _lib.stream_handle(self.device)

    def _prep(self) -> None:
        _lib.call("wb_cnn_prep", _lib.ptr(self.params), _lib.ptr(self.kw), self._stream())

    def set_parameters(self, params: list[np.ndarray]) -> None:
        """Load six arrays (w1, b1, w2, b2, w3, b3) and reset the optimiser."""
        self.params.copy_(torch.from_numpy(pack(params)))
        self.m.zero_()
        self.v.zero_()
        self.steps = 0
        self._prep()

    def parameters(self) -> list[np.ndarray]:
        return unpack(self.params.cpu().numpy())

    def gradients(self) -> list[np.ndarray]:
        return unpack(self.grad.cpu().numpy())

    # -- inputs -------------------------------------------------------------------------
    def _state_args(self, model):
        if tuple(model.n) != self.shape:
            raise ValueError(f"model grid {tuple(model.n)} is not the indicator's {self.shape}")
        stream = self._stream()
        pmask = sum(int(p) << a for a, p in enumerate(model.periodic))
        nx, ny, nz = self.shape
        _lib.call("wb_mhd3d_fill_halos", _lib.ptr(model.state), nx, ny, nz, pmask, stream)
        return (_lib.ptr(model.state), nx, ny, nz, *model.spacing, *model.mean_field, model.gamma)

    def fit_normalisation(self, model) -> None:
        """Per-channel mean and 1/std of the primitives of ``model`` (kept from now on)."""
        wg = 256
        partial = torch.empty(wg * 16, dtype=torch.float64, device=self.device)
        _lib.call("wb_cnn_moments", *self._state_args(model), _lib.ptr(partial), wg,
                  _lib.ptr(self.norm), self._stream())

    def features(self, model) -> torch.Tensor:
        """The network input of ``model``'s state: ``n x 8`` bf16 (as int16), new tensor."""
        x = torch.empty(self.n * 8, dtype=torch.int16, device=self.device)
        _lib.call("wb_cnn_input", *self._state_args(model), _lib.ptr(self.norm), _lib.ptr(x),
                  self._stream())
        return x

    # -- network ------------------------------------------------------------------------
    def forward(self, x: torch.Tensor) -> torch.Tensor:
        """Logits of every cell, ``(nx, ny, nz)`` FP32 (a view of the indicator's buffer)."""
        if x.numel() != self.n * 8 or x.dtype != torch.int16:
            raise ValueError("x must be the n x 8 bf16 features of this grid")
        nx, ny, nz = self.shape
        _lib.call("wb_cnn_forward", _lib.ptr(x), _lib.ptr(self.params), _lib.ptr(self.kw),
                  nx, ny, nz, _lib.ptr(self.a1), _lib.ptr(self.a2), _lib.ptr(self.logit),
                  self.workgroups, self._stream())
        return self.logit.view(self.shape)

    def backward(self, x: torch.Tensor, label: torch.Tensor) -> torch.Tensor:
        """Gradients of the mean BCE of the last forward's logits; returns the loss (device)."""
        if label.numel() != self.n or label.dtype != torch.float32:
            raise ValueError("label must be n FP32 values")
        nx, ny, nz = self.shape
        _lib.call("wb_cnn_backward", _lib.ptr(x), _lib.ptr(self.a1), _lib.ptr(self.a2),
                  _lib.ptr(self.logit), _lib.ptr(label), _lib.ptr(self.kw), nx, ny, nz,
                  _lib.ptr(self.g3), _lib.ptr(self.g2), _lib.ptr(self.g1), _lib.ptr(self.partial),
                  self.workgroups, _lib.ptr(self.grad), _lib.ptr(self.loss), self._stream())
        return self.loss / self.n

    def optimise(self) -> None:
        """One Adam step from the current gradients, then the BF16 kernel weights."""
        self.steps += 1
        b1, b2 = self.betas
        _lib.call("wb_cnn_adam", _lib.ptr(self.params), _lib.ptr(self.grad), _lib.ptr(self.m),
                  _lib.ptr(self.v), NPARAMS, self.lr, b1, b2, self.eps, self.steps, self._stream())
        self._prep()

    def train_step(self, x: torch.Tensor, label: torch.Tensor) -> torch.Tensor:
        self.forward(x)
        loss = self.backward(x, label)
        self.optimise()
        return loss

    def fit(self, samples, iterations: int) -> list[float]:
        """Adam over ``samples`` (pairs of features and FP32 labels) in turn; losses per pass."""
        losses = []
        for it in range(iterations):
            total = torch.zeros((), dtype=torch.float32, device=self.device)
            for x, label in samples:
                total = total + self.train_step(x, label)
            losses.append(total / len(samples))
        return [float(v) for v in torch.stack(losses).cpu()]

    def predict(self, model) -> torch.Tensor:
        """Logits of ``model``'s current state, ``(nx, ny, nz)`` (a copy)."""
        return self.forward(self.features(model)).clone()
