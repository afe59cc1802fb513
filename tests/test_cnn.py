"""The refinement-indicator CNN: the oracle pinned by finite differences (CPU), and
the MFMA kernels against the oracle (GPU)."""

import numpy as np
import pytest

from oracle import cnn as C


def _data(n=8, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, n, n, 8))
    y = (rng.uniform(size=(n, n, n)) < 0.3).astype(np.float64)
    return x, y


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-8, -2.5, 3.0e38], np.float32)
    r = C.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0 + 4 * 2**-8)
    assert r[3] == -2.5


def test_backward_matches_finite_differences():
    """float64, no rounding: every parameter's gradient against central differences."""
    x, y = _data(6)
    params = [p.astype(np.float64) * 3 for p in C.init(1)]
    params[1] += 0.05
    params[3] += 0.05
    loss, grads, _ = C.loss_and_grads(x, y, params, rnd=C.same)
    rng = np.random.default_rng(2)
    for k, (p, g) in enumerate(zip(params, grads)):
        for _ in range(4):
            idx = tuple(rng.integers(0, s) for s in p.shape)
            h = 1e-6
            old = p[idx]
            p[idx] = old + h
            lp = C.bce(C.forward(x, params, C.same)[3], y)
            p[idx] = old - h
            lm = C.bce(C.forward(x, params, C.same)[3], y)
            p[idx] = old
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[idx]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (k, idx, fd, g[idx])


def test_conv_transpose_is_the_adjoint():
    """<conv(x, w, 0), g> = <x, conv_transpose(g, w)>."""
    rng = np.random.default_rng(3)
    x = rng.standard_normal((5, 6, 4, 8))
    g = rng.standard_normal((5, 6, 4, 16))
    w = rng.standard_normal((27, 8, 16))
    lhs = np.sum(C.conv(x, w, np.zeros(16)) * g)
    rhs = np.sum(x * C.conv_transpose(g, w))
    np.testing.assert_allclose(lhs, rhs, rtol=1e-12)


def test_conv_follows_the_tap_definition():
    """conv against a direct loop: tap t = 9(dx+1) + 3(dy+1) + (dz+1) reads v + (dx, dy, dz)."""
    rng = np.random.default_rng(4)
    n = (3, 4, 5)
    x = rng.standard_normal(n + (2,))
    w = rng.standard_normal((27, 2, 3))
    b = rng.standard_normal(3)
    got = C.conv(x, w, b)
    for v in [(0, 0, 0), (2, 3, 4), (1, 0, 3)]:
        ref = b.copy()
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    t = 9 * (dx + 1) + 3 * (dy + 1) + (dz + 1)
                    src = ((v[0] + dx) % n[0], (v[1] + dy) % n[1], (v[2] + dz) % n[2])
                    ref += x[src] @ w[t]
        np.testing.assert_allclose(got[v], ref, rtol=1e-12)


def test_bf16_matches_torch():
    """The oracle's rounding equals torch's float32 -> bfloat16 conversion (nearest even)."""
    import torch

    rng = np.random.default_rng(5)
    v = (rng.standard_normal(100000) * np.exp(rng.uniform(-20, 20, 100000))).astype(np.float32)
    ties = (np.arange(1000, dtype=np.uint32) << 16 | 0x8000).view(np.float32)
    v = np.concatenate([v, ties, -ties])
    ref = torch.from_numpy(v).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(C.bf16(v), ref)


# ---- the MFMA kernels against the oracle (tolerances: FP32 sums in another order can
# move a BF16 rounding by one unit; everything else is stated per test) ---------------

def _bf16_bits(a):
    """float32 values (already bf16) to their 16-bit patterns as int16."""
    u = np.ascontiguousarray(C.bf16(np.asarray(a, np.float32))).view(np.uint32) >> 16
    return u.astype(np.uint16).view(np.int16)


def _from_bits(t):
    """int16 torch tensor of bf16 patterns to float32 numpy."""
    u = t.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
    return u.view(np.float32)


def _close_bf16(got, ref, frac=1e-3):
    """Equal except for a fraction ``frac`` of values one BF16 unit apart."""
    got, ref = np.asarray(got, np.float32), np.asarray(ref, np.float32)
    diff = got != ref
    assert diff.mean() <= frac, diff.mean()
    np.testing.assert_array_less(np.abs(got - ref)[diff], 2.0**-7 * np.abs(ref)[diff] + 1e-30)


def _net_inputs(shape, seed):
    rng = np.random.default_rng(seed)
    x = C.bf16(rng.standard_normal(shape + (8,)).astype(np.float32))
    y = (rng.uniform(size=shape) < 0.3).astype(np.float32)
    params = C.init(seed)
    for k in (1, 3, 5):
        params[k] = (0.1 * rng.standard_normal(params[k].shape)).astype(np.float32)
    return x, y, params


@pytest.mark.gpu
@pytest.mark.parametrize("cin", [8, 16])
def test_gpu_layer_matches_oracle(gpu, cin):
    """One forward layer (mode 0): the MFMA operand layout and tap order."""
    import torch
    from wellbalanced_amd import _lib

    shape = (6, 5, 32)
    rng = np.random.default_rng(cin)
    x = C.bf16(rng.standard_normal(shape + (cin,)).astype(np.float32))
    w = C.bf16(0.2 * rng.standard_normal((27, cin, 16)).astype(np.float32))
    b = (0.1 * rng.standard_normal(16)).astype(np.float32)
    ref = C.bf16(np.maximum(C.conv(x, w, b), 0.0))
    kw = torch.from_numpy(_bf16_bits(np.transpose(w, (0, 2, 1)).ravel())).to(gpu)
    xin = torch.from_numpy(_bf16_bits(x.ravel())).to(gpu)
    out = torch.empty(np.prod(shape) * 16, dtype=torch.int16, device=gpu)
    bias = torch.from_numpy(b).to(gpu)
    _lib.call("wb_cnn_conv", _lib.ptr(xin), cin, _lib.ptr(kw), _lib.ptr(bias), *shape, 0, None,
              _lib.ptr(out), None, 64, _lib.stream_handle(gpu))
    _close_bf16(_from_bits(out).reshape(shape + (16,)), ref)


@pytest.mark.gpu
def test_gpu_logit_layer_matches_oracle(gpu):
    """The 1-channel layer (mode 1: 16 voxels along y in N, banded weights) on a grid
    smaller than a tile in x and ragged in y: FP32 sums of the same BF16 products."""
    import torch
    from wellbalanced_amd import _lib

    shape = (3, 21, 32)
    rng = np.random.default_rng(5)
    x = C.bf16(rng.standard_normal(shape + (16,)).astype(np.float32))
    w = C.bf16(0.2 * rng.standard_normal((27, 16, 1)).astype(np.float32))
    b = np.array([0.3] + [0.0] * 15, np.float32)
    ref = C.conv(x, w, b[:1])[..., 0]
    kw = np.zeros((27, 16, 16), np.float32)
    kw[:, 0, :] = w[:, :, 0]  # the forward layout K[t][co][ci], row co = 0
    kw = torch.from_numpy(_bf16_bits(kw.ravel())).to(gpu)
    xin = torch.from_numpy(_bf16_bits(x.ravel())).to(gpu)
    logit = torch.empty(int(np.prod(shape)), dtype=torch.float32, device=gpu)
    bias = torch.from_numpy(b).to(gpu)
    _lib.call("wb_cnn_conv", _lib.ptr(xin), 16, _lib.ptr(kw), _lib.ptr(bias), *shape, 1, None,
              None, _lib.ptr(logit), 1, _lib.stream_handle(gpu))
    got = logit.cpu().numpy().reshape(shape)
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-5 * np.abs(ref).max())


@pytest.mark.gpu
@pytest.mark.parametrize("cin", [8, 16])
def test_gpu_weight_gradient_matches_oracle(gpu, cin):
    """One layer's weight and bias gradients (every tap, a persistent grid of 5
    workgroups over a grid ragged in x and y)."""
    import torch
    from wellbalanced_amd import _lib

    shape = (6, 11, 32)
    rng = np.random.default_rng(20 + cin)
    x = C.bf16(rng.standard_normal(shape + (cin,)).astype(np.float32))
    g = C.bf16(rng.standard_normal(shape + (16,)).astype(np.float32))
    ref_w = C.weight_grad(x, g)
    ref_b = g.reshape(-1, 16).sum(axis=0)
    xin = torch.from_numpy(_bf16_bits(x.ravel())).to(gpu)
    gin = torch.from_numpy(_bf16_bits(g.ravel())).to(gpu)
    wg = 5
    partial = torch.empty(wg * (27 * 256 + 16), dtype=torch.float32, device=gpu)
    out = torch.empty(27 * 256 + 16, dtype=torch.float32, device=gpu)
    _lib.call("wb_cnn_wgrad", _lib.ptr(xin), cin, _lib.ptr(gin), *shape, _lib.ptr(partial), wg,
              _lib.ptr(out), _lib.stream_handle(gpu))
    o = out.cpu().numpy()
    got_w = o[: 27 * 256].reshape(27, 16, 16)[:, :cin, :]
    np.testing.assert_allclose(got_w, ref_w, rtol=1e-4, atol=1e-4 * np.abs(ref_w).max())
    np.testing.assert_allclose(o[27 * 256 :], ref_b, rtol=1e-4, atol=1e-4 * np.abs(ref_b).max())
    assert not o[: 27 * 256].reshape(27, 16, 16)[:, cin:, :].any()


@pytest.mark.gpu
def test_gpu_network_matches_oracle(gpu):
    """Forward, loss and every gradient of the whole network against the oracle.
    Tolerances: logits 2e-3 absolute + 1e-2 relative, loss 1e-4 relative, each
    gradient 1e-2 in relative L2 norm."""
    import torch
    from wellbalanced_amd import cnn

    shape = (8, 6, 32)
    x, y, params = _net_inputs(shape, 11)
    net = cnn.Indicator(shape, device=gpu, workgroups=7)
    net.set_parameters(params)
    xin = torch.from_numpy(_bf16_bits(x.ravel())).to(gpu)
    label = torch.from_numpy(y.ravel()).to(gpu)
    logit = net.forward(xin).cpu().numpy()
    loss = float(net.backward(xin, label))
    grads = net.gradients()
    ref_loss, ref_grads, ref_logit = C.loss_and_grads(x, y, params)
    np.testing.assert_allclose(logit, ref_logit, rtol=1e-2, atol=2e-3)
    assert abs(loss - ref_loss) <= 1e-4 * abs(ref_loss)
    for k, (g, r) in enumerate(zip(grads, ref_grads)):
        err = np.linalg.norm(g - r) / np.linalg.norm(r)
        assert err <= 1e-2, (k, err)


@pytest.mark.gpu
def test_gpu_adam_matches_numpy(gpu):
    import torch
    from wellbalanced_amd import _lib, cnn

    net = cnn.Indicator((1, 1, 16), device=gpu, lr=1e-2)
    rng = np.random.default_rng(6)
    p = net.params.cpu().numpy().astype(np.float64)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for step in range(1, 4):
        g = rng.standard_normal(p.size).astype(np.float32)
        net.grad.copy_(torch.from_numpy(g))
        net.optimise()
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g.astype(np.float64) ** 2
        p -= 1e-2 * (m / (1 - 0.9**step)) / (np.sqrt(v / (1 - 0.999**step)) + 1e-8)
    np.testing.assert_allclose(net.params.cpu().numpy(), p, rtol=1e-5, atol=1e-6)
    del _lib


@pytest.mark.gpu
def test_gpu_features_are_the_normalised_primitives(gpu):
    from wellbalanced_amd import MHD3D, cnn, initial

    model = MHD3D.from_problem(initial.blast_turbulence_3d(16, seed=1))
    model.advance(2)
    net = cnn.Indicator(model.n, device=gpu)
    net.fit_normalisation(model)
    data = model.dataset()
    names = ["density", "velocity_x", "velocity_y", "velocity_z", "pressure",
             "magnetic_field_x", "magnetic_field_y", "magnetic_field_z"]
    prim = np.stack([data[k] for k in names], axis=-1)
    mean = prim.reshape(-1, 8).mean(axis=0)
    std = prim.reshape(-1, 8).std(axis=0)
    norm = net.norm.cpu().numpy()
    np.testing.assert_allclose(norm[:8], mean, rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(norm[8:], 1.0 / std, rtol=1e-5)
    feats = _from_bits(net.features(model)).reshape(prim.shape)
    ref = C.bf16(((prim.astype(np.float32) - norm[:8]) * norm[8:]).astype(np.float32))
    _close_bf16(feats, ref)


@pytest.mark.gpu
def test_gpu_training_lowers_the_loss(gpu):
    """A rule the network can represent (label = a smoothed channel > 0) is learned."""
    import torch
    from wellbalanced_amd import cnn

    shape = (16, 16, 32)
    rng = np.random.default_rng(8)
    x = C.bf16(rng.standard_normal(shape + (8,)).astype(np.float32))
    s = sum(np.roll(x[..., 0], k, axis=a) for a in range(3) for k in (-1, 0, 1))
    y = (s > 0.0).astype(np.float32)
    net = cnn.Indicator(shape, device=gpu, lr=3e-3, seed=2)
    xin = torch.from_numpy(_bf16_bits(x.ravel())).to(gpu)
    label = torch.from_numpy(y.ravel()).to(gpu)
    losses = net.fit([(xin, label)], 150)
    assert losses[-1] < 0.5 * losses[0], (losses[0], losses[-1])
    pred = net.forward(xin).cpu().numpy() > 0
    assert (pred == (y > 0)).mean() > 0.9


@pytest.mark.gpu
def test_gpu_learned_indicator_gain(gpu):
    """What the network gains over the plain estimator: trained on a blast in
    turbulence (64^3) to predict the Löhner flags 8 steps ahead, on later snapshots it
    never saw it recovers more of those future flags than the current plain flags do,
    at the same number of flagged cells.  Measured: 0.80-0.83 against 0.68-0.75."""
    from wellbalanced_amd import MHD3D, cnn, indicator, initial

    model = MHD3D.from_problem(initial.blast_turbulence_3d(64, seed=0))
    net = cnn.Indicator(model.n, device=gpu, lr=3e-3, seed=0)
    model.advance(4)
    net.fit_normalisation(model)
    samples = indicator.lookahead_samples(model, net, snapshots=14, every=4, lead=8)
    train, test = samples[:9], samples[9:]
    losses = net.fit([(x, (later > 0.8).float().reshape(-1).contiguous())
                      for _, x, _, later in train], 200)
    assert losses[-1] < 0.05 * losses[0]
    gains = []
    for _, x, now, later in test:
        target = later > 0.8
        count = int((now > 0.8).sum())
        learned = indicator.recall_at_count(net.forward(x), target, count)
        plain = indicator.recall_at_count(now, target, count)
        assert learned > plain, (learned, plain)
        gains.append(learned - plain)
    assert np.mean(gains) >= 0.05, gains
    # the fallback: without a network the flags are the estimator's
    assert bool((indicator.refine(model) == (indicator.lohner(model) > 0.8)).all())
