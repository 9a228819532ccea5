"""CPU checks of the PyTorch reference network used to judge the device
U-Net: parameter layout round trip, and its autograd gradient against
central finite differences in FP64 (which pins the reference itself)."""
import numpy as np
import torch

import unet_ref as R


def test_param_count_and_roundtrip():
    n = R.num_params()
    # 14 convolutions of the 16-32-64-128 U-Net + the 1x1 output
    assert n == 487_153
    p = np.random.default_rng(0).standard_normal(n).astype(np.float32)
    layers, out = R.unpack(p)
    assert layers[0][0].shape == (16, 2, 3, 3)
    assert layers[8][0].shape == (64, 192, 3, 3)
    assert np.array_equal(R.pack(layers, out).astype(np.float32), p)


def _problem(H=8, B=2, seed=1):
    rng = np.random.default_rng(seed)
    r = rng.standard_normal((B, H, H))
    f = rng.standard_normal((B, H, H)) * 0.5
    mask = np.ones((H, H))
    mask[-1, :] = 0
    t = rng.standard_normal((B, H, H)) * mask
    t /= np.linalg.norm(t.reshape(B, -1), axis=1)[:, None, None]
    return r, f, t, mask


def _init(seed=0):
    rng = np.random.default_rng(seed)
    parts = []
    for (m, cout, t, t2), cin in zip(R.SPEC, R.layer_cin()):
        parts.append(rng.standard_normal(cout * 9 * cin) * np.sqrt(2.0 / (9 * cin)))
        parts.append(rng.standard_normal(cout) * 0.1)
    parts.append(rng.standard_normal(16) * 0.25)
    parts.append(np.array([0.05]))
    return np.concatenate(parts)


def test_reference_gradient_matches_finite_differences():
    r, f, t, mask = _problem()
    p = _init()
    g, per, z = R.grads(p, r, f, t, mask, bf16_points=False, dtype=torch.float64)

    def L(pp):
        layers, out = R.unpack(pp, torch.float64)
        zz = R.forward((layers, out), r, f, bf16_points=False)
        return float(R.loss_fn(zz, t, mask)[0])

    rng = np.random.default_rng(3)
    idx = rng.choice(p.size, 40, replace=False)
    # include the output layer and a deep-layer bias
    idx = np.concatenate([idx, [p.size - 1, p.size - 5]])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (L(p + e) - L(p - e)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_loss_is_scale_invariant_and_zero_at_target():
    r, f, t, mask = _problem()
    z = torch.as_tensor(t * 3.7)
    L, per = R.loss_fn(z, t, mask)
    assert float(L) < 1e-12
    L2, _ = R.loss_fn(-z, t, mask)
    assert abs(float(L2) - 4 * t.shape[0]) < 1e-9


def _problem3(S=8, B=1, seed=2):
    rng = np.random.default_rng(seed)
    r = rng.standard_normal((B, S, S, S))
    f = rng.standard_normal((B, S, S, S)) * 0.5
    mask = np.ones((S, S, S))
    mask[-1] = 0
    t = rng.standard_normal((B, S, S, S)) * mask
    t /= np.linalg.norm(t.reshape(B, -1), axis=1)[:, None, None, None]
    return r, f, t, mask


def _init3(seed=0):
    rng = np.random.default_rng(seed)
    parts = []
    for (m, cout, t, t2), cin in zip(R.SPEC, R.layer_cin()):
        parts.append(rng.standard_normal(cout * 27 * cin) * np.sqrt(2.0 / (27 * cin)))
        parts.append(rng.standard_normal(cout) * 0.1)
    parts.append(rng.standard_normal(16) * 0.25)
    parts.append(np.array([0.05]))
    return np.concatenate(parts)


def test_3d_layout_roundtrip_and_tap_order():
    """[cout][27][cin] with taps (dz, dy, dx) row-major: tap 0 is the
    (-1, -1, -1) corner of the 3x3x3 kernel, tap 13 its centre."""
    n = R.num_params(3)
    assert n == 17 + sum(c * 27 * ci + c for (m, c, t, t2), ci in zip(R.SPEC, R.layer_cin()))
    p = np.zeros(n, np.float32)
    cin0 = R.layer_cin()[0]
    p[0 * cin0 + 1] = 1.0            # cout 0, tap 0, ci 1
    p[13 * cin0 + 0] = 2.0           # cout 0, tap 13 (centre), ci 0
    layers, out = R.unpack(p, dim=3)
    w = layers[0][0]
    assert w.shape == (16, 2, 3, 3, 3)
    assert float(w[0, 1, 0, 0, 0]) == 1.0 and float(w[0, 0, 1, 1, 1]) == 2.0
    assert float(w.abs().sum()) == 3.0
    assert np.array_equal(R.pack(layers, out).astype(np.float32), p)


def test_3d_reference_gradient_matches_finite_differences():
    r, f, t, mask = _problem3()
    p = _init3()
    g, per, z = R.grads(p, r, f, t, mask, bf16_points=False, dtype=torch.float64)

    def L(pp):
        layers, out = R.unpack(pp, torch.float64, dim=3)
        zz = R.forward((layers, out), r, f, bf16_points=False)
        return float(R.loss_fn(zz, t, mask)[0])

    rng = np.random.default_rng(4)
    idx = np.concatenate([rng.choice(p.size, 24, replace=False), [p.size - 1, p.size - 5]])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (L(p + e) - L(p - e)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])
