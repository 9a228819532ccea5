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
