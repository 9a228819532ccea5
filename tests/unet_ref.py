"""Plain PyTorch (CPU, FP32) reference of the device U-Net, for the numerics
tests of the BF16 MFMA kernels, on 2D grids (3x3 convolutions) and on 3D
grids (3x3x3 convolutions, 2x2x2 pooling); ``dim`` selects which.

``bf16_points=True`` rounds to BF16 exactly where the device kernels do
(network input, weights, every convolution output after bias+ReLU, every
pooled value, and the ReLU-masked gradient that feeds the weight- and
input-gradient products), so that the device result must agree to FP32
accumulation-order noise; ``bf16_points=False`` is the pure FP32 network.
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F

# (mode, cout, input tensor, skip tensor); tensor i+1 = output of conv i
SPEC = [("plain", 16, 0, -1), ("plain", 16, 1, -1),
        ("pool", 32, 2, -1), ("plain", 32, 3, -1),
        ("pool", 64, 4, -1), ("plain", 64, 5, -1),
        ("pool", 128, 6, -1), ("plain", 128, 7, -1),
        ("upcat", 64, 8, 6), ("plain", 64, 9, -1),
        ("upcat", 32, 10, 4), ("plain", 32, 11, -1),
        ("upcat", 16, 12, 2), ("plain", 16, 13, -1)]


def layer_cin():
    ch = [2]
    cins = []
    for mode, cout, t, t2 in SPEC:
        cins.append(ch[t] + (ch[t2] if mode == "upcat" else 0))
        ch.append(cout)
    return cins


def num_params(dim: int = 2):
    taps = 3 ** dim
    return sum(c * taps * ci + c for (m, c, t, t2), ci in zip(SPEC, layer_cin())) + 17


def unpack(params: np.ndarray, dtype=torch.float32, dim: int = 2):
    """Flat parameter array (device layout [cout][taps][cin], taps (dz,) dy, dx
    row-major) -> list of (W[cout,cin,(3,)3,3], b), (ow, ob)."""
    p = torch.as_tensor(np.asarray(params), dtype=dtype)
    off = 0
    layers = []
    taps = 3 ** dim
    for (mode, cout, t, t2), cin in zip(SPEC, layer_cin()):
        n = cout * taps * cin
        w = p[off:off + n].reshape(cout, *([3] * dim), cin)
        w = w.permute(0, dim + 1, *range(1, dim + 1)).contiguous()
        off += n
        b = p[off:off + cout].clone()
        off += cout
        layers.append((w, b))
    ow = p[off:off + 16].clone()
    ob = p[off + 16:off + 17].clone()
    assert off + 17 == p.numel()
    return layers, (ow, ob)


def pack(layers, out) -> np.ndarray:
    parts = []
    for w, b in layers:
        dim = w.dim() - 2
        parts.append(w.permute(0, *range(2, dim + 2), 1).reshape(-1))
        parts.append(b.reshape(-1))
    parts += [out[0].reshape(-1), out[1].reshape(-1)]
    return torch.cat(parts).detach().numpy()


def q(x):
    return x.to(torch.bfloat16).to(x.dtype)


class _QF(torch.autograd.Function):
    """round to BF16 in the forward pass, identity backward"""
    @staticmethod
    def forward(ctx, x):
        return q(x)

    @staticmethod
    def backward(ctx, g):
        return g


class _QB(torch.autograd.Function):
    """identity forward, round the gradient to BF16"""
    @staticmethod
    def forward(ctx, x):
        return x.clone()

    @staticmethod
    def backward(ctx, g):
        return q(g)


def forward(params, r, feat, bf16_points=True, keep=False):
    """r, feat: (B, H, W) or (B, D, H, W).  Returns z of the same shape and,
    with keep, the list of tensors (B, C, ...) in the device's numbering."""
    QF = _QF.apply if bf16_points else (lambda x: x)
    QB = _QB.apply if bf16_points else (lambda x: x)
    layers, (ow, ob) = params
    dt = ow.dtype
    dim = np.ndim(r) - 1
    conv = F.conv3d if dim == 3 else F.conv2d
    pool = F.avg_pool3d if dim == 3 else F.avg_pool2d
    x0 = QF(torch.stack([torch.as_tensor(r, dtype=dt), torch.as_tensor(feat, dtype=dt)], 1))
    acts = [x0]
    for (mode, cout, t, t2), (w, b) in zip(SPEC, layers):
        x = acts[t]
        if mode == "pool":
            x = QF(pool(x, 2))
        elif mode == "upcat":
            x = torch.cat([F.interpolate(x, scale_factor=2, mode="nearest"), acts[t2]], 1)
        y = conv(x, QF(w), b, padding=1)
        y = QB(y)
        acts.append(QF(F.relu(y)))
    z = conv(acts[-1], ow.view(1, 16, *([1] * dim)), ob)[:, 0]
    return (z, acts) if keep else z


def loss_fn(z, target, mask):
    """sum over samples of |z m/|z m| - t|^2 = 2 - 2 (z m . t)/|z m| (t unit norm)."""
    m = torch.as_tensor(mask, dtype=z.dtype)
    t = torch.as_tensor(target, dtype=z.dtype)
    zm = z * m
    ax = tuple(range(1, z.dim()))
    nz = torch.sqrt((zm * zm).sum(dim=ax))
    c = (zm * t).sum(dim=ax) / nz
    return (2 - 2 * c).sum(), 2 - 2 * c


def grads(params_flat, r, feat, target, mask, bf16_points=True, dtype=torch.float32):
    layers, out = unpack(params_flat, dtype, dim=np.ndim(r) - 1)
    leaves = []
    for w, b in layers:
        w.requires_grad_(True)
        b.requires_grad_(True)
        leaves += [w, b]
    out[0].requires_grad_(True)
    out[1].requires_grad_(True)
    z = forward((layers, out), r, feat, bf16_points)
    L, per = loss_fn(z, target, mask)
    L.backward()
    g = pack([(w.grad, b.grad) for w, b in layers], (out[0].grad, out[1].grad))
    return g, per.detach().numpy(), z.detach().numpy()
