"""Device U-Net (BF16 MFMA, FP32 accumulation) against the PyTorch CPU
reference of the same network (tests/unet_ref.py, itself checked against
finite differences in test_unet_ref.py).

Tolerances: with the reference rounding to BF16 at the same points as the
kernels, the only differences are FP32 accumulation order and the rare BF16
rounding decision it flips: relative L2 error <= 5e-3 for the first
convolution (which tests the MFMA operand/accumulator lane layout directly),
<= 2e-2 for every later tensor and for each layer's gradient.  Against the
pure FP32 network the BF16 effect is stated, not hidden: <= 0.1 relative
for the output, and for the gradient through all 15 layers <= 0.2 (measured
0.11 at 32^2, random initialisation) or the BF16 reference's own distance
from FP32 plus 0.05, whichever is larger.  Two shapes: 32^2 at batch 2 (direct-load
convolutions) and 128^2 at batch 4, where the level-0 convolutions and their
adjoints take the LDS-staged kernel."""
import numpy as np
import pytest
import torch

import unet_ref as R

pytestmark = pytest.mark.gpu

SHAPES = [(32, 2), (128, 4)]


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def nhwc(t):
    return t.detach().permute(0, 2, 3, 1).contiguous().numpy()


@pytest.fixture(scope="module", params=SHAPES, ids=lambda s: f"{s[0]}x{s[1]}")
def net(gpu, request):
    from getdp_amd.learned import UNet
    H, B = request.param
    n = UNet(H, H, batch=B, seed=7)
    n.Bt = B
    # nonzero biases so that the bias path is exercised
    p = n.get_params()
    rng = np.random.default_rng(5)
    off = 0
    for (m, cout, t, t2), cin in zip(R.SPEC, R.layer_cin()):
        off += cout * 9 * cin
        p[off:off + cout] = rng.standard_normal(cout) * 0.05
        off += cout
    p[-1] = 0.03
    n.set_params(p)
    return n


def inputs(net, seed=0):
    H, B = net.H, net.Bt
    rng = np.random.default_rng(seed)
    r = rng.standard_normal((B, H, H)).astype(np.float32)
    f = (rng.standard_normal((B, H, H)) * 0.5).astype(np.float32)
    mask = np.ones((H, H), np.float32)
    mask[-1, :] = 0
    mask[:, -1] = 0
    t = rng.standard_normal((B, H, H)) * mask
    t /= np.linalg.norm(t.reshape(B, -1), axis=1)[:, None, None]
    return r, f, t.astype(np.float32), mask


def test_num_params_and_roundtrip(net):
    assert net.num_params == R.num_params()
    p = net.get_params()
    net.set_params(p)
    assert np.array_equal(net.get_params(), p)


def test_forward_matches_reference(net):
    r, f, t, mask = inputs(net)
    z = net.forward(r, f)
    params = R.unpack(net.get_params())
    zr, acts = R.forward(params, r, f, bf16_points=True, keep=True)
    # input tensor: 8 stored channels, 2 used
    H, B = net.H, net.Bt
    a0 = net.activation(0, B).reshape(B, H, H, 8)
    assert np.array_equal(a0[..., :2], nhwc(acts[0]))
    assert not a0[..., 2:].any()
    for i in range(1, 15):
        d = net.activation(i, B)
        ref = nhwc(acts[i]).ravel()
        e = rel(d, ref)
        assert e <= (5e-3 if i == 1 else 2e-2), (i, e)
    assert rel(z, zr.detach().numpy()) <= 2e-2
    z32 = R.forward(params, r, f, bf16_points=False).detach().numpy()
    assert rel(z, z32) <= 0.1


def test_first_conv_exact_layout(net):
    """A single 3x3 convolution tested element by element: each output may
    differ from the FP32 sum of the same BF16 products by at most one BF16
    rounding step (2^-8 relative) -- a wrong MFMA lane mapping moves values
    between pixels/channels and fails by O(1)."""
    r, f, t, mask = inputs(net, 1)
    net.forward(r, f)
    params = R.unpack(net.get_params())
    _, acts = R.forward(params, r, f, bf16_points=True, keep=True)
    layers, _ = params
    w, b = layers[0]
    y = torch.relu(torch.nn.functional.conv2d(acts[0], R.q(w), b, padding=1))
    ref = nhwc(y)
    d = net.activation(1, net.Bt).reshape(ref.shape)
    assert np.all(np.abs(d - ref) <= 2.0 ** -8 * np.abs(ref) + 1e-6)


def test_gradients_match_reference(net):
    r, f, t, mask = inputs(net, 2)
    p0 = net.get_params()
    loss = net.train_step(r, f, t, mask, lr=0.0)
    g = net.get_grads()
    assert np.array_equal(net.get_params(), p0)  # lr = 0 leaves the weights
    gr, per, _ = R.grads(p0, r, f, t, mask, bf16_points=True)
    g64, _, _ = R.grads(p0, r, f, t, mask, bf16_points=True, dtype=torch.float64)
    assert abs(loss - per.mean()) <= 1e-3 * max(1.0, abs(per.mean()))
    # per layer: 2e-2, or twice the reference's own FP32/FP64 distance where
    # the BF16 rounding points make the gradient that sensitive (larger grids:
    # more ReLU and rounding decisions flip between accumulation orders)
    off = 0
    for li, ((m, cout, tt, t2), cin) in enumerate(zip(R.SPEC, R.layer_cin())):
        nw = cout * 9 * cin
        ew = rel(g[off:off + nw], gr[off:off + nw])
        eb = rel(g[off + nw:off + nw + cout], gr[off + nw:off + nw + cout])
        sw = max(2e-2, 2.0 * rel(gr[off:off + nw], g64[off:off + nw]))
        sb = max(2e-2, 2.0 * rel(gr[off + nw:off + nw + cout], g64[off + nw:off + nw + cout]))
        assert ew <= sw and eb <= sb, (li, ew, eb, sw, sb)
        off += nw + cout
    assert rel(g[off:], gr[off:]) <= 2e-2
    cosine = lambda u, v: float(np.dot(u, v) / np.linalg.norm(u) / np.linalg.norm(v))
    assert cosine(g, gr) >= 0.999 or 1 - cosine(g, gr) <= 4 * (1 - cosine(gr, g64)) + 1e-4
    # the BF16 effect against the pure FP32 network: what the BF16 reference
    # itself shows, plus the device's distance from that reference
    g32, _, _ = R.grads(p0, r, f, t, mask, bf16_points=False)
    assert rel(g, g32) <= max(0.2, rel(gr, g32) + 0.05), (rel(g, g32), rel(gr, g32))


def test_training_reduces_loss(gpu):
    """Adam on the device fits a fixed small set of pairs (smooth target of a
    known linear operator: e = 5-point Laplacian^{-1} r)."""
    from getdp_amd.learned import UNet
    n = UNet(16, 16, batch=8, seed=3)
    rng = np.random.default_rng(0)
    N = 15
    L = (np.diag(4 * np.ones(N * N)) - np.diag(np.ones(N * N - 1), 1) - np.diag(np.ones(N * N - 1), -1)
         - np.diag(np.ones(N * N - N), N) - np.diag(np.ones(N * N - N), -N))
    for i in range(1, N):
        L[i * N, i * N - 1] = L[i * N - 1, i * N] = 0
    mask = np.zeros((16, 16), np.float32)
    mask[:N, :N] = 1
    rs, es = [], []
    for k in range(8):
        rr = rng.standard_normal(N * N)
        ee = np.linalg.solve(L, rr)
        g_r = np.zeros((16, 16))
        g_e = np.zeros((16, 16))
        g_r[:N, :N] = rr.reshape(N, N)
        g_e[:N, :N] = ee.reshape(N, N)
        rs.append(g_r / np.sqrt((g_r ** 2).sum() / mask.sum()))
        es.append(g_e / np.linalg.norm(g_e))
    r = np.asarray(rs, np.float32)
    e = np.asarray(es, np.float32)
    f = np.zeros_like(r)
    losses = [n.train_step(r, f, e, mask, 2e-3) for _ in range(150)]
    assert np.isfinite(losses).all()
    assert np.mean(losses[-10:]) < 0.5 * np.mean(losses[:5]), (losses[:5], losses[-10:])
