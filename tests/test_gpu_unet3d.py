"""The device U-Net on 3D grids (3x3x3 convolutions, 2x2x2 pooling and
upsampling, BF16 MFMA with FP32 accumulation) against the PyTorch CPU
reference of the same network (tests/unet_ref.py with conv3d, itself checked
against finite differences in FP64 in test_unet_ref.py).  With the reference
rounding to BF16 at the same points: forward tensors within relative L2
5e-3 (first convolution) and 2e-2 (later ones), as in 2D; the first
convolution element by element within one BF16 rounding step (which pins the
27-tap order and the MFMA lane layout in 3D).  Gradients: the 27-tap sums
and 2x2x2 pools make the gradient sensitive to which BF16 roundings flip:
the reference itself moves by 4-8 % per layer between FP32 and FP64
accumulation with the same BF16 rounding points.  A device as noisy as the
FP32 reference is at most twice that from it (triangle inequality through
the FP64 reference), so that is the bound (per layer, at least 2e-2, and
for the cosine over all parameters); a wrong adjoint or tap order would be
O(1) on some layer.  Finally the learned
two-stage FGMRES on the 3D problem against the oracle's solution."""
import numpy as np
import pytest
import torch

import unet_ref as R

pytestmark = pytest.mark.gpu

SHAPES = [(16, 2), (32, 2)]  # at 32^3 the level-0 convolutions take the LDS-staged kernel


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def ndhwc(t):
    return t.detach().permute(0, 2, 3, 4, 1).contiguous().numpy()


@pytest.fixture(scope="module", params=SHAPES, ids=lambda s: f"{s[0]}x{s[1]}")
def net(gpu, request):
    from getdp_amd.learned import UNet
    S, B = request.param
    n = UNet(S, S, batch=B, seed=11, D=S)
    n.Bt = B
    p = n.get_params()
    rng = np.random.default_rng(5)
    off = 0
    for (m, cout, t, t2), cin in zip(R.SPEC, R.layer_cin()):
        off += cout * 27 * cin
        p[off:off + cout] = rng.standard_normal(cout) * 0.05
        off += cout
    p[-1] = 0.03
    n.set_params(p)
    return n


def inputs(net, seed=0):
    S, B = net.H, net.Bt
    rng = np.random.default_rng(seed)
    r = rng.standard_normal((B, S, S, S)).astype(np.float32)
    f = (rng.standard_normal((B, S, S, S)) * 0.5).astype(np.float32)
    mask = np.ones((S, S, S), np.float32)
    mask[0] = 0
    mask[:, :, -1] = 0
    t = rng.standard_normal((B, S, S, S)) * mask
    t /= np.linalg.norm(t.reshape(B, -1), axis=1)[:, None, None, None]
    return r, f, t.astype(np.float32), mask


def test_num_params_and_roundtrip(net):
    assert net.num_params == R.num_params(3)
    p = net.get_params()
    net.set_params(p)
    assert np.array_equal(net.get_params(), p)


def test_forward_matches_reference(net):
    r, f, t, mask = inputs(net)
    z = net.forward(r, f)
    params = R.unpack(net.get_params(), dim=3)
    zr, acts = R.forward(params, r, f, bf16_points=True, keep=True)
    S, B = net.H, net.Bt
    a0 = net.activation(0, B).reshape(B, S, S, S, 8)
    assert np.array_equal(a0[..., :2], ndhwc(acts[0]))
    for i in range(1, 15):
        e = rel(net.activation(i, B), ndhwc(acts[i]))
        assert e <= (5e-3 if i == 1 else 2e-2), (i, e)
    assert rel(z, zr.detach().numpy()) <= 2e-2
    z32 = R.forward(params, r, f, bf16_points=False).detach().numpy()
    assert rel(z, z32) <= 0.1


def test_first_conv_exact_layout(net):
    r, f, t, mask = inputs(net, 1)
    net.forward(r, f)
    params = R.unpack(net.get_params(), dim=3)
    _, acts = R.forward(params, r, f, bf16_points=True, keep=True)
    w, b = params[0][0]
    y = torch.relu(torch.nn.functional.conv3d(acts[0], R.q(w), b, padding=1))
    ref = ndhwc(y)
    d = net.activation(1, net.Bt).reshape(ref.shape)
    assert np.all(np.abs(d - ref) <= 2.0 ** -8 * np.abs(ref) + 1e-6)


def test_gradients_match_reference(net):
    r, f, t, mask = inputs(net, 2)
    p0 = net.get_params()
    loss = net.train_step(r, f, t, mask, lr=0.0)
    g = net.get_grads()
    assert np.array_equal(net.get_params(), p0)
    gr, per, _ = R.grads(p0, r, f, t, mask, bf16_points=True)
    g64, _, _ = R.grads(p0, r, f, t, mask, bf16_points=True, dtype=torch.float64)
    assert abs(loss - per.mean()) <= 1e-3 * max(1.0, abs(per.mean()))
    off = 0
    for li, ((m, cout, tt, t2), cin) in enumerate(zip(R.SPEC, R.layer_cin())):
        nw = cout * 27 * cin
        ew = rel(g[off:off + nw], gr[off:off + nw])
        eb = rel(g[off + nw:off + nw + cout], gr[off + nw:off + nw + cout])
        sw = max(2e-2, 2.0 * rel(gr[off:off + nw], g64[off:off + nw]))
        sb = max(2e-2, 2.0 * rel(gr[off + nw:off + nw + cout], g64[off + nw:off + nw + cout]))
        assert ew <= sw and eb <= sb, (li, ew, eb, sw, sb)
        off += nw + cout
    assert rel(g[off:], gr[off:]) <= 2e-2
    cosine = lambda u, v: float(np.dot(u, v) / np.linalg.norm(u) / np.linalg.norm(v))
    # 1 - cos grows as the square of the relative difference: twice the
    # reference's FP32/FP64 distance is four times its 1 - cos
    assert 1 - cosine(g, gr) <= 4 * (1 - cosine(gr, g64)) + 1e-4


def test_learned_fgmres_3d_matches_oracle(gpu, oracle):
    """An untrained network is still a valid (if poor) right preconditioner:
    the two-stage FGMRES on poisson3d(16) must reach 1e-10 (with the ILU(0)
    fallback if needed) and agree with the oracle's ILU(0)-GMRES solution;
    FP64 and the mixed-precision (FP32 inner) variant, with and without the
    least-squares scaling of the network's output."""
    from getdp_amd import learned as L2
    from getdp_amd import learned3d as L3
    from getdp_amd import problems as P
    from getdp_amd.learned import UNet
    pr = P.poisson3d(16, seed=4)
    ia, ja, a, b, tab = oracle.assemble(pr)
    xo, io, _ = oracle.solve(ia, ja, a, b, algorithm=8, preconditioner=6, krylov=30)
    fe, A, bv = gpu.assemble(pr)
    maps = L3.grid_maps(pr, fe.dof_table(0), A)
    net = UNet(16, 16, batch=1, seed=2, D=16)
    for prec in (0, 1):
        for lsq in (False, True):
            net.set_mode(lsq=lsq)
            x = gpu.Vector(fe.N)
            out = L2.solve_learned(A, bv, x, net, maps, maxits=500, two_stage=True, precision=prec)
            assert out["rel_res"] <= 1.5e-10, (prec, lsq, out)
            assert np.abs(x.get() - xo).max() <= 1e-8 * np.abs(xo).max()
