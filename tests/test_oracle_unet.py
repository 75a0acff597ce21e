"""The U-Net oracle (oracle/unet.c) pinned on the CPU.

The U-Net has no counterpart in the reference.  Each operator of the oracle
is checked against PyTorch's CPU implementation (an independent library) in
float64: conv3d and both of its adjoints, avg_pool3d (3, stride 2, padding
1, divisor 27) and trilinear x2 interpolation with aligned corners, plus
their adjoints by autograd.  The whole network (loss and every parameter
gradient) is checked against the same network written in torch, and its
gradient against central finite differences.  bfloat16 rounding is checked
against torch's float32 -> bfloat16 conversion (nearest even)."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

import oracle as o

torch.set_default_dtype(torch.float64)


def to_t(x):
    """[z][y][x][c] -> [1][c][z][y][x]"""
    return torch.from_numpy(np.ascontiguousarray(x)).permute(3, 0, 1, 2)[None].contiguous()


def from_t(t):
    return t[0].permute(1, 2, 3, 0).detach().numpy()


def w_t(w):
    """[cout][27][cin] with t = (dx+1) + 3(dy+1) + 9(dz+1) -> [cout][cin][kz][ky][kx]"""
    cout, _, cin = w.shape
    return torch.from_numpy(w.reshape(cout, 3, 3, 3, cin)).permute(0, 4, 1, 2, 3).contiguous()


SHAPE = (7, 6, 5)   # nz, ny, nx: distinct extents catch transposed axes


def test_conv3d_matches_torch():
    rng = np.random.default_rng(1)
    x = rng.standard_normal(SHAPE + (8,))
    w = rng.standard_normal((16, 27, 8))
    b = rng.standard_normal(16)
    want = from_t(F.conv3d(to_t(x), w_t(w), torch.from_numpy(b), padding=1))
    assert np.allclose(o.conv3d(x, w, b), want, rtol=1e-12, atol=1e-12)
    assert np.allclose(o.conv3d(x, w, b, relu=True), np.maximum(want, 0), rtol=1e-12, atol=1e-12)


def test_conv3d_adjoints_match_torch():
    rng = np.random.default_rng(2)
    x = rng.standard_normal(SHAPE + (8,))
    w = rng.standard_normal((16, 27, 8))
    gy = rng.standard_normal(SHAPE + (16,))
    xt = to_t(x).requires_grad_(True)
    wt = w_t(w).requires_grad_(True)
    bt = torch.zeros(16, requires_grad=True)
    (F.conv3d(xt, wt, bt, padding=1) * to_t(gy)).sum().backward()
    assert np.allclose(o.conv3d_bwd_data(gy, w), from_t(xt.grad), rtol=1e-12, atol=1e-12)
    gw, gb = o.conv3d_bwd_weight(x, gy)
    want_w = wt.grad.permute(0, 2, 3, 4, 1).reshape(16, 27, 8).numpy()
    assert np.allclose(gw, want_w, rtol=1e-12, atol=1e-11)
    assert np.allclose(gb, bt.grad.numpy(), rtol=1e-12, atol=1e-12)


def test_pool_and_adjoint_match_torch():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((9, 7, 5, 4))
    xt = to_t(x).requires_grad_(True)
    yt = F.avg_pool3d(xt, 3, stride=2, padding=1, count_include_pad=True)
    assert np.allclose(o.pool(x), from_t(yt), rtol=1e-13, atol=1e-14)
    g = rng.standard_normal(yt.shape[2:] + (4,))
    (yt * to_t(g)).sum().backward()
    assert np.allclose(o.pool_bwd(g, x.shape[:3]), from_t(xt.grad), rtol=1e-13, atol=1e-14)


def test_upsample_and_adjoint_match_torch():
    rng = np.random.default_rng(4)
    x = rng.standard_normal((5, 4, 3, 4))
    xt = to_t(x).requires_grad_(True)
    yt = F.interpolate(xt, size=(9, 7, 5), mode="trilinear", align_corners=True)
    assert np.allclose(o.up(x), from_t(yt), rtol=1e-13, atol=1e-14)
    g = rng.standard_normal((9, 7, 5, 4))
    (yt * to_t(g)).sum().backward()
    assert np.allclose(o.up_bwd(g), from_t(xt.grad), rtol=1e-13, atol=1e-14)


def test_upsample_reproduces_linear_fields():
    """Trilinear interpolation is exact on affine functions of the index."""
    nz, ny, nx = 5, 4, 6
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    f = (0.5 + 2 * x - 3 * y + 0.25 * z)[..., None]
    fz, fy, fx = np.meshgrid(np.arange(2 * nz - 1), np.arange(2 * ny - 1), np.arange(2 * nx - 1), indexing="ij")
    want = (0.5 + fx - 1.5 * fy + 0.125 * fz)[..., None]
    assert np.allclose(o.up(f), want, rtol=0, atol=1e-13)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(5)
    v = np.concatenate([rng.standard_normal(5000) * 10.0 ** rng.integers(-8, 8, 5000),
                        # exact ties: bf16 mantissa + half an ulp, both parities
                        np.array([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -(1.0 + 2 ** -8), 0.0, -0.0])])
    v = v.astype(np.float32)
    want = torch.from_numpy(v).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(o.bf16(v), want)
    assert all(o._lib.ou_bf16(float(a)) == b for a, b in zip(v[-5:], want[-5:]))


class TorchUNet(torch.nn.Module):
    """The U-Net written with torch operators, for the whole-network pin."""

    def __init__(self, params):
        super().__init__()
        self.W, self.B = [], []
        off = 0
        for cin, cout in o.UNET_LAYERS:
            n = cout * 27 * cin
            w = torch.tensor(params[off:off + n].reshape(cout, 27, cin).reshape(cout, 3, 3, 3, cin)
                             .transpose(0, 4, 1, 2, 3).copy(), requires_grad=True)
            b = torch.tensor(params[off + n:off + n + cout].copy(), requires_grad=True)
            self.W.append(w)
            self.B.append(b)
            off += n + cout
        self.scale = float(params[off])

    def forward(self, x):
        c = lambda i, t, relu=True: (F.relu if relu else (lambda v: v))(
            F.conv3d(t, self.W[i], self.B[i], padding=1))
        pool = lambda t: F.avg_pool3d(t, 3, stride=2, padding=1, count_include_pad=True)
        up = lambda t: F.interpolate(t, size=tuple(2 * s - 1 for s in t.shape[2:]), mode="trilinear",
                                     align_corners=True)
        a1 = c(0, x)
        a2 = c(1, a1)
        a3 = c(2, pool(a2))
        a4 = c(3, a3)
        a5 = c(4, pool(a4))
        a6 = c(5, a5)
        a7 = c(6, torch.cat([up(a6), a4], 1))
        a8 = c(7, a7)
        a9 = c(8, torch.cat([up(a8), a2], 1))
        a10 = c(9, a9)
        return c(10, a10, relu=False)


def unet_case(seed, shape=(9, 5, 5)):
    rng = np.random.default_rng(seed)
    params = o.unet_init(seed, scale=0.7)
    params[:-1] += 0.01 * rng.standard_normal(params.size - 1)   # nonzero biases
    feat = rng.standard_normal(shape + (8,))
    target = rng.standard_normal(shape + (3,))
    return params, feat, target


def test_unet_matches_torch_network():
    params, feat, target = unet_case(6)
    net = TorchUNet(params)
    h = net(to_t(feat))
    loss_t = ((h - to_t(target) / net.scale) ** 2).mean()
    loss_t.backward()
    loss, y, grad = o.unet_loss(params, feat, target)
    assert abs(loss - loss_t.item()) <= 1e-12 * abs(loss)
    assert np.allclose(y, net.scale * from_t(h), rtol=1e-11, atol=1e-12)
    want = np.concatenate([t for w, b in zip(net.W, net.B)
                           for t in (w.grad.permute(0, 2, 3, 4, 1).reshape(-1).numpy(), b.grad.numpy())])
    assert np.allclose(grad, want, rtol=1e-9, atol=1e-13 * np.abs(want).max())


def test_unet_gradient_finite_differences():
    params, feat, target = unet_case(7, shape=(5, 5, 5))
    _, _, grad = o.unet_loss(params, feat, target)
    rng = np.random.default_rng(8)
    offs = np.cumsum([0] + [cout * 27 * cin + cout for cin, cout in o.UNET_LAYERS])
    idx = [int(rng.integers(offs[l], offs[l + 1])) for l in range(len(o.UNET_LAYERS))]
    idx += [int(offs[l + 1] - 1) for l in range(len(o.UNET_LAYERS))]   # a bias of each layer
    h = 1e-6
    bad = []
    for i in idx:
        pp, pm = params.copy(), params.copy()
        pp[i] += h
        pm[i] -= h
        fd = (o.unet_loss(pp, feat, target, grad=False)[0] - o.unet_loss(pm, feat, target, grad=False)[0]) / (2 * h)
        if abs(fd - grad[i]) > 1e-6 * abs(grad).max() + 1e-4 * abs(fd):
            bad.append((i, fd, grad[i]))
    assert not bad, bad


def test_unet_rounded_close_to_exact():
    """bfloat16 storage changes the output by a few bf16 ulps of its size."""
    params, feat, target = unet_case(9)
    l0, y0, g0 = o.unet_loss(params, feat, target)
    l1, y1, g1 = o.unet_loss(params, feat, target, round_bf16=True)
    assert np.max(np.abs(y1 - y0)) < 0.05 * np.max(np.abs(y0))
    assert abs(l1 - l0) < 0.02 * l0
    assert np.max(np.abs(g1 - g0)) < 0.1 * np.max(np.abs(g0))


def test_features_sample_the_fine_grid():
    from conftest import small_braid
    import glemur_amd as g
    p = small_braid(17, 13, 21)
    B0, _ = g.init_state(p)
    f = o.unet_features(B0, (17, 13, 21), 2, (9, 7, 11))
    b = B0.reshape(23, 15, 19, 3)[1:-1:2, 1:-1:2, 1:-1:2]
    assert np.array_equal(f[..., 0], b[..., 0]) and np.array_equal(f[..., 2], b[..., 2] - 1.0)
    assert f[0, 0, 0, 3] == -1.0 and f[-1, -1, -1, 3] == 1.0 and f[-1, 0, 0, 5] == 1.0
    assert np.all(f[..., 6:] == 0)
