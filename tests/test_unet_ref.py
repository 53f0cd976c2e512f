"""The CPU reference of the learned-proposal network, pinned without the GPU.

* finite differences: the hand-derived backward pass of tests/unet_ref.py
  against central differences of its own float64 forward pass;
* an independent autograd cross-check: the same network written with
  torch.nn.functional on the CPU (float64) must give the same loss and
  gradients.
Small widths keep these fast; the structure (levels, skips, pooling,
upsampling, head) is the library's.
"""
import numpy as np
import pytest

import unet_ref as ur

WIDTHS = (8, 8, 16, 16)


def sample(seed, B=2, H=16, W=16):
    rng = np.random.default_rng(seed)
    terrain = rng.random((B, H, W))
    params = rng.uniform(-1, 1, (B, 3))
    target = rng.normal(0, 1, (B, H, W))
    return terrain, params, target


def test_param_count_matches_library_layout():
    assert ur.param_count() == sum(9 * ci * co + co for ci, co, _ in ur.layer_shapes()) + 33
    shapes = ur.layer_shapes()
    assert len(shapes) == 14
    assert shapes[0] == (8, 32, 0) and shapes[7] == (160, 160, 3)
    assert shapes[8] == (160 + 128, 128, 2) and shapes[13] == (32, 32, 0)
    assert 1_000_000 < ur.param_count() < 2_000_000


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -2.5, 3.0e-39], dtype=np.float32)
    r = ur.bf16_round(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa wins (1.0)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125 and r[3] == 1.0078125
    assert r[4] == -2.5
    assert (r.view(np.uint32) & 0xFFFF == 0).all()


def test_backward_matches_finite_differences():
    terrain, params, target = sample(1)
    p = ur.init_params(3, WIDTHS)
    p += np.random.default_rng(4).normal(0, 0.05, p.size)   # nonzero biases too
    net = ur.Net(p, WIDTHS)
    loss, g = net.loss_grad(terrain, params, target)
    rng = np.random.default_rng(5)
    # a few indices in every layer's weights and biases, and the head
    idx = []
    off = 0
    for ci, co, _ in ur.layer_shapes(WIDTHS):
        nw = 9 * ci * co
        idx += list(off + rng.choice(nw, 3, replace=False))
        idx.append(off + nw + int(rng.integers(co)))
        off += nw + co
    idx += [off + 1, off + WIDTHS[0]]
    h = 1e-6
    bad = 0
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (ur.Net(pp, WIDTHS).loss_grad(terrain, params, target)[0] -
              ur.Net(pm, WIDTHS).loss_grad(terrain, params, target)[0]) / (2 * h)
        err = abs(fd - g[i]) / max(abs(fd), abs(g[i]), 1e-8)
        if err > 1e-5:
            bad += 1
        assert err < 1e-2, (i, fd, g[i])
    assert bad <= 2, bad


def test_matches_torch_autograd():
    torch = pytest.importorskip("torch")
    F = torch.nn.functional
    terrain, params, target = sample(2, B=2, H=32, W=32)
    p = ur.init_params(7, WIDTHS) + np.random.default_rng(8).normal(0, 0.05, ur.param_count(WIDTHS))
    loss, g = ur.Net(p, WIDTHS).loss_grad(terrain, params, target)

    pt = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    convs, off = [], 0
    for ci, co, _ in ur.layer_shapes(WIDTHS):
        w = pt[off:off + 9 * co * ci].view(3, 3, co, ci).permute(2, 3, 0, 1)
        off += 9 * co * ci
        convs.append((w, pt[off:off + co]))
        off += co
    hw, hb = pt[off:off + WIDTHS[0]], pt[off + WIDTHS[0]]
    B, H, W = terrain.shape
    x = torch.zeros(B, 8, H, W, dtype=torch.float64)
    x[:, 0] = torch.tensor(terrain)
    x[:, 1:4] = torch.tensor(params)[:, :, None, None]

    def conv(i, t):
        return F.relu(F.conv2d(t, convs[i][0], convs[i][1], padding=1))

    L = len(WIDTHS)
    e, inp = [], x
    for l in range(L):
        t = conv(2 * l + 1, conv(2 * l, inp))
        e.append(t)
        if l < L - 1:
            inp = F.avg_pool2d(t, 2)
    lo = e[-1]
    for l in range(L - 2, -1, -1):
        k = 2 * L + 2 * (L - 2 - l)
        u = torch.cat([F.interpolate(lo, scale_factor=2, mode="nearest"), e[l]], dim=1)
        lo = conv(k + 1, conv(k, u))
    y = torch.einsum("bchw,c->bhw", lo, hw) + hb
    lt = ((y - torch.tensor(target)) ** 2).mean()
    lt.backward()
    assert abs(lt.item() - loss) <= 1e-12 * max(1.0, abs(loss))
    np.testing.assert_allclose(g, pt.grad.numpy(), rtol=1e-9, atol=1e-12)
