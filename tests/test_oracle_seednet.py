"""Pins the float64 oracle of the learned seed network (oracle/seednet.py).

The network has no reference implementation; its layers are textbook
operations.  Each backward is checked against central finite differences of
its forward, the convolution against a direct loop, the upsampling against
closed forms, and the whole network's parameter gradient against finite
differences of its loss.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import pytest

from oracle import seednet as net


def _fd(f, x, idx, eps=1e-6):
    xp, xm = x.copy(), x.copy()
    xp[idx] += eps
    xm[idx] -= eps
    return (f(xp) - f(xm)) / (2 * eps)


def _some(shape, rng, k=6):
    flat = rng.choice(int(np.prod(shape)), size=min(k, int(np.prod(shape))), replace=False)
    return [np.unravel_index(i, shape) for i in flat]


@pytest.mark.parametrize("stride", [1, 2])
def test_conv_forward_matches_direct_loop(stride):
    rng = np.random.default_rng(stride)
    x = rng.normal(size=(2, 5, 7, 3))
    w = rng.normal(size=(4, 3, 3, 3))
    b = rng.normal(size=4)
    y = net.conv_forward(x, w, b, stride, False)
    Ho, Wo = (5 - 1) // stride + 1, (7 - 1) // stride + 1
    assert y.shape == (2, Ho, Wo, 4)
    for n in range(2):
        for oy in range(Ho):
            for ox in range(Wo):
                for co in range(4):
                    s = b[co]
                    for ky in range(3):
                        for kx in range(3):
                            iy, ix = stride * oy + ky - 1, stride * ox + kx - 1
                            if 0 <= iy < 5 and 0 <= ix < 7:
                                s += np.dot(w[co, ky, kx], x[n, iy, ix])
                    assert abs(y[n, oy, ox, co] - s) < 1e-12


@pytest.mark.parametrize("stride,leaky", [(1, True), (2, True), (2, False)])
def test_conv_backward_matches_finite_differences(stride, leaky):
    rng = np.random.default_rng(10 + stride)
    x = rng.normal(size=(2, 6, 5, 3))
    w = rng.normal(size=(4, 3, 3, 3))
    b = rng.normal(size=4)
    y = net.conv_forward(x, w, b, stride, leaky)
    g = rng.normal(size=y.shape)
    d = net.leaky_backward(y, g) if leaky else g
    dx, dw, db = net.conv_backward(x, w, d, stride)
    for idx in _some(x.shape, rng):
        assert abs(_fd(lambda v: np.sum(g * net.conv_forward(v, w, b, stride, leaky)), x, idx) - dx[idx]) < 1e-6
    for idx in _some(w.shape, rng):
        assert abs(_fd(lambda v: np.sum(g * net.conv_forward(x, v, b, stride, leaky)), w, idx) - dw[idx]) < 1e-6
    for idx in _some(b.shape, rng, 4):
        assert abs(_fd(lambda v: np.sum(g * net.conv_forward(x, w, v, stride, leaky)), b, idx) - db[idx]) < 1e-6


def test_correlation_forward_and_backward():
    rng = np.random.default_rng(3)
    f1 = rng.normal(size=(2, 5, 6, 4))
    f2 = rng.normal(size=(2, 5, 6, 4))
    r, sc = 2, 0.25
    c = net.corr_forward(f1, f2, r, sc)
    # Offset (dy, dx) = (1, -2) at pixel (2, 3): f2 at (3, 1).
    d = (1 + r) * (2 * r + 1) + (-2 + r)
    assert abs(c[1, 2, 3, d] - sc * np.dot(f1[1, 2, 3], f2[1, 3, 1])) < 1e-12
    assert c[0, 4, 0, (2 + r) * (2 * r + 1) + (0 + r)] == 0.0     # below the image
    g = rng.normal(size=c.shape)
    d1, d2 = net.corr_backward(f1, f2, g, r, sc)
    for idx in _some(f1.shape, rng):
        assert abs(_fd(lambda v: np.sum(g * net.corr_forward(v, f2, r, sc)), f1, idx) - d1[idx]) < 1e-6
    for idx in _some(f2.shape, rng):
        assert abs(_fd(lambda v: np.sum(g * net.corr_forward(f1, v, r, sc)), f2, idx) - d2[idx]) < 1e-6


def test_warp_backward_and_integer_shifts():
    rng = np.random.default_rng(4)
    f = rng.normal(size=(1, 6, 7, 3))
    flow = np.zeros((1, 6, 7, 2))
    flow[..., 0] = 1.0    # integer shift: exact copy of the right neighbour
    out = net.warp_forward(f, flow)
    np.testing.assert_allclose(out[0, :, :-1], f[0, :, 1:], atol=1e-15)
    assert np.all(out[0, :, -1] == 0)
    flow = rng.uniform(-1.5, 1.5, size=(1, 6, 7, 2))
    g = rng.normal(size=f.shape)
    df = net.warp_backward(f.shape, flow, g)
    for idx in _some(f.shape, rng):
        assert abs(_fd(lambda v: np.sum(g * net.warp_forward(v, flow)), f, idx) - df[idx]) < 1e-6


def test_upsample_is_exact_on_linear_fields():
    yy, xx = np.meshgrid(np.arange(5.0), np.arange(6.0), indexing="ij")
    coarse = np.stack([0.3 * xx - 0.2 * yy + 1.0, 0.1 * xx + 0.4 * yy], axis=-1)[None]
    fine = net.upsample_flow(coarse, 9, 11)
    fy, fx = np.meshgrid(np.arange(9.0), np.arange(11.0), indexing="ij")
    expect = 2.0 * np.stack([0.3 * fx / 2 - 0.2 * fy / 2 + 1.0, 0.1 * fx / 2 + 0.4 * fy / 2], axis=-1)
    np.testing.assert_allclose(fine[0], expect, atol=1e-12)


def test_huber_gradient():
    rng = np.random.default_rng(5)
    f = rng.normal(scale=2.0, size=(2, 3, 4, 2))
    t = rng.normal(size=f.shape)
    loss, g = net.huber(f, t, 0.7)
    for idx in _some(f.shape, rng):
        assert abs(_fd(lambda v: net.huber(v, t, 0.7)[0], f, idx) - g[idx]) < 1e-6


@dataclass(frozen=True)
class _Stage:
    level: int
    radius: int
    head: tuple


@dataclass(frozen=True)
class _Arch:
    encoder: tuple = ((4, 2), (6, 2))
    stages: tuple = (_Stage(1, 1, ((5, True), (2, False))), _Stage(0, 1, ((5, True), (2, False))))


def test_network_gradient_matches_finite_differences():
    rng = np.random.default_rng(6)
    arch = _Arch()
    params = {}
    cin = 1
    for i, (cout, _) in enumerate(arch.encoder):
        params[f"enc{i}"] = (rng.normal(scale=0.5, size=(cout, 3, 3, cin)), rng.normal(scale=0.1, size=cout))
        cin = cout
    for si, st in enumerate(arch.stages):
        cin = (2 * st.radius + 1) ** 2 + arch.encoder[st.level][0] + (2 if si else 0)
        for hi, (cout, _) in enumerate(st.head):
            params[f"s{si}h{hi}"] = (rng.normal(scale=0.3, size=(cout, 3, 3, cin)), rng.normal(scale=0.1, size=cout))
            cin = cout
    ref = rng.normal(size=(2, 12, 12, 1))
    tgt = rng.normal(size=(2, 12, 12, 1))
    truths = [rng.normal(size=(2, 3, 3, 2)), rng.normal(size=(2, 6, 6, 2))]
    weights = (0.5, 0.25)
    loss, grads, flow = net.forward_backward(params, arch, ref, tgt, truths, weights)
    ups = net.forward_backward.last_ups
    assert flow.shape == (2, 6, 6, 2)
    for name in ("enc0", "enc1", "s0h0", "s1h0", "s1h1"):
        w, b = params[name]
        for idx in _some(w.shape, rng, 3):
            def f(v):
                p = dict(params)
                p[name] = (v, b)
                return net.forward_backward(p, arch, ref, tgt, truths, weights, ups=ups)[0]
            assert abs(_fd(f, w, idx) - grads[name][0][idx]) < 2e-6, (name, idx)
