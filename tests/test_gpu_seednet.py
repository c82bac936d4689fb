"""GPU: the learned seed network's kernels and gradients against the float64 oracle
(oracle/seednet.py, pinned by finite differences in tests/test_oracle_seednet.py).

Operands are bfloat16-representable, so the kernels and the oracle see the same
inputs; the kernels round activations to bfloat16 and accumulate in float32,
which sets the tolerances below.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import pytest

from oracle import seednet as onet

pytestmark = pytest.mark.gpu


def _say(*args):
    print("[case]", *args, flush=True)


def bf16(a):
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)
    u = (u + np.uint32(0x7FFF) + ((u >> np.uint32(16)) & np.uint32(1))) & np.uint32(0xFFFF0000)
    return u.view(np.float32)


@pytest.fixture(scope="module")
def env():
    from live_dic import _native

    return _native, _native.library()


def _dev(a, dtype=None):
    """Device copy of a host array (dtype BF16: rounded to bfloat16)."""
    from live_dic import _native

    return _native.DeviceArray.from_numpy(np.ascontiguousarray(a), dtype)


def _zeros(shape, dtype):
    from live_dic import _native

    return _native.DeviceArray.zeros(shape, dtype)


def _sync():
    from live_dic import _native

    _native.check(_native.library().ldic_device_synchronize())


BF16 = "bfloat16"


def _kp(k):
    return (k + 15) // 16 * 16


def _pack_weights(w):
    """(Cout, 3, 3, Cin) -> (Cout, KP) with column tap*Cin + ci."""
    cout, _, _, cin = w.shape
    out = np.zeros((cout, _kp(9 * cin)), np.float32)
    out[:, :9 * cin] = w.reshape(cout, 9 * cin)
    return out


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("n,h,w,cin,cout,stride,leaky,stored,offset", [
    (2, 19, 23, 1, 16, 2, True, 1, 0),        # first encoder layer: scalar gather
    (2, 17, 12, 16, 32, 2, True, 16, 0),      # vector gather
    (1, 9, 11, 53, 40, 1, True, 60, 4),       # channel slice of a wider buffer
    (2, 8, 9, 24, 2, 1, False, 24, 0),        # flow head: float32 output, 2 channels
])
def test_conv_forward_and_gradients(env, n, h, w, cin, cout, stride, leaky, stored, offset):
    native, lib = env
    _say("conv", n, h, w, cin, cout, stride, leaky, stored, offset)
    rng = np.random.default_rng(cin * 7 + cout)
    x_full = np.zeros((n, h, w, stored), np.float32)
    x_full[..., offset:offset + cin] = bf16(rng.normal(size=(n, h, w, cin)))
    x = x_full[..., offset:offset + cin].astype(np.float64)
    wt = bf16(rng.normal(scale=1 / math.sqrt(9 * cin), size=(cout, 3, 3, cin)))
    b = rng.normal(scale=0.1, size=cout).astype(np.float32)
    ho, wo = (h - 1) // stride + 1, (w - 1) // stride + 1
    g = native.NNConv(n, h, w, cin, stored, offset, ho, wo, cout, cout, 0, stride)
    xd = _dev(x_full, BF16)
    wd = _dev(_pack_weights(wt), BF16)
    bd = _dev(b)
    f32_out = cout == 2
    y16 = _zeros((n, ho, wo, cout), BF16)
    y32 = _zeros((n, ho, wo, cout), np.float32)
    native.check(lib.ldic_nn_conv_forward(ctypes.byref(g), xd.data_ptr(), wd.data_ptr(), bd.data_ptr(),
                                          1 if leaky else 0, None if f32_out else y16.data_ptr(),
                                          y32.data_ptr() if f32_out else None, None))
    _sync()
    y = (y32 if f32_out else y16).numpy()
    y_ref = onet.conv_forward(x, wt.astype(np.float64), b.astype(np.float64), stride, leaky)
    assert _rel(y, y_ref) < (1e-6 if f32_out else 4e-3)
    # Split-K: the GEMM columns in three slices through a float32 workspace.
    work = _zeros(3 * n * ho * wo * cout, np.float32)
    y16.zero()
    y32.zero()
    native.check(lib.ldic_nn_conv_forward_split(ctypes.byref(g), xd.data_ptr(), wd.data_ptr(), bd.data_ptr(),
                                                1 if leaky else 0, None if f32_out else y16.data_ptr(),
                                                y32.data_ptr() if f32_out else None, work.data_ptr(), 3, None))
    _sync()
    ys = (y32 if f32_out else y16).numpy()
    assert _rel(ys, y_ref) < (1e-6 if f32_out else 4e-3)
    # The slices are added in a fixed order: a second run gives the same bits.
    native.check(lib.ldic_nn_conv_forward_split(ctypes.byref(g), xd.data_ptr(), wd.data_ptr(), bd.data_ptr(),
                                                1 if leaky else 0, None if f32_out else y16.data_ptr(),
                                                y32.data_ptr() if f32_out else None, work.data_ptr(), 3, None))
    np.testing.assert_array_equal((y32 if f32_out else y16).numpy(), ys)

    dy = rng.normal(size=(n, ho, wo, cout)).astype(np.float32)
    dy_b = bf16(dy).astype(np.float64)          # the kernels feed dy to the matrix cores as bfloat16
    dx_ref, dw_ref, db_ref = onet.conv_backward(x, wt.astype(np.float64), dy_b, stride)
    wtt = _zeros((cin, _kp(9 * cout)), BF16)
    native.check(lib.ldic_nn_transpose_weights(cin, cout, wd.data_ptr(), wtt.data_ptr(), None))
    dyd = _dev(dy)
    dx = _dev(np.full((n, h, w, stored), 7.0, np.float32))
    dwd = _zeros((cout, _kp(9 * cin)), np.float32)
    dbd = _zeros(cout, np.float32)
    native.check(lib.ldic_nn_conv_backward(ctypes.byref(g), xd.data_ptr(), wtt.data_ptr(), dyd.data_ptr(),
                                           dx.data_ptr(), stored, offset, 1, dwd.data_ptr(), dbd.data_ptr(), None))
    _sync()
    dxh = dx.numpy()
    # Accumulated into the slice, channels outside it untouched.
    assert _rel(dxh[..., offset:offset + cin] - 7.0, dx_ref) < 1e-5
    if stored > cin:
        assert np.all(np.delete(dxh, np.s_[offset:offset + cin], axis=-1) == 7.0)
    dwh = dwd.numpy()
    assert np.all(dwh[:, 9 * cin:] == 0.0)
    assert _rel(dwh[:, :9 * cin].reshape(dw_ref.shape), dw_ref) < 1e-5
    assert _rel(dbd.numpy(), dy.sum(axis=(0, 1, 2))) < 1e-5


def test_correlation_warp_upsample_loss_adam(env):
    native, lib = env
    _say("correlation, warp, upsample, loss, adam")
    rng = np.random.default_rng(11)
    n, h, w, c, r = 2, 9, 10, 12, 3
    f1 = bf16(rng.normal(size=(n, h, w, c)))
    f2 = bf16(rng.normal(size=(n, h, w, c)))
    D2 = (2 * r + 1) ** 2
    stored = D2 + 3
    spec = native.NNCorr(n, h, w, c, r, c, 0, c, 0, stored, 0, 1.0 / c)
    d1, d2 = _dev(f1, BF16), _dev(f2, BF16)
    out = _zeros((n, h, w, stored), BF16)
    native.check(lib.ldic_nn_corr_forward(ctypes.byref(spec), d1.data_ptr(), d2.data_ptr(), out.data_ptr(), None))
    ref = onet.corr_forward(f1.astype(np.float64), f2.astype(np.float64), r, 1.0 / c)
    _sync()
    assert _rel(out.numpy()[..., :D2], ref) < 4e-3
    dc = rng.normal(size=(n, h, w, stored)).astype(np.float32)
    g1_ref, g2_ref = onet.corr_backward(f1.astype(np.float64), f2.astype(np.float64), dc[..., :D2].astype(np.float64),
                                        r, 1.0 / c)
    g1 = _zeros((n, h, w, c), np.float32)
    g2 = _zeros((n, h, w, c), np.float32)
    dcd = _dev(dc)
    native.check(lib.ldic_nn_corr_backward(ctypes.byref(spec), d1.data_ptr(), d2.data_ptr(), dcd.data_ptr(), stored, 0,
                                           g1.data_ptr(), c, 0, g2.data_ptr(), c, 0, None))
    _sync()
    assert _rel(g1.numpy(), g1_ref) < 1e-5 and _rel(g2.numpy(), g2_ref) < 1e-5

    flow = rng.uniform(-2.5, 2.5, size=(n, h, w, 2)).astype(np.float32)
    fd = _dev(flow)
    wout = _zeros((n, h, w, c), BF16)
    native.check(lib.ldic_nn_warp_forward(n, h, w, c, d2.data_ptr(), fd.data_ptr(), wout.data_ptr(), None))
    gout = rng.normal(size=(n, h, w, c)).astype(np.float32)
    gw = _zeros((n, h, w, c), np.float32)
    native.check(lib.ldic_nn_warp_backward(n, h, w, c, fd.data_ptr(), _dev(gout).data_ptr(), gw.data_ptr(),
                                           None))
    _sync()
    assert _rel(wout.numpy(), onet.warp_forward(f2.astype(np.float64), flow.astype(np.float64))) < 4e-3
    assert _rel(gw.numpy(), onet.warp_backward(f2.shape, flow.astype(np.float64), gout.astype(np.float64))) < 1e-5

    coarse = rng.normal(size=(n, 5, 5, 2)).astype(np.float32)
    fine = _zeros((n, h, w, 2), np.float32)
    native.check(lib.ldic_nn_upsample_flow(n, 5, 5, h, w, _dev(coarse).data_ptr(), fine.data_ptr(), None, 0, 0,
                                           None))
    _sync()
    assert _rel(fine.numpy(), onet.upsample_flow(coarse.astype(np.float64), h, w)) < 1e-6

    truth = rng.normal(size=flow.shape).astype(np.float32)
    grad = _zeros(fd.shape, np.float32)
    loss = _zeros(1, np.float32)
    native.check(lib.ldic_nn_flow_loss(n * h * w, fd.data_ptr(), _dev(truth).data_ptr(), 0.5, grad.data_ptr(),
                                       loss.data_ptr(), None))
    l_ref, g_ref = onet.huber(flow.astype(np.float64), truth.astype(np.float64), 0.5)
    _sync()
    assert abs(float(loss.item()) - l_ref) < 1e-5 * max(1.0, l_ref)
    assert _rel(grad.numpy(), g_ref) < 1e-6

    # Adam against Algorithm 1 of Kingma & Ba (2015), in float64, over three steps.
    p0 = rng.normal(size=257).astype(np.float32)
    grads = [rng.normal(size=257).astype(np.float32) for _ in range(3)]
    p = _dev(p0)
    m = _zeros(p.shape, np.float32)
    v = _zeros(p.shape, np.float32)
    want, mr, vr = p0.astype(np.float64), np.zeros(257), np.zeros(257)
    for k, gk in enumerate(grads):
        native.check(lib.ldic_nn_adam(257, p.data_ptr(), _dev(gk).data_ptr(), m.data_ptr(), v.data_ptr(), 1e-2,
                                      0.9, 0.999, 1e-8, k + 1, None, None))
        g64 = gk.astype(np.float64)
        mr = 0.9 * mr + 0.1 * g64
        vr = 0.999 * vr + 0.001 * g64 * g64
        mh, vh = mr / (1 - 0.9 ** (k + 1)), vr / (1 - 0.999 ** (k + 1))
        want = want - 1e-2 * mh / (np.sqrt(vh) + 1e-8)
    _sync()
    np.testing.assert_allclose(p.numpy(), want, rtol=1e-5, atol=1e-6)


def test_network_gradients_match_the_oracle(env):
    native, lib = env
    from live_dic import seednet

    _say("network gradient check, small architecture, 2 pairs of 48x40")
    arch = seednet.Architecture(encoder=((8, 2), (12, 2), (16, 2)),
                                stages=(seednet.Stage(2, 2, ((24, True), (2, False))),
                                        seednet.Stage(1, 2, ((24, True), (2, False)))), input_level=0)
    net = seednet.SeedNet(arch, seed=3)
    rng = np.random.default_rng(8)
    n, h, w = 2, 48, 40
    ref = rng.integers(0, 256, size=(n, h, w), dtype=np.uint8)
    tgt = np.roll(ref, (1, 2), (1, 2))
    rd, td = _dev(ref), _dev(tgt)
    flows = net.forward(rd.data_ptr(), td.data_ptr(), n=n, h=h, w=w, training=True)
    truths = [rng.normal(scale=0.5, size=tuple(f.shape)).astype(np.float32) for f in flows]
    weights = (0.6, 0.3)
    net.backward(truths, weights)
    _sync()
    params = net.oracle_parameters()
    x = lambda a: ((a.astype(np.float64) - 128.0) / 64.0)[..., None]
    loss, grads, flow = onet.forward_backward(params, arch, x(ref), x(tgt), [t.astype(np.float64)
                                                                            for t in truths], weights)
    got = net.gradients()
    _say("loss gpu/oracle", float(net.loss.item()), loss)
    assert abs(float(net.loss.item()) - loss) < 0.02 * loss
    assert _rel(flows[-1].numpy(), flow) < 0.03
    for name, (dw, db) in grads.items():
        gw, gb = got[name]
        cos = float(np.sum(gw * dw) / (np.linalg.norm(gw) * np.linalg.norm(dw) + 1e-30))
        _say(name, "rel", round(_rel(gw, dw), 4), "cos", round(cos, 5))
        assert cos > 0.995 and _rel(gw, dw) < 0.08, name
        assert _rel(gb, db) < 0.08, name
