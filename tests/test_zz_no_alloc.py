"""GPU: the allocation-free split-K convolution (ldic_nn_conv_forward_split_static) against the
workspace entry point (ldic_nn_conv_forward_split), through the C ABI, on two calls in a row, the
second smaller; and its size limit."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bf16(values: np.ndarray) -> np.ndarray:
    """float32 -> bfloat16 bit patterns (round to nearest even)."""
    u = np.ascontiguousarray(values, np.float32).view(np.uint32)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def _case(batch, height, width, cin, cout, stride, seed):
    from live_dic import _native

    rng = np.random.default_rng(seed)
    oh, ow = (height - 1) // stride + 1, (width - 1) // stride + 1
    conv = _native.NNConv(batch, height, width, cin, cin, 0, oh, ow, cout, cout, 0, stride, 0)
    kp = (9 * cin + 15) & ~15
    x = _bf16(rng.normal(0, 1, (batch, height, width, cin)))
    w = np.zeros((cout, kp), np.float32)
    w[:, :9 * cin] = rng.normal(0, 0.2, (cout, 9 * cin))
    w = _bf16(w)
    bias = rng.normal(0, 0.1, cout).astype(np.float32)
    return conv, x, w, bias, batch * oh * ow * cout


def _run(lib, entry, conv, x, w, bias, count, splits, workspace=None):
    from live_dic import _native

    dx, dw, db = (_native.DeviceBuffer(a.nbytes) for a in (x, w, bias))
    dx.upload(x)
    dw.upload(w)
    db.upload(bias)
    dy = _native.DeviceBuffer(count * 4)
    if entry == "workspace":
        rc = lib.ldic_nn_conv_forward_split(ctypes.byref(conv), dx.address, dw.address, db.address, 1, None,
                                            dy.address, workspace.address, splits, None)
    else:
        rc = lib.ldic_nn_conv_forward_split_static(ctypes.byref(conv), dx.address, dw.address, db.address, 1, None,
                                                   dy.address, splits, None)
    _native.check(rc)
    _native.check(lib.ldic_device_synchronize())
    out = np.empty(count, np.float32)
    dy.download(out)
    for b in (dx, dw, db, dy):
        b.close()
    return out


def test_static_scratch_forward_matches_workspace_forward_twice():
    from live_dic import _native

    lib = _native.library()
    for batch, h, w, cin, cout, stride, splits in ((2, 33, 29, 24, 48, 1, 4), (1, 12, 9, 16, 32, 2, 3)):
        print("[case] split-K conv", batch, h, w, cin, cout, stride, splits, flush=True)
        conv, x, wt, bias, count = _case(batch, h, w, cin, cout, stride, 100 + h)
        ws = _native.DeviceBuffer(splits * count * 4)
        want = _run(lib, "workspace", conv, x, wt, bias, count, splits, ws)
        ws.close()
        got = _run(lib, "static", conv, x, wt, bias, count, splits)
        # The same kernels add the split-K slices in the same order: identical bits.
        np.testing.assert_array_equal(got, want)
        assert np.abs(want).max() > 0.1


def test_static_scratch_refuses_oversized_partial_sums():
    from live_dic import _native

    lib = _native.library()
    conv = _native.NNConv(1, 128, 128, 16, 16, 0, 128, 128, 32, 32, 0, 1, 0)   # 3 x 524,288 partial sums
    buf = _native.DeviceBuffer(1 << 22)
    rc = lib.ldic_nn_conv_forward_split_static(ctypes.byref(conv), buf.address, buf.address, buf.address, 0, None,
                                               buf.address, 3, None)
    assert rc == -1
    assert b"static scratch" in lib.ldic_last_error()
    buf.close()
