"""float64 NumPy restatement of the learned seed network's layers and of the
network itself, with their gradients.  TEST INFRASTRUCTURE ONLY.

The network has no counterpart in the reference; its layers are textbook
operations, and this module is pinned by central finite differences
(tests/test_oracle_seednet.py) before it checks the HIP kernels.
Conventions (shared with the kernels): NHWC tensors; 3x3 convolutions with
padding 1 and stride s, ``y[n,oy,ox,co] = b[co] + sum w[co,ky,kx,ci] x[n, s*oy+ky-1,
s*ox+kx-1, ci]``; leaky ReLU slope 0.1; local correlation over offsets
(dy, dx) in [-r, r]^2, dy outer; bilinear warps read zero outside the image.
"""

from __future__ import annotations

import numpy as np

LEAKY = 0.1


def _pad(x):
    return np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))


def out_size(n: int, stride: int) -> int:
    return (n - 1) // stride + 1


def im2col(x: np.ndarray, stride: int) -> np.ndarray:
    """(N*Ho*Wo, 9*Cin) with column tap*Cin + ci, tap = 3 ky + kx."""
    N, H, W, C = x.shape
    Ho, Wo = out_size(H, stride), out_size(W, stride)
    xp = _pad(x)
    cols = np.empty((N, Ho, Wo, 9, C), dtype=np.float64)
    for ky in range(3):
        for kx in range(3):
            cols[:, :, :, 3 * ky + kx, :] = xp[:, ky:ky + stride * (Ho - 1) + 1:stride,
                                               kx:kx + stride * (Wo - 1) + 1:stride, :]
    return cols.reshape(N * Ho * Wo, 9 * C)


def col2im(cols: np.ndarray, shape, stride: int) -> np.ndarray:
    N, H, W, C = shape
    Ho, Wo = out_size(H, stride), out_size(W, stride)
    xp = np.zeros((N, H + 2, W + 2, C))
    c = cols.reshape(N, Ho, Wo, 9, C)
    for ky in range(3):
        for kx in range(3):
            xp[:, ky:ky + stride * (Ho - 1) + 1:stride, kx:kx + stride * (Wo - 1) + 1:stride, :] += c[:, :, :, 3 * ky + kx]
    return xp[:, 1:-1, 1:-1, :]


def conv_forward(x, w, b, stride, leaky):
    """w: (Cout, 3, 3, Cin).  Returns the output (after the activation when leaky)."""
    N, H, W, C = x.shape
    Cout = w.shape[0]
    y = im2col(x, stride) @ w.reshape(Cout, 9 * C).T + b
    if leaky:
        y = np.where(y > 0, y, LEAKY * y)
    return y.reshape(N, out_size(H, stride), out_size(W, stride), Cout)


def leaky_backward(y, dy):
    return np.where(y > 0, dy, LEAKY * dy)


def conv_backward(x, w, dy_pre, stride):
    """Gradients of x, w, b from the pre-activation gradient."""
    Cout = w.shape[0]
    C = x.shape[-1]
    d = dy_pre.reshape(-1, Cout)
    cols = im2col(x, stride)
    dw = (d.T @ cols).reshape(w.shape)
    db = d.sum(axis=0)
    dx = col2im(d @ w.reshape(Cout, 9 * C), x.shape, stride)
    return dx, dw, db


def _shifted(f, dy, dx):
    """g[n, y, x] = f[n, y + dy, x + dx], zero outside."""
    N, H, W, C = f.shape
    g = np.zeros_like(f)
    ys, ye = max(0, -dy), min(H, H - dy)
    xs, xe = max(0, -dx), min(W, W - dx)
    if ys < ye and xs < xe:
        g[:, ys:ye, xs:xe] = f[:, ys + dy:ye + dy, xs + dx:xe + dx]
    return g


def corr_forward(f1, f2, radius, scale):
    D = 2 * radius + 1
    out = np.empty(f1.shape[:3] + (D * D,))
    for d in range(D * D):
        dy, dx = d // D - radius, d % D - radius
        out[..., d] = scale * np.sum(f1 * _shifted(f2, dy, dx), axis=-1)
    return out


def corr_backward(f1, f2, dc, radius, scale):
    D = 2 * radius + 1
    df1 = np.zeros_like(f1)
    df2 = np.zeros_like(f2)
    for d in range(D * D):
        dy, dx = d // D - radius, d % D - radius
        g = dc[..., d:d + 1] * scale
        df1 += g * _shifted(f2, dy, dx)
        df2 += _shifted(g * f1, -dy, -dx)
    return df1, df2


def _bilinear_taps(flow, H, W):
    N = flow.shape[0]
    yy, xx = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    sx = xx[None] + flow[..., 0]
    sy = yy[None] + flow[..., 1]
    x0, y0 = np.floor(sx).astype(int), np.floor(sy).astype(int)
    ax, ay = sx - x0, sy - y0
    taps = []
    for j in (0, 1):
        for i in (0, 1):
            xi, yi = x0 + i, y0 + j
            wgt = (ax if i else 1 - ax) * (ay if j else 1 - ay)
            ok = (xi >= 0) & (xi < W) & (yi >= 0) & (yi < H)
            taps.append((np.clip(xi, 0, W - 1), np.clip(yi, 0, H - 1), np.where(ok, wgt, 0.0)))
    return np.arange(N)[:, None, None], taps


def warp_forward(f, flow):
    N, H, W, C = f.shape
    nn_, taps = _bilinear_taps(flow, H, W)
    out = np.zeros_like(f)
    for xi, yi, wgt in taps:
        out += wgt[..., None] * f[nn_, yi, xi]
    return out


def warp_backward(f_shape, flow, dout):
    N, H, W, C = f_shape
    nn_, taps = _bilinear_taps(flow, H, W)
    df = np.zeros(f_shape)
    for xi, yi, wgt in taps:
        np.add.at(df, (np.broadcast_to(nn_, xi.shape), yi, xi), wgt[..., None] * dout)
    return df


def bilinear_flow(flow, x, y):
    """Clamped bilinear sample of a (H, W, 2) flow at (x, y) arrays."""
    H, W, _ = flow.shape
    x = np.clip(x, 0, W - 1)
    y = np.clip(y, 0, H - 1)
    x0 = np.minimum(np.floor(x).astype(int), max(W - 2, 0))
    y0 = np.minimum(np.floor(y).astype(int), max(H - 2, 0))
    x1, y1 = np.minimum(x0 + 1, W - 1), np.minimum(y0 + 1, H - 1)
    ax, ay = (x - x0)[..., None], (y - y0)[..., None]
    return ((1 - ay) * ((1 - ax) * flow[y0, x0] + ax * flow[y0, x1])
            + ay * ((1 - ax) * flow[y1, x0] + ax * flow[y1, x1]))


def upsample_flow(coarse, fine_h, fine_w):
    yy, xx = np.meshgrid(np.arange(fine_h), np.arange(fine_w), indexing="ij")
    return np.stack([2.0 * bilinear_flow(c, 0.5 * xx, 0.5 * yy) for c in coarse])


def huber(flow, truth, weight):
    """Loss and gradient of the Huber (delta 1) flow loss, mean over vectors."""
    count = flow.size // 2
    d = flow - truth
    ad = np.abs(d)
    loss = weight / count * np.sum(np.where(ad < 1, 0.5 * d * d, ad - 0.5))
    return loss, weight / count * np.clip(d, -1, 1)


# -- the network -------------------------------------------------------------------------------
def forward_backward(params: dict, arch, ref, tgt, truths, weights, ups=None):
    """Loss, parameter gradients and final flow of the network (float64).

    ``params``: name -> (w (Cout,3,3,Cin), b); ``arch``: the architecture
    description of ``live_dic.seednet.Architecture``; ``ref``/``tgt``: (N, H, W, 1)
    normalised images; ``truths``: per stage flow targets (N, h, w, 2).  The
    upsampled flows are constants of the gradient (stage-wise training); pass
    ``ups`` (as returned in ``forward_backward.last_ups``) to hold them fixed,
    e.g. for finite differences.
    """
    enc = {}
    for side, x in (("ref", ref), ("tgt", tgt)):
        feats = []
        for i, (cout, stride) in enumerate(arch.encoder):
            w, b = params[f"enc{i}"]
            x = conv_forward(x, w, b, stride, True)
            feats.append(x)
        enc[side] = feats
    grads = {k: (np.zeros_like(v[0]), np.zeros_like(v[1])) for k, v in params.items()}
    denc = {s: [np.zeros_like(f) for f in enc[s]] for s in enc}
    loss = 0.0
    flow = None
    tape = []
    used_ups = []
    for si, stage in enumerate(arch.stages):
        lvl = stage.level
        f1, f2 = enc["ref"][lvl], enc["tgt"][lvl]
        N, H, W, C = f1.shape
        if flow is None:
            up = None
            f2w = f2
        else:
            up = upsample_flow(flow, H, W) if ups is None else ups[si]
            f2w = warp_forward(f2, up)
        used_ups.append(up)
        scale = 1.0 / C
        corr = corr_forward(f1, f2w, stage.radius, scale)
        parts = [corr, f1] + ([up] if up is not None else [])
        x = np.concatenate(parts, axis=-1)
        acts = [x]
        for hi, (cout, act) in enumerate(stage.head):
            w, b = params[f"s{si}h{hi}"]
            x = conv_forward(x, w, b, 1, act)
            acts.append(x)
        out = acts[-1] + (up if up is not None else 0.0)
        l, g = huber(out, truths[si], weights[si])
        loss += l
        tape.append((si, stage, acts, f1, f2, f2w, up, scale, g))
        flow = out
    for si, stage, acts, f1, f2, f2w, up, scale, g in reversed(tape):
        d = g
        for hi in reversed(range(len(stage.head))):
            cout, act = stage.head[hi]
            w, b = params[f"s{si}h{hi}"]
            if act:
                d = leaky_backward(acts[hi + 1], d)
            dx, dw, db = conv_backward(acts[hi], w, d, 1)
            grads[f"s{si}h{hi}"][0][...] += dw
            grads[f"s{si}h{hi}"][1][...] += db
            d = dx
        D2 = (2 * stage.radius + 1) ** 2
        C = f1.shape[-1]
        dcorr, df1 = d[..., :D2], d[..., D2:D2 + C]
        g1, g2w = corr_backward(f1, f2w, dcorr, stage.radius, scale)
        denc["ref"][stage.level] += df1 + g1
        if up is None:
            denc["tgt"][stage.level] += g2w
        else:
            denc["tgt"][stage.level] += warp_backward(f2.shape, up, g2w)
    for side, x0 in (("ref", ref), ("tgt", tgt)):
        feats = enc[side]
        for i in reversed(range(len(arch.encoder))):
            w, b = params[f"enc{i}"]
            d = leaky_backward(feats[i], denc[side][i])
            x_in = x0 if i == 0 else feats[i - 1]
            dx, dw, db = conv_backward(x_in, w, d, arch.encoder[i][1])
            grads[f"enc{i}"][0][...] += dw
            grads[f"enc{i}"][1][...] += db
            if i > 0:
                denc[side][i - 1] += dx
    forward_backward.last_ups = used_ups
    return loss, grads, flow
