"""FP64 NumPy restatement of the reference image preparation.  TEST INFRASTRUCTURE.

Follows ``_fir_weights``, ``_gaussian_sigma_one``, ``_spline_filter_axis``,
``_cubic_bspline_prefilter`` and ``_reference_arrays`` in
``live_dic/wgpu_backend.py:527-823`` and ``_reduce_half`` at ``:318-334``.
The product computes the same quantities on the GPU; this module only checks it.
"""

from __future__ import annotations

import numpy as np

FIR_RADIUS = 15
POLE = np.sqrt(3.0) - 2.0


def fir_weights() -> np.ndarray:
    """Gaussian(sigma=1, 7 taps) convolved with sqrt(3) z^|k|, |k| <= 12.

    wgpu_backend.py:527-541.  31 float32 weights, index = offset + 15.
    """
    offsets = np.arange(-3, 4, dtype=np.float64)
    gaussian = np.exp(-0.5 * offsets * offsets)
    gaussian /= gaussian.sum()
    weights = np.zeros(2 * FIR_RADIUS + 1, dtype=np.float64)
    for gi, off in enumerate(range(-3, 4)):
        for k in range(-12, 13):
            weights[off + k + FIR_RADIUS] += gaussian[gi] * np.sqrt(3.0) * POLE ** abs(k)
    return weights.astype(np.float32)


def gaussian_sigma_one(image: np.ndarray) -> np.ndarray:
    """Separable mirrored 7-tap Gaussian in float64 (wgpu_backend.py:544-561)."""
    offsets = np.arange(-3, 4, dtype=np.float64)
    w = np.exp(-0.5 * offsets * offsets)
    w /= w.sum()
    v = image.astype(np.float64)
    px = np.pad(v, ((0, 0), (3, 3)), mode="reflect")
    horizontal = 0.0
    for i in range(7):
        horizontal = horizontal + w[i] * px[:, i : i + v.shape[1]]
    py = np.pad(horizontal, ((3, 3), (0, 0)), mode="reflect")
    out = 0.0
    for i in range(7):
        out = out + w[i] * py[i : i + v.shape[0], :]
    return out


def spline_axis(c: np.ndarray, axis: int) -> None:
    """Causal then anti-causal cubic B-spline recursion in place (:564-579)."""
    v = np.swapaxes(c, axis, -1)
    gain = (1.0 - POLE) * (1.0 - 1.0 / POLE)
    v *= gain
    n = v.shape[-1]
    horizon = min(n, int(np.ceil(np.log(1e-12) / np.log(abs(POLE)))))
    powers = POLE ** np.arange(horizon, dtype=np.float64)
    v[..., 0] = np.sum(v[..., :horizon] * powers, axis=-1)
    for i in range(1, n):
        v[..., i] += POLE * v[..., i - 1]
    v[..., -1] = POLE / (POLE * POLE - 1.0) * (POLE * v[..., -2] + v[..., -1])
    for i in range(n - 2, -1, -1):
        v[..., i] = POLE * (v[..., i + 1] - v[..., i])


def bspline_coefficients(intensity: np.ndarray) -> np.ndarray:
    c = np.array(intensity, dtype=np.float64, copy=True, order="C")
    spline_axis(c, 1)
    spline_axis(c, 0)
    return c


def reduce_half_float(image: np.ndarray) -> np.ndarray:
    """Separable float32 [1,4,6,4,1]/16 reduction with truncation (:318-334)."""
    src = np.asarray(image, dtype=np.float32)
    h, w = src.shape
    rh, rw = h // 2, w // 2
    taps = np.asarray((1.0, 4.0, 6.0, 4.0, 1.0), dtype=np.float32) / 16.0
    px = np.pad(src, ((0, 0), (2, 2)), mode="reflect")
    hor = sum(taps[i] * px[:, i : i + 2 * rw : 2] for i in range(5))
    py = np.pad(hor, ((2, 2), (0, 0)), mode="reflect")
    red = sum(taps[i] * py[i : i + 2 * rh : 2] for i in range(5))
    return np.ascontiguousarray(np.clip(red, 0.0, 255.0).astype(np.uint8))


def grid_shape(width: int, height: int, size: int, stride: int) -> tuple[int, int]:
    """subset.py:51-65."""
    if width < size or height < size:
        return (0, 0)
    return (1 + (width - size) // stride, 1 + (height - size) // stride)


def _param_count(model: str) -> int:
    return 12 if model == "quadratic" else 6


def reference_arrays(
    image: np.ndarray,
    *,
    interpolation: str = "bspline",
    motion_model: str = "affine",
    subset_size: int = 16,
    subset_stride: int | None = None,
    spatial_weighting: str = "uniform",
    gaussian_sigma: float | None = None,
):
    """Reference plane (H, W, 4) float32, inverse (N, P*P) and metadata (N, M).

    Restates wgpu_backend.py:589-823: Gaussian + B-spline coefficients (or the
    raw image for direct interpolants), B-spline derivative gradients, per-subset
    anchored mean, energy, J^T f, sum J, H = sum w J J^T and its inverse.
    """
    if interpolation == "bspline":
        intensity = gaussian_sigma_one(image)
        source = bspline_coefficients(intensity)
    else:
        intensity = image.astype(np.float64)
        source = intensity
    p = np.pad(source, 1, mode="reflect")
    gx = 0.5 * (p[1:-1, 2:] - p[1:-1, :-2])
    gy = 0.5 * (p[2:, 1:-1] - p[:-2, 1:-1])
    if interpolation == "bspline":
        gx = (0.5 * (p[:-2, 2:] - p[:-2, :-2]) + 4.0 * gx + 0.5 * (p[2:, 2:] - p[2:, :-2])) / 6.0
        gy = (0.5 * (p[2:, :-2] - p[:-2, :-2]) + 4.0 * gy + 0.5 * (p[2:, 2:] - p[:-2, 2:])) / 6.0
    return subset_arrays(intensity, gx, gy, motion_model=motion_model, subset_size=subset_size,
                         subset_stride=subset_stride, spatial_weighting=spatial_weighting,
                         gaussian_sigma=gaussian_sigma)


def subset_arrays(intensity, gx, gy, *, motion_model="affine", subset_size=16, subset_stride=None,
                  spatial_weighting="uniform", gaussian_sigma=None):
    """The per-subset part of ``reference_arrays`` (wgpu_backend.py:650-823) for given float64
    intensity and gradient planes: the reference plane, inverse Hessians and metadata."""
    h, w = intensity.shape
    S = int(subset_size)
    D = S if subset_stride is None else int(subset_stride)
    half = 0.5 * (S - 1)
    gw, gh = grid_shape(w, h, S, D)
    sigma = 0.25 * S if gaussian_sigma is None else float(gaussian_sigma)
    ly, lx = np.meshgrid(np.arange(S, dtype=np.float64) - half, np.arange(S, dtype=np.float64) - half,
                         indexing="ij")
    lx = lx.ravel()
    ly = ly.ravel()
    if spatial_weighting == "gaussian":
        wts = np.exp(-0.5 * (lx * lx + ly * ly) / (sigma * sigma))
    else:
        wts = np.ones(lx.size, dtype=np.float64)
    wsum = float(wts.sum(dtype=np.float64))

    reference = np.zeros((h, w, 4), dtype=np.float32)
    independent = D != S
    win = np.lib.stride_tricks.sliding_window_view
    if independent:
        reference[..., 0] = intensity
    else:
        ch, cw = gh * S, gw * S
        tiled = intensity[:ch, :cw].reshape(gh, S, gw, S)
        means = np.einsum("aybx,yx->ab", tiled, wts.reshape(S, S), optimize=True) / wsum
        centered = intensity.copy()
        centered[:ch, :cw] -= np.repeat(np.repeat(means, S, axis=0), S, axis=1)
        reference[..., 0] = centered
    reference[..., 1] = gx
    reference[..., 2] = gy

    P = _param_count(motion_model)
    M = 28 if motion_model == "quadratic" else 16
    inverse = np.zeros((gw * gh, P * P), dtype=np.float32)
    metadata = np.zeros((gw * gh, M), dtype=np.float32)
    for row in range(gh):
        if independent:
            fwin = win(intensity, (S, S))[row * D][::D][:gw].reshape(gw, -1)
            means = np.einsum("np,p->n", fwin, wts, optimize=True) / wsum
            f = fwin - means[:, None]
            gxt = win(gx, (S, S))[row * D][::D][:gw].reshape(gw, -1)
            gyt = win(gy, (S, S))[row * D][::D][:gw].reshape(gw, -1)
        else:
            sl = slice(row * S, (row + 1) * S)
            f = centered[sl, : gw * S].reshape(S, gw, S).transpose(1, 0, 2).reshape(gw, -1)
            gxt = gx[sl, : gw * S].reshape(S, gw, S).transpose(1, 0, 2).reshape(gw, -1)
            gyt = gy[sl, : gw * S].reshape(S, gw, S).transpose(1, 0, 2).reshape(gw, -1)
            means = np.zeros(gw, dtype=np.float64)
        energy = np.einsum("np,np,p->n", f, f, wts, optimize=True)
        if motion_model == "translation":
            hxx = np.einsum("np,np,p->n", gxt, gxt, wts, optimize=True)
            hxy = np.einsum("np,np,p->n", gxt, gyt, wts, optimize=True)
            hyy = np.einsum("np,np,p->n", gyt, gyt, wts, optimize=True)
            det = hxx * hyy - hxy * hxy
            scale = np.maximum(np.abs(hxx * hyy), 1.0)
            valid = np.isfinite(det) & (np.abs(det) > np.finfo(np.float64).eps * scale) & (energy > 1e-12)
            idet = np.zeros(gw, dtype=np.float64)
            idet[valid] = 1.0 / det[valid]
            inv = np.zeros((gw, 6, 6), dtype=np.float64)
            inv[:, 0, 0] = hyy * idet
            inv[:, 0, 3] = -hxy * idet
            inv[:, 3, 0] = -hxy * idet
            inv[:, 3, 3] = hxx * idet
            jtf = np.zeros((gw, 6))
            jtf[:, 0] = np.einsum("np,np,p->n", gxt, f, wts, optimize=True)
            jtf[:, 3] = np.einsum("np,np,p->n", gyt, f, wts, optimize=True)
            sj = np.zeros((gw, 6))
            sj[:, 0] = np.einsum("np,p->n", gxt, wts, optimize=True)
            sj[:, 3] = np.einsum("np,p->n", gyt, wts, optimize=True)
        else:
            comps = [gxt, gxt * lx, gxt * ly, gyt, gyt * lx, gyt * ly]
            if motion_model == "quadratic":
                hx2, xy, hy2 = 0.5 * lx * lx, lx * ly, 0.5 * ly * ly
                comps += [gxt * hx2, gxt * xy, gxt * hy2, gyt * hx2, gyt * xy, gyt * hy2]
            J = np.stack(comps, axis=2)
            H = np.einsum("npi,npj,p->nij", J, J, wts, optimize=True)
            jtf = np.einsum("npi,np,p->ni", J, f, wts, optimize=True)
            sj = np.einsum("npi,p->ni", J, wts, optimize=True)
            try:
                inv = np.linalg.inv(H)
                valid = np.isfinite(inv).all(axis=(1, 2)) & (energy > 1e-12)
            except np.linalg.LinAlgError:
                inv = np.zeros_like(H)
                valid = np.zeros(gw, dtype=bool)
                for k in range(gw):
                    try:
                        inv[k] = np.linalg.inv(H[k])
                        valid[k] = energy[k] > 1e-12
                    except np.linalg.LinAlgError:
                        pass
        with np.errstate(invalid="ignore"):
            representable = np.max(np.abs(inv), axis=(1, 2)) <= np.finfo(np.float32).max
        valid &= representable
        inv[~valid] = 0.0
        o = row * gw
        inverse[o : o + gw] = inv.reshape(gw, P * P)
        metadata[o : o + gw, 0] = np.sqrt(energy)
        metadata[o : o + gw, 1] = valid.astype(np.float32)
        metadata[o : o + gw, 2 : 2 + P] = jtf
        metadata[o : o + gw, 2 + P : 2 + 2 * P] = sj
        metadata[o : o + gw, 2 + 2 * P] = means
        metadata[o : o + gw, 3 + 2 * P] = wsum
    return reference, inverse, metadata
