"""FP64 restatement of partial-support preparation, and an independent IC-GN.  TEST INFRASTRUCTURE.

``partial_reference`` follows ``prepare_partial_reference`` (live_dic/partial.py:26-130):
for every partial subset the support-weighted statistics (energy, validity,
J^T f, sum J, mean, weight sum) and the inverse Hessian replace the subset's
full-window metadata row.  It is written per parameter column (weighted sums
over the support) rather than as matrix products, and is pinned by
``tests/test_oracle_partial.py``: a support covering the whole window must give
back the full-window preparation of ``reference_prep.reference_arrays``; the
admission rule must reproduce the reference's thin-band case
(tests/test_partial_roi.py:158-188); and the partial solve built on it must
agree with ``double_icgn``.

``double_icgn`` is a matrix-form float64 inverse-compositional Gauss-Newton
over one support (the method of tests/test_partial_roi.py:52-86: solve with the
support Hessian, compose the inverse increment, stop on the corner change).  It
uses no metadata or stored inverse and so checks the whole partial pipeline
from an independent direction.
"""

from __future__ import annotations

import numpy as np

# Admission limits of partial supports (roi.py:20-21).
MAX_CONDITION = 1e8
MIN_CONTRAST = 1.0


def unpack(words, size: int) -> np.ndarray:
    """Row-major support bits, pixel k in bit k % 32 of little-endian word k // 32."""
    w = np.asarray(words, dtype=np.uint64)
    k = np.arange(size * size)
    return ((w[k // 32] >> (k % 32).astype(np.uint64)) & 1).astype(bool).reshape(size, size)


def partial_reference(reference, metadata, inverse, indices, words, *, grid_width: int, size: int, stride: int,
                      motion_model: str, weighting: str = "uniform", sigma: float | None = None):
    """Updates ``metadata`` (n, 16) and ``inverse`` (n, 36) in place for the listed subsets.

    ``reference`` is the (H, W, 4) prepared plane (f, gx, gy, 0) as stored by
    the preparation.  Returns per-subset (rank, condition, contrast).
    """
    plane = np.asarray(reference, dtype=np.float64)
    h = (size - 1) / 2.0
    sig = 0.25 * size if sigma is None else sigma
    cols = 2 if motion_model == "translation" else 6
    keep = (0, 3) if motion_model == "translation" else (0, 1, 2, 3, 4, 5)
    scale = np.array([1.0, 1.0 / size, 1.0 / size, 1.0, 1.0 / size, 1.0 / size])
    diagnostics = []
    for k, index in enumerate(np.asarray(indices, dtype=np.int64)):
        gy, gx = divmod(int(index), grid_width)
        mask = unpack(words[k], size)
        ly, lx = np.nonzero(mask)
        f = plane[gy * stride + ly, gx * stride + lx, 0]
        gxv = plane[gy * stride + ly, gx * stride + lx, 1]
        gyv = plane[gy * stride + ly, gx * stride + lx, 2]
        x = lx - h
        y = ly - h
        wt = np.exp(-(x * x + y * y) / (2.0 * sig * sig)) if weighting == "gaussian" else np.ones(x.size)
        total = float(np.sum(wt))
        mean = float(np.sum(wt * f) / total)
        fc = f - mean
        columns = [gxv, gxv * x, gxv * y, gyv, gyv * x, gyv * y]
        hess = np.empty((cols, cols))
        for a, pa in enumerate(keep):
            for b, pb in enumerate(keep):
                hess[a, b] = np.sum(wt * columns[pa] * columns[pb])
        s = scale[list(keep)]
        lam = np.linalg.eigvalsh(hess * np.outer(s, s))
        top = float(lam[-1])
        rank = int(np.sum(lam > np.finfo(np.float64).eps * cols * max(top, 0.0)))
        condition = top / lam[0] if lam[0] > 0 else 0.0
        contrast = float(np.sqrt(np.sum(wt * fc * fc) / total))
        ok = rank == cols and condition <= MAX_CONDITION and contrast >= MIN_CONTRAST and mask.sum() > cols
        full = np.zeros((6, 6))
        if ok:
            if cols == 2:
                det = hess[0, 0] * hess[1, 1] - hess[0, 1] * hess[1, 0]
                block = np.array([[hess[1, 1], -hess[0, 1]], [-hess[1, 0], hess[0, 0]]]) / det
            else:
                block = np.linalg.inv(hess)
            ok = bool(np.all(np.isfinite(block)) and np.abs(block).max() <= np.finfo(np.float32).max)
            if ok:
                for a, pa in enumerate(keep):
                    for b, pb in enumerate(keep):
                        full[pa, pb] = block[a, b]
        row = np.zeros(16)
        row[0] = np.sqrt(np.sum(wt * fc * fc))
        row[1] = -4.0 if ok else 0.0
        for p in range(6):
            row[2 + p] = np.sum(wt * fc * columns[p])
            row[8 + p] = np.sum(wt * columns[p])
        row[14] = mean
        row[15] = total
        metadata[index] = row
        inverse[index] = full.ravel()
        diagnostics.append((rank, condition if np.isfinite(condition) else 0.0, contrast))
    return diagnostics


def _bilinear(image, x, y):
    ix, iy = np.floor(x).astype(int), np.floor(y).astype(int)
    fx, fy = x - ix, y - iy
    top = image[iy, ix] * (1 - fx) + image[iy, ix + 1] * fx
    bottom = image[iy + 1, ix] * (1 - fx) + image[iy + 1, ix + 1] * fx
    return top * (1 - fy) + bottom * fy


def double_icgn(tile, target, support, *, origin, model, weights, tolerance=1e-4, iterations=30):
    """Float64 IC-GN of one support against a uint8 target, bilinear sampling.

    ``tile`` is the (S, S, 4) block of the prepared reference plane at
    ``origin`` (x, y of its top-left pixel).  Returns (tx, ux, uy, ty, vx, vy).
    """
    size = support.shape[0]
    h = (size - 1) / 2.0
    yy, xx = np.nonzero(support)
    x = xx - h
    y = yy - h
    w = weights[yy, xx]
    f = tile[yy, xx, 0].astype(np.float64)
    f = f - np.sum(w * f) / np.sum(w)
    gx = tile[yy, xx, 1].astype(np.float64)
    gy = tile[yy, xx, 2].astype(np.float64)
    jac = np.column_stack((gx, gx * x, gx * y, gy, gy * x, gy * y))
    keep = [0, 3] if model == "translation" else list(range(6))
    j = jac[:, keep]
    hess = (j * w[:, None]).T @ j
    norm_f = np.sqrt(np.sum(w * f * f))
    centre = np.array([origin[0] + h, origin[1] + h])
    corners = np.array([[-h, -h, h, h], [-h, h, -h, h]])
    img = np.asarray(target, dtype=np.float64)
    A = np.eye(2)
    t = np.zeros(2)
    for _ in range(iterations):
        pos = A @ np.vstack((x, y)) + (centre + t)[:, None]
        g = _bilinear(img, pos[0], pos[1])
        g = g - np.sum(w * g) / np.sum(w)
        ratio = norm_f / np.sqrt(np.sum(w * g * g))
        d = np.zeros(6)
        d[keep] = np.linalg.solve(hess, j.T @ (w * (ratio * g - f)))
        inc = np.array([[1.0 + d[1], d[2]], [d[4], 1.0 + d[5]]])
        A_new = A @ np.linalg.inv(inc)
        t_new = t - A_new @ np.array([d[0], d[3]])
        moved = (A_new - A) @ corners + (t_new - t)[:, None]
        A, t = A_new, t_new
        if np.max(np.hypot(moved[0], moved[1])) < tolerance:
            break
    return np.array([t[0], A[0, 0] - 1.0, A[0, 1], t[1], A[1, 0], A[1, 1] - 1.0])
