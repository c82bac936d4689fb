"""Float64 statistics of partial-support subsets (reference: live_dic/partial.py:26-130).

Computed once per reference on the host from the device-prepared reference plane
(f, gx, gy per pixel): for each partial subset, the support-weighted mean,
energy, J^T f, sum J and weight sum, the Hessian's rank, normalised condition
and weighted contrast, and its inverse when the support admits a solve.
Admitted subsets get metadata validity -4 (solved by the partial pass, skipped
by the full-window pass); the others validity 0 and a zero inverse.
"""

from __future__ import annotations

from dataclasses import replace

import numpy as np

from .roi import PARTIAL_MAX_NORMALIZED_CONDITION, PARTIAL_MIN_WEIGHTED_CONTRAST, RoiPlan, immutable, unpack_support


def partial_rows(reference4: np.ndarray, plan: RoiPlan, *, width: int, size: int, stride: int, motion_model: str,
                 weighting: str, sigma: float):
    """Returns (metadata rows (k, 16), inverse rows (k, 36), updated plan) for the plan's partial subsets."""
    pixels = np.asarray(reference4, dtype=np.float32).reshape(-1, width, 4)
    count = 0 if plan.partial_indices is None else int(plan.partial_indices.size)
    rows = np.zeros((count, 16), np.float32)
    inverses = np.zeros((count, 36), np.float32)
    grid_width = plan.selected.shape[1]
    yy, xx = np.indices((size, size), dtype=np.float64)
    lx = xx.ravel() - (size - 1) / 2
    ly = yy.ravel() - (size - 1) / 2
    spatial = np.exp(-0.5 * (lx ** 2 + ly ** 2) / sigma ** 2) if weighting == "gaussian" else np.ones(size * size)
    active = [0, 3] if motion_model == "translation" else list(range(6))
    norm = np.array([1, 1 / size, 1 / size, 1, 1 / size, 1 / size])[active]
    conditions = np.zeros(plan.selected.shape)
    ranks = np.zeros(plan.selected.shape, np.uint32)
    contrasts = np.zeros(plan.selected.shape)
    for k in range(count):
        index = int(plan.partial_indices[k])
        gy, gx = divmod(index, grid_width)
        support = unpack_support(plan.support_words[k], size).ravel()
        w = spatial * support
        wsum = w.sum(dtype=np.float64)
        tile = pixels[gy * stride:gy * stride + size, gx * stride:gx * stride + size].reshape(-1, 4).astype(np.float64)
        mean = np.dot(w, tile[:, 0]) / wsum
        f = tile[:, 0] - mean
        gxv, gyv = tile[:, 1], tile[:, 2]
        jac = np.stack((gxv, gxv * lx, gxv * ly, gyv, gyv * lx, gyv * ly), axis=1)
        j = jac[:, active]
        hess = j.T @ (w[:, None] * j)
        eig = np.linalg.eigvalsh(hess * norm[:, None] * norm[None, :])
        largest = float(eig[-1])
        rank = int(np.count_nonzero(eig > np.finfo(np.float64).eps * len(active) * max(largest, 0.0)))
        contrast = float(np.sqrt(np.dot(w, f * f) / wsum))
        condition = largest / eig[0] if eig[0] > 0 else 0.0
        conditions[gy, gx] = condition if np.isfinite(condition) else 0.0
        ranks[gy, gx] = rank
        contrasts[gy, gx] = contrast
        valid = (rank == len(active) and condition <= PARTIAL_MAX_NORMALIZED_CONDITION
                 and contrast >= PARTIAL_MIN_WEIGHTED_CONTRAST and support.sum() > len(active))
        inv = np.zeros((6, 6))
        if valid:
            if motion_model == "translation":
                det = hess[0, 0] * hess[1, 1] - hess[0, 1] ** 2
                block = np.array([[hess[1, 1], -hess[0, 1]], [-hess[0, 1], hess[0, 0]]]) / det
            else:
                block = np.linalg.inv(hess)
            valid = bool(np.isfinite(block).all() and np.max(np.abs(block)) <= np.finfo(np.float32).max)
            if valid:
                inv[np.ix_(active, active)] = block
        inverses[k] = inv.ravel()
        rows[k, 0] = np.sqrt(np.dot(w, f * f))
        rows[k, 1] = -4.0 if valid else 0.0
        rows[k, 2:8] = jac.T @ (w * f)
        rows[k, 8:14] = jac.T @ w
        # The stored intensities are centred on the full-window mean (or raw); this
        # is the support's remaining offset.
        rows[k, 14] = mean
        rows[k, 15] = wsum
    plan = replace(plan, support_condition=immutable(conditions), support_rank=immutable(ranks),
                   support_contrast=immutable(contrasts))
    return rows, inverses, plan
