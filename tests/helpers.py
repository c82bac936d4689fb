"""Shared comparison helpers for the parity tests."""

from __future__ import annotations

from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"


def load_gray(name: str) -> np.ndarray:
    from PIL import Image

    with Image.open(GOLDEN / name) as im:
        return np.asarray(im.convert("L"), dtype=np.uint8).copy()


def compare_fields(candidate: np.ndarray, expected: np.ndarray, params: int) -> dict:
    """Status identity and parameter agreement of two packed (gh, gw, stride) fields.

    The bar: identical status on >= 99.9 % of subsets and |du| <= 1e-3 px on the
    subsets both accept (after the reference's own implementation-vs-implementation
    comparator, wgpu_backend.py:222-299, tightened to the north-star tolerance).
    """
    cs, es = candidate[..., params], expected[..., params]
    both = (cs > 0.5) & (es > 0.5)
    diff = np.abs(candidate[..., :params] - expected[..., :params])[both]
    disp = np.maximum(diff[:, 0], diff[:, 3]) if diff.size else np.zeros(0)
    return {
        "status_equal": float(np.mean(cs == es)),
        "accepted_candidate": int(np.count_nonzero(cs > 0.5)),
        "accepted_expected": int(np.count_nonzero(es > 0.5)),
        "disp_within_1e3": float(np.mean(disp <= 1e-3)) if disp.size else 1.0,
        "disp_max": float(disp.max()) if disp.size else 0.0,
        "param_rmse": float(np.sqrt(np.mean(diff.astype(np.float64) ** 2))) if diff.size else 0.0,
    }


def oracle_rois(size: int, stride: int, mask: np.ndarray, levels: int, *, policy: str = "full",
                truncation: str = "none", median_filter_size: int = 3) -> list[dict]:
    """Per-level ROI tables for ``oracle.Pipeline`` from the oracle's planner
    (selection, link bits, seed links and partial supports)."""
    from oracle.roi_plan import prepare_roi, reduce_mask, seed_permissions

    masks = [mask]
    for _ in range(1, levels):
        masks.append(reduce_mask(masks[-1]))
    plans = [prepare_roi(m, size, stride, policy=policy, truncation=truncation,
                         median_filter_size=median_filter_size) for m in masks]
    out = []
    for k, plan in enumerate(plans):
        d = dict(selected=plan.selected, neighbour_bits=plan.neighbour_bits, effective=plan.effective,
                 median_words=plan.median_neighbour_words, partial_indices=plan.partial_indices,
                 support_words=plan.support_words, support_boxes=plan.support_boxes, plan=plan)
        if k + 1 < levels and plan.effective:
            d["seed_links"] = seed_permissions(plans[k + 1], plan, source_scale=1 << (k + 1),
                                               destination_scale=1 << k, size=size, stride=stride,
                                               reference_mask=mask)
        out.append(d)
    return out


def bilinear_shift(source: np.ndarray, dx: float, dy: float) -> np.ndarray:
    """Target of a known translation (dx, dy): each target pixel samples the source bilinearly at
    (x - dx, y - dy) with mirrored borders (period 2(n-1)), rounded to uint8.  This is the image
    formation of the reference's fractional-shift fixture (python/tests/uhd_partial_cases.py:24-52)."""
    h, w = source.shape

    def mirror(i, n):
        period = 2 * (n - 1)
        i = np.mod(i, period)
        return np.where(i < n, i, period - i)

    x = np.arange(w, dtype=np.float64) - dx
    x0 = np.floor(x).astype(np.int64)
    fx = x - x0
    xa, xb = mirror(x0, w), mirror(x0 + 1, w)
    out = np.empty_like(source)
    src = source.astype(np.float64)
    for start in range(0, h, 256):
        y = np.arange(start, min(start + 256, h), dtype=np.float64) - dy
        y0 = np.floor(y).astype(np.int64)
        fy = (y - y0)[:, None]
        ya, yb = mirror(y0, h), mirror(y0 + 1, h)
        top = src[ya][:, xa] * (1 - fx) + src[ya][:, xb] * fx
        bottom = src[yb][:, xa] * (1 - fx) + src[yb][:, xb] * fx
        out[start:start + len(y)] = np.clip(np.rint(top * (1 - fy) + bottom * fy), 0, 255).astype(np.uint8)
    return out


def tensile_truth() -> dict:
    import json

    return json.loads((GOLDEN / "uhd_tensile_truth.json").read_text(encoding="utf-8"))


def interior_centres(shape: tuple[int, int], size: int, stride: int, width: int, height: int, margin: int):
    """Subset centres (x, y) and the interior mask of a margin in pixels
    (python/tests/test_uhd_fixtures.py:43-49)."""
    gy, gx = np.indices(shape)
    x = gx * stride + (size - 1) / 2
    y = gy * stride + (size - 1) / 2
    inside = (x >= margin) & (x < width - margin) & (y >= margin) & (y < height - margin)
    return inside, x, y
