"""NumPy restatement of the reference's ROI planner.  TEST INFRASTRUCTURE ONLY.

Follows live_dic/roi.py of the reference: full-window selection by prefix sums
(:63-97), neighbour links and median link words (:178-263), partial supports
with the even-centre rule and visibility truncation (:266-355) and seed
transfer links (:358-437).  The library plans ROIs on the device
(csrc/ldic_roi.hip); the GPU tests compare its tables with these, and
tests/test_oracle_roi.py pins these against brute force.
"""

from __future__ import annotations

import hashlib
import struct
from dataclasses import dataclass
from functools import lru_cache

import numpy as np


def immutable(array: np.ndarray) -> np.ndarray:
    return np.frombuffer(array.tobytes(order="C"), dtype=array.dtype).reshape(array.shape)


def subset_grid_shape(width, height, *, subset_size, subset_stride):
    if width < subset_size or height < subset_size:
        return (0, 0)
    return (1 + (width - subset_size) // subset_stride, 1 + (height - subset_size) // subset_stride)


def integral(mask: np.ndarray) -> np.ndarray:
    """Exclusive 2-D prefix count of forbidden pixels, shape (h+1, w+1), int64."""
    forbidden = (~mask).astype(np.int64)
    out = np.zeros((mask.shape[0] + 1, mask.shape[1] + 1), dtype=np.int64)
    out[1:, 1:] = forbidden.cumsum(0).cumsum(1)
    return out


def boxes_clear(prefix, x0, y0, x1, y1) -> np.ndarray:
    """True where the half-open box [x0, x1) x [y0, y1) holds no forbidden pixel."""
    return (prefix[y1, x1] - prefix[y0, x1] - prefix[y1, x0] + prefix[y0, x0]) == 0


def select_windows(mask: np.ndarray, size: int, stride: int, *, summed_area=None) -> np.ndarray:
    h, w = mask.shape
    gw, gh = subset_grid_shape(w, h, subset_size=size, subset_stride=stride)
    ys, xs = np.mgrid[:gh, :gw] * stride
    prefix = integral(mask) if summed_area is None else summed_area
    return boxes_clear(prefix, xs, ys, xs + size, ys + size)


def reduce_mask(mask: np.ndarray) -> np.ndarray:
    """Coarse pixel allowed iff all 25 reflected source pixels at (2x, 2y) are (roi.py:89-97)."""
    h, w = mask.shape
    padded = np.pad(mask, 2, mode="reflect")
    out = np.ones((h // 2, w // 2), dtype=bool)
    for dy in range(5):
        for dx in range(5):
            out &= padded[dy : dy + 2 * (h // 2) : 2, dx : dx + 2 * (w // 2) : 2]
    return immutable(out)


def _segment_clear(prefix, ax, ay, bx, by) -> np.ndarray:
    """Closed pixel box enclosing the segment (ax, ay)-(bx, by) is mask-free."""
    h, w = prefix.shape[0] - 1, prefix.shape[1] - 1
    x0 = np.maximum(0, np.ceil(np.minimum(ax, bx) - 0.5).astype(np.int64))
    y0 = np.maximum(0, np.ceil(np.minimum(ay, by) - 0.5).astype(np.int64))
    x1 = np.minimum(w, np.floor(np.maximum(ax, bx) + 0.5).astype(np.int64) + 1)
    y1 = np.minimum(h, np.floor(np.maximum(ay, by) + 0.5).astype(np.int64) + 1)
    return boxes_clear(prefix, x0, y0, x1, y1)


@dataclass(frozen=True)
class OraclePlan:
    """The per-level tables of one planned mask."""

    mask: np.ndarray
    selected: np.ndarray
    neighbour_bits: np.ndarray
    mask_hash: str
    effective: bool
    selected_count: int
    policy: str = "full"
    truncation: str = "none"
    partial_indices: np.ndarray | None = None
    median_neighbour_words: np.ndarray | None = None
    support_words: np.ndarray | None = None
    support_boxes: np.ndarray | None = None
    coverage: np.ndarray | None = None

    def neighbour_words(self) -> np.ndarray:
        if self.median_neighbour_words is not None:
            return np.ascontiguousarray(self.median_neighbour_words, dtype=np.float32)
        words = np.zeros((self.selected.size, 4), dtype=np.float32)
        words[:, 0] = self.neighbour_bits.ravel()
        return words


def pack_support(support: np.ndarray) -> np.ndarray:
    """Row-major pixels as little-endian uint32 words, pixel 0 in bit 0 (roi.py:266-270)."""
    bits = np.packbits(np.asarray(support, bool).ravel(), bitorder="little")
    bits = np.pad(bits, (0, (-bits.size) % 4))
    return bits.view("<u4").astype(np.uint32)


def unpack_support(words: np.ndarray, size: int) -> np.ndarray:
    raw = np.asarray(words, dtype="<u4").view(np.uint8)
    return np.unpackbits(raw, bitorder="little")[: size * size].reshape(size, size).astype(bool)


def central_component(support: np.ndarray) -> np.ndarray:
    """4-connected component of allowed pixels grown from the central pixel; for an even width
    all four central pixels must be allowed, otherwise the component is empty (roi.py:280-299)."""
    size = support.shape[0]
    lo, hi = (size - 1) // 2, size // 2
    out = np.zeros_like(support, dtype=bool)
    if not support[lo:hi + 1, lo:hi + 1].all():
        return out
    out[lo, lo] = True
    stack = [(lo, lo)]
    while stack:
        y, x = stack.pop()
        for ny, nx in ((y - 1, x), (y + 1, x), (y, x - 1), (y, x + 1)):
            if 0 <= ny < size and 0 <= nx < size and support[ny, nx] and not out[ny, nx]:
                out[ny, nx] = True
                stack.append((ny, nx))
    return out


@lru_cache(maxsize=7)
def visibility_paths(size: int) -> np.ndarray:
    """(S*S, words) packed sets: the closed pixel cells that the segment from the window centre
    to each pixel centre touches (slab test per axis; touching a cell edge or corner counts),
    the visibility rule of roi.py:302-330."""
    n = size * size
    centre = (size - 1) / 2
    yy, xx = np.divmod(np.arange(n, dtype=np.float64), size)
    t_lo = np.zeros((n, n))
    t_hi = np.ones((n, n))
    for coord in (xx, yy):                       # rows: target pixel, columns: cell
        d = coord - centre
        for i in range(n):
            if d[i] == 0.0:
                outside = np.abs(coord - centre) > 0.5
                t_hi[i, outside] = -1.0
                continue
            a = (coord - 0.5 - centre) / d[i]
            b = (coord + 0.5 - centre) / d[i]
            t_lo[i] = np.maximum(t_lo[i], np.minimum(a, b))
            t_hi[i] = np.minimum(t_hi[i], np.maximum(a, b))
    paths = t_lo <= t_hi
    out = np.stack([pack_support(row) for row in paths])
    out.setflags(write=False)
    return out


def partial_windows(mask: np.ndarray, complete: np.ndarray, size: int, stride: int, *, truncation: str):
    """Selection, indices, support words, support boxes and coverage of the partial policy
    (roi.py:333-355): every incomplete window whose central component is not empty."""
    if truncation not in ("none", "visibility"):
        raise ValueError("roi_truncation must be 'none' or 'visibility'")
    selected = complete.copy()
    coverage = complete.astype(np.float32)
    indices, words, boxes = [], [], []
    paths = visibility_paths(size) if truncation == "visibility" else None
    for y, x in np.argwhere(~complete):
        local = mask[y * stride:y * stride + size, x * stride:x * stride + size]
        support = central_component(local)
        if support.any() and paths is not None:
            packed = pack_support(support)
            visible = ~np.any(paths & ~packed, axis=1)
            support = central_component(support & visible.reshape(size, size))
        if not support.any():
            continue
        selected[y, x] = True
        indices.append(y * complete.shape[1] + x)
        words.append(pack_support(support))
        sy, sx = np.nonzero(support)
        boxes.append((sx.min(), sy.min(), sx.max(), sy.max()))
        coverage[y, x] = support.sum() / (size * size)
    nwords = (size * size + 31) // 32
    return (selected, immutable(np.asarray(indices, np.uint32)),
            immutable(np.asarray(words, np.uint32).reshape(-1, nwords)),
            immutable(np.asarray(boxes, np.float32).reshape(-1, 4)), immutable(coverage))


def prepare_roi(mask, size: int, stride: int, *, policy: str = "full", truncation: str = "none",
                median_filter_size: int = 3) -> OraclePlan | None:
    """Selection, 3x3 neighbour bits and optional 9x9 median words (roi.py:178-263)."""
    if mask is None:
        return None
    if policy not in ("full", "partial") or (policy == "full" and truncation != "none"):
        raise ValueError("Unknown ROI policy/truncation")
    h, w = mask.shape
    every = bool(mask.all())
    prefix = None if every else integral(mask)
    gw, gh = subset_grid_shape(w, h, subset_size=size, subset_stride=stride)
    selected = np.ones((gh, gw), dtype=bool) if every else select_windows(mask, size, stride, summed_area=prefix)
    partial = (None,) * 4
    if policy == "partial":
        selected, *partial = partial_windows(mask, selected, size, stride, truncation=truncation)
    gy, gx = np.indices(selected.shape)
    cx = gx * stride + (size - 1) / 2
    cy = gy * stride + (size - 1) / 2

    def linked(dx: int, dy: int) -> np.ndarray:
        nx, ny = gx + dx, gy + dy
        inside = (nx >= 0) & (ny >= 0) & (nx < gw) & (ny < gh)
        ok = inside & selected & selected[ny.clip(0, gh - 1), nx.clip(0, gw - 1)]
        if prefix is not None:
            ok &= _segment_clear(prefix, cx, cy, cx + dx * stride, cy + dy * stride)
        return ok

    bits = np.zeros(selected.shape, dtype=np.uint32)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            bits |= linked(dx, dy).astype(np.uint32) << np.uint32((dy + 1) * 3 + dx + 1)
    words = None
    if median_filter_size > 3:
        r = median_filter_size // 2
        words = np.zeros((gh, gw, 4), dtype=np.float32)
        for dy in range(-r, r + 1):
            for dx in range(-r, r + 1):
                bit = (dy + 4) * 9 + dx + 4
                words[..., bit // 24] += linked(dx, dy).astype(np.float32) * float(1 << (bit % 24))
        words = immutable(words.reshape(-1, 4))
    digest = hashlib.sha256(b"live-dic-roi-v1\0" + struct.pack("<QQ", *mask.shape) + mask.tobytes(order="C")).hexdigest()
    indices, support_words, boxes, coverage = partial
    return OraclePlan(mask, immutable(selected), immutable(bits), digest, not every, int(selected.sum()),
                      policy, truncation, indices, words, support_words, boxes, coverage)


def seed_permissions(source: OraclePlan, destination: OraclePlan, *, source_scale: int, destination_scale: int,
                     size: int, stride: int, reference_mask: np.ndarray, reference_prefix=None) -> np.ndarray:
    """Transfer link words (roi.py:358-437): bits 0-3 corners, bit 4 selection,
    bits 5.. one-based index of a fallback coarse subset."""
    sel_s, sel_d = source.selected, destination.selected
    sh, sw = sel_s.shape
    y, x = np.indices(sel_d.shape)
    c = size / 2
    sx = (destination_scale * (c + x * stride) + (source_scale - destination_scale) / 2) / source_scale
    sy = (destination_scale * (c + y * stride) + (source_scale - destination_scale) / 2) / source_scale
    lx = np.clip(np.floor((sx - c) / stride), 0, max(0, sw - 2)).astype(int)
    ly = np.clip(np.floor((sy - c) / stride), 0, max(0, sh - 2)).astype(int)
    prefix = reference_prefix if reference_prefix is not None else integral(reference_mask)
    half = (size - 1) / 2
    bits = sel_d.astype(np.uint32) << np.uint32(4)
    for corner in range(4):
        nx, ny = lx + (corner & 1), ly + (corner >> 1)
        inside = (nx < sw) & (ny < sh)
        nx, ny = nx.clip(0, sw - 1), ny.clip(0, sh - 1)
        ok = inside & sel_s[ny, nx] & sel_d
        ok &= _segment_clear(prefix, destination_scale * (half + x * stride), destination_scale * (half + y * stride),
                             source_scale * (half + nx * stride), source_scale * (half + ny * stride))
        bits |= ok.astype(np.uint32) << np.uint32(corner)
    if source_scale > destination_scale and sel_s.size < (1 << 27) - 1:
        my, mx = np.nonzero(sel_d & ((bits & 15) == 0))
        best = np.full(mx.shape, np.inf)
        chosen = np.zeros(mx.shape, dtype=np.uint32)
        for dy in range(-2, 4):
            for dx in range(-2, 4):
                nx, ny = lx[my, mx] + dx, ly[my, mx] + dy
                inside = (nx >= 0) & (ny >= 0) & (nx < sw) & (ny < sh)
                cxn, cyn = nx.clip(0, sw - 1), ny.clip(0, sh - 1)
                dist = (sx[my, mx] - (c + cxn * stride)) ** 2 + (sy[my, mx] - (c + cyn * stride)) ** 2
                ok = inside & sel_s[cyn, cxn] & (dist < best)
                ok &= _segment_clear(prefix, destination_scale * (half + mx * stride),
                                     destination_scale * (half + my * stride), source_scale * (half + cxn * stride),
                                     source_scale * (half + cyn * stride))
                best[ok] = dist[ok]
                chosen[ok] = (cyn[ok] * sw + cxn[ok] + 1).astype(np.uint32)
        bits[my, mx] |= chosen << np.uint32(5)
    return immutable(bits)
