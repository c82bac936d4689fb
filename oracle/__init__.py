"""CPU oracle for the live-dic IC-GN hot path.  TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package, and only as the checker.  The product
(``live_dic`` and ``liblivedic.so``) never imports, links or runs it.

The C part (``oracle.c``) restates the per-frame WGSL kernels; the NumPy part
(``reference_prep``) restates the FP64 host reference preparation.  ``Pipeline``
reproduces the reference's coarse-to-fine per-frame command sequence
(``live_dic/full_pyramid_backend.py:1436-1569``).
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

from . import reference_prep as prep

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "_build" / "liboracle.so"
_lib = None


def build() -> Path:
    """Compile the C oracle with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(_LIB_PATH))
        _lib.ora_solve.restype = ctypes.c_int
    return _lib


class SolveCfg(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int32), ("height", ctypes.c_int32),
        ("grid_width", ctypes.c_int32), ("grid_height", ctypes.c_int32),
        ("subset_size", ctypes.c_int32), ("subset_stride", ctypes.c_int32),
        ("motion_model", ctypes.c_int32), ("interpolation", ctypes.c_int32),
        ("spatial_weighting", ctypes.c_int32), ("roi_enabled", ctypes.c_int32),
        ("gaussian_sigma", ctypes.c_float), ("iterations", ctypes.c_int32),
        ("tolerance", ctypes.c_float), ("zncc_min", ctypes.c_float),
        ("lanes", ctypes.c_int32), ("preserve_disabled", ctypes.c_int32),
        ("divergence_patience", ctypes.c_int32), ("divergence_min_iterations", ctypes.c_int32),
    ]


class SeedCfg(ctypes.Structure):
    _fields_ = [
        ("source_grid_width", ctypes.c_int32), ("source_grid_height", ctypes.c_int32),
        ("destination_grid_width", ctypes.c_int32), ("destination_grid_count", ctypes.c_int32),
        ("source_scale", ctypes.c_float), ("destination_scale", ctypes.c_float),
        ("subset_size", ctypes.c_float), ("subset_stride", ctypes.c_float),
        ("motion_model", ctypes.c_int32), ("track_seed_availability", ctypes.c_int32),
    ]


class MedianCfg(ctypes.Structure):
    _fields_ = [
        ("grid_width", ctypes.c_int32), ("grid_count", ctypes.c_int32),
        ("subset_stride", ctypes.c_int32), ("motion_model", ctypes.c_int32),
        ("outlier_policy", ctypes.c_int32), ("median_filter_size", ctypes.c_int32),
        ("median_filter_scope", ctypes.c_int32), ("roi_enabled", ctypes.c_int32),
        ("median_seed_available", ctypes.c_int32),
    ]


MODELS = {"affine": 0, "translation": 1, "quadratic": 2}
INTERP = {"bspline": 0, "linear": 1, "cubic": 2}
POLICIES = {"flag": 0, "replace": 1, "disabled": 2, "median": 3}
DEFAULT_ITERATIONS = {1: (30,), 2: (8, 16), 3: (6, 8, 16), 4: (12, 8, 8, 16)}
DEFAULT_TOLERANCES = (1e-2, 5e-3, 5e-3)


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def reduce(image: np.ndarray) -> np.ndarray:
    src = np.ascontiguousarray(image, dtype=np.uint8)
    h, w = src.shape
    out = np.empty((h // 2, w // 2), dtype=np.uint8)
    lib().ora_reduce(_p(src), ctypes.c_int(w), ctypes.c_int(h), _p(out))
    return out


def prefilter(image: np.ndarray, fir: np.ndarray | None = None) -> np.ndarray:
    src = np.ascontiguousarray(image, dtype=np.uint8)
    h, w = src.shape
    fir = prep.fir_weights() if fir is None else np.ascontiguousarray(fir, dtype=np.float32)
    out = np.empty((h, w), dtype=np.float32)
    scratch = np.empty((h, w), dtype=np.float32)
    lib().ora_prefilter(_p(src), ctypes.c_int(w), ctypes.c_int(h), _p(fir), _p(out), _p(scratch))
    return out


def solve(cfg: SolveCfg, reference, metadata, inverse, coeff, target, seed=None, active=None,
          state=None) -> np.ndarray:
    """One solver pass; with ``cfg.preserve_disabled`` and ``state`` (updated in place) the
    inactive subsets keep their previous state (retry_seeded, icgn.wgsl:1010-1017)."""
    stride = 16 if cfg.motion_model == 2 else 8
    n = cfg.grid_width * cfg.grid_height
    if state is None:
        state = np.zeros((n, stride), dtype=np.float32)
    assert state.shape == (n, stride) and state.dtype == np.float32 and state.flags.c_contiguous
    ref = np.ascontiguousarray(reference, dtype=np.float32)
    meta = np.ascontiguousarray(metadata, dtype=np.float32)
    inv = np.ascontiguousarray(inverse, dtype=np.float32)
    co = None if coeff is None else np.ascontiguousarray(coeff, dtype=np.float32)
    tg = None if target is None else np.ascontiguousarray(target, dtype=np.uint8)
    sd = None if seed is None else np.ascontiguousarray(seed, dtype=np.float32)
    ac = None if active is None else np.ascontiguousarray(active, dtype=np.uint32)
    rc = lib().ora_solve(ctypes.byref(cfg), _p(ref), _p(meta), _p(inv), _p(co), _p(tg), _p(sd), _p(ac), _p(state))
    if rc != 0:
        raise ValueError("oracle solve rejected its configuration")
    return state


def solve_partial(cfg: SolveCfg, reference, metadata, inverse, coeff, target, indices, words, boxes, state,
                  seed=None, active=None) -> np.ndarray:
    """The partial-support pass (``ora_solve_partial``) over ``state`` (n, 8), in place."""
    idx = np.ascontiguousarray(indices, dtype=np.uint32)
    if idx.size == 0:
        return state
    ref = np.ascontiguousarray(reference, dtype=np.float32)
    meta = np.ascontiguousarray(metadata, dtype=np.float32)
    inv = np.ascontiguousarray(inverse, dtype=np.float32)
    co = None if coeff is None else np.ascontiguousarray(coeff, dtype=np.float32)
    tg = None if target is None else np.ascontiguousarray(target, dtype=np.uint8)
    sd = None if seed is None else np.ascontiguousarray(seed, dtype=np.float32)
    ac = None if active is None else np.ascontiguousarray(active, dtype=np.uint32)
    wd = np.ascontiguousarray(words, dtype=np.uint32)
    bx = np.ascontiguousarray(boxes, dtype=np.float32)
    assert state.flags.c_contiguous and state.dtype == np.float32
    rc = lib().ora_solve_partial(ctypes.byref(cfg), _p(ref), _p(meta), _p(inv), _p(co), _p(tg), _p(sd), _p(ac),
                                 ctypes.c_int(idx.size), _p(idx), _p(wd), _p(bx), _p(state))
    if rc != 0:
        raise ValueError("oracle partial solve rejected its configuration")
    return state


def transfer_seed(cfg: SeedCfg, source_state, permissions=None, return_weight=False):
    stride = 16 if cfg.motion_model == 2 else 8
    seed = np.zeros((cfg.destination_grid_count, stride), dtype=np.float32)
    active = np.zeros(cfg.destination_grid_count, dtype=np.uint32)
    weight = np.zeros(cfg.destination_grid_count, dtype=np.float32)
    src = np.ascontiguousarray(source_state, dtype=np.float32)
    perm = None if permissions is None else np.ascontiguousarray(permissions, dtype=np.uint32)
    lib().ora_transfer_seed(ctypes.byref(cfg), _p(src), _p(perm), _p(seed), _p(active), _p(weight))
    return (seed, active, weight) if return_weight else (seed, active)


NEVER = 0xFFFFFFFF


def temporal_merge(history, ages, raw, output, use_filtered, max_age, return_expired=False):
    """update_temporal_state (full_pyramid.wgsl:519-565) on float32 (count, stride) arrays and uint64
    ages; returns the number of merged subsets (and, with ``return_expired``, of entries whose age
    has just passed ``max_age``).  ``raw is None``: no measurement, every entry ages."""
    P = 12 if history.shape[1] == 16 else 6
    if raw is None:
        upd = np.zeros(len(ages), bool)
    else:
        rs, fs = raw[:, P], output[:, P]
        upd = ((fs == 1.0) | (fs == 2.0)) & (bool(use_filtered) | (rs == 1.0) | (rs == 2.0))
        history[upd] = (output if use_filtered else raw)[upd]
    ages[upd] = 0
    aging = ~upd & (ages != NEVER)
    ages[aging] = np.minimum(ages[aging], 0xFFFFFFFD) + 1
    history[aging & (ages > max_age), P] = -1.0
    expired = int(np.sum(aging & (ages == max_age + 1)))
    return (int(upd.sum()), expired) if return_expired else int(upd.sum())


NO_SOURCE = 0xFFFFFFFF


def jump_flood_nearest(sources, gw, gh):
    """Nearest source subset of every grid point by jump flooding (steps N/2 .. 1, then one more
    step 1; candidates are the 3x3 points at the step; ties to the smaller index), as the
    library's neighbour re-seeding computes it."""
    n = gw * gh
    near = np.where(sources, np.arange(n, dtype=np.int64), -1).reshape(gh, gw)
    yy, xx = np.indices((gh, gw))
    steps = []
    step = 1
    while step * 2 < max(gw, gh):
        step *= 2
    while step >= 1:
        steps.append(step)
        step //= 2
    steps.append(1)
    for step in steps:
        best = near.copy()
        bd = np.where(best >= 0, (best % gw - xx) ** 2 + (best // gw - yy) ** 2, np.iinfo(np.int64).max)
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                if dx == 0 and dy == 0:
                    continue
                ny, nx = yy + dy * step, xx + dx * step
                ok = (nx >= 0) & (ny >= 0) & (nx < gw) & (ny < gh)
                c = np.full((gh, gw), -1, np.int64)
                c[ok] = near[ny[ok], nx[ok]]
                d = np.where(c >= 0, (c % gw - xx) ** 2 + (c // gw - yy) ** 2, np.iinfo(np.int64).max)
                take = (c >= 0) & ((d < bd) | ((d == bd) & (c < best)))
                best = np.where(take, c, best)
                bd = np.where(take, d, bd)
        near = best
    return near.ravel()


def reseed(state, metadata, gw, gh, stride_px, selected=None):
    """Neighbour re-seeding (the library's addition after the seeded fallback): sources are
    accepted subsets with at least two accepted 8-neighbours; every usable unaccepted subset
    takes its nearest source's warp extrapolated to its centre.  Returns (seed, active)."""
    n = gw * gh
    st = np.asarray(state, np.float32).reshape(n, -1)
    P = 12 if st.shape[1] == 16 else 6
    acc = (st[:, P] >= 0.5).reshape(gh, gw)
    pad = np.pad(acc, 1)
    nb = sum(pad[1 + dy:1 + dy + gh, 1 + dx:1 + dx + gw].astype(int) for dy in (-1, 0, 1) for dx in (-1, 0, 1)
             if dx or dy)
    near = jump_flood_nearest((acc & (nb >= 2)).ravel(), gw, gh)
    meta1 = np.asarray(metadata, np.float32).reshape(n, -1)[:, 1]
    usable = ((meta1 >= 0.5) | (meta1 == -4.0)) & (True if selected is None else np.asarray(selected).ravel() != 0)
    active = usable & ~acc.ravel() & (near >= 0)
    seed = np.zeros_like(st)
    i = np.flatnonzero(active)
    src = near[i]
    q = st[src]
    dx = np.float32(stride_px) * (i % gw - src % gw).astype(np.float32)
    dy = np.float32(stride_px) * (i // gw - src // gw).astype(np.float32)
    seed[i, 0] = (q[:, 0] + q[:, 1] * dx) + q[:, 2] * dy
    seed[i, 3] = (q[:, 3] + q[:, 4] * dx) + q[:, 5] * dy
    seed[i, 1], seed[i, 2], seed[i, 4], seed[i, 5] = q[:, 1], q[:, 2], q[:, 4], q[:, 5]
    return seed, active.astype(np.uint32)


def temporal_predict(history, previous, ages, previous_ages):
    """predict_temporal_state (full_pyramid.wgsl:83-113): returns (seed history, predicted count).
    Subsets whose current and previous measurements are fresh (age 0) and accepted get
    2 * current - previous for every parameter; status and iteration words are copied."""
    P = 12 if history.shape[1] == 16 else 6
    go = (ages == 0) & (previous_ages == 0) & (history[:, P] >= 0.5) & (previous[:, P] >= 0.5)
    out = history.copy()
    out[go, :P] = 2.0 * history[go, :P] - previous[go, :P]
    return out, int(go.sum())


def temporal_save_previous(history, ages, raw, output, previous, previous_ages):
    """save_previous_temporal_state (full_pyramid.wgsl:115-137), in place, before the merge."""
    P = 12 if history.shape[1] == 16 else 6
    rs, fs = raw[:, P], output[:, P]
    keep = ((rs == 1.0) | (rs == 2.0)) & ((fs == 1.0) | (fs == 2.0))
    previous[keep] = history[keep]
    previous_ages[keep] = ages[keep]
    aging = ~keep & (previous_ages != NEVER)
    previous_ages[aging] = np.minimum(previous_ages[aging], 0xFFFFFFFD) + 1


def temporal_refine(cfg: SeedCfg, coarse_history, coarse_ages, fine_history, fine_ages, max_age, links, seed,
                    active):
    """refine_temporal_seed (full_pyramid.wgsl:451-515), in place on the fine ``seed``.

    Where the fine subset is active, its history is fresh enough, its absolute seed is marked
    available, every permitted coarse corner has the fine subset's age and the coarse history
    interpolates with full weight: seed = fine history + (absolute seed - interpolated coarse
    history) for every parameter."""
    P = 12 if seed.shape[1] == 16 else 6
    n = cfg.destination_grid_count
    perm = np.full(n, 31, np.uint32) if links is None else np.asarray(links, np.uint32).ravel()
    predicted, _, weight = transfer_seed(cfg, coarse_history, perm, return_weight=True)
    i = np.arange(n)
    S, D = cfg.subset_size, cfg.subset_stride
    ss, ds = cfg.source_scale, cfg.destination_scale
    cx = (ds * (0.5 * S + (i % cfg.destination_grid_width) * D) + (ss - ds) * 0.5) / ss
    cy = (ds * (0.5 * S + (i // cfg.destination_grid_width) * D) + (ss - ds) * 0.5) / ss
    gw, gh = cfg.source_grid_width, cfg.source_grid_height
    lx = np.clip(np.floor((np.float32(cx) - 0.5 * S) / D).astype(np.int64), 0, max(0, gw - 2))
    ly = np.clip(np.floor((np.float32(cy) - 0.5 * S) / D).astype(np.int64), 0, max(0, gh - 2))
    age = fine_ages
    same = np.ones(n, bool)
    for corner in range(4):
        sx, sy = lx + (corner & 1), ly + (corner >> 1)
        use = ((perm >> corner) & 1).astype(bool) & (sx < gw) & (sy < gh)
        idx = np.where(use, sy * gw + sx, 0)
        same &= ~use | (coarse_ages[idx] == age)
    go = ((np.asarray(active) != 0) & (age != NEVER) & (age <= max_age) & (seed[:, P] >= -0.5) & same
          & (weight >= 0.99999))
    seed[go, :P] = fine_history[go, :P] + (seed[go, :P] - predicted[go, :P])
    return int(go.sum())


def median(cfg: MedianCfg, state, metadata, roi_words=None):
    st = np.ascontiguousarray(state, dtype=np.float32)
    out = np.empty_like(st)
    meta = np.ascontiguousarray(metadata, dtype=np.float32)
    words = None if roi_words is None else np.ascontiguousarray(roi_words, dtype=np.float32)
    lib().ora_median(ctypes.byref(cfg), _p(st), _p(meta), _p(words), _p(out))
    return out


def affine_zncc(width, height, grid_w, grid_h, size, stride, interpolation, reference, state, coeff, target):
    out = np.empty(grid_w * grid_h, dtype=np.float32)
    ref = np.ascontiguousarray(reference, dtype=np.float32)
    st = np.ascontiguousarray(state, dtype=np.float32)
    co = None if coeff is None else np.ascontiguousarray(coeff, dtype=np.float32)
    tg = np.ascontiguousarray(target, dtype=np.uint8)
    lib().ora_affine_zncc(
        ctypes.c_int(width), ctypes.c_int(height), ctypes.c_int(grid_w), ctypes.c_int(grid_h),
        ctypes.c_int(size), ctypes.c_int(stride), ctypes.c_int(interpolation),
        _p(ref), _p(st), _p(co), _p(tg), _p(out))
    out[out < -1.0] = np.nan
    return out


class Level:
    """Prepared reference of one pyramid level."""

    def __init__(self, image, opts, roi=None):
        self.image = image
        self.height, self.width = image.shape
        self.grid_width, self.grid_height = prep.grid_shape(
            self.width, self.height, opts["subset_size"], opts["subset_stride"])
        self.reference, self.inverse, self.metadata = prep.reference_arrays(
            image, interpolation=opts["interpolation"], motion_model=opts["motion_model"],
            subset_size=opts["subset_size"], subset_stride=opts["subset_stride"],
            spatial_weighting=opts["spatial_weighting"], gaussian_sigma=opts["gaussian_sigma"])
        self.roi = roi
        self.roi_words = None
        self.seed_links = None
        # reference_valid (wgpu_backend.py:2428-2433): numerically admissible subsets, before
        # the ROI marks the unselected ones -3.
        self.reference_valid = self.metadata[:, 1] >= 0.5
        if roi is not None and roi.get("effective", False):
            self.metadata[~roi["selected"].ravel(), 1] = -3.0
            self.roi_words = np.zeros((self.grid_width * self.grid_height, 4), dtype=np.float32)
            if roi.get("median_words") is not None:
                self.roi_words[:] = roi["median_words"]
            else:
                self.roi_words[:, 0] = roi["neighbour_bits"].ravel()
            self.seed_links = roi.get("seed_links")
        # Partial supports (roi_policy="partial"): FP64 statistics replace the rows.
        self.partial = None
        if roi is not None and roi.get("partial_indices") is not None and len(roi["partial_indices"]):
            from .partial import partial_reference

            plane = self.reference.reshape(self.height, self.width, 4)
            meta = self.metadata.reshape(-1, 16)
            inv = self.inverse.reshape(-1, 36)
            partial_reference(plane, meta, inv, roi["partial_indices"], roi["support_words"],
                              grid_width=self.grid_width, size=opts["subset_size"], stride=opts["subset_stride"],
                              motion_model=opts["motion_model"], weighting=opts["spatial_weighting"],
                              sigma=opts["gaussian_sigma"])
            self.partial = (np.asarray(roi["partial_indices"], np.uint32), np.asarray(roi["support_words"], np.uint32),
                            np.asarray(roi["support_boxes"], np.float32))
            idx = self.partial[0].astype(np.int64)
            self.reference_valid[idx] = meta[idx, 1] == -4.0


class Pipeline:
    """Coarse-to-fine oracle registration with the reference's defaults.

    Orders levels fine (index 0) to coarse; schedules are given coarse to fine
    as in ``ICGNWGPUFullPyramid`` (full_pyramid_backend.py:51-62, 1135-1368).
    ``rois`` optionally holds one dict per level with ``selected``,
    ``neighbour_bits``, ``effective`` and, for finer levels, ``seed_links``.
    """

    def __init__(self, *, pyramid_levels=2, iterations_per_level=None, tolerance_per_level=None,
                 zncc_min=0.8, motion_model="affine", interpolation="bspline", subset_size=16,
                 subset_stride=None, spatial_weighting="uniform", gaussian_sigma=None,
                 outlier_policy="replace", median_filter_size=3, median_filter_scope="rejected",
                 lanes=16, initialization="zero", temporal_max_age=None, temporal_prediction="none",
                 temporal_fine_correction=False, divergence_patience=None, divergence_min_iterations=4,
                 pyramid_policy="fixed", adaptive_acceptance=0.95):
        L = pyramid_levels
        self.levels_count = L
        self.iterations = tuple(iterations_per_level or DEFAULT_ITERATIONS[L])
        default_tol = (1e-2, *DEFAULT_TOLERANCES) if L == 4 else DEFAULT_TOLERANCES[-L:]
        self.tolerances = tuple(tolerance_per_level or default_tol)
        self.zncc_min = float(zncc_min)
        stride = subset_size if subset_stride is None else subset_stride
        self.opts = dict(motion_model=motion_model, interpolation=interpolation, subset_size=subset_size,
                         subset_stride=stride, spatial_weighting=spatial_weighting,
                         gaussian_sigma=0.25 * subset_size if gaussian_sigma is None else gaussian_sigma)
        self.outlier_policy = outlier_policy
        self.median_filter_size = median_filter_size
        self.median_filter_scope = median_filter_scope
        self.lanes = lanes
        self.fir = prep.fir_weights()
        self.levels: list[Level] = []
        # Temporal initialisation (full_pyramid_backend.py:433-881, default options).
        self.temporal = initialization == "temporal"
        self.temporal_max_age = 0xFFFFFFFE if temporal_max_age is None else int(temporal_max_age)
        self.temporal_links = None
        self.last_counts = (0, 0)
        self.velocity = temporal_prediction == "velocity"
        self.correct = bool(temporal_fine_correction) and L == 2
        self.patience = (0 if divergence_patience is None else int(divergence_patience),
                         int(divergence_min_iterations))
        self.last_predicted = 0
        self.adaptive = pyramid_policy == "adaptive" and self.temporal and L == 2
        self.adaptive_acceptance = float(adaptive_acceptance)
        self.next_levels = L
        self.last_levels = L
        self.last_expired = 0
        self.fine_links = None
        self.population = 0

    def set_reference(self, image, rois=None, temporal_links=None):
        images = [np.ascontiguousarray(image, dtype=np.uint8)]
        for _ in range(1, self.levels_count):
            images.append(reduce(images[-1]))
        self.levels = [Level(img, self.opts, None if rois is None else rois[k]) for k, img in enumerate(images)]
        self.temporal_links = temporal_links
        fine = self.levels[0]
        self.population = int(np.sum(fine.reference_valid))
        self.fine_links = None
        if fine.roi is not None and fine.roi.get("effective", False):
            self.fine_links = np.where(np.asarray(fine.roi["selected"]).ravel(), 31, 0).astype(np.uint32)
        self.reset_temporal_state()
        return self

    def reset_temporal_state(self):
        """reset_temporal_state (full_pyramid_backend.py:660-677)."""
        self.next_levels = self.levels_count
        fine = self.levels[0]
        stride = 16 if self.opts["motion_model"] == "quadratic" else 8
        self.history = np.zeros((fine.grid_width * fine.grid_height, stride), np.float32)
        self.history[:, 12 if stride == 16 else 6] = -1.0
        self.ages = np.full(fine.grid_width * fine.grid_height, NEVER, np.uint64)
        self.previous = self.history.copy()
        self.previous_ages = self.ages.copy()
        coarse = self.levels[-1]
        self.coarse_history = np.zeros((coarse.grid_width * coarse.grid_height, stride), np.float32)
        self.coarse_history[:, 12 if stride == 16 else 6] = -1.0
        self.coarse_ages = np.full(coarse.grid_width * coarse.grid_height, NEVER, np.uint64)

    def _solve_cfg(self, k):
        lv = self.levels[k]
        o = self.opts
        sched = self.levels_count - 1 - k
        return SolveCfg(
            lv.width, lv.height, lv.grid_width, lv.grid_height, o["subset_size"], o["subset_stride"],
            MODELS[o["motion_model"]], INTERP[o["interpolation"]], 1 if o["spatial_weighting"] == "gaussian" else 0,
            1 if lv.roi_words is not None else 0, o["gaussian_sigma"], self.iterations[sched],
            self.tolerances[sched], self.zncc_min, self.lanes, 0, *(self.patience if k == 0 else (0, 4)))

    def _median(self, k, state):
        if self.outlier_policy == "disabled":
            return state
        lv = self.levels[k]
        o = self.opts
        cfg = MedianCfg(
            lv.grid_width, lv.grid_width * lv.grid_height, o["subset_stride"], MODELS[o["motion_model"]],
            POLICIES[self.outlier_policy], self.median_filter_size,
            1 if self.median_filter_scope == "all" else 0, 1 if lv.roi_words is not None else 0,
            1 if (self.outlier_policy == "median" and k > 0) else 0)
        return median(cfg, state, lv.metadata, lv.roi_words)

    def register(self, target, return_levels=False):
        # The adaptive pyramid solves the finest level alone after a good frame
        # (full_pyramid_backend.py:1436-1569, 1775-1791).
        Lu = self.next_levels if self.adaptive else self.levels_count
        self.last_levels = Lu
        targets = [np.ascontiguousarray(target, dtype=np.uint8)]
        for _ in range(1, Lu):
            targets.append(reduce(targets[-1]))
        outputs = [None] * self.levels_count
        raws = [None] * self.levels_count
        o = self.opts
        for k in range(Lu - 1, -1, -1):
            lv = self.levels[k]
            coeff = prefilter(targets[k], self.fir) if o["interpolation"] == "bspline" else None
            cfg = self._solve_cfg(k)
            if k == Lu - 1 and self.temporal:
                # project_temporal_seed (full_pyramid.wgsl:437-449): finest history -> the
                # coarsest grid solved this frame (the finest one itself in an adaptive
                # finest-only frame, with links "selected or not").
                fine = self.levels[0]
                scfg = SeedCfg(fine.grid_width, fine.grid_height, lv.grid_width, lv.grid_width * lv.grid_height,
                               1.0, float(1 << k), float(o["subset_size"]), float(o["subset_stride"]),
                               MODELS[o["motion_model"]], 1 if self.correct else 0)
                source = self.history
                if self.velocity:
                    source, self.last_predicted = temporal_predict(self.history, self.previous, self.ages,
                                                                   self.previous_ages)
                plinks = self.temporal_links if k == self.levels_count - 1 else self.fine_links
                seed, active, weight = transfer_seed(scfg, source, plinks, return_weight=True)
                links = np.full(lv.grid_width * lv.grid_height, 31, np.uint32) if plinks is None \
                    else np.asarray(plinks, np.uint32).ravel()
                counted = (weight > 0) & ((links & 16) != 0)
                seeded = int(np.sum(counted))
                if self.correct:
                    # project_temporal_seed marks every seed it does not count.
                    seed[~counted, 12 if o["motion_model"] == "quadratic" else 6] = -1.0
                state = solve(cfg, lv.reference, lv.metadata, lv.inverse, coeff, targets[k], seed, active)
            elif k == Lu - 1:
                seed = active = None
                state = solve(cfg, lv.reference, lv.metadata, lv.inverse, coeff, targets[k])
            else:
                src = self.levels[k + 1]
                scfg = SeedCfg(src.grid_width, src.grid_height, lv.grid_width, lv.grid_width * lv.grid_height,
                               float(1 << (k + 1)), float(1 << k), float(o["subset_size"]),
                               float(o["subset_stride"]), MODELS[o["motion_model"]], 1 if self.correct else 0)
                seed, active = transfer_seed(scfg, outputs[k + 1], lv.seed_links)
                if k == 0 and self.correct:
                    temporal_refine(scfg, self.coarse_history, self.coarse_ages, self.history, self.ages,
                                    self.temporal_max_age, lv.seed_links, seed, active)
                state = solve(cfg, lv.reference, lv.metadata, lv.inverse, coeff, targets[k], seed, active)
            if lv.partial is not None:
                solve_partial(cfg, lv.reference, lv.metadata, lv.inverse, coeff, targets[k], *lv.partial, state,
                              seed, active)
            outputs[k] = self._median(k, state)
            raws[k] = state
        if self.temporal:
            filtered = self.outlier_policy == "replace" and not self.velocity
            if self.velocity:
                temporal_save_previous(self.history, self.ages, raws[0], outputs[0], self.previous,
                                       self.previous_ages)
            merged, expired = temporal_merge(self.history, self.ages, raws[0], outputs[0], filtered,
                                             self.temporal_max_age, return_expired=True)
            if self.correct:
                # A frame that skipped the coarse level leaves no coarse measurement.
                temporal_merge(self.coarse_history, self.coarse_ages, raws[-1], outputs[-1], filtered,
                               self.temporal_max_age)
            self.last_counts = (seeded, merged)
            self.last_expired = expired if self.adaptive else 0
            if self.adaptive:
                good = self.population > 0 and merged / self.population >= self.adaptive_acceptance and not expired
                self.next_levels = 1 if good else 2
        stride = 16 if o["motion_model"] == "quadratic" else 8
        fine = outputs[0].reshape(self.levels[0].grid_height, self.levels[0].grid_width, stride)
        return (fine, outputs) if return_levels else fine


def cpu_threads() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def zncc_search(ref, tgt, size, stride, radius, prior=None, mask=None):
    """Exhaustive integer ZNCC search (``ora_zncc_search``).

    ``prior`` is a packed state array (translation at components 0 and 3) or
    None; ``mask`` selects the subsets to search.  Returns (gh, gw, 2) shifts
    and (gh, gw) scores.
    """
    r = np.ascontiguousarray(ref, dtype=np.uint8)
    t = np.ascontiguousarray(tgt, dtype=np.uint8)
    h, w = r.shape
    gw, gh = prep.grid_shape(w, h, size, stride)
    shifts = np.zeros((gh * gw, 2), np.float32)
    scores = np.zeros(gh * gw, np.float32)
    pr = None if prior is None else np.ascontiguousarray(prior, dtype=np.float32).reshape(gh * gw, -1)
    mk = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8).reshape(-1)
    lib().ora_zncc_search(_p(r), _p(t), ctypes.c_int(w), ctypes.c_int(h), ctypes.c_int(gw), ctypes.c_int(gh),
                          ctypes.c_int(size), ctypes.c_int(stride), ctypes.c_int(radius), _p(pr),
                          ctypes.c_int(0 if pr is None else pr.shape[1]), _p(mk), _p(shifts), _p(scores))
    return shifts.reshape(gh, gw, 2), scores.reshape(gh, gw)
