"""Seeded analytic speckle pairs rendered on the GPU.

The workloads that measure this engine are synthetic: Gaussian-blob speckle
(background 30 grey, blob sigma 2.5 px, amplitudes U[80, 200]) rendered
analytically at inverse-mapped coordinates for a known displacement field,
with independent Gaussian noise on every image, rounded and clipped to uint8.
Blob positions come from a seeded NumPy generator; rendering runs in the
``ldic_render_speckle`` kernel of liblivedic.so.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from . import _native

BACKGROUND = 30.0
BLOB_SIGMA = 2.5
AMPLITUDE_RANGE = (80.0, 200.0)


@dataclass
class Field:
    """u(X) = quadratic polynomial about ``origin`` + one sinusoid per component.

    ``poly`` uses the quadratic state order (tx, du/dx, du/dy, ty, dv/dx, dv/dy,
    d2u/dx2, d2u/dxdy, d2u/dy2, d2v/dx2, d2v/dxdy, d2v/dy2); ``wave_u`` and
    ``wave_v`` are (amplitude, kx, ky, phase) of amplitude*sin(kx X + ky Y + phase).
    """

    origin: tuple[float, float] = (0.0, 0.0)
    poly: tuple[float, ...] = (0.0,) * 12
    wave_u: tuple[float, float, float, float] = (0.0, 0.0, 0.0, 0.0)
    wave_v: tuple[float, float, float, float] = (0.0, 0.0, 0.0, 0.0)

    def scaled(self, s: float) -> Field:
        return Field(self.origin, tuple(s * p for p in self.poly),
                     (s * self.wave_u[0], *self.wave_u[1:]), (s * self.wave_v[0], *self.wave_v[1:]))

    def native(self) -> _native.Field:
        f = _native.Field()
        f.enabled = 1
        f.origin_x, f.origin_y = self.origin
        for i, p in enumerate(self.poly):
            f.poly[i] = p
        for i, p in enumerate(self.wave_u):
            f.wave[0][i] = p
        for i, p in enumerate(self.wave_v):
            f.wave[1][i] = p
        return f

    def displacement(self, X: np.ndarray, Y: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        """u, v at reference coordinates (float64)."""
        p = self.poly
        dx, dy = X - self.origin[0], Y - self.origin[1]
        u = p[0] + p[1] * dx + p[2] * dy + 0.5 * p[6] * dx * dx + p[7] * dx * dy + 0.5 * p[8] * dy * dy
        v = p[3] + p[4] * dx + p[5] * dy + 0.5 * p[9] * dx * dx + p[10] * dx * dy + 0.5 * p[11] * dy * dy
        a, kx, ky, ph = self.wave_u
        u = u + a * np.sin(kx * X + ky * Y + ph)
        a, kx, ky, ph = self.wave_v
        v = v + a * np.sin(kx * X + ky * Y + ph)
        return u, v

    def gradient(self, X: np.ndarray, Y: np.ndarray) -> np.ndarray:
        """(..., 4) du/dx, du/dy, dv/dx, dv/dy."""
        p = self.poly
        dx, dy = X - self.origin[0], Y - self.origin[1]
        g = np.stack([p[1] + p[6] * dx + p[7] * dy, p[2] + p[7] * dx + p[8] * dy,
                      p[4] + p[9] * dx + p[10] * dy, p[5] + p[10] * dx + p[11] * dy], axis=-1)
        a, kx, ky, ph = self.wave_u
        c = a * np.cos(kx * X + ky * Y + ph)
        g[..., 0] += c * kx
        g[..., 1] += c * ky
        a, kx, ky, ph = self.wave_v
        c = a * np.cos(kx * X + ky * Y + ph)
        g[..., 2] += c * kx
        g[..., 3] += c * ky
        return g


def blobs(width: int, height: int, count: int, seed: int) -> np.ndarray:
    """(count, 3) float64 blob x, y, amplitude from a seeded generator."""
    rng = np.random.default_rng(seed)
    xy = rng.uniform((0.0, 0.0), (float(width), float(height)), size=(count, 2))
    amp = rng.uniform(*AMPLITUDE_RANGE, size=count)
    return np.column_stack([xy, amp])


class SpeckleTexture:
    """Cell-binned blob list resident on the device."""

    def __init__(self, width: int, height: int, blob_table: np.ndarray, *, sigma: float = BLOB_SIGMA,
                 background: float = BACKGROUND, margin: int = 64):
        self.width, self.height = width, height
        self.sigma, self.background = float(sigma), float(background)
        self.cell = int(math.ceil(4.5 * sigma))
        # Cells cover the image plus a margin so inverse-mapped points near the
        # border still see every contributing blob.
        self.origin = -margin
        self.cells_x = (width + 2 * margin) // self.cell + 1
        self.cells_y = (height + 2 * margin) // self.cell + 1
        cx = np.clip(((blob_table[:, 0] - self.origin) // self.cell).astype(np.int64), 0, self.cells_x - 1)
        cy = np.clip(((blob_table[:, 1] - self.origin) // self.cell).astype(np.int64), 0, self.cells_y - 1)
        key = cy * self.cells_x + cx
        order = np.argsort(key, kind="stable")
        packed = np.zeros((len(order), 4), np.float32)
        packed[:, 0] = blob_table[order, 0]
        packed[:, 1] = blob_table[order, 1]
        packed[:, 2] = blob_table[order, 2]
        starts = np.searchsorted(key[order], np.arange(self.cells_x * self.cells_y + 1)).astype(np.int32)
        self._blobs = _native.DeviceBuffer(max(packed.nbytes, 16))
        self._blobs.upload(packed)
        self._starts = _native.DeviceBuffer(starts.nbytes)
        self._starts.upload(starts)

    def render_device(self, out: _native.DeviceBuffer, fld: Field | None, noise_sigma: float, noise_seed: int,
                      float_out: _native.DeviceBuffer | None = None) -> None:
        sp = _native.Speckle(self._blobs.address, self._starts.address, self.cells_x, self.cells_y, self.cell,
                             float(self.origin), self.sigma, self.background)
        nf = None if fld is None else ctypes.byref(fld.native())
        lib = _native.library()
        _native.check(lib.ldic_render_speckle(ctypes.byref(sp), nf, self.width, self.height, float(noise_sigma),
                                              int(noise_seed) & 0xFFFFFFFFFFFFFFFF, out.address,
                                              None if float_out is None else float_out.address, None))
        _native.check(lib.ldic_device_synchronize())

    def render(self, fld: Field | None = None, *, noise_sigma: float = 1.0, noise_seed: int = 0) -> np.ndarray:
        buf = _native.DeviceBuffer(self.width * self.height)
        try:
            self.render_device(buf, fld, noise_sigma, noise_seed)
            out = np.empty((self.height, self.width), np.uint8)
            buf.download(out)
            return out
        finally:
            buf.close()


# -- The measured workloads ---------------------------------------------------------------------
SEED = 20261203


def small_affine_pair(seed: int = SEED) -> tuple[np.ndarray, np.ndarray, Field]:
    """512x512, 4000 blobs, u = (0.3, -0.7) + 0.5 % x-strain about the image centre, noise 1 grey."""
    w = h = 512
    tex = SpeckleTexture(w, h, blobs(w, h, 4000, seed))
    fld = Field(origin=((w - 1) / 2, (h - 1) / 2), poly=(0.3, 0.005, 0.0, -0.7) + (0.0,) * 8)
    ref = tex.render(None, noise_sigma=1.0, noise_seed=seed * 2 + 1)
    tgt = tex.render(fld, noise_sigma=1.0, noise_seed=seed * 2 + 2)
    return ref, tgt, fld


def affine_field(width: int, height: int, tx: float, ty: float, theta: float = 0.0, exx: float = 0.0,
                 eyy: float = 0.0) -> Field:
    """u(X) = t + (R(theta) diag(1 + exx, 1 + eyy) - I)(X - c) about the image centre c."""
    c, s_ = math.cos(theta), math.sin(theta)
    a00, a01 = c * (1 + exx), -s_ * (1 + eyy)
    a10, a11 = s_ * (1 + exx), c * (1 + eyy)
    return Field(origin=((width - 1) / 2, (height - 1) / 2),
                 poly=(tx, a00 - 1.0, a01, ty, a10, a11 - 1.0) + (0.0,) * 6)


def uhd_sequence_field(width: int = 3840, height: int = 2160) -> Field:
    """u_x = 2 + 0.002 x + 1.5 sin(2 pi y / H), u_y = -1 + 0.001 y + 0.8 cos(2 pi x / W)."""
    return Field(origin=(0.0, 0.0), poly=(2.0, 0.002, 0.0, -1.0, 0.0, 0.001) + (0.0,) * 6,
                 wave_u=(1.5, 0.0, 2 * math.pi / height, 0.0),
                 wave_v=(0.8, 2 * math.pi / width, 0.0, math.pi / 2))


# Blob density of the 512x512 pair (4000 blobs).  A summed texture with 1.2 M blobs
# at 3840x2160 saturates: its mean is 254.9 grey of 255 and the window statistics
# vanish, so the 4K sequence keeps the 512x512 density instead (126,563 blobs).
BLOB_DENSITY = 4000 / (512 * 512)


class UhdSequence:
    """3840x2160 speckle at the 512x512 pair's blob density; frame k shows the field
    scaled by s = k / (frames - 1)."""

    def __init__(self, frames: int = 60, seed: int = SEED, width: int = 3840, height: int = 2160,
                 blob_count: int | None = None):
        if blob_count is None:
            blob_count = int(round(BLOB_DENSITY * width * height))
        self.width, self.height, self.frames, self.seed = width, height, frames, seed
        self.texture = SpeckleTexture(width, height, blobs(width, height, blob_count, seed))
        self.field = uhd_sequence_field(width, height)

    def frame_field(self, k: int) -> Field | None:
        s = k / max(1, self.frames - 1)
        return None if s == 0 else self.field.scaled(s)

    def render(self, k: int) -> np.ndarray:
        return self.texture.render(self.frame_field(k), noise_sigma=1.0, noise_seed=self.seed * 1000 + k)

    def render_device(self, k: int, out: _native.DeviceBuffer) -> None:
        self.texture.render_device(out, self.frame_field(k), 1.0, self.seed * 1000 + k)


def subset_centres(width: int, height: int, size: int, stride: int) -> tuple[np.ndarray, np.ndarray]:
    """Reference-frame centres (grid*stride + (S-1)/2) of the subset grid."""
    gw = 1 + (width - size) // stride
    gh = 1 + (height - size) // stride
    gy, gx = np.mgrid[:gh, :gw].astype(np.float64)
    return gx * stride + 0.5 * (size - 1), gy * stride + 0.5 * (size - 1)


def quadratic_plate_field(width: int = 3840, height: int = 2160) -> Field:
    """u_x = 1e-6 (x - W/2)^2 + 0.001 y, u_y = -8e-7 (y - H/2)^2 - 0.0005 x (about the image centre)."""
    cx, cy = width / 2, height / 2
    return Field(origin=(cx, cy),
                 poly=(0.001 * cy, 0.0, 0.001, -0.0005 * cx, -0.0005, 0.0, 2e-6, 0.0, 0.0, 0.0, 0.0, -1.6e-6))


def plate_with_holes(width: int = 3840, height: int = 2160, seed: int = SEED, margin: int = 96,
                     holes: int = 3, radius: tuple[int, int] = (200, 400)) -> np.ndarray:
    """Rectangle minus `holes` disjoint circular holes of seeded radii and centres."""
    rng = np.random.default_rng(seed + 7)
    mask = np.zeros((height, width), dtype=bool)
    mask[margin:height - margin, margin:width - margin] = True
    yy, xx = np.mgrid[:height, :width]
    placed: list[tuple[float, float, float]] = []
    scale = min(width / 3840, height / 2160)
    rmin, rmax = radius[0] * scale, radius[1] * scale
    while len(placed) < holes:
        r = rng.uniform(rmin, rmax)
        cx = rng.uniform(margin + r, width - margin - r)
        cy = rng.uniform(margin + r, height - margin - r)
        if all((cx - px) ** 2 + (cy - py) ** 2 > (r + pr + 40) ** 2 for px, py, pr in placed):
            placed.append((cx, cy, r))
            mask &= (xx - cx) ** 2 + (yy - cy) ** 2 > r * r
    return mask


class QuadraticPairs:
    """Consecutive-pair sequence: frame k shows the quadratic field scaled by s_k = k / (frames - 1)."""

    def __init__(self, frames: int = 30, seed: int = SEED, width: int = 3840, height: int = 2160):
        self.width, self.height, self.frames, self.seed = width, height, frames, seed
        count = int(round(BLOB_DENSITY * width * height))
        self.texture = SpeckleTexture(width, height, blobs(width, height, count, seed + 3))
        self.field = quadratic_plate_field(width, height)

    def scale(self, k: int) -> float:
        return k / max(1, self.frames - 1)

    def render(self, k: int) -> np.ndarray:
        s = self.scale(k)
        return self.texture.render(None if s == 0 else self.field.scaled(s), noise_sigma=1.0,
                                   noise_seed=self.seed * 7919 + k)

    def render_device(self, k: int, out: _native.DeviceBuffer) -> None:
        s = self.scale(k)
        self.texture.render_device(out, None if s == 0 else self.field.scaled(s), 1.0, self.seed * 7919 + k)

    def pair_truth(self, k: int, X: np.ndarray, Y: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        """Displacement from frame k to frame k+1 at frame-k pixel coordinates (X, Y)."""
        s0, s1 = self.scale(k), self.scale(k + 1)
        f0 = self.field.scaled(s0)
        mx, my = X.astype(np.float64).copy(), Y.astype(np.float64).copy()
        for _ in range(30):  # material point M with M + u_k(M) = (X, Y)
            u, v = f0.displacement(mx, my)
            g = f0.gradient(mx, my)
            rx, ry = mx + u - X, my + v - Y
            j11, j12, j21, j22 = 1 + g[..., 0], g[..., 1], g[..., 2], 1 + g[..., 3]
            det = j11 * j22 - j12 * j21
            mx -= (j22 * rx - j12 * ry) / det
            my -= (-j21 * rx + j11 * ry) / det
        u, v = self.field.displacement(mx, my)
        return (s1 - s0) * u, (s1 - s0) * v


# The 16-pair heterogeneous batch: 4 each of 640x480, 1280x720, 1920x1080 and 3840x2160;
# subset sizes 15/21/31/41 and steps 4/8/16 in rotation, affine and quadratic warps
# alternating; random affine fields (translation U[-5, 5] px, rotation U[-1, 1] deg,
# normal strains U[-1 %, 1 %]); each pair's pyramid depth follows its displacement range.
# Rotation and strain about the frame centre move the corners of a 1920x1080 frame by up to
# ~28 px: a zero start then needs four levels (with three, corner subsets of one 1080p pair
# stayed near zero and 13 % of its subsets failed; with four, 6 %).
BATCH_SIZES = [(640, 480)] * 4 + [(1280, 720)] * 4 + [(1920, 1080)] * 4 + [(3840, 2160)] * 4
BATCH_LEVELS = {640: 2, 1280: 4, 1920: 4, 3840: 4}


@dataclass(frozen=True)
class BatchPair:
    width: int
    height: int
    subset: int
    stride: int
    model: str
    levels: int
    texture: SpeckleTexture
    field: Field
    seed: int

    def label(self) -> str:
        return f"{self.width}x{self.height} S{self.subset} D{self.stride} {self.model} L{self.levels}"

    def reference(self) -> np.ndarray:
        return self.texture.render(None, noise_sigma=1.0, noise_seed=2 * self.seed + 1)

    def target_device(self, out: _native.DeviceBuffer) -> None:
        self.texture.render_device(out, self.field, 1.0, 2 * self.seed + 2)


def heterogeneous_batch(seed: int = 4242) -> list[BatchPair]:
    from .seednet import random_affine

    rng = np.random.default_rng(seed)
    pairs = []
    for i, (w, h) in enumerate(BATCH_SIZES):
        S = (15, 21, 31, 41)[i % 4]
        D = min((4, 8, 16)[i % 3], S)
        fld = random_affine(rng, w, h, translation=5.0, rotation_deg=1.0, strain=0.01)
        tex = SpeckleTexture(w, h, blobs(w, h, int(round(BLOB_DENSITY * w * h)), 77 + i))
        pairs.append(BatchPair(w, h, S, D, "quadratic" if i % 2 else "affine", BATCH_LEVELS[w], tex, fld, i))
    return pairs
