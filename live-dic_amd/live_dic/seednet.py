"""Learned seed network: a bfloat16 correlation CNN trained and run on the
gfx950 matrix cores by this library's own kernels (``ldic_nn_*``).

No counterpart in the reference, whose coarsest-level initialisers are zeros,
a temporal prediction, a Fourier estimate or an external CPM-HOG matcher
(full_pyramid_backend.py:1693-1716).  The network replaces the pyramid and
the ZNCC seed search: from the reference and target reduced twice by the
integer pyramid it predicts a displacement field, coarse to fine, at 1/64,
1/32 and 1/16 of the finest resolution; affine seeds for every subset are
read from the last field and refined by a single finest-level IC-GN solve, with
the ZNCC search as the fallback where the solve does not converge or its final
ZNCC stays below 0.9.

Architecture (``Architecture``):
- encoder, shared by both images: four 3x3 stride-2 convolutions with leaky
  ReLU, 1 -> 16 -> 32 -> 48 -> 64 channels;
- stage A at the last encoder level: correlation of the two feature maps over
  +-8 feature pixels, concatenated with the reference features, into a head of
  three 3x3 convolutions (128, 96, 2 channels) that outputs the flow;
- stages B and C one and two levels finer: the previous flow, upsampled,
  warps the target features; their +-4 correlation, the reference features
  and the upsampled flow feed a head of the same shape, whose output is added
  to the upsampled flow.
Every stage is supervised (Huber loss against the analytic field); gradients
do not flow through the upsampled flow (stage-wise training).

Device memory comes from the library (``_native.DeviceArray``) and work is
queued on HIP streams given as handles (0: the null stream); every arithmetic
kernel is in liblivedic.so, and no other framework is loaded.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import _native
from ._native import BF16, DeviceArray

WEIGHTS_PATH = Path(__file__).resolve().parent / "weights" / "seednet.npz"


@dataclass(frozen=True)
class Stage:
    level: int                       # encoder output the stage works on
    radius: int                      # correlation radius in feature pixels
    head: tuple = ((128, True), (96, True), (2, False))   # (out channels, leaky ReLU)


@dataclass(frozen=True)
class Architecture:
    encoder: tuple = ((16, 2), (32, 2), (48, 2), (64, 2))  # (out channels, stride)
    stages: tuple = (Stage(3, 8), Stage(2, 4), Stage(1, 4))
    input_level: int = 2             # network input: the finest image reduced this many times

    def scale(self, level: int) -> int:
        """Finest-level pixels per pixel of encoder output ``level``."""
        return 1 << (self.input_level + level + 1)

    def stage_channels(self, si: int) -> tuple[int, int]:
        """(channels used, channels stored) of a stage's head input."""
        st = self.stages[si]
        real = (2 * st.radius + 1) ** 2 + self.encoder[st.level][0] + (2 if si > 0 else 0)
        return real, (real + 3) // 4 * 4

    def convs(self) -> list[tuple[str, int, int, int, int, bool]]:
        """(name, input channels used, input channels stored, outputs, stride, leaky) per convolution."""
        out = []
        cin = 1
        for i, (cout, stride) in enumerate(self.encoder):
            out.append((f"enc{i}", cin, cin, cout, stride, True))
            cin = cout
        for si, st in enumerate(self.stages):
            real, stored = self.stage_channels(si)
            for hi, (cout, act) in enumerate(st.head):
                out.append((f"s{si}h{hi}", real, stored, cout, 1, act))
                real = stored = cout
        return out

    def parameter_count(self) -> int:
        return sum(cout * (9 * real + 1) for _, real, _, cout, _, _ in self.convs())

    def macs(self, height: int, width: int) -> int:
        """Multiply-accumulates of one forward pass on one input pair of height x width."""
        total, h, w = 0, height, width
        sizes = []
        for i, (cout, stride) in enumerate(self.encoder):
            cin = 1 if i == 0 else self.encoder[i - 1][0]
            h, w = (h - 1) // stride + 1, (w - 1) // stride + 1
            sizes.append((h, w))
            total += 2 * h * w * cout * 9 * cin
        for si, st in enumerate(self.stages):
            h, w = sizes[st.level]
            real, _ = self.stage_channels(si)
            d2, c = (2 * st.radius + 1) ** 2, self.encoder[st.level][0]
            total += h * w * d2 * c
            cin = real
            for cout, _ in st.head:
                total += h * w * cout * 9 * cin
                cin = cout
        return total


def _kp(k: int) -> int:
    return (k + 15) // 16 * 16


def _ptr(t) -> int:
    """Device address of a DeviceArray (or of any array object exposing data_ptr())."""
    return t.data_ptr()


def _zeros(shape, dtype, stream: int = 0) -> DeviceArray:
    return DeviceArray.zeros(shape, dtype, stream)


class SeedNet:
    """Parameters, activations and training state of the network on one device."""

    def __init__(self, arch: Architecture | None = None, *, seed: int = 0, device: str | int = "cuda"):
        if device not in ("cuda", "cuda:0", 0):
            raise ValueError("the seed network runs on the library's current device (device 0)")
        self.arch = arch or Architecture()
        self.lib = _native.library()
        self._stream = 0
        rng = np.random.default_rng(seed)
        self.params = {}
        for name, real, stored, cout, stride, act in self.arch.convs():
            kp = _kp(9 * stored)
            w = np.zeros((cout, kp), np.float32)
            std = math.sqrt(2.0 / (9 * real)) if act else 0.1 * math.sqrt(1.0 / (9 * real))
            core = rng.normal(0.0, std, size=(cout, 9, real)).astype(np.float32)
            w[:, :9 * stored].reshape(cout, 9, stored)[:, :, :real] = core
            self.params[name] = self._new_param(w, np.zeros(cout, np.float32), real, stored, cout, stride, act)
        self._refresh_weights()
        self.step_count = 0
        self._shape = None

    # -- parameters ------------------------------------------------------------------------------
    def _new_param(self, w, b, real, stored, cout, stride, act):
        f32 = np.float32
        return {
            "real": real, "cin": stored, "cout": cout, "stride": stride, "act": act,
            "w": DeviceArray.from_numpy(w), "b": DeviceArray.from_numpy(b),
            "w16": _zeros(w.shape, BF16), "wt16": _zeros((stored, _kp(9 * cout)), BF16),
            "dw": _zeros(w.shape, f32), "db": _zeros(b.shape, f32),
            "mw": _zeros(w.shape, f32), "vw": _zeros(w.shape, f32),
            "mb": _zeros(b.shape, f32), "vb": _zeros(b.shape, f32),
        }

    def _refresh_weights(self) -> None:
        s = self.stream_handle
        for p in self.params.values():
            _native.check(self.lib.ldic_nn_to_bf16(p["w"].numel(), _ptr(p["w"]), _ptr(p["w16"]), s))
            _native.check(self.lib.ldic_nn_transpose_weights(p["cin"], p["cout"], _ptr(p["w16"]), _ptr(p["wt16"]), s))

    def oracle_parameters(self) -> dict:
        """float64 (w (Cout, 3, 3, Cin), b) per convolution, in the oracle's layout (from the bf16 copies)."""
        out = {}
        for name, p in self.params.items():
            w = p["w16"].numpy().astype(np.float64)
            cout, cin, real = p["cout"], p["cin"], p["real"]
            w = w[:, :9 * cin].reshape(cout, 9, cin)[:, :, :real].reshape(cout, 3, 3, real)
            out[name] = (w, p["b"].numpy().astype(np.float64))
        return out

    def gradients(self) -> dict:
        """float64 (dw, db) per convolution in the oracle's layout."""
        out = {}
        for name, p in self.params.items():
            dw = p["dw"].numpy().astype(np.float64)
            cout, cin, real = p["cout"], p["cin"], p["real"]
            dw = dw[:, :9 * cin].reshape(cout, 9, cin)[:, :, :real].reshape(cout, 3, 3, real)
            out[name] = (dw, p["db"].numpy().astype(np.float64))
        return out

    def save(self, path: str | Path) -> None:
        arrays = {}
        for name, p in self.params.items():
            arrays[f"{name}.w"] = p["w"].numpy()
            arrays[f"{name}.b"] = p["b"].numpy()
        arrays["step"] = np.array(self.step_count)
        Path(path).parent.mkdir(parents=True, exist_ok=True)
        np.savez_compressed(path, **arrays)

    def load(self, path: str | Path) -> None:
        with np.load(path, allow_pickle=False) as data:
            for name, p in self.params.items():
                w, b = data[f"{name}.w"], data[f"{name}.b"]
                if w.shape != p["w"].shape or b.shape != p["b"].shape:
                    raise ValueError(f"weight file does not match the architecture at {name}")
                _native.check(self.lib.ldic_device_synchronize())
                p["w"].upload(w)
                p["b"].upload(b)
            self.step_count = int(data["step"]) if "step" in data else 0
        self._refresh_weights()
        self.synchronize()

    # -- streams ---------------------------------------------------------------------------------
    @property
    def stream_handle(self) -> int:
        return self._stream

    def use_stream(self, handle: int | None) -> None:
        """Queue all further work on a HIP stream given by its handle (None or 0: the null stream).
        Work already queued on the previous stream completes first, so the new stream (which may be
        a non-blocking one) never reads weights that are still being written."""
        self.synchronize()
        self._stream = int(handle or 0)

    def synchronize(self) -> None:
        """Wait for the work queued on the network's stream."""
        _native.check(self.lib.ldic_stream_synchronize(self._stream or None))

    # -- buffers ---------------------------------------------------------------------------------
    def _ensure(self, n: int, h: int, w: int, training: bool) -> None:
        key = (n, h, w, training)
        if self._shape == key:
            return
        bf, f32 = BF16, np.float32
        a = self.arch
        self.n, self.h, self.w = n, h, w
        _native.check(self.lib.ldic_device_synchronize())   # the old buffers may still be in use
        self.img16 = _zeros((2 * n, h, w, 1), bf)
        self.enc, self.denc, self.enc_shapes = [], [], []
        hh, ww = h, w
        for cout, stride in a.encoder:
            hh, ww = (hh - 1) // stride + 1, (ww - 1) // stride + 1
            self.enc_shapes.append((hh, ww, cout))
            self.enc.append(_zeros((2 * n, hh, ww, cout), bf))
            self.denc.append(_zeros((2 * n, hh, ww, cout), f32) if training else None)
        self.st = []
        for si, stage in enumerate(a.stages):
            hh, ww, c = self.enc_shapes[stage.level]
            _, stored = a.stage_channels(si)
            b = {"x": _zeros((n, hh, ww, stored), bf)}
            b["acts"] = [_zeros((n, hh, ww, co), bf) for co, _ in stage.head[:-1]]
            b["res"] = _zeros((n, hh, ww, 2), f32)
            b["flow"] = _zeros((n, hh, ww, 2), f32)
            if si > 0:
                b["up"] = _zeros((n, hh, ww, 2), f32)
                b["warped"] = _zeros((n, hh, ww, c), bf)
            if training:
                b["dx"] = _zeros((n, hh, ww, stored), f32)
                b["dacts"] = [_zeros((n, hh, ww, co), f32) for co, _ in stage.head[:-1]]
                b["dflow"] = _zeros((n, hh, ww, 2), f32)
                b["truth"] = _zeros((n, hh, ww, 2), f32)
                if si > 0:
                    b["dwarped"] = _zeros((n, hh, ww, c), f32)
            self.st.append(b)
        self.loss = _zeros(1, f32)
        # Split-K workspace: one slice per split (at most 8) of the largest output of any
        # convolution at this shape.
        largest = max(2 * n * hh_ * ww_ * c for hh_, ww_, c in self.enc_shapes)
        for si, stage in enumerate(a.stages):
            hh, ww, _ = self.enc_shapes[stage.level]
            largest = max(largest, n * hh * ww * max(co for co, _ in stage.head))
        self.workspace = _zeros(8 * largest, f32)
        _native.check(self.lib.ldic_device_synchronize())
        self._shape = key

    # -- layers ----------------------------------------------------------------------------------
    def _conv(self, p, batch, h, w, x, x_stride, x_offset, y, y_stride, y_f32=None):
        g = _native.NNConv(batch, h, w, p["cin"], x_stride, x_offset, (h - 1) // p["stride"] + 1,
                           (w - 1) // p["stride"] + 1, p["cout"], y_stride, 0, p["stride"], p.get("tile", 0))
        splits = p.get("splits", 1)
        _native.check(self.lib.ldic_nn_conv_forward_split(
            ctypes.byref(g), _ptr(x), _ptr(p["w16"]), _ptr(p["b"]), 1 if p["act"] else 0,
            None if y is None else _ptr(y), None if y_f32 is None else _ptr(y_f32),
            _ptr(self.workspace) if splits > 1 else None, splits, self.stream_handle))
        return g

    def _conv_back(self, p, g, x, dy, dx, dx_stride, accumulate):
        _native.check(self.lib.ldic_nn_conv_backward(
            ctypes.byref(g), _ptr(x), _ptr(p["wt16"]), _ptr(dy), None if dx is None else _ptr(dx), dx_stride, 0,
            1 if accumulate else 0, _ptr(p["dw"]), _ptr(p["db"]), self.stream_handle))

    # -- forward / backward ----------------------------------------------------------------------
    def forward(self, ref16_src, tgt16_src=None, *, n: int, h: int, w: int, training: bool = False,
                images_u8: bool = True):
        """Run the network on n pairs of h x w images (device pointers to uint8 planes, or with
        images_u8=False a (2n, h, w, 1) bfloat16 tensor in ``ref16_src``).  Returns the flow of every
        stage as (n, h_s, w_s, 2) float32 tensors, in pixels of that stage."""
        self._ensure(n, h, w, training)
        lib, s = self.lib, self.stream_handle
        a = self.arch
        if images_u8:
            half = n * h * w
            _native.check(lib.ldic_nn_images_to_bf16(half, ref16_src, _ptr(self.img16), s))
            _native.check(lib.ldic_nn_images_to_bf16(half, tgt16_src, _ptr(self.img16) + 2 * half, s))
        else:
            self.img16.copy_from(ref16_src, s)
        self._gconv = {}
        x, hh, ww, cin = self.img16, h, w, 1
        for i, (cout, stride) in enumerate(a.encoder):
            p = self.params[f"enc{i}"]
            self._gconv[f"enc{i}"] = self._conv(p, 2 * n, hh, ww, x, cin, 0, self.enc[i], cout)
            x, (hh, ww, cin) = self.enc[i], self.enc_shapes[i]
        prev = None
        for si, stage in enumerate(a.stages):
            b = self.st[si]
            hh, ww, c = self.enc_shapes[stage.level]
            feats = self.enc[stage.level]
            f1, f2 = _ptr(feats), _ptr(feats) + 2 * n * hh * ww * c   # bf16: 2 bytes per element
            d2 = (2 * stage.radius + 1) ** 2
            _, stored = a.stage_channels(si)
            if prev is not None:
                ph, pw = prev.shape[1], prev.shape[2]
                _native.check(lib.ldic_nn_upsample_flow(n, ph, pw, hh, ww, _ptr(prev), _ptr(b["up"]),
                                                        _ptr(b["x"]), stored, d2 + c, s))
                _native.check(lib.ldic_nn_warp_forward(n, hh, ww, c, f2, _ptr(b["up"]), _ptr(b["warped"]), s))
                f2 = _ptr(b["warped"])
                f2_stride = c
            else:
                f2_stride = c
            corr = _native.NNCorr(n, hh, ww, c, stage.radius, c, 0, f2_stride, 0, stored, 0, 1.0 / c)
            b["corr"] = corr
            _native.check(lib.ldic_nn_corr_forward(ctypes.byref(corr), f1, f2, _ptr(b["x"]), s))
            _native.check(lib.ldic_nn_copy_slice(n * hh * ww, c, f1, c, 0, _ptr(b["x"]), stored, d2, s))
            x, xs = b["x"], stored
            b["g"] = []
            for hi, (cout, act) in enumerate(stage.head):
                p = self.params[f"s{si}h{hi}"]
                last = hi == len(stage.head) - 1
                if last:
                    g = self._conv(p, n, hh, ww, x, xs, 0, None, 2, y_f32=b["res"])
                else:
                    g = self._conv(p, n, hh, ww, x, xs, 0, b["acts"][hi], cout)
                    x, xs = b["acts"][hi], cout
                b["g"].append(g)
            b["flow"].copy_from(b["res"], s)
            if prev is not None:
                _native.check(lib.ldic_nn_add_slice(n * hh * ww, 2, _ptr(b["up"]), 2, 0, _ptr(b["flow"]), 2, 0,
                                                    s))
            prev = b["flow"]
        return [b["flow"] for b in self.st]

    def backward(self, truths, weights) -> None:
        """Accumulate parameter gradients of the supervised loss (call after forward(training=True));
        ``truths`` are (n, h_s, w_s, 2) float32 tensors per stage."""
        lib, s = self.lib, self.stream_handle
        a = self.arch
        n = self.n
        if any(isinstance(t, np.ndarray) for t in truths):
            # Host targets are copied synchronously into buffers that the previous step
            # may still be reading on the stream.
            self.synchronize()
        self.loss.zero(s)
        for p in self.params.values():
            p["dw"].zero(s)
            p["db"].zero(s)
        for d in self.denc:
            d.zero(s)
        for si in reversed(range(len(a.stages))):
            stage, b = a.stages[si], self.st[si]
            hh, ww, c = self.enc_shapes[stage.level]
            d2 = (2 * stage.radius + 1) ** 2
            _, stored = a.stage_channels(si)
            b["truth"].copy_from(truths[si], s)
            _native.check(lib.ldic_nn_flow_loss(n * hh * ww, _ptr(b["flow"]), _ptr(b["truth"]),
                                                float(weights[si]), _ptr(b["dflow"]), _ptr(self.loss), s))
            dy = b["dflow"]
            for hi in reversed(range(len(stage.head))):
                p = self.params[f"s{si}h{hi}"]
                x = b["x"] if hi == 0 else b["acts"][hi - 1]
                if p["act"]:
                    _native.check(lib.ldic_nn_leaky_backward(dy.numel(), _ptr(b["acts"][hi]), _ptr(dy), s))
                if hi == 0:
                    dx, dxs = b["dx"], stored
                else:
                    dx, dxs = b["dacts"][hi - 1], stage.head[hi - 1][0]
                self._conv_back(p, b["g"][hi], x, dy, dx, dxs, False)
                dy = dx
            feats, dfeats = self.enc[stage.level], self.denc[stage.level]
            f1, f2 = _ptr(feats), _ptr(feats) + 2 * n * hh * ww * c
            df1, df2 = _ptr(dfeats), _ptr(dfeats) + 4 * n * hh * ww * c   # float32: 4 bytes
            if si > 0:
                b["dwarped"].zero(s)
                _native.check(lib.ldic_nn_corr_backward(ctypes.byref(b["corr"]), f1, _ptr(b["warped"]),
                                                        _ptr(b["dx"]), stored, 0, df1, c, 0,
                                                        _ptr(b["dwarped"]), c, 0, s))
                _native.check(lib.ldic_nn_warp_backward(n, hh, ww, c, _ptr(b["up"]), _ptr(b["dwarped"]), df2, s))
            else:
                _native.check(lib.ldic_nn_corr_backward(ctypes.byref(b["corr"]), f1, f2, _ptr(b["dx"]), stored, 0,
                                                        df1, c, 0, df2, c, 0, s))
            _native.check(lib.ldic_nn_add_slice(n * hh * ww, c, _ptr(b["dx"]), stored, d2, df1, c, 0, s))
        for i in reversed(range(len(a.encoder))):
            p = self.params[f"enc{i}"]
            _native.check(lib.ldic_nn_leaky_backward(self.denc[i].numel(), _ptr(self.enc[i]),
                                                     _ptr(self.denc[i]), s))
            x = self.img16 if i == 0 else self.enc[i - 1]
            dx = None if i == 0 else self.denc[i - 1]
            self._conv_back(p, self._gconv[f"enc{i}"], x, self.denc[i], dx,
                            0 if i == 0 else a.encoder[i - 1][0], True)

    def adam_step(self, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> None:
        self.step_count += 1
        lib, s = self.lib, self.stream_handle
        for p in self.params.values():
            _native.check(lib.ldic_nn_adam(p["w"].numel(), _ptr(p["w"]), _ptr(p["dw"]), _ptr(p["mw"]), _ptr(p["vw"]),
                                           lr, beta1, beta2, eps, self.step_count, _ptr(p["w16"]), s))
            _native.check(lib.ldic_nn_adam(p["b"].numel(), _ptr(p["b"]), _ptr(p["db"]), _ptr(p["mb"]), _ptr(p["vb"]),
                                           lr, beta1, beta2, eps, self.step_count, None, s))
            _native.check(lib.ldic_nn_transpose_weights(p["cin"], p["cout"], _ptr(p["w16"]), _ptr(p["wt16"]), s))

    def autotune(self, h: int, w: int, n: int = 1, reps: int = 5) -> dict:
        """Pick each forward convolution's output-channel tile (32, 64, 128 per workgroup) and its
        split of the GEMM columns (1, 2, 4, 8) by timing whole forward passes on device-resident
        inputs of this shape, one layer at a time."""
        import time

        rng = np.random.default_rng(0)
        imgs = DeviceArray.from_numpy(rng.integers(0, 256, (2 * n, h, w), dtype=np.uint8))
        ref, tgt = _ptr(imgs), _ptr(imgs) + n * h * w

        def timed():
            # Wall time of `reps` queued forwards between two waits on the stream: the
            # kernels of one forward keep the device busy longer than their launches take.
            self.forward(ref, tgt, n=n, h=h, w=w)
            self.synchronize()
            start = time.perf_counter()
            for _ in range(reps):
                self.forward(ref, tgt, n=n, h=h, w=w)
            self.synchronize()
            return (time.perf_counter() - start) * 1e3 / reps

        for p in self.params.values():
            p["tile"], p["splits"] = 0, 1
        choice = {}
        for name, p in self.params.items():
            best = None
            for tile in (32, 64, 128):
                for splits in (1, 2, 4, 8):
                    if splits > 1 and 9 * p["cin"] < 128 * splits:
                        continue
                    p["tile"], p["splits"] = tile, splits
                    ms = timed()
                    if best is None or ms < best[0]:
                        best = (ms, tile, splits)
            p["tile"], p["splits"] = best[1], best[2]
            choice[name] = [best[1], best[2]]
        choice["forward_ms"] = timed()
        return choice

    def output_scale(self) -> int:
        """Finest-level pixels per pixel of the last stage's flow."""
        return self.arch.scale(self.arch.stages[-1].level)


# -- training data ------------------------------------------------------------------------------
FRAME = (3840, 2160)


def random_affine(rng: np.random.Generator, width: int = FRAME[0], height: int = FRAME[1],
                  translation: float = 64.0, rotation_deg: float = 10.0, strain: float = 0.03):
    """t ~ U[-64, 64] px, rotation ~ U[-10, 10] deg, normal strains ~ U[-3 %, 3 %] about the frame centre."""
    from .synthetic import affine_field

    return affine_field(width, height, rng.uniform(-translation, translation), rng.uniform(-translation, translation),
                        math.radians(rng.uniform(-rotation_deg, rotation_deg)), rng.uniform(-strain, strain),
                        rng.uniform(-strain, strain))


class TrainingPairs:
    """Batches of crops of a 3840x2160 speckle frame and its random affine deformation, rendered
    on the device by the engine's own renderer, with per-stage flow targets."""

    def __init__(self, net: SeedNet, *, batch: int = 8, crop: int = 1024, seed: int = 1, noise: float = 2.0,
                 texture_refresh: int = 200):
        from . import synthetic

        self.net, self.batch, self.crop, self.noise = net, batch, crop, noise
        self.rng = np.random.default_rng(seed)
        self.refresh = texture_refresh
        self.synthetic = synthetic
        self.full = _zeros((2 * batch, crop, crop), np.uint8)
        c = crop >> net.arch.input_level
        self.half = _zeros((2 * batch, crop // 2, crop // 2), np.uint8)
        self.small = _zeros((2 * batch, c, c), np.uint8)
        self.fields = DeviceArray(2 * batch * ctypes.sizeof(_native.Field), np.uint8)
        self.offsets = DeviceArray((2 * batch, 2), np.int32)
        self.size = c
        self.pairs = 0
        self._texture = None
        self._new_texture()

    def _new_texture(self):
        w, h = FRAME
        count = int(round(self.synthetic.BLOB_DENSITY * w * h))
        table = self.synthetic.blobs(w, h, count, int(self.rng.integers(1 << 31)))
        self._texture = self.synthetic.SpeckleTexture(w, h, table)

    def next(self):
        """Render one batch; returns (device pointers of reference and target inputs, truths per stage)."""
        net, lib = self.net, self.net.lib
        if self.pairs and self.pairs // self.batch % self.refresh == 0:
            self._new_texture()
        B, C = self.batch, self.crop
        fields = (_native.Field * (2 * B))()
        offsets = np.zeros((2 * B, 2), np.int32)
        flds = []
        for b in range(B):
            fld = random_affine(self.rng)
            flds.append(fld)
            fields[B + b] = fld.native()
            ox = int(self.rng.integers(0, FRAME[0] - C + 1))
            oy = int(self.rng.integers(0, FRAME[1] - C + 1))
            offsets[b] = offsets[B + b] = (ox, oy)
        # The previous batch's render may still read the field and offset tables.
        net.synchronize()
        self.fields.upload(np.frombuffer(bytes(fields), np.uint8))
        self.offsets.upload(offsets)
        fbuf, obuf = self.fields, self.offsets
        t = self._texture
        sp = _native.Speckle(t._blobs.address, t._starts.address, t.cells_x, t.cells_y, t.cell, float(t.origin),
                             t.sigma, t.background)
        s = net.stream_handle
        _native.check(lib.ldic_render_speckle_batch(ctypes.byref(sp), _ptr(fbuf), _ptr(obuf), 2 * B, C, C,
                                                    self.noise, int(self.rng.integers(1 << 62)), _ptr(self.full),
                                                    s))
        _native.check(lib.ldic_reduce_device(_ptr(self.full), C, C, 2 * B, _ptr(self.half), s))
        _native.check(lib.ldic_reduce_device(_ptr(self.half), C // 2, C // 2, 2 * B, _ptr(self.small), s))
        truths = []
        a = net.arch
        for st in a.stages:
            sc = a.scale(st.level)
            hh = ww = self._level_size(st.level)
            yy, xx = np.meshgrid(np.arange(hh, dtype=np.float64), np.arange(ww, dtype=np.float64), indexing="ij")
            tr = np.zeros((B, hh, ww, 2), np.float32)
            for b, fld in enumerate(flds):
                u, v = fld.displacement(xx * sc + offsets[b, 0], yy * sc + offsets[b, 1])
                tr[b, ..., 0], tr[b, ..., 1] = u / sc, v / sc
            truths.append(tr)
        self.pairs += B
        half = B * self.size * self.size
        return _ptr(self.small), _ptr(self.small) + half, truths

    def _level_size(self, level: int) -> int:
        n = self.size
        for _, stride in self.net.arch.encoder[:level + 1]:
            n = (n - 1) // stride + 1
        return n


def train(net: SeedNet, *, pairs: int = 50_000, batch: int = 8, crop: int = 1024, lr: float = 1e-3,
          weights=(0.32, 0.16, 0.08), log=None, seed: int = 1):
    """Train on `pairs` freshly rendered pairs; returns the loss history (one value per step)."""
    data = TrainingPairs(net, batch=batch, crop=crop, seed=seed)
    steps = max(1, pairs // batch)
    history = []
    for k in range(steps):
        ref, tgt, truths = data.next()
        net.forward(ref, tgt, n=batch, h=data.size, w=data.size, training=True)
        net.backward(truths, weights)
        rate = lr * (0.5 * (1 + math.cos(math.pi * k / steps)) * 0.95 + 0.05)
        net.adam_step(rate)
        if k % 50 == 0 or k == steps - 1:
            value = net.loss.item()
            history.append((k, value))
            if log:
                log(f"step {k}/{steps} loss {value:.5f}")
    net.synchronize()
    return history


class LearnedSeeder:
    """Finest-level seeds of a single-level engine from the network, queued on the engine's slot
    stream: the target is reduced twice by the integer pyramid, the network runs on the pair,
    and ``ldic_nn_seed_from_flow`` reads translation and gradient seeds at every subset centre.
    The seeded solve, with the ZNCC fallback and ``reseed_passes`` passes of neighbour
    re-seeding (unaccepted subsets restart from their nearest accepted subset's warp), then
    runs on the same stream."""

    def __init__(self, engine, net: SeedNet, *, slot: int = 0, fallback_zncc: float = 0.9,
                 fallback_radius: int = 8, reseed_passes: int = 2, gradients: bool = True):
        self.engine, self.net, self.slot = engine, net, slot
        lib = engine._lib
        handle = ctypes.c_void_p()
        _native.check(lib.ldic_slot_stream(engine._handle, slot, ctypes.byref(handle)))
        self.stream = int(handle.value)
        net.use_stream(self.stream)
        engine._set_initialization(_native.INIT_ZERO, 0, fallback_zncc, fallback_radius, reseed_passes=reseed_passes)
        self.width, self.height, self.gw, self.gh = engine._level_shapes[0]
        L = net.arch.input_level
        self.sizes = [(self.width >> k, self.height >> k) for k in range(L + 1)]
        self.ref_levels = [_zeros((h, w), np.uint8) for w, h in self.sizes[1:]]
        self.tgt_levels = [_zeros((h, w), np.uint8) for w, h in self.sizes[1:]]
        n = self.gw * self.gh
        self.seeds = _zeros((n, 8), np.float32)
        self.active = _zeros(n, np.int32)
        self.gradients = gradients
        self.scale = net.output_scale()
        w, h = self.sizes[-1]
        self.tiles = net.autotune(h, w)

    def _reduce(self, device_image: int, levels) -> None:
        lib = self.engine._lib
        src = device_image
        for k, dst in enumerate(levels):
            w, h = self.sizes[k]
            _native.check(lib.ldic_reduce_device(src, w, h, 1, _ptr(dst), self.stream))
            src = _ptr(dst)

    def set_reference(self, device_reference: int) -> None:
        """Reduce the (already prepared) reference once; device_reference holds the finest plane."""
        self._reduce(device_reference, self.ref_levels)

    def submit_pair(self, device_reference: int, device_target: int) -> None:
        """Consecutive-pair mode: queue the reference's preparation (finest level) and its
        reduction, then the target's registration, all on the slot stream."""
        _native.check(self.engine._lib.ldic_prepare_reference_resident(self.engine._handle, self.slot,
                                                                        device_reference, self.width))
        self.set_reference(device_reference)
        self.submit(device_target)

    def submit(self, device_target: int, row_stride: int | None = None) -> None:
        """Queue reduction, network, seeds and the seeded solve of one resident target."""
        lib = self.engine._lib
        self._reduce(device_target, self.tgt_levels)
        w, h = self.sizes[-1]
        flows = self.net.forward(_ptr(self.ref_levels[-1]), _ptr(self.tgt_levels[-1]), n=1, h=h, w=w)
        flow = flows[-1]
        spec = _native.NNSeed(self.gw, self.gh, self.engine.subset_size, self.engine.subset_stride,
                              flow.shape[2], flow.shape[1], float(self.scale), 8, 1 if self.gradients else 0, None)
        _native.check(lib.ldic_nn_seed_from_flow(ctypes.byref(spec), _ptr(flow), _ptr(self.seeds),
                                                 _ptr(self.active), self.stream))
        _native.check(lib.ldic_register_seeded_resident(self.engine._handle, self.slot, device_target,
                                                        row_stride or self.width, _ptr(self.seeds),
                                                        _ptr(self.active)))
