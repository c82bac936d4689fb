"""GPU: the learned seed network end to end.

- Training on engine-rendered pairs, by the library's own forward and backward
  kernels, lowers the loss.
- The committed trained network, seeding one finest-level IC-GN solve with the
  ZNCC fallback, registers a large-displacement 4K pair (translation, rotation
  and strain of the training distribution) that the same solve without the
  network does not.
"""

from __future__ import annotations

import math
import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

W, H, S, D = 3840, 2160, 31, 8


def _say(*args):
    print("[case]", *args, flush=True)


def test_training_steps_fit_a_fixed_batch():
    """The library's forward, backward and Adam kernels, run on one fixed batch of four
    rendered pairs, drive its loss down: the training loop works end to end."""
    from live_dic import seednet

    net = seednet.SeedNet(seed=5)
    batch = seednet.TrainingPairs(net, batch=4, crop=1024, seed=12345)
    ref, tgt, truths = batch.next()
    weights = (0.32, 0.16, 0.08)
    losses = []
    for step in range(150):
        net.forward(ref, tgt, n=4, h=batch.size, w=batch.size, training=True)
        net.backward(truths, weights)
        if step % 25 == 0 or step == 149:
            losses.append(float(net.loss.item()))
        net.adam_step(1e-3)
    _say("fixed-batch loss", [round(v, 4) for v in losses])
    assert np.isfinite(losses[-1]) and losses[-1] < 0.5 * losses[0]


def _pair(seed):
    from live_dic import _native, seednet, synthetic

    rng = np.random.default_rng(seed)
    fld = seednet.random_affine(rng)
    tex = synthetic.SpeckleTexture(W, H, synthetic.blobs(W, H, int(round(synthetic.BLOB_DENSITY * W * H)), seed))
    ref = _native.DeviceBuffer(W * H)
    tgt = _native.DeviceBuffer(W * H)
    tex.render_device(ref, None, 2.0, 1)
    tex.render_device(tgt, fld, 2.0, 2)
    host = np.empty((H, W), np.uint8)
    ref.download(host)
    return ref, tgt, host, fld


def _score(state, fld):
    from live_dic import synthetic

    cx, cy = synthetic.subset_centres(W, H, S, D)
    u, v = fld.displacement(cx, cy)
    inside = (cx + u > S) & (cx + u < W - S) & (cy + v > S) & (cy + v < H - S)
    ok = (state[..., 6] > 0.5) & inside
    err = np.hypot(state[..., 0] - u, state[..., 3] - v)[ok]
    return ok.sum() / inside.sum(), float(np.median(err)) if err.size else np.inf


def test_trained_network_seeds_large_displacements():
    from live_dic import ICGNWGPU, _native, seednet

    if not seednet.WEIGHTS_PATH.exists():
        pytest.skip("no trained weights committed")
    _say("learned seeds vs zero seeds, 4K pair of the training distribution")
    ref, tgt, host, fld = _pair(20261204)
    lib = _native.library()
    with ICGNWGPU(subset_size=S, subset_stride=D) as eng:
        eng.set_reference(host, iterations=20, tolerance=1e-3)
        gw, gh = eng.grid_width, eng.grid_height
        zero = eng.register_seeded(_download(tgt), np.zeros((gh, gw, 6), np.float32), np.ones((gh, gw), bool),
                                   fallback_zncc=0.9, fallback_radius=8)
        net = seednet.SeedNet()
        net.load(seednet.WEIGHTS_PATH)
        seeder = seednet.LearnedSeeder(eng, net)
        seeder.set_reference(ref.address)
        seeder.submit(tgt.address)
        _native.check(lib.ldic_slot_wait(eng._handle, 0, None))
        state = np.empty((gh, gw, 8), np.float32)
        _native.check(lib.ldic_slot_output(eng._handle, 0, state.ctypes.data))
    without = _score(np.concatenate([zero.parameters, zero.status[..., None], zero.iterations[..., None]], -1), fld)
    learned = _score(state, fld)
    _say("accepted fraction / median error: without network", without, "with network", learned)
    assert learned[0] > 0.9 and learned[1] < 0.02
    assert without[0] < 0.5 * learned[0]


def _download(buf):
    out = np.empty((H, W), np.uint8)
    buf.download(out)
    return out


_NO_TORCH_RUN = r"""
import sys
import numpy as np
from live_dic import ICGNWGPU, _native, seednet, synthetic

W, H, S, D = 1920, 1080, 31, 8
rng = np.random.default_rng(7)
fld = seednet.random_affine(rng, W, H, translation=12.0, rotation_deg=2.0, strain=0.01)
tex = synthetic.SpeckleTexture(W, H, synthetic.blobs(W, H, int(round(synthetic.BLOB_DENSITY * W * H)), 7))
ref, tgt = _native.DeviceBuffer(W * H), _native.DeviceBuffer(W * H)
tex.render_device(ref, None, 2.0, 1)
tex.render_device(tgt, fld, 2.0, 2)
host = np.empty((H, W), np.uint8)
ref.download(host)
lib = _native.library()
with ICGNWGPU(subset_size=S, subset_stride=D) as eng:
    eng.set_reference(host, iterations=20, tolerance=1e-3)
    net = seednet.SeedNet()
    net.load(seednet.WEIGHTS_PATH)
    seeder = seednet.LearnedSeeder(eng, net)
    seeder.set_reference(ref.address)
    seeder.submit(tgt.address)
    _native.check(lib.ldic_slot_wait(eng._handle, 0, None))
    state = np.empty((eng.grid_height, eng.grid_width, 8), np.float32)
    _native.check(lib.ldic_slot_output(eng._handle, 0, state.ctypes.data))
cx, cy = synthetic.subset_centres(W, H, S, D)
u, v = fld.displacement(cx, cy)
ok = state[..., 6] > 0.5
err = np.hypot(state[..., 0] - u, state[..., 3] - v)[ok]
print("accepted", ok.mean(), "median", float(np.median(err)), flush=True)
assert ok.mean() > 0.9 and np.median(err) < 0.02
assert "torch" not in sys.modules, "torch was imported"
"""


def test_learned_path_runs_without_torch():
    """The seed network, its seeder and the seeded solve run in a process that never imports
    a tensor framework (a 1080p pair of modest displacement)."""
    if not _trained():
        pytest.skip("no trained weights committed")
    _say("learned seeder in a torch-free process")
    root = Path(__file__).resolve().parents[1]
    proc = subprocess.run([sys.executable, "-u", "-c", _NO_TORCH_RUN], capture_output=True, text=True, timeout=240,
                          env={**os.environ, "PYTHONPATH": str(root / "live-dic_amd")})
    print(proc.stdout[-1000:])
    assert proc.returncode == 0, proc.stderr[-3000:]


def _trained() -> bool:
    from live_dic import seednet

    return seednet.WEIGHTS_PATH.exists()
