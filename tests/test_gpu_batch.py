"""GPU: batched heterogeneous registration (persistent solves, autotuner, HIP graph).

Every pair of the batch must come out bit-identical to the same engine
registering its pair alone, in eager mode and on every replay of the captured
graph, with the program's own eager work between the replays.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (width, height, subset, step, model, levels)
PAIRS = [
    (320, 240, 15, 4, "affine", 1),
    (480, 360, 21, 8, "quadratic", 2),
    (640, 480, 31, 16, "affine", 2),
    (800, 600, 41, 8, "affine", 3),
    (400, 300, 21, 4, "affine", 2),
]


def _say(*args):
    print("[case]", *args, flush=True)


def _render(w, h, seed):
    from live_dic import _native, seednet, synthetic

    rng = np.random.default_rng(seed)
    fld = seednet.random_affine(rng, w, h, translation=5.0, rotation_deg=1.0, strain=0.01)
    tex = synthetic.SpeckleTexture(w, h, synthetic.blobs(w, h, int(round(synthetic.BLOB_DENSITY * w * h)), seed))
    ref = tex.render(None, noise_sigma=1.0, noise_seed=seed * 3 + 1)
    tgt = _native.DeviceBuffer(w * h)
    tex.render_device(tgt, fld, 1.0, seed * 3 + 2)
    host = np.empty((h, w), np.uint8)
    tgt.download(host)
    return ref, tgt, host, fld


def test_batch_matches_single_engines_eager_and_graph():
    from live_dic import ICGNWGPUFullPyramid
    from live_dic.batch import BatchRegistration

    engines, targets, hosts, expected = [], [], [], []
    for i, (w, h, S, D, model, levels) in enumerate(PAIRS):
        _say("pair", i, w, h, S, D, model, levels)
        ref, tgt, host, _ = _render(w, h, 40 + i)
        eng = ICGNWGPUFullPyramid(pyramid_levels=levels, subset_size=S, subset_stride=D, motion_model=model)
        eng.set_reference(ref)
        res = eng.register(host)
        stride = 16 if model == "quadratic" else 8
        exp = np.empty((*res.status.shape, stride), np.float32)
        from live_dic import _native

        _native.check(_native.library().ldic_slot_output(eng._handle, 0, exp.ctypes.data))
        engines.append(eng)
        targets.append(tgt)
        hosts.append(host)
        expected.append(exp)
        assert np.mean(res.status > 0.5) > 0.5
    # An engine outside the batch does the program's own eager work between graph replays.
    other = ICGNWGPUFullPyramid(pyramid_levels=1, subset_size=21, subset_stride=8)
    other.set_reference(np.ascontiguousarray(np.flipud(hosts[0])))
    with BatchRegistration(engines, [t.address for t in targets]) as batch:
        _say("batch eager")
        batch.launch(use_graph=False)
        batch.wait()
        for got, exp in zip(batch.outputs(), expected):
            np.testing.assert_array_equal(got, exp)
        _say("autotune and capture")
        table = batch.tune()
        assert all(g["grid"] > 0 and g["candidates_ms"] for g in table)
        for replay in range(3):
            _say("graph replay", replay)
            # Eager work of the program between replays: a registration on another engine.
            other.register(hosts[0])
            batch.launch(use_graph=True)
            ms = batch.wait()
            assert ms > 0
            for got, exp in zip(batch.outputs(), expected):
                np.testing.assert_array_equal(got, exp)
    other.close()
    for e in engines:
        e.close()


def test_batch_holds_its_engines_until_closed():
    """While a batch exists its engines refuse registrations, a new reference, new level
    settings and close(); after the batch closes they work again."""
    import ctypes

    from live_dic import ICGNWGPUFullPyramid, _native
    from live_dic.batch import BatchRegistration
    from live_dic.engine import ICGNError

    lib = _native.library()
    ref, tgt, host, _ = _render(320, 240, 71)
    eng = ICGNWGPUFullPyramid(pyramid_levels=1, subset_size=21, subset_stride=8)
    eng.set_reference(ref)
    before = eng.register(host).parameters.copy()
    batch = BatchRegistration([eng], [tgt.address])
    with pytest.raises(ICGNError):
        eng.register(host)
    with pytest.raises(ICGNError):
        eng.set_reference(ref)
    with pytest.raises(ICGNError):
        eng.close()
    cfg = _native.LevelConfig(10, 1e-3, 0.8)
    assert lib.ldic_set_level_config(eng._handle, 0, ctypes.byref(cfg)) == -2
    with pytest.raises(ValueError):
        BatchRegistration([eng, eng], [tgt.address, tgt.address])
    batch.launch(use_graph=False)
    batch.wait()
    batch.close()
    np.testing.assert_array_equal(eng.register(host).parameters, before)
    eng.close()
    tgt.close()


def test_reference_change_waits_for_other_slots():
    """A resident reference preparation on one slot is refused while another slot has queued
    frames that read the current reference, and accepted once they are done."""
    from live_dic import ICGNWGPU, _native

    lib = _native.library()
    ref, tgt, host, _ = _render(320, 240, 72)
    eng = ICGNWGPU(subset_size=21, subset_stride=8)
    eng.set_reference(ref)
    _native.check(lib.ldic_set_slots(eng._handle, 2))
    dref = _native.DeviceBuffer(320 * 240)
    dref.upload(ref)
    _native.check(lib.ldic_register_resident(eng._handle, 1, tgt.address, 320))
    assert lib.ldic_prepare_reference_resident(eng._handle, 0, dref.address, 320) == -2
    assert lib.ldic_register_pair_resident(eng._handle, 0, dref.address, tgt.address, 320) == -2
    _native.check(lib.ldic_slot_wait(eng._handle, 1, None))
    _native.check(lib.ldic_prepare_reference_resident(eng._handle, 0, dref.address, 320))
    _native.check(lib.ldic_slot_wait(eng._handle, 0, None))
    eng.close()
    dref.close()
    tgt.close()
