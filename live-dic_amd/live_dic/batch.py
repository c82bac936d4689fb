"""Batched heterogeneous registration: many image pairs, one frame sequence.

No counterpart in the reference, which registers one pair per engine.  Each
engine (``ICGNWGPU`` or ``ICGNWGPUFullPyramid``, prepared with
``set_reference``) contributes one pair with its own size, subset size, step,
warp model and pyramid depth; ``ldic_batch_*`` runs all of them per frame with
one persistent solve launch per level and solver variant, grid sizes picked by
the on-device autotuner, the whole frame replayed as one HIP graph.
"""

from __future__ import annotations

import ctypes
import json

import numpy as np

from . import _native
from .engine import ICGNError


class BatchRegistration:
    """Holds each member engine's registration lock and the library's batch reference from
    creation to close(): meanwhile the engines refuse registrations (``ICGNError``), and the
    library refuses to re-prepare, reconfigure or destroy them."""

    def __init__(self, engines, device_targets):
        if len(engines) != len(device_targets) or not engines:
            raise ValueError("one device target per engine is required")
        if len({id(e) for e in engines}) != len(engines):
            raise ValueError("an engine can contribute only one pair to a batch")
        self.engines = list(engines)
        self._lib = _native.library()
        self._handle = ctypes.c_void_p()
        self._held = []
        try:
            for e in self.engines:
                if not e._lock.acquire(blocking=False):
                    raise ICGNError("Another registration or registration stream is active.")
                self._held.append(e)
            handles = (ctypes.c_void_p * len(engines))(*[e._handle.value for e in engines])
            targets = (ctypes.c_void_p * len(engines))(*[int(t) for t in device_targets])
            _native.check(self._lib.ldic_batch_create(handles, len(engines), targets, ctypes.byref(self._handle)))
        except BaseException:
            self._release_engines()
            raise
        for e in self.engines:
            e._batch_holders += 1
        self.tuned = False

    def _release_engines(self) -> None:
        while self._held:
            self._held.pop()._lock.release()

    def tune(self) -> list[dict]:
        """Run the autotuner and capture the frame graph; returns the launch table."""
        _native.check(self._lib.ldic_batch_tune(self._handle))
        self.tuned = True
        return self.describe()

    def describe(self) -> list[dict]:
        buf = ctypes.create_string_buffer(1 << 16)
        _native.check(self._lib.ldic_batch_describe(self._handle, buf, len(buf)))
        return json.loads(buf.value.decode())

    def launch(self, use_graph: bool = True) -> None:
        if use_graph and not self.tuned:
            raise ICGNError("call tune() before launching the graph")
        _native.check(self._lib.ldic_batch_launch(self._handle, 1 if use_graph else 0))

    def wait(self) -> float:
        """Wait for the queued frames; returns the device milliseconds of the last one."""
        ms = ctypes.c_double()
        _native.check(self._lib.ldic_batch_wait(self._handle, ctypes.byref(ms)))
        return ms.value

    def outputs(self) -> list[np.ndarray]:
        """Packed finest-level states of every pair (after wait())."""
        out = []
        for e in self.engines:
            _, _, gw, gh = e._level_shapes[0]
            st = np.empty((gh, gw, 16 if e.motion_model == "quadratic" else 8), np.float32)
            _native.check(self._lib.ldic_slot_output(e._handle, 0, st.ctypes.data))
            out.append(st)
        return out

    def close(self) -> None:
        if self._handle:
            self._lib.ldic_batch_destroy(self._handle)
            self._handle = ctypes.c_void_p()
            for e in self.engines:
                e._batch_holders -= 1
        self._release_engines()

    def __enter__(self):
        return self

    def __exit__(self, *_):
        self.close()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass
