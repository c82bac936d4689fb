"""The C ABI library loads and exports every symbol include/livedic.h declares (no GPU calls)."""

from __future__ import annotations

import ctypes
import os
import re
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols() -> set[str]:
    text = (ROOT / "include" / "livedic.h").read_text()
    return set(re.findall(r"\b(ldic_[a-z_0-9]+)\s*\(", text))


def test_library_exports_every_declared_symbol():
    from live_dic import _native

    lib = ctypes.CDLL(str(_native.LIBRARY_PATH))
    names = declared_symbols()
    assert len(names) >= 25
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_the_header():
    from live_dic import _native

    assert declared_symbols() == set(_native.SIGNATURES)


def test_abi_version_and_error_channel_without_device():
    from live_dic import _native

    lib = _native.library()
    assert lib.ldic_abi_version() == 1
    opts = _native.Options(0, 0, 3, 3, 0, 1.0, 1, 3, 0, 2, 0)  # subset 3 is invalid
    handle = ctypes.c_void_p()
    assert lib.ldic_create(ctypes.byref(opts), ctypes.byref(handle)) == -1
    assert b"subset_size" in lib.ldic_last_error()
    assert not handle.value


def test_engines_and_the_seed_network_load_without_torch():
    """Neither the library nor the seed-network module imports a tensor framework: device
    memory and streams come from the library itself."""
    code = ("import sys; from live_dic import ICGNWGPUFullPyramid, ICGNWGPU, seednet, batch, _native; "
            "_native.library(); assert 'torch' not in sys.modules, 'torch was imported'")
    env = {**os.environ, "PYTHONPATH": str(ROOT / "live-dic_amd")}
    proc = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stderr[-2000:]
