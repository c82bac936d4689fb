"""Build liblivedic.so for gfx950 with hipcc (in-tree, no JIT cache).

Usage: python live-dic_amd/build.py [--jobs N]
"""

from __future__ import annotations

import argparse
import concurrent.futures
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
CSRC = HERE / "csrc"
OUT_DIR = HERE / "live_dic"
LIBRARY = OUT_DIR / "liblivedic.so"
SOURCES = (
    "ldic_filters.hip",
    "ldic_solver.hip",
    "ldic_post.hip",
    "ldic_prep.hip",
    "ldic_render.hip",
    "ldic_seed.hip",
    "ldic_seednet.hip",
    "ldic_roi.hip",
    "ldic_runtime.hip",
)
ARCH = os.environ.get("LDIC_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wall",
    "-Wno-unused-result",
    "-I", str(ROOT / "include"),
    "-I", str(CSRC),
]


def _compile(source: str, obj_dir: Path) -> Path:
    obj = obj_dir / (Path(source).stem + ".o")
    cmd = ["hipcc", *FLAGS, "-c", str(CSRC / source), "-o", str(obj)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed for {source}:\n{proc.stdout}\n{proc.stderr}")
    return obj


def build(jobs: int | None = None) -> Path:
    obj_dir = HERE / "_build"
    obj_dir.mkdir(exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 1)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as pool:
        objects = list(pool.map(lambda s: _compile(s, obj_dir), SOURCES))
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = ["hipcc", f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objects)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{proc.stdout}\n{proc.stderr}")
    os.replace(tmp, LIBRARY)
    return LIBRARY


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("--jobs", type=int, default=None)
    args = parser.parse_args()
    print(build(args.jobs))
    return 0


if __name__ == "__main__":
    sys.exit(main())
