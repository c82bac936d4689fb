"""Extract the reference's own fixtures and WGSL golden fields used by the parity tests.

Reads (as data, never executes) files of the reference checkout:
- DIC/LucasKanadeICGN/assets/frame_000.png, frame_010.png (512x512 translation fixtures,
  assets/source.json: 0.2/-0.1 px per frame);
- DIC/LucasKanadeICGN/assets/uhd_tensile/reference.png, frame_040.png (native UHD tensile pair);
- DIC/LucasKanadeICGN/benchmarks/results/binding-budget-20261031/before.npz, the WGSL fields
  captured by benchmarks/qualify_binding_budget.py (ICGNWGPUFullPyramid, outlier_policy="median",
  16/16, B-spline; ROI mask [160:2000,160:3680] minus [880:1200,1800:2040]).
- DIC/LucasKanadeICGN/assets/uhd_tensile/machine_truth.csv and manifest.json: the analytic
  stretches of frame 40 and the manifest's regression_limits;
- DIC/LucasKanadeICGN/benchmarks/results/browser-gpu-reference-20261101/affine-uhd.json: the
  reference's published accuracy of the frame-40 affine field (2 levels, 16/16, median).
Writes tests/golden/*.png (copied bytes), tests/golden/uhd_goldens.npz (selected cases) and
tests/golden/uhd_tensile_truth.json.
The assets and golden fields are GPL-2.0-or-later data of the reference project.
"""

from __future__ import annotations

import csv
import json
import shutil
from pathlib import Path

import numpy as np

SRC = Path("/root/reference/DIC/LucasKanadeICGN")
OUT = Path(__file__).resolve().parent
CASES = (
    "affine-l1-none-zero",
    "translation-l1-none-zero",
    "quadratic-l1-none-zero",
    "affine-l2-none-zero",
    "affine-l3-none-zero",
    "quadratic-l2-none-zero",
    "affine-l1-full-zero",
    "translation-l2-full-zero",
)
# Partial-support ROI cases: frame 1 only (the tensile pair, or the (-1, 2) roll for translation).
PARTIAL_CASES = (
    "translation-l1-partial-zero",
    "affine-l1-partial-zero",
    "translation-l2-partial-zero",
    "affine-l2-partial-zero",
)


# Temporal-initialisation cases (translation: frames are rolls of the reference; affine: frame 1
# is the tensile pair), frames 1 and/or 2 (frame 0 equals the zero start).
TEMPORAL_CASES = (
    ("translation-l2-none-temporal", (1, 2)),
    ("translation-l3-none-temporal", (1, 2)),
    ("affine-l2-none-temporal", (1,)),
    # Latency profile: velocity prediction, fine correction, divergence patience 3 from iteration 4.
    ("translation-l2-none-latency", (1, 2)),
    ("translation-l2-full-latency", (2,)),
    ("translation-l2-partial-latency", (2,)),
)


def main() -> None:
    shutil.copyfile(SRC / "assets/frame_000.png", OUT / "frame_000.png")
    shutil.copyfile(SRC / "assets/frame_010.png", OUT / "frame_010.png")
    shutil.copyfile(SRC / "assets/uhd_tensile/reference.png", OUT / "uhd_reference.png")
    shutil.copyfile(SRC / "assets/uhd_tensile/frame_040.png", OUT / "uhd_frame_040.png")
    fields = {}
    with np.load(SRC / "benchmarks/results/binding-budget-20261031/before.npz", allow_pickle=False) as data:
        for case in CASES:
            for frame in (0, 1):
                for key in ("parameters", "status", "iterations"):
                    name = f"{case}-f{frame}-{key}"
                    fields[name] = data[name]
        for case, frames in TEMPORAL_CASES:
            for frame in frames:
                for key in ("parameters", "status", "iterations"):
                    name = f"{case}-f{frame}-{key}"
                    fields[name] = data[name]
        for case in PARTIAL_CASES:
            for key in ("parameters", "status", "iterations"):
                name = f"{case}-f1-{key}"
                fields[name] = data[name]
    np.savez_compressed(OUT / "uhd_goldens.npz", **fields)
    print("wrote", len(fields), "arrays")
    tensile = SRC / "assets/uhd_tensile"
    with (tensile / "machine_truth.csv").open(newline="", encoding="utf-8") as stream:
        rows = {int(row["frame"]): row for row in csv.DictReader(stream)}
    manifest = json.loads((tensile / "manifest.json").read_text(encoding="utf-8"))
    truth = {
        "source": "assets/uhd_tensile/machine_truth.csv (frame 40) and manifest.json",
        "forward_map": manifest["forward_map"],
        "frame_040": {key: float(rows[40][key]) for key in ("stretch_x", "stretch_y")},
        "regression_limits": manifest["regression_limits"],
    }
    report = json.loads((SRC / "benchmarks/results/browser-gpu-reference-20261101/affine-uhd.json")
                        .read_text(encoding="utf-8"))
    truth["published_frame_040_affine"] = {
        "source": "benchmarks/results/browser-gpu-reference-20261101/affine-uhd.json",
        "engine_options": report["engineOptions"],
        "accepted": report["quality"]["cpuAccepted"],
        "analytical": report["analytical"],
    }
    (OUT / "uhd_tensile_truth.json").write_text(json.dumps(truth, indent=2) + "\n", encoding="utf-8")


if __name__ == "__main__":
    main()
