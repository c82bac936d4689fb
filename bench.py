"""Throughput of the 4K fixed-reference affine IC-GN sequence on one MI355X.

Workload (one step = one frame registered against the fixed reference):
3840x2160 uint8 analytic speckle (Gaussian blobs, sigma 2.5 px, amplitudes
U[80, 200], 30 grey background, noise sigma 1 grey; 126,563 blobs, the density of
the 512x512 pair, because 1.2 M summed blobs saturate the image at 255), a 60-frame sequence with
u_x = s(2 + 0.002x + 1.5 sin(2 pi y/2160)), u_y = s(-1 + 0.001y + 0.8 cos(2 pi x/3840)),
s = k/59; affine IC-GN, subset 31, step 8 (127,359 subsets), B-spline, FP32,
two-level pyramid, GPU reference preparation.  Frames are rendered on the GPU
and stay resident in HBM; the timed region covers the per-frame work only
(pyramid, prefilters, solves, outlier passes, seed transfer).

Prints one JSON line (see the repository's measurement contract).
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "live-dic_amd"))

METRIC = "4K frames/s + IC-GN subsets/s; learned vs ZNCC seed at ≤0.01 px RMS error"
HBM_PEAK_GBPS = 8000.0
# HBM bytes of one fine-level k_solve launch of this workload, measured with rocprofv3
# FETCH_SIZE/WRITE_SIZE and scaled by a 256 MiB copy calibration (reads 2.0x FETCH_SIZE,
# writes 1.0x WRITE_SIZE): profiles/k_solve_affine_4k_s31_d8_pmc.md.
MEASURED_FINE_SOLVE_TRAFFIC_BYTES = (2 * 117_755 + 3_986) * 1024
SUBSET, STRIDE = 31, 8
WIDTH, HEIGHT = 3840, 2160
FRAMES = 60
LEVELS = 2
ITERATIONS = (10, 20)     # coarse, fine
TOLERANCES = (1e-2, 1e-3)


def subset_bytes(S: int, params: int = 6, meta: int = 16, state: int = 8) -> int:
    """Algorithmic HBM bytes of one subset solve at one level (independent of iterations):
    reference f, gx, gy (12 S^2), target coefficient footprint 4 (S+3)^2, metadata and
    inverse Hessian 4 (M + P^2), seed read and state write 2*4*stride."""
    return 12 * S * S + 4 * (S + 3) ** 2 + 4 * (meta + params * params) + 2 * 4 * state


def cpu_baseline(seq, frame_index: int, budget_s: float = 25.0) -> dict:
    """The C oracle (OpenMP, all host threads) on the same workload, on a bounded sample."""
    import oracle

    ref = seq.render(0)
    tgt = seq.render(frame_index)
    pipe = oracle.Pipeline(pyramid_levels=LEVELS, iterations_per_level=ITERATIONS, tolerance_per_level=TOLERANCES,
                           subset_size=SUBSET, subset_stride=STRIDE, outlier_policy="replace")
    pipe.set_reference(ref)
    t0 = time.perf_counter()
    pipe.register(tgt)
    per_frame = time.perf_counter() - t0
    frames = 1
    while time.perf_counter() - t0 + per_frame < budget_s and frames < 3:
        pipe.register(tgt)
        frames += 1
    elapsed = time.perf_counter() - t0
    threads = oracle.cpu_threads()
    return {
        "value": frames / elapsed,
        "unit": "frames/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{frames} full 4K frame(s) (frame {frame_index} of the 60-frame sequence, both pyramid levels) "
                  f"through the C oracle with {threads} OpenMP threads; reference prepared beforehand",
    }


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=120)
    parser.add_argument("--warmup", type=int, default=10)
    parser.add_argument("--drain", type=int, default=32, help="frames queued between host waits")
    parser.add_argument("--no-cpu-baseline", action="store_true")
    parser.add_argument("--calibrate-copy", action="store_true",
                        help="only copy 256 MiB with one plain device copy kernel (PMC byte-counter calibration)")
    args = parser.parse_args()
    if args.calibrate_copy:
        import torch

        src = torch.empty(256 << 20, dtype=torch.uint8, device="cuda").fill_(1)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copied_bytes": src.numel(), "read_bytes": src.numel(), "written_bytes": dst.numel()}))
        return 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (see DESIGN.md for the "
              "sharding design); run with --gpus 1", file=sys.stderr)
        return 2

    from live_dic import ICGNWGPUFullPyramid, _native, synthetic

    lib = _native.library()
    print(f"[bench] rendering {FRAMES} frames {WIDTH}x{HEIGHT}", file=sys.stderr, flush=True)
    seq = synthetic.UhdSequence(frames=FRAMES)
    frames = []
    for k in range(FRAMES):
        buf = _native.DeviceBuffer(WIDTH * HEIGHT)
        seq.render_device(k, buf)
        frames.append(buf)
    reference = seq.render(0)

    eng = ICGNWGPUFullPyramid(pyramid_levels=LEVELS, iterations_per_level=ITERATIONS,
                              tolerance_per_level=TOLERANCES, subset_size=SUBSET, subset_stride=STRIDE)
    print("[bench] preparing the reference on the GPU", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    info = eng.set_reference(reference)
    prep_ms = (time.perf_counter() - t0) * 1e3
    gw, gh = info.grid
    n_fine = gw * gh
    handle = eng._handle

    def run(steps: int, start: int) -> tuple[float, int]:
        """Queue frames on one stream; the host waits every `drain` frames.  Returns the
        summed finest-level solver time (HIP events around each launch) and its count."""
        total, count = 0.0, 0
        for i in range(steps):
            k = 1 + (start + i) % (FRAMES - 1)
            _native.check(lib.ldic_register_resident(handle, 0, frames[k].address, WIDTH))
            if (i + 1) % args.drain == 0 or i + 1 == steps:
                t = _native.Timing()
                _native.check(lib.ldic_slot_wait(handle, 0, ctypes.byref(t)))
                total += t.solve_ms
                count += t.frames
        return total, count

    print(f"[bench] warmup {args.warmup}", file=sys.stderr, flush=True)
    if args.warmup:
        run(args.warmup, 0)
    _native.check(lib.ldic_device_synchronize())
    print(f"[bench] timing {args.steps} frames", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    solve_total, solve_count = run(args.steps, args.warmup)
    _native.check(lib.ldic_device_synchronize())
    elapsed = time.perf_counter() - t0
    fps = args.steps / elapsed

    # Accuracy of the last timed frame against the analytic field.
    last = 1 + (args.warmup + args.steps - 1) % (FRAMES - 1)
    out = np.empty((gh, gw, 8), np.float32)
    _native.check(lib.ldic_slot_output(handle, 0, out.ctypes.data))
    cx, cy = synthetic.subset_centres(WIDTH, HEIGHT, SUBSET, STRIDE)
    u, v = seq.field.scaled(last / (FRAMES - 1)).displacement(cx, cy)
    ok = out[..., 6] > 0.5
    rmse = float(np.sqrt(np.mean((out[..., 0][ok] - u[ok]) ** 2 + (out[..., 3][ok] - v[ok]) ** 2)))
    eng.close()

    avg_solve_ms = solve_total / solve_count if solve_count else float("nan")
    traffic_bytes = n_fine * subset_bytes(SUBSET)
    achieved = traffic_bytes / (avg_solve_ms * 1e-3) / 1e9
    line = {
        "metric": METRIC,
        "value": fps,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (GPU-rendered analytic speckle, seeded)",
        "config": {
            "workload": "4K fixed-reference affine IC-GN sequence, 3840x2160, subset 31, step 8, "
                        "2-level pyramid, B-spline, 60 frames",
            "subsets_per_frame": n_fine,
            "subsets_per_s": fps * n_fine,
            "accepted_fraction_last_frame": float(ok.mean()),
            "rmse_px_last_frame": rmse,
            "reference_prep_ms": prep_ms,
            "iterations_per_level": list(ITERATIONS),
            "tolerance_per_level": list(TOLERANCES),
            "solver_launches_timed": solve_count,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": MEASURED_FINE_SOLVE_TRAFFIC_BYTES,
            "kernel": "k_solve<affine, bspline> (finest level)",
            "kernel_avg_ms": avg_solve_ms,
            "algorithmic_bytes_per_launch": traffic_bytes,
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        print("[bench] CPU oracle baseline", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(seq, 30)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
