"""Throughput of the 4K fixed-reference affine IC-GN sequence on one MI355X.

Workload (one step = one frame registered against the fixed reference):
3840x2160 uint8 analytic speckle (Gaussian blobs, sigma 2.5 px, amplitudes
U[80, 200], 30 grey background, noise sigma 1 grey; 126,563 blobs, the density of
the 512x512 pair, because 1.2 M summed blobs saturate the image at 255), a 60-frame sequence with
u_x = s(2 + 0.002x + 1.5 sin(2 pi y/2160)), u_y = s(-1 + 0.001y + 0.8 cos(2 pi x/3840)),
s = k/59; affine IC-GN, subset 31, step 8 (127,359 subsets), B-spline, FP32,
two-level pyramid, GPU reference preparation.

`value` is measured at the reference's product boundary (LK/BENCHMARK.md:11-15): resident,
contiguous uint8 NumPy frames in host memory go through
ICGNWGPUFullPyramid.register_stream(pipeline_depth=2) and come back as materialised NumPy fields
and statistics; every upload and readback is inside the timed region.  The same frames resident in
HBM (`resident` key) time the device work alone.

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
# writes 1.0x WRITE_SIZE), and its L2 hit rate TCC_HIT / (TCC_HIT + TCC_MISS):
# profiles/k_solve_affine_4k_s31_d8_pmc.md.
MEASURED_FINE_SOLVE_TRAFFIC_BYTES = (2 * 116_892 + 4_286) * 1024
MEASURED_FINE_SOLVE_L2_HIT_RATE = 4.3226e7 / (4.3226e7 + 1.6606e6)
# FP32 arithmetic of the same launch (rocprofv3 SQ_INSTS_VALU_FMA_F32 + _MUL_F32 + _ADD_F32 wave
# instructions and SQ_INSTS_VALU_FLOPS_FP32, same profile), set against the published
# non-packed FP32 vector peak, 78.6 TFLOP/s (this build emits no packed FP32): one wave64 FP32
# FMA, multiply or add per 4 cycles per SIMD, 6.0e8 wave-instructions/s per SIMD over
# 256 CUs x 4 SIMDs at 2.4 GHz.  The launch's other VALU instructions (integer, conversion,
# move, DPP, lane reads: 25 % of SQ_INSTS_VALU) issue faster and are reported beside it.
MEASURED_FINE_SOLVE_VALU_INSTRUCTIONS = 3.6887e8
MEASURED_FINE_SOLVE_FP32_INSTRUCTIONS = 1.6092e8 + 7.3194e7 + 4.3653e7
MEASURED_FINE_SOLVE_FP32_FLOPS = 4.4039e8 * 64
SIMDS = 256 * 4
FP32_VECTOR_PEAK_FLOPS = 78.6e12
FP32_ISSUE_PEAK_PER_SIMD = FP32_VECTOR_PEAK_FLOPS / (SIMDS * 64 * 2)
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
                  f"through the C oracle with {threads} OpenMP threads, NumPy frame to NumPy field; reference "
                  "prepared beforehand",
    }


# Learned seed vs ZNCC seed at large displacement: 3840x2160 pairs, translation U[-64, 64] px,
# rotation U[-10, 10] deg, normal strains U[-3 %, 3 %] about the frame centre, noise 2 grey.
LARGE_PAIRS = 6
LARGE_LEVELS = 4                       # ZNCC path: 1/8-resolution search covers +-512 px
LARGE_SEARCH_RADIUS = 504              # finest-level pixels (coarsest-level radius 64)
LARGE_ITERATIONS = (12, 8, 8, 16)
LARGE_RESEED = 2                        # neighbour re-seeding passes after the finest solve, both paths
LARGE_SLOTS = 2                         # frames in flight (engine slots), both paths


def _large_targets(count: int, seed: int = 5150):
    """One reference and `count` device-resident targets of the large-displacement workload,
    each under its own random affine field."""
    from live_dic import _native, seednet, synthetic

    rng = np.random.default_rng(seed)
    tex = synthetic.SpeckleTexture(WIDTH, HEIGHT, synthetic.blobs(
        WIDTH, HEIGHT, int(round(synthetic.BLOB_DENSITY * WIDTH * HEIGHT)), seed))
    ref = _native.DeviceBuffer(WIDTH * HEIGHT)
    tex.render_device(ref, None, 2.0, 1)
    host = np.empty((HEIGHT, WIDTH), np.uint8)
    ref.download(host)
    targets = []
    for k in range(count):
        fld = seednet.random_affine(rng)
        tgt = _native.DeviceBuffer(WIDTH * HEIGHT)
        tex.render_device(tgt, fld, 2.0, 2 * k + 2)
        targets.append((tgt, fld))
    return ref, host, targets


def _accuracy(states, targets) -> dict:
    """Over subsets whose deformed window stays in the frame: accepted fraction, and over the
    accepted ones the RMSE, median and 99th percentile of the displacement error (px) and the
    fraction within 0.01 px."""
    from live_dic import synthetic

    cx, cy = synthetic.subset_centres(WIDTH, HEIGHT, SUBSET, STRIDE)
    acc, errs = [], []
    for st, (_, fld) in zip(states, targets):
        u, v = fld.displacement(cx, cy)
        inside = (cx + u > SUBSET) & (cx + u < WIDTH - SUBSET) & (cy + v > SUBSET) & (cy + v < HEIGHT - SUBSET)
        ok = (st[..., 6] > 0.5) & inside
        acc.append(ok.sum() / max(inside.sum(), 1))
        errs.append(np.hypot(st[..., 0] - u, st[..., 3] - v)[ok])
    e = np.concatenate(errs)
    return {"accepted_fraction": float(np.mean(acc)), "rmse_px": float(np.sqrt(np.mean(e * e))),
            "median_px": float(np.median(e)), "p99_px": float(np.percentile(e, 99)),
            "within_0.01px": float(np.mean(e <= 0.01)), "over_1px": int(np.sum(e > 1.0))}


def learned_vs_zncc(reps: int = 10) -> dict:
    """Throughput and accuracy of the two seeding paths on the same fixed reference and targets."""
    from live_dic import ICGNWGPU, ICGNWGPUFullPyramid, _native, seednet

    if not seednet.WEIGHTS_PATH.exists():
        return {"skipped": f"no trained weights at {seednet.WEIGHTS_PATH.relative_to(ROOT)}"}
    lib = _native.library()
    ref, host, targets = _large_targets(LARGE_PAIRS)
    result = {"workload": f"3840x2160 reference and {LARGE_PAIRS} targets, each t U[-64,64] px, rotation "
                          "U[-10,10] deg, strains U[-3%,3%] about the centre, noise 2; affine IC-GN subset 31 "
                          "step 8; reference prepared once; targets resident in HBM"}

    def timed(submits, handle, out_shape):
        """submits[k](target) queues one frame on slot k.  Accuracy from slot 0 frame by frame;
        throughput with consecutive frames alternating over the slots (frames are independent
        with a fixed reference, so one frame's small kernels overlap another's solves)."""
        states = []
        for tgt, _ in targets:                          # warm-up and accuracy
            submits[0](tgt)
            _native.check(lib.ldic_slot_wait(handle, 0, None))
            st = np.empty(out_shape, np.float32)
            _native.check(lib.ldic_slot_output(handle, 0, st.ctypes.data))
            states.append(st)
        for k in range(1, len(submits)):                # warm-up of the other slots
            submits[k](targets[0][0])
            _native.check(lib.ldic_slot_wait(handle, k, None))
        _native.check(lib.ldic_device_synchronize())
        t0 = time.perf_counter()
        i = 0
        for _ in range(reps):
            for tgt, _ in targets:
                submits[i % len(submits)](tgt)
                i += 1
        for k in range(len(submits)):
            _native.check(lib.ldic_slot_wait(handle, k, None))
        _native.check(lib.ldic_device_synchronize())
        rate = reps * len(targets) / (time.perf_counter() - t0)
        return {"frames_per_s": rate, "slots": len(submits), **_accuracy(states, targets)}

    zeng = ICGNWGPUFullPyramid(pyramid_levels=LARGE_LEVELS, iterations_per_level=LARGE_ITERATIONS,
                               subset_size=SUBSET, subset_stride=STRIDE, initialization="zncc",
                               seed_search_radius=LARGE_SEARCH_RADIUS, reseed_passes=LARGE_RESEED)
    info = zeng.set_reference(host)
    gw, gh = info.grid
    _native.check(lib.ldic_set_slots(zeng._handle, LARGE_SLOTS))
    zncc = timed([lambda t, k=k: _native.check(lib.ldic_register_resident(zeng._handle, k, t.address, WIDTH))
                  for k in range(LARGE_SLOTS)], zeng._handle, (gh, gw, 8))
    zeng.close()

    # The finest-level solve uses the ZNCC path's finest-level controls (16 iterations, 5e-3).
    # Each slot has its own network instance (activation buffers) and seeder on its stream.
    leng = ICGNWGPU(subset_size=SUBSET, subset_stride=STRIDE)
    leng.set_reference(host, iterations=LARGE_ITERATIONS[-1], tolerance=5e-3)
    _native.check(lib.ldic_set_slots(leng._handle, LARGE_SLOTS))
    seeders = []
    for k in range(LARGE_SLOTS):
        net = seednet.SeedNet()
        net.load(seednet.WEIGHTS_PATH)
        seeders.append(seednet.LearnedSeeder(leng, net, slot=k, reseed_passes=LARGE_RESEED))
        seeders[-1].set_reference(ref.address)
    seeder = seeders[0]
    learned = timed([lambda t, s=s: s.submit(t.address) for s in seeders], leng._handle, (gh, gw, 8))
    leng.close()
    zncc["path"] = f"{LARGE_LEVELS}-level pyramid {list(LARGE_ITERATIONS)}, coarsest-level ZNCC search " \
                   f"+-{LARGE_SEARCH_RADIUS // 8 + 1} px, {LARGE_RESEED} neighbour re-seeding passes"
    learned.update({"path": "seed network (BF16 MFMA) + finest-level IC-GN (16 iterations, tolerance 5e-3), "
                            f"ZNCC fallback (+-8 px) below 0.9, {LARGE_RESEED} neighbour re-seeding passes",
                    "network_parameters": net.arch.parameter_count(),
                    "network_gmac_per_frame": net.arch.macs(HEIGHT >> 2, WIDTH >> 2) / 1e9,
                    "network_forward_ms": seeder.tiles.get("forward_ms")})
    result.update({"zncc": zncc, "learned": learned, "speedup": learned["frames_per_s"] / zncc["frames_per_s"]})
    return result


def batched_heterogeneous(frames: int = 20) -> dict:
    """Aggregate subsets/s of the 16-pair batch (synthetic.heterogeneous_batch) replayed as one HIP
    graph per frame, and the p50/p99 latency of a frame (every pair of a frame completes with it)."""
    from live_dic import ICGNWGPUFullPyramid, _native, synthetic
    from live_dic.batch import BatchRegistration

    engines, targets, fields, configs = [], [], [], []
    for pair in synthetic.heterogeneous_batch():
        tgt = _native.DeviceBuffer(pair.width * pair.height)
        pair.target_device(tgt)
        eng = ICGNWGPUFullPyramid(pyramid_levels=pair.levels, subset_size=pair.subset, subset_stride=pair.stride,
                                  motion_model=pair.model)
        eng.set_reference(pair.reference())
        engines.append(eng)
        targets.append(tgt)
        fields.append(pair.field)
        configs.append(pair.label())
    subsets = sum(e.grid_width * e.grid_height for e in engines)
    with BatchRegistration(engines, [t.address for t in targets]) as batch:
        table = batch.tune()
        batch.launch(use_graph=True)
        batch.wait()
        lat = []
        for _ in range(frames):
            batch.launch(use_graph=True)
            lat.append(batch.wait())
        t0 = time.perf_counter()
        for _ in range(frames):
            batch.launch(use_graph=True)
        batch.wait()
        per_frame = (time.perf_counter() - t0) / frames
        outs = batch.outputs()
    medians, accepted, accepted_inside = [], [], []
    for e, st, fld in zip(engines, outs, fields):
        w, h = e._level_shapes[0][:2]
        S = e.subset_size
        cx, cy = synthetic.subset_centres(w, h, S, e.subset_stride)
        u, v = fld.displacement(cx, cy)
        P = 12 if e.motion_model == "quadratic" else 6
        ok = st[..., P] > 0.5
        # A 1-degree rotation about the centre moves a 4K corner by ~38 px: subsets whose
        # warped window leaves the frame cannot be registered and are left out here.
        inside = (cx + u > S) & (cx + u < w - S) & (cy + v > S) & (cy + v < h - S)
        accepted.append(float(ok.mean()))
        accepted_inside.append(float(ok[inside].mean()))
        medians.append(float(np.median(np.hypot(st[..., 0] - u, st[..., 3] - v)[ok])))
    for e in engines:
        e.close()
    return {
        "workload": "16 pairs: 4x each of 640x480, 1280x720, 1920x1080, 3840x2160; subsets 15/21/31/41, steps "
                    "4/8/16, affine and quadratic; random affine fields; one HIP graph per frame",
        "pairs": configs,
        "subsets_per_frame": subsets,
        "subsets_per_s": subsets / per_frame,
        "frame_ms": per_frame * 1e3,
        "latency_ms_p50": float(np.percentile(lat, 50)),
        "latency_ms_p99": float(np.percentile(lat, 99)),
        "accepted_fraction_min": min(accepted),
        "accepted_fraction_inside_frame_min": min(accepted_inside),
        "accepted_fraction_inside_frame_per_pair": [round(a, 4) for a in accepted_inside],
        "median_error_px_max": max(medians),
        "persistent_launches": table,
    }


PAIR_FRAMES = 30
PAIR_SUBSET, PAIR_STRIDE = 25, 4
PAIR_ITERATIONS, PAIR_TOLERANCE = 30, 1e-3


def consecutive_quadratic() -> dict:
    """Quadratic IC-GN in consecutive-pair mode at 4K: frame k is the reference of frame k + 1,
    prepared on the GPU in the pair's own sequence (Gaussian, B-spline scans, gradients, FP64
    subset preparation), then the target prefilter, the solve and the outlier pass.  ROI: the
    plate rectangle minus 3 circular holes of radius 200-400 px; subset 25, step 4."""
    from live_dic import ICGNWGPUFullPyramid, _native, synthetic

    lib = _native.library()
    seq = synthetic.QuadraticPairs(frames=PAIR_FRAMES)
    mask = synthetic.plate_with_holes()
    frames = []
    for k in range(PAIR_FRAMES):
        buf = _native.DeviceBuffer(WIDTH * HEIGHT)
        seq.render_device(k, buf)
        frames.append(buf)
    eng = ICGNWGPUFullPyramid(pyramid_levels=1, iterations_per_level=(PAIR_ITERATIONS,),
                              tolerance_per_level=(PAIR_TOLERANCE,), motion_model="quadratic",
                              subset_size=PAIR_SUBSET, subset_stride=PAIR_STRIDE)
    try:
        info = eng.set_reference(seq.render(0), roi_mask=mask)
        gw, gh = info.grid
        selected = int(np.count_nonzero(eng.roi.selected))
        h = eng._handle
        _native.check(lib.ldic_register_pair_resident(h, 0, frames[0].address, frames[1].address, WIDTH))
        _native.check(lib.ldic_slot_wait(h, 0, ctypes.byref(_native.Timing())))
        _native.check(lib.ldic_device_synchronize())
        pairs = PAIR_FRAMES - 1
        t0 = time.perf_counter()
        for k in range(pairs):
            _native.check(lib.ldic_register_pair_resident(h, 0, frames[k].address, frames[k + 1].address, WIDTH))
        _native.check(lib.ldic_slot_wait(h, 0, ctypes.byref(_native.Timing())))
        _native.check(lib.ldic_device_synchronize())
        per_pair = (time.perf_counter() - t0) / pairs
        single = []
        for k in (pairs - 2, pairs - 1):
            t = _native.Timing()
            _native.check(lib.ldic_register_pair_resident(h, 0, frames[k].address, frames[k + 1].address, WIDTH))
            _native.check(lib.ldic_slot_wait(h, 0, ctypes.byref(t)))
            single.append((t.gpu_ms, t.solve_ms))
        state = np.empty((gh, gw, 16), np.float32)
        _native.check(lib.ldic_slot_output(h, 0, state.ctypes.data))
    finally:
        eng.close()
    cx, cy = synthetic.subset_centres(WIDTH, HEIGHT, PAIR_SUBSET, PAIR_STRIDE)
    u, v = seq.pair_truth(PAIR_FRAMES - 2, cx, cy)
    ok = state[..., 12] > 0.5
    err = np.hypot(state[..., 0] - u, state[..., 3] - v)[ok]
    return {
        "workload": f"4K quadratic consecutive pairs, plate with 3 holes, subset {PAIR_SUBSET}, step {PAIR_STRIDE}, "
                    f"{PAIR_FRAMES} frames ({pairs} pairs), {PAIR_ITERATIONS} iterations, tolerance {PAIR_TOLERANCE}",
        "grid": [gw, gh],
        "roi_subsets": selected,
        "pairs_per_s": 1.0 / per_pair,
        "pair_ms": per_pair * 1e3,
        "roi_subsets_per_s": selected / per_pair,
        "last_pair_gpu_ms": single[-1][0],
        "last_pair_solve_ms": single[-1][1],
        "accepted_fraction": float(ok.sum()) / max(1, selected),
        "rmse_px": float(np.sqrt(np.mean(err**2))) if err.size else None,
        "mean_iterations": float(state[..., 13][ok].mean()) if ok.any() else None,
    }


# The reference's own CPU figure for this path (its WGSL on Mesa lavapipe through wgpu-py, which
# cannot run on this machine): docs/how-to/run-on-cpu.rst:140-165.
PUBLISHED_REFERENCE_CPU = {
    "value": 4.56,
    "unit": "frames/s",
    "ms_per_frame_median": 218.2,
    "cores": 8,
    "kind": "reference (published, not re-measured)",
    "hardware": "Intel Core Ultra 5 228V, Mesa 25.2.8 llvmpipe, wgpu-py 0.32.0",
    "workload": "3840x2160 rigid translation, translation model 16/16, B-spline, median, 2 levels, noise 1; "
                "NumPy image to NumPy field",
    "source": "LK/docs/how-to/run-on-cpu.rst:140-165",
}

SMALL_SUBSET, SMALL_STRIDE = 21, 8


def valu_limiter(avg_ms: float, alone_ms: float | None) -> dict:
    """FP32 VALU issue of the fine solve per SIMD against the non-packed FP32 issue peak, and its
    FP32 FLOP rate against the matching 78.6 TFLOP/s."""
    def rate(ms: float) -> float:
        return MEASURED_FINE_SOLVE_FP32_INSTRUCTIONS / SIMDS / (ms * 1e-3)

    out = {
        "bound": "valu",
        "unit": "FP32 FMA/MUL/ADD wave-instructions/s per SIMD",
        "achieved": rate(avg_ms),
        "peak": FP32_ISSUE_PEAK_PER_SIMD,
        "frac": rate(avg_ms) / FP32_ISSUE_PEAK_PER_SIMD,
        "fp32_instructions_per_launch": MEASURED_FINE_SOLVE_FP32_INSTRUCTIONS,
        "valu_instructions_per_launch": MEASURED_FINE_SOLVE_VALU_INSTRUCTIONS,
        "fp32_tflops": MEASURED_FINE_SOLVE_FP32_FLOPS / (avg_ms * 1e-3) / 1e12,
        "fp32_tflops_peak": FP32_VECTOR_PEAK_FLOPS / 1e12,
    }
    if alone_ms:
        out["frac_alone"] = rate(alone_ms) / FP32_ISSUE_PEAK_PER_SIMD
    return out


def small_pair_line(frames: int = 400, warmup: int = 20) -> dict:
    """The 512x512 pair (subset 21, step 8, 20 iterations, tolerance 1e-4, one level): GPU frames/s
    at the NumPy boundary through register_stream(pipeline_depth=2), accuracy against the analytic
    field, and the C oracle on the host at the same boundary."""
    import oracle
    from live_dic import ICGNWGPU, synthetic

    ref, tgt, fld = synthetic.small_affine_pair()
    with ICGNWGPU(subset_size=SMALL_SUBSET, subset_stride=SMALL_STRIDE) as eng:
        eng.set_reference(ref, iterations=20, tolerance=1e-4)
        for _ in eng.register_stream([tgt] * warmup, pipeline_depth=2):
            pass
        t0 = time.perf_counter()
        for res in eng.register_stream([tgt] * frames, pipeline_depth=2):
            last = res
        gpu_rate = frames / (time.perf_counter() - t0)
    cx, cy = synthetic.subset_centres(512, 512, SMALL_SUBSET, SMALL_STRIDE)
    u, v = fld.displacement(cx, cy)
    ok = last.status == 1
    rmse = float(np.sqrt(np.mean((last.parameters[..., 0][ok] - u[ok]) ** 2 + (last.parameters[..., 3][ok] - v[ok]) ** 2)))
    pipe = oracle.Pipeline(pyramid_levels=1, iterations_per_level=(20,), tolerance_per_level=(1e-4,),
                           subset_size=SMALL_SUBSET, subset_stride=SMALL_STRIDE).set_reference(ref)
    pipe.register(tgt)
    count, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < 3.0:
        pipe.register(tgt)
        count += 1
    cpu_rate = count / (time.perf_counter() - t0)
    return {
        "workload": "512x512 speckle pair, u = (0.3, -0.7) + 0.5 % x-strain, noise 1; affine subset 21 step 8 "
                    "(3,844 subsets), 20 iterations, tolerance 1e-4, one level, B-spline",
        "frames_per_s": gpu_rate,
        "subsets_per_s": gpu_rate * last.status.size,
        "boundary": "host NumPy uint8 -> NumPy field, register_stream(pipeline_depth=2)",
        "accepted": int(ok.sum()),
        "rmse_px": rmse,
        "cpu_oracle_frames_per_s": cpu_rate,
        "cpu_oracle_threads": oracle.cpu_threads(),
    }


def numpy_boundary(eng, host_frames, steps: int, warmup: int, depth: int = 2) -> dict:
    """register_stream over host NumPy frames: frames/s and the median per-frame split of the
    library's timing keys (the reference's call-time split, performance.rst:516-527)."""
    order = [host_frames[1 + i % (FRAMES - 1)] for i in range(warmup + steps)]
    for _ in eng.register_stream(order[:warmup], pipeline_depth=depth):
        pass
    timings = []
    last = None
    t0 = time.perf_counter()
    for res in eng.register_stream(order[warmup:], pipeline_depth=depth):
        timings.append(res.timing)
        last = res
    elapsed = time.perf_counter() - t0

    def median(key):
        values = [t[key] for t in timings if key in t]
        return float(np.median(values)) if values else None

    solve = [t["solveGpuMs"] for t in timings if t.get("solveGpuMs")]
    return {
        "frames_per_s": steps / elapsed,
        "ms_per_frame": elapsed / steps * 1e3,
        "fine_solve_ms_mean": float(np.mean(solve)) if solve else None,
        "pipeline_depth": depth,
        "split_ms_median": {key: median(key) for key in (
            "inputPreparationWallMs", "uploadWallMs", "gpuMs", "solveGpuMs", "completionWaitWallMs",
            "resultAgeWallMs")},
        "last": last,
    }


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=120)
    parser.add_argument("--warmup", type=int, default=10)
    parser.add_argument("--drain", type=int, default=32, help="frames queued between host waits")
    parser.add_argument("--slots", type=int, default=2,
                        help="engine slots (HIP streams) that consecutive frames alternate over")
    parser.add_argument("--alone-frames", type=int, default=FRAMES - 1,
                        help="frames queued on one slot after the timed region, for the solver's duration "
                             "without a concurrent frame (the default covers frames 1-59 once)")
    parser.add_argument("--no-cpu-baseline", action="store_true")
    parser.add_argument("--no-learned", action="store_true", help="skip the learned-vs-ZNCC seeding comparison")
    parser.add_argument("--no-batch", action="store_true", help="skip the batched heterogeneous workload")
    parser.add_argument("--no-pairs", action="store_true",
                        help="skip the quadratic consecutive-pair workload with ROI holes")
    parser.add_argument("--only", choices=("learned", "batch", "pairs"),
                        help="run only this secondary workload and print its JSON object (for profiling)")
    parser.add_argument("--calibrate-copy", action="store_true",
                        help="only copy 256 MiB with one plain device copy kernel (PMC byte-counter calibration)")
    args = parser.parse_args()
    if args.calibrate_copy:
        from live_dic import _native

        lib = _native.library()
        src = _native.DeviceArray.zeros(256 << 20, np.uint8)
        dst = _native.DeviceArray(256 << 20, np.uint8)
        _native.check(lib.ldic_device_copy_async(dst.ptr, src.ptr, src.nbytes, None))
        _native.check(lib.ldic_device_synchronize())
        print(json.dumps({"copied_bytes": src.nbytes, "read_bytes": src.nbytes, "written_bytes": dst.nbytes}))
        return 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (frames would shard by "
              "bands of the subset grid); run with --gpus 1", file=sys.stderr)
        return 2
    if args.only:
        legs = {"learned": learned_vs_zncc, "batch": batched_heterogeneous, "pairs": consecutive_quadratic}
        print(json.dumps(legs[args.only]()))
        return 0

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
    if args.slots > 1:
        _native.check(lib.ldic_set_slots(handle, args.slots))

    def run(steps: int, start: int, slots: int) -> tuple[float, int]:
        """Queue frames round-robin on the first `slots` slots' streams (frames are
        independent with a fixed reference, so one frame's small kernels and grid tails
        overlap another's solves); the host waits every `drain` frames.  Returns the
        summed finest-level solver time (HIP events around each launch) and its count."""
        total, count = 0.0, 0
        for i in range(steps):
            k = 1 + (start + i) % (FRAMES - 1)
            _native.check(lib.ldic_register_resident(handle, (start + i) % slots, frames[k].address, WIDTH))
            if (i + 1) % args.drain == 0 or i + 1 == steps:
                for slot in range(slots):
                    t = _native.Timing()
                    _native.check(lib.ldic_slot_wait(handle, slot, ctypes.byref(t)))
                    total += t.solve_ms
                    count += t.frames
        return total, count

    print(f"[bench] warmup {args.warmup}", file=sys.stderr, flush=True)
    if args.warmup:
        run(args.warmup, 0, args.slots)
    _native.check(lib.ldic_device_synchronize())
    print(f"[bench] timing {args.steps} frames", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    solve_total, solve_count = run(args.steps, args.warmup, args.slots)
    _native.check(lib.ldic_device_synchronize())
    elapsed = time.perf_counter() - t0
    fps = args.steps / elapsed

    # Accuracy of the last timed frame against the analytic field.
    last = 1 + (args.warmup + args.steps - 1) % (FRAMES - 1)
    last_slot = (args.warmup + args.steps - 1) % args.slots
    out = np.empty((gh, gw, 8), np.float32)
    _native.check(lib.ldic_slot_output(handle, last_slot, out.ctypes.data))
    cx, cy = synthetic.subset_centres(WIDTH, HEIGHT, SUBSET, STRIDE)
    u, v = seq.field.scaled(last / (FRAMES - 1)).displacement(cx, cy)
    ok = out[..., 6] > 0.5
    rmse = float(np.sqrt(np.mean((out[..., 0][ok] - u[ok]) ** 2 + (out[..., 3][ok] - v[ok]) ** 2)))
    mean_iterations = []  # coarse to fine, accepted subsets of the last frame
    for level in reversed(range(LEVELS)):
        lgw, lgh = eng._level_shapes[level][2:4]
        st = np.empty((lgh, lgw, 8), np.float32)
        _native.check(lib.ldic_export_level_state(handle, last_slot, level, st.ctypes.data))
        acc = st[..., 6] > 0.5
        mean_iterations.append(float(st[..., 7][acc].mean()) if acc.any() else None)
    # The same kernel with no concurrent frame (one slot), after the timed region.
    alone_total, alone_count = run(args.alone_frames, 0, 1) if args.alone_frames > 0 else (0.0, 0)
    resident_fps, resident_elapsed = fps, elapsed

    # The headline: the same frames as host NumPy arrays through register_stream.
    print("[bench] NumPy boundary: register_stream(pipeline_depth=2) over host frames", file=sys.stderr,
          flush=True)
    host_frames = []
    for buf in frames:
        arr = np.empty((HEIGHT, WIDTH), np.uint8)
        buf.download(arr)
        host_frames.append(arr)
    boundary = numpy_boundary(eng, host_frames, args.steps, args.warmup, depth=args.slots)
    fps = boundary["frames_per_s"]
    elapsed = args.steps / fps
    last_result = boundary.pop("last")
    last_numpy = 1 + (args.warmup + args.steps - 1) % (FRAMES - 1)
    u2, v2 = seq.field.scaled(last_numpy / (FRAMES - 1)).displacement(cx, cy)
    ok2 = last_result.accepted_mask
    boundary["rmse_px_last_frame"] = float(np.sqrt(np.mean(
        (last_result.parameters[..., 0][ok2] - u2[ok2]) ** 2 + (last_result.parameters[..., 3][ok2] - v2[ok2]) ** 2)))
    boundary["accepted_fraction_last_frame"] = float(ok2.mean())
    eng.close()

    resident_solve_ms = solve_total / solve_count if solve_count else float("nan")
    # The dominant kernel's duration over the headline's timed region: the fine solve's HIP
    # events on its slot stream, one per NumPy-boundary frame.
    avg_solve_ms = boundary["fine_solve_ms_mean"] or resident_solve_ms
    traffic_bytes = n_fine * subset_bytes(SUBSET)
    achieved = traffic_bytes / (avg_solve_ms * 1e-3) / 1e9
    alone_ms = alone_total / alone_count if alone_count else None
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
        "data": "synthetic (GPU-rendered analytic speckle, seeded; frames held as host NumPy uint8 arrays)",
        "config": {
            "workload": "4K fixed-reference affine IC-GN sequence, 3840x2160, subset 31, step 8, "
                        "2-level pyramid, B-spline, 60 frames",
            "boundary": "host NumPy uint8 frame -> materialised NumPy field and statistics through "
                        "ICGNWGPUFullPyramid.register_stream(pipeline_depth=2); uploads and readbacks timed",
            "subsets_per_frame": n_fine,
            "subsets_per_s": fps * n_fine,
            "accepted_fraction_last_frame": boundary["accepted_fraction_last_frame"],
            "rmse_px_last_frame": boundary["rmse_px_last_frame"],
            "reference_prep_ms": prep_ms,
            "iterations_per_level": list(ITERATIONS),
            "mean_iterations_per_level": mean_iterations,
            "tolerance_per_level": list(TOLERANCES),
            "solver_launches_timed": solve_count,
            "slots": args.slots,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": MEASURED_FINE_SOLVE_TRAFFIC_BYTES,
            "l2_hit_rate": MEASURED_FINE_SOLVE_L2_HIT_RATE,
            "kernel": "k_solve<affine, bspline> (finest level)",
            "kernel_avg_ms": avg_solve_ms,
            "kernel_avg_ms_resident_run": resident_solve_ms,
            "algorithmic_bytes_per_launch": traffic_bytes,
            # Each array the launch touches, read or written once: the float4 reference plane and
            # the coefficient plane (W x H x 20 B), metadata and inverses (4 (M + P^2) B per
            # subset) and the seed read plus state write (64 B per subset).
            "compulsory_bytes_per_launch": WIDTH * HEIGHT * 20 + n_fine * (4 * (16 + 36) + 64),
            # The timed region runs two frames concurrently, so each launch shares the GPU;
            # the same kernel on one stream, measured after the timed region:
            "kernel_avg_ms_alone": alone_ms,
            "frac_alone": traffic_bytes / (alone_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS if alone_ms else None,
            # What actually limits the kernel: VALU instruction issue.  HBM moves 0.12x the
            # algorithmic bytes (the rest is served by L2 and LDS), while the SIMDs issue VALU
            # instructions at the FP32 issue limit.
            "limiter": valu_limiter(avg_solve_ms, alone_ms),
        },
        "cpu_baseline": None,
        "numpy_boundary": boundary,
        "resident": {
            "what": "the same frames resident in HBM (ldic_register_resident), device work only",
            "frames_per_s": resident_fps,
            "ms_per_frame": resident_elapsed / args.steps * 1e3,
            "subsets_per_s": resident_fps * n_fine,
            "accepted_fraction_last_frame": float(ok.mean()),
            "rmse_px_last_frame": rmse,
            "slots": args.slots,
        },
    }
    if not args.no_learned:
        print("[bench] learned vs ZNCC seeding at large displacement", file=sys.stderr, flush=True)
        line["learned_vs_zncc"] = learned_vs_zncc()
    if not args.no_batch:
        print("[bench] batched heterogeneous pairs", file=sys.stderr, flush=True)
        line["batched_heterogeneous"] = batched_heterogeneous()
    if not args.no_pairs:
        print("[bench] quadratic consecutive pairs with ROI holes", file=sys.stderr, flush=True)
        line["consecutive_quadratic"] = consecutive_quadratic()
    if not args.no_cpu_baseline:
        print("[bench] CPU oracle baseline", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(seq, 30)
        line["cpu_baseline"]["published_reference"] = PUBLISHED_REFERENCE_CPU
        print("[bench] 512x512 pair (GPU and CPU oracle)", file=sys.stderr, flush=True)
        line["small_pair_512"] = small_pair_line()
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
