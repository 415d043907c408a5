"""Throughput benchmark of the raw -> output stage chain on one MI355X.

A step is one pass of the stage list over one synthetic 14-bit RGGB frame that
is already resident in HBM as u16: ingest (black/white normalisation) and
every stage, ending with the f32 RGB output in HBM.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

W, H = 6016, 4016
BLACK, WHITE = 600.0, 16383.0
HBM_PEAK_GBS = 8000.0

# Stage list of the measured workload (reference stage names).
STAGES = """
[[pipeline_modules]]
name = "Demosaic"
algorithm = "amaze"

[[pipeline_modules]]
name = "CST"
target_color_space = "Srgb"
"""
STAGE_NAMES = ["Demosaic{amaze}", "CST{Srgb}"]


def calibration_matrix(seed: int = 7) -> np.ndarray:
    """Seeded diagonally dominant XYZ->camera matrix with condition number < 10."""
    rng = np.random.default_rng(seed)
    while True:
        m = rng.uniform(-0.25, 0.25, (3, 3)) + np.diag(rng.uniform(0.8, 1.3, 3))
        if np.linalg.cond(m) < 10:
            return m.astype(np.float32)


def cpu_baseline(raw: np.ndarray, cal: np.ndarray, budget_s: float = 12.0) -> dict:
    """The oracle (C++ restatement, all host cores via OpenMP) on a crop of the same frame."""
    import oracle
    cw, ch = 1536, 1024
    crop = (0, 0, cw, ch)
    times = []
    t_end = time.perf_counter() + budget_s
    while time.perf_counter() < t_end or len(times) < 2:
        t0 = time.perf_counter()
        plane = oracle.normalize_raw(raw, BLACK, WHITE, crop=crop)
        rgb = oracle.amaze(plane, "RGGB")
        oracle.camera_cst(rgb, cal.ravel(), oracle.CS_SRGB)
        times.append(time.perf_counter() - t0)
        if len(times) >= 50:
            break
    t = float(np.median(times[1:] if len(times) > 2 else times))
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(cw * ch / 1e6 / t, 3), "unit": "MP/s", "cores": cores, "kind": "port",
            "sample": f"{cw}x{ch} crop of the same frame, {len(times)} runs, median {t * 1e3:.1f} ms"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(frames shard with no exchange; see the sharding notes)", file=sys.stderr)
        return 2

    import torch
    import imops_amd as im

    ctx = im.Context(0)
    print(f"[bench] generating {W}x{H} raw", file=sys.stderr, flush=True)
    raw = im.synth_raw_u16(W, H, "RGGB", 0)
    cal = calibration_matrix()
    d_raw = torch.from_numpy(raw.view(np.int16)).to("cuda")
    torch.cuda.synchronize()

    img = im.Image(ctx).set_calibration_matrix(cal.ravel())
    cfg = im.parse_config(STAGES)
    stream = torch.cuda.ExternalStream(ctx.stream)

    def step():
        img.set_raw_u16(d_raw, "RGGB", BLACK, WHITE)
        im.run_pixel_pipeline(img, cfg)

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    torch.cuda.synchronize()

    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    ev1.synchronize()
    ctx.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = ev0.elapsed_time(ev1) / args.steps
    print(f"[bench] {args.steps} steps, {ms:.3f} ms/step (wall {wall / args.steps * 1e3:.3f} ms)",
          file=sys.stderr, flush=True)

    # Dominant kernel: AMaZE, timed alone on the library stream (one launch per call).
    img.set_raw_u16(d_raw, "RGGB", BLACK, WHITE)
    img.demosaic("amaze")
    ctx.synchronize()
    reps = max(5, args.steps)
    ev0.record(stream)
    for _ in range(reps):
        img.demosaic("amaze")
    ev1.record(stream)
    ev1.synchronize()
    amaze_ms = ev0.elapsed_time(ev1) / reps
    amaze_bytes = W * H * (4 + 12)  # f32 CFA plane in + interleaved f32 RGB out
    achieved = amaze_bytes / (amaze_ms * 1e-3) / 1e9

    out = {
        "metric": "24MP raw->sRGB MP/s (plain FP32)",
        "value": round(W * H / 1e6 / (ms * 1e-3), 3),
        "unit": "MP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene, Poisson + read noise, u16 14-bit in HBM)",
        "config": {"workload": f"{W}x{H} RGGB raw u16 -> " + " -> ".join(STAGE_NAMES) + " (f32 out)",
                   "width": W, "height": H, "cfa": "RGGB", "black": BLACK, "white": WHITE},
        "roofline": {"bound": "hbm", "kernel": "amaze_kernel", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel_ms": round(amaze_ms, 4),
                     "algorithmic_bytes": amaze_bytes},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(raw, cal)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
