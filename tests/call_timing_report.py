"""Call timing: backprojection() with host buffers beside device-resident steps (test tooling).

On configuration C2 (2048 x 2048 pixels from 2000 pulses x 2048 samples, N_fft 8192) the script
times, in one process and in this order:
  1. `--steps` device-resident steps (range profiles, then the accumulation into device memory), each
     synchronised on its own, the accumulation timed by HIP events;
  2. `--calls` backprojection() calls with host buffers (the reference's entry point, as
     sarbp.cpp:276-280 times it), wall clock around each call;
  3. an idle pause of `--idle-ms`, then the resident steps again.
Each phase prints one JSON line with every time it took, so that the accumulation's time inside the
calls (from a kernel trace of the same command) can be set beside the resident step's time at the
same point of the run.

Usage (on a GPU machine, from the repository root):
    python tests/call_timing_report.py [--steps 8] [--calls 8] [--idle-ms 50]
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))  # repository root
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--calls", type=int, default=8)
    ap.add_argument("--idle-ms", type=float, default=50.0)
    args = ap.parse_args()

    import torch

    sar = importlib.import_module("halide-sar-app_amd")
    scenes = importlib.import_module("halide-sar-app_amd.scenes")
    pipe = importlib.import_module("halide-sar-app_amd.pipeline")
    if not torch.cuda.is_available():
        raise SystemExit("call_timing_report needs a GPU")
    sc = scenes.config_scene("C2")
    grid = pipe.scene_grid(sc, with_pixel_locs=False)
    ds = pipe.DeviceScene(sc, grid)
    plan = ds.plan()
    out = ds.output()
    stream = torch.cuda.current_stream()

    def resident(phase: str) -> None:
        times = []
        for _ in range(args.steps):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            plan.range_profiles(ds.phs, ds.k_r)
            e0.record(stream)
            plan.accumulate(ds.k_r, ds.pos, ds.r, out)
            e1.record(stream)
            torch.cuda.synchronize()
            times.append(round(e0.elapsed_time(e1), 3))
        print(json.dumps({"phase": phase, "what": "resident step, accumulation (HIP events)",
                          "ms": times}), flush=True)

    resident("resident, from the start of the run")
    phs = ds.phs.cpu().numpy()
    plane = sar.image_grid(grid["nu"], grid["nv"], grid["d_u"], grid["d_v"], grid["u_hat"],
                           grid["v_hat"])
    out_h = np.empty((grid["nv"], grid["nu"]), np.complex128)
    call = (phs, sc.k_r, sc.taylor, ds.n_fft, sc.delta_r, plane["u"], plane["v"], sc.pos,
            plane["pixel_locs"], out_h)
    times = []
    for _ in range(args.calls):
        t = time.perf_counter()
        rv = sar.backprojection(*call)
        times.append(round((time.perf_counter() - t) * 1e3, 3))
        if rv != 0:
            raise SystemExit(f"backprojection() failed: {rv} {sar.last_error()}")
    print(json.dumps({"phase": "calls", "what": "backprojection() with host buffers (wall clock; "
                      "the first call creates the plan and staging)", "ms": times}), flush=True)
    resident("resident, right after the calls")
    time.sleep(args.idle_ms * 1e-3)
    resident(f"resident, after {args.idle_ms:g} ms idle")


if __name__ == "__main__":
    main()
