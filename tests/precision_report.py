"""Precision form: per-variant error of the HIP path against the CPU FP64 oracle (test tooling).

For a seeded point-scatterer scene (configuration C3 by default: 8192 x 8192 pixels from 8000 pulses
x 4096 samples, N_fft 16384) each variant of the device pipeline forms the image once, timed, and is
compared with the oracle on
  - a uniform random sample of pixels: relative L2 error and phase RMS error (radians, over sampled
    pixels within 40 dB of the brightest one);
  - windows around the brightest point targets: peak position and peak-sidelobe ratio (PSLR, the
    highest sidelobe outside the 3x3-pixel main lobe relative to the peak, in dB) of the variant
    minus that of the oracle.
Variants: range profiles stored as complex f32 (c32) or complex f16 with a per-profile power-of-two
scale (c16), each with f64 row coordinates; and c32 rows with double-float row coordinates (df32:
range_mode SARBP_RANGE_DF32, an exact integer part per pulse and column and an f32 cubic per
pixel).

Usage (on a GPU machine, from the repository root):
    python tests/precision_report.py --config C3 [--sample 2048] [--targets 4]
Prints one JSON line per variant. The oracle is loaded only as the checker.
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


def pslr_db(win: np.ndarray) -> tuple[float, tuple[int, int]]:
    """Peak-sidelobe ratio of a window around a point response: peak = max |z|; sidelobes = max |z|
    outside the 3 x 3 pixels around the peak; returns (20 log10(sidelobe / peak), peak index)."""
    a = np.abs(win)
    iy, ix = np.unravel_index(np.argmax(a), a.shape)
    peak = a[iy, ix]
    mask = np.ones_like(a, bool)
    mask[max(0, iy - 1):iy + 2, max(0, ix - 1):ix + 2] = False
    side = a[mask].max() if mask.any() else 0.0
    return float(20 * np.log10(max(side, 1e-300) / peak)), (int(iy), int(ix))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="C3")
    ap.add_argument("--sample", type=int, default=2048)
    ap.add_argument("--targets", type=int, default=4)
    ap.add_argument("--half", type=int, default=16, help="target window half-size (pixels)")
    ap.add_argument("--variants", default="c32,df32,c16")
    args = ap.parse_args(argv)

    import torch
    sar = importlib.import_module("halide-sar-app_amd")
    scenes = importlib.import_module("halide-sar-app_amd.scenes")
    pipe = importlib.import_module("halide-sar-app_amd.pipeline")
    import oracle

    sc = scenes.config_scene(args.config)
    grid = pipe.scene_grid(sc, with_pixel_locs=False)
    nu, nv = grid["nu"], grid["nv"]
    print(f"precision: {args.config} np={sc.npulses} ns={sc.nsamples} image {nu}x{nv}",
          file=sys.stderr, flush=True)
    ds = pipe.DeviceScene(sc, grid)
    phs = ds.phs.cpu().numpy()

    # sample pixels: uniform random + windows around the brightest targets inside the image
    rng = np.random.default_rng(2024)
    ys = list(rng.integers(0, nv, args.sample))
    xs = list(rng.integers(0, nu, args.sample))
    amps = np.abs(sc.amps)
    order = np.argsort(-amps)
    windows = []
    for t in order:
        if len(windows) >= args.targets:
            break
        uu = sc.targets[t] @ grid["u_hat"]
        vv = sc.targets[t] @ grid["v_hat"]
        ix = int(np.argmin(np.abs(grid["u"] - uu)))
        iy = nv - 1 - int(np.argmin(np.abs(grid["v"] - vv)))
        h = args.half
        if h <= ix < nu - h and h <= iy < nv - h:
            windows.append((iy, ix))
            for y in range(iy - h, iy + h + 1):
                for x in range(ix - h, ix + h + 1):
                    ys.append(y)
                    xs.append(x)
    ys = np.array(ys, np.int32)
    xs = np.array(xs, np.int32)

    t0 = time.perf_counter()
    Q = oracle.range_profiles(phs, sc.k_r, sc.taylor, ds.n_fft)
    _, nr0 = oracle.norm_r0(sc.pos)
    pts = sar.grid_points(grid, xs, ys)   # a one-row geometry holding just the sampled pixels
    cnt = len(xs)
    ref = oracle.backproject_points(Q, sc.nsamples, sc.delta_r, sc.k_r[sc.nsamples // 2], sc.pos,
                                    nr0, pts, cnt, 1, np.arange(cnt, dtype=np.int32),
                                    np.zeros(cnt, np.int32),
                                    nthreads=max(1, min(16, os.cpu_count() or 1)))
    del Q
    t_oracle = time.perf_counter() - t0
    ns_ = args.sample
    wsz = (2 * args.half + 1) ** 2

    yi = torch.from_numpy(ys.astype(np.int64)).to(ds.dev)
    xi = torch.from_numpy(xs.astype(np.int64)).to(ds.dev)
    for v in args.variants.split(","):
        qf = {"c32": sar.Q_C32, "c16": sar.Q_C16, "df32": sar.Q_C32}[v]
        rm = sar.RANGE_DF32 if v == "df32" else sar.RANGE_F64
        plan = ds.plan(q_format=qf, range_mode=rm)
        out = ds.output()
        ds.run(plan, out)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ds.run(plan, out)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t1) * 1e3
        got = out[yi, xi].cpu().numpy()
        g, r = got[:ns_], ref[:ns_]
        rel = float(np.linalg.norm(g - r) / np.linalg.norm(r))
        strong = np.abs(r) >= 1e-2 * np.abs(r).max()
        ph = np.angle(g[strong] * np.conj(r[strong]))
        dpslr, moved = [], 0
        for k in range(len(windows)):
            a = got[ns_ + k * wsz: ns_ + (k + 1) * wsz].reshape(2 * args.half + 1, -1)
            b = ref[ns_ + k * wsz: ns_ + (k + 1) * wsz].reshape(2 * args.half + 1, -1)
            pa, ia = pslr_db(a)
            pb, ib = pslr_db(b)
            dpslr.append(abs(pa - pb))
            moved += int(ia != ib)
        line = {"precision_variant": v, "config": args.config, "pixels": nu * nv,
                "pulses": sc.npulses, "ms": round(ms, 2),
                "gpixel_pulses_per_s": round(nu * nv * sc.npulses / (ms * 1e-3) / 1e9, 1),
                "rel_l2": rel, "phase_rms_rad": float(np.sqrt(np.mean(ph ** 2))),
                "pslr_err_db_max": float(max(dpslr)) if dpslr else None,
                "peaks_moved": moved, "targets": len(windows), "sample_pixels": ns_,
                "oracle_s": round(t_oracle, 1)}
        print(json.dumps(line), flush=True)
        del plan, out
        torch.cuda.empty_cache()


if __name__ == "__main__":
    main()
