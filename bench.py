"""Benchmark of the backprojection hot path on one MI355X.

A step is one pass of the pipeline (range profiles: window, pad, fftshift, FFT; then the per-pixel
accumulation over all pulses with the mid-pulse phase and row flip) over one synthetic scene whose
inputs are already resident in HBM: configuration C2, a 2048 x 2048 image from 2000 pulses x 2048
samples (N_fft 8192, Taylor 30, linear spotlight path at 15 km). Prints one JSON line.

  value     = pixel-pulses / step time (Gpixel*pulses/s)
  roofline  = the accumulation kernel: algorithmic FLOP (30 per pixel-pulse) / its mean duration,
              against the FP32 vector peak; HBM traffic per launch and the binding resource (the
              texture data path, with its busy fraction) from the committed rocprofv3 counter
              summary when present, with the gathers' bytes per second measured live
  rel_l2    = relative L2 error of the timed image vs the CPU FP64 oracle on sampled pixels
  cpu_baseline = the oracle (a C restatement of the reference CPU schedule's loop nest) timed on
              the host cores on a bounded row sample, extrapolated to the whole image
"""
import argparse
import importlib
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


METRIC = "Gpixel·pulses/s at ≤1e-4 rel-L2 vs FP64; serving p50/p99 latency/job"
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    # untimed steps: the GPU's clock rises over the first ~25 ms of continuous load (C2's
    # accumulation takes ~4.8, 4.6, 4.4, 4.3 ms on the first steps of a run, then ~4.2 ms)
    ap.add_argument("--warmup", type=int, default=6)
    ap.add_argument("--config", default="C2")
    ap.add_argument("--cpu-rows", type=int, default=256,
                    help="image rows the CPU baseline computes (0 disables it)")
    ap.add_argument("--check-pixels", type=int, default=512)
    ap.add_argument("--q-format", default="c32", choices=["c32", "c16"])
    ap.add_argument("--tile-ppt", type=int, default=0,
                    help="pixels per thread of the accumulation tile (0: library default)")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="after the steps, run one plain device-to-device copy of this many MiB "
                         "(reference kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--abi-call", type=int, default=1,
                    help="also time one backprojection() call with host buffers (1) or not (0)")
    ap.add_argument("--serving", default="C4,C5",
                    help="comma-separated serving job streams (C4, C5) whose latency/throughput "
                         "figures the line carries, or none")
    ap.add_argument("--serving-jobs", type=int, default=0,
                    help="cut the serving stream to its first N jobs (0: the whole stream)")
    ap.add_argument("--serving-load", type=float, default=0.7,
                    help="Poisson arrival rate as a fraction of the measured capacity")
    ap.add_argument("--published-case", type=int, default=1,
                    help="also time the sizes of the reference's published GPU run (1) or not (0)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this library runs on one GPU only (asked for --gpus {args.gpus}); "
              "sharding over several GPUs is not built", file=sys.stderr)
        sys.exit(2)

    import torch
    sar = importlib.import_module("halide-sar-app_amd")
    scenes = importlib.import_module("halide-sar-app_amd.scenes")
    pipe = importlib.import_module("halide-sar-app_amd.pipeline")
    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a GPU")

    sc = scenes.config_scene(args.config)
    grid = pipe.scene_grid(sc, with_pixel_locs=False)
    log(f"bench: {args.config} np={sc.npulses} ns={sc.nsamples} image {grid['nu']}x{grid['nv']}")
    ds = pipe.DeviceScene(sc, grid)
    qf = sar.Q_C16 if args.q_format == "c16" else sar.Q_C32
    plan = ds.plan(q_format=qf, tile_ppt=args.tile_ppt)
    out = ds.output()
    stream = torch.cuda.current_stream()
    pp = pipe.pixel_pulses(sc, grid)

    for _ in range(args.warmup):
        ds.run(plan, out)
    torch.cuda.synchronize()

    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record(stream)
        plan.range_profiles(ds.phs, ds.k_r)
        ev[k][1].record(stream)
        plan.accumulate(ds.k_r, ds.pos, ds.r, out)
        ev[k][2].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    fft_ms = float(np.mean([ev[k][0].elapsed_time(ev[k][1]) for k in range(args.steps)]))
    acc_ms = float(np.mean([ev[k][1].elapsed_time(ev[k][2]) for k in range(args.steps)]))
    step_ms = wall * 1e3 / args.steps
    value = pp / (step_ms * 1e-3) / 1e9
    log(f"bench: step {step_ms:.3f} ms (range profiles {fft_ms:.3f} ms, accumulate {acc_ms:.3f} ms)"
        f" -> {value:.1f} Gpixel-pulses/s")

    # the same accumulation with row windows staged in LDS (sarbp_plan_desc.row_windows = 1):
    # the same image bit for bit; timed beside the default form, not part of the step
    win_ms = win_paths = None
    if ds.grid_mode and qf == sar.Q_C32:
        wplan = ds.plan(q_format=qf, tile_ppt=args.tile_ppt, row_windows=1)
        wout = ds.output()
        ds.run(wplan, wout)
        torch.cuda.synchronize()
        same = bool(torch.equal(wout, out))
        ew = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ew[0].record(stream)
        for _ in range(5):
            wplan.accumulate(ds.k_r, ds.pos, ds.r, wout)
        ew[1].record(stream)
        torch.cuda.synchronize()
        win_ms = ew[0].elapsed_time(ew[1]) / 5
        win_paths = dict(wplan.tile_paths(), same_bits_as_default=same)
        log(f"bench: accumulate with row windows in LDS {win_ms:.3f} ms (same bits: {same})")
        del wplan, wout

    if args.copy_calibration_mb > 0:
        n = args.copy_calibration_mb * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=ds.dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"bench: calibration copy of {args.copy_calibration_mb} MiB done")
        del src, dst

    # dB + u8 output of the same image (reported beside the step, not part of it)
    npix = grid["nu"] * grid["nv"]
    dB = torch.empty(npix, dtype=torch.float64, device=ds.dev)
    u8 = torch.empty(npix, dtype=torch.uint8, device=ds.dev)
    # one pass (max |z| found by the accumulation: sarbp_plan_img_output), and the two-pass form
    # for an arbitrary image (sarbp_img_output_device)
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    plan.img_output(out, dB, u8, sc.db_min, sc.db_max)
    sar.img_output_device(out, npix, dB, u8, sc.db_min, sc.db_max)
    e0.record(stream)
    for _ in range(5):
        plan.img_output(out, dB, u8, sc.db_min, sc.db_max)
    e1.record(stream)
    for _ in range(5):
        sar.img_output_device(out, npix, dB, u8, sc.db_min, sc.db_max)
    e2.record(stream)
    torch.cuda.synchronize()
    img_ms = e0.elapsed_time(e1) / 5
    img2_ms = e1.elapsed_time(e2) / 5

    # accuracy of the timed image vs the oracle, and the CPU baseline on a row sample
    import oracle
    img = out.cpu().numpy()
    phs = ds.phs.cpu().numpy()
    cpu_info = host_cpu()
    nthreads = cpu_info["threads"]
    t_q0 = time.perf_counter()
    Q = oracle.range_profiles(phs, sc.k_r, sc.taylor, ds.n_fft)
    _, nr0 = oracle.norm_r0(sc.pos)
    t_q = time.perf_counter() - t_q0
    kc = float(sc.k_r[sc.nsamples // 2])
    rng = np.random.default_rng(12345)
    ys = rng.integers(0, grid["nv"], args.check_pixels)
    xs = rng.integers(0, grid["nu"], args.check_pixels)
    pts = sar.grid_points(grid, xs, ys)  # one-row geometry holding the sampled pixels
    cnt = len(xs)
    ref = oracle.backproject_points(Q, sc.nsamples, sc.delta_r, kc, sc.pos, nr0, pts, cnt, 1,
                                    np.arange(cnt, dtype=np.int32), np.zeros(cnt, np.int32),
                                    nthreads=nthreads)
    got = img[ys, xs]
    rel = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    log(f"bench: rel_l2 vs FP64 oracle on {args.check_pixels} pixels = {rel:.3e}")

    # host-side pixel locations of the whole image: input of the Halide-buffer entry point and of
    # the CPU baseline (built outside every timed region)
    band = sar.image_grid(grid["nu"], grid["nv"], grid["d_u"], grid["d_v"], grid["u_hat"],
                          grid["v_hat"]) if grid.get("pixel_locs") is None else grid

    # one call of the reference's entry point with host buffers: what `sarbp` measures, PCIe
    # copies of phs, pixel_locs (24 B/pixel) and the f64 image included (not the line's value)
    abi = None
    if args.abi_call:
        out_h = np.empty((grid["nv"], grid["nu"]), np.complex128)
        call = (phs, sc.k_r, sc.taylor, ds.n_fft, sc.delta_r, band["u"], band["v"], sc.pos,
                band["pixel_locs"], out_h)
        # five warm-up calls, then the median of five. The first call creates the plan and the
        # staging and touches the host pages; the GPU clock then takes a few calls to reach its
        # level under this duty cycle (each call leaves the GPU idle while phs is uploaded): the
        # calls ran ~6.7, 6.5, 6.4, 6.2, 5.9 ms, then ~5.8 ms, on one MI355X
        warm, times = [], []
        rv = 0
        for i in range(10):
            t_a = time.perf_counter()
            rv = rv or sar.backprojection(*call)
            (warm if i < 5 else times).append((time.perf_counter() - t_a) * 1e3)
        if rv != 0:
            raise SystemExit(f"backprojection() failed: {rv} {sar.last_error()}")
        abi_ms = sorted(times)[2]
        plane_bands = sar.abi_last_call_bands()
        abi = {"what": ("one backprojection() call with host buffers, as sarbp.cpp:276-280 times "
                        "it: PCIe copies of phs and the f64 image included; pixel_locs "
                        "(24 B/pixel) checked on the host against the image plane of u, v and "
                        "uploaded only for bands off it; median of 5 after 5 warm-up calls"),
               "ms": round(abi_ms, 3), "gpixel_pulses_per_s": round(pp / abi_ms / 1e6, 3),
               "runs_ms": [round(t, 3) for t in times],
               "warmup_runs_ms": [round(t, 3) for t in warm],
               "bands_on_image_plane": f"{plane_bands[0]} of {plane_bands[1]}",
               "rel_l2_vs_resident": float(np.linalg.norm(out_h - img) / np.linalg.norm(img))}
        log(f"bench: backprojection() with host buffers {abi_ms:.2f} ms (median of 5)")
        del out_h

    cpu = None
    if args.cpu_rows > 0:
        rows = min(args.cpu_rows, grid["nv"])
        y0 = grid["nv"] // 2 - rows // 2

        def run_rows(nrows, threads):
            t1 = time.perf_counter()
            oracle.backproject(Q, sc.nsamples, sc.delta_r, kc, sc.pos, nr0, band["pixel_locs"],
                               grid["nu"], grid["nv"], region=(0, y0, grid["nu"], nrows),
                               nthreads=threads)
            return time.perf_counter() - t1

        runs = sorted(run_rows(rows, nthreads) for _ in range(3))
        t_rows = runs[1]   # median of three
        t_full = t_q + t_rows * grid["nv"] / rows
        rows1 = max(4, rows // 16)
        t_rows1 = run_rows(rows1, 1)
        t_full1 = t_q + t_rows1 * grid["nv"] / rows1
        cpu = {"value": round(pp / t_full / 1e9, 5), "unit": "Gpixel·pulses/s",
               "cores": nthreads, "kind": "port",
               "code": ("oracle/sarbp_oracle.c: the reference CPU schedule's loop nest "
                        "(backprojection.cpp:303-314: 4-pixel groups over threads, "
                        "norm_rr0[np] staged per pixel of the group) in f64 C with libm sin/cos, "
                        "built -O3 -mavx2: the staged norm_rr0 loop (sqrt included) runs four "
                        "f64 lanes wide, as the reference vectorises norm_rr0 over pulses "
                        "(:303-305); its accumulation update (img.inner.update, :309) is not "
                        "vectorised, and runs scalar here as there"),
               "usable_cores": cpu_info["usable"], "cpu_model": cpu_info["model"],
               "sample": (f"{rows} of {grid['nv']} rows ({rows * grid['nu']} pixels) x "
                          f"{sc.npulses} pulses, median of 3 runs {t_rows:.2f} s "
                          f"(runs {runs[0]:.2f}/{runs[1]:.2f}/{runs[2]:.2f} s) + all "
                          f"{sc.npulses} range profiles (FFT, one thread) in {t_q:.1f} s, "
                          f"extrapolated to the full image"),
               "value_1_thread": round(pp / t_full1 / 1e9, 5),
               "sample_1_thread": (f"{rows1} rows in {t_rows1:.2f} s on one thread + the range "
                                   f"profiles, extrapolated")}
        log(f"bench: cpu baseline {cpu['value']} Gpp/s on {nthreads} threads, "
            f"{cpu['value_1_thread']} Gpp/s on one")

    flop = pp * pipe.algorithmic_flops_per_pp()
    achieved = flop / (acc_ms * 1e-3) / 1e12
    ppt_used = plan.tile_ppt
    kernel = (f"accumulate_kernel<{'QPairF16' if qf == sar.Q_C16 else 'QPairF32'}, {ppt_used}, "
              f"{'grid' if ds.grid_mode else 'pixel_locs'}>")
    # HBM traffic per launch and the binding resource are not measured by this run: they come from
    # the committed rocprofv3 counter summary of the same kernel on the same workload (the line
    # names the file); nothing is reported when the summary covers another kernel.
    traffic = limiter = None
    traffic_source = None
    pmc_rel = os.path.join("profiles", f"accumulate_{args.config}_hbm_traffic.json")
    pmc = os.path.join(ROOT, pmc_rel)
    if os.path.exists(pmc):
        with open(pmc) as f:
            rec = json.load(f)
        if rec.get("kernel") == kernel:   # counters of this kernel on this workload only
            traffic = rec.get("bytes_per_launch")
            limiter = rec.get("limiter")
            traffic_source = (f"{pmc_rel}: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (and the "
                              f"limiter's counters) of this kernel on this workload, scaled by a "
                              f"copy kernel; not measured by this run")
    if limiter is not None:
        limiter = dict(limiter, source=pmc_rel)
        if "gather_bytes_per_pixel_pulse" in limiter:
            # the gathers' bytes through the texture path at this run's rate
            limiter["gather_tb_per_s"] = round(
                pp * limiter["gather_bytes_per_pixel_pulse"] / (acc_ms * 1e-3) / 1e12, 3)
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gpixel·pulses/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": (f"{args.config}: backprojection of {grid['nu']}x{grid['nv']} pixels from "
                         f"{sc.npulses} pulses x {sc.nsamples} samples, N_fft {ds.n_fft}, Taylor "
                         f"{sc.taylor}, inputs resident in HBM"),
            "pixels": grid["nu"] * grid["nv"], "pulses": sc.npulses, "samples": sc.nsamples,
            "n_fft": ds.n_fft, "q_format": args.q_format, "parallelism": "single-gpu",
        },
        "rel_l2": rel,
        "stage_ms": {"range_profiles": round(fft_ms, 4), "accumulate": round(acc_ms, 4),
                     "img_output_dB_u8": round(img_ms, 4),
                     "img_output_dB_u8_two_pass": round(img2_ms, 4),
                     "accumulate_row_windows": round(win_ms, 4) if win_ms else None},
        "row_windows_tiles": win_paths,
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                     "traffic": traffic, "traffic_source": traffic_source, "kernel": kernel,
                     "flop_per_launch": flop,
                     "achieved_from": (f"{pp:.4g} pixel-pulses x 30 FLOP / mean accumulation "
                                       f"duration {acc_ms:.4f} ms (HIP events on the launch "
                                       f"stream, {args.steps} steps)"),
                     "measured_limiter": limiter},
        "cpu_baseline": cpu,
        "abi_call": abi,
    }
    if args.published_case:
        del out, dB, u8, plan, ds
        torch.cuda.empty_cache()
        line["published_case"] = run_published_case(sar, scenes, pipe, torch)
    streams = [c for c in args.serving.split(",") if c and c != "none"]
    if streams:
        line["serving"] = run_serving(args, streams[0])
        for extra in streams[1:]:
            line[f"serving_{extra}"] = run_serving(args, extra)
    print(json.dumps(line), flush=True)


def host_cpu() -> dict:
    """The host CPU: model name, the cores this process may run on (affinity mask) and the threads
    the CPU baseline uses: all of them, or the CPU share OMP_NUM_THREADS grants when it is set."""
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    share = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(usable, int(share)) if share.isdigit() and int(share) > 0 else usable
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "usable": usable, "threads": max(1, threads)}


# The reference's published GPU figure: `sarbp -s cuda_distributed -t 30 -u 2 -r 0.25` on its
# Sandia 8192-pulse x 8192-sample data (8192 x 8192 image, N_fft 32768) over 4 Summit nodes with one
# V100 each; rank 0's backprojection() call returned in 22 014 ms
# (exp/summit/darpa-demo2/logs/SARBP-Sandia-4.763156:182-197).
PUBLISHED_GPU = {"value": round(8192.0 ** 3 / 22.014 / 1e9, 2), "unit": "Gpixel·pulses/s",
                 "hardware": "4 x NVIDIA V100 (Summit, one per node), Halide cuda_distributed",
                 "call_ms": 22014,
                 "source": "exp/summit/darpa-demo2/logs/SARBP-Sandia-4.763156:182-197"}


def run_published_case(sar, scenes, pipe, torch) -> dict:
    """The sizes of the reference's published GPU run on one MI355X: 8192 pulses x 8192 samples,
    an 8192 x 8192 image, Taylor 30, N_fft 32768 (four-step FFT), on a synthetic scene of the same
    sizes (the reference's Sandia data is not here). Times the device-resident step (HIP events)
    and backprojection() with host buffers (wall clock, median of 3 after a warm-up call), and
    compares the two images."""
    sc = scenes.config_scene("P8192_S8192")
    grid = pipe.scene_grid(sc, with_pixel_locs=False)
    ds = pipe.DeviceScene(sc, grid)
    pp = pipe.pixel_pulses(sc, grid)
    plan = ds.plan()
    out = ds.output()
    stream = torch.cuda.current_stream()
    for _ in range(2):
        ds.run(plan, out)
    torch.cuda.synchronize()
    res_ms = []
    for _ in range(3):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        ds.run(plan, out)
        e1.record(stream)
        torch.cuda.synchronize()
        res_ms.append(e0.elapsed_time(e1))
    img = out.cpu().numpy()
    del out, plan
    phs = ds.phs.cpu().numpy()
    n_fft = ds.n_fft
    del ds
    torch.cuda.empty_cache()
    plane = sar.image_grid(grid["nu"], grid["nv"], grid["d_u"], grid["d_v"], grid["u_hat"],
                           grid["v_hat"])
    out_h = np.empty((grid["nv"], grid["nu"]), np.complex128)
    call = (phs, sc.k_r, sc.taylor, n_fft, sc.delta_r, plane["u"], plane["v"], sc.pos,
            plane["pixel_locs"], out_h)
    rv = sar.backprojection(*call)
    times = []
    for _ in range(3):
        t_a = time.perf_counter()
        rv = rv or sar.backprojection(*call)
        times.append((time.perf_counter() - t_a) * 1e3)
    if rv != 0:
        raise SystemExit(f"backprojection() failed on {sc.name}: {rv} {sar.last_error()}")
    call_ms = sorted(times)[1]
    res_med = sorted(res_ms)[1]
    agree = float(np.linalg.norm(out_h - img) / np.linalg.norm(img))
    log(f"bench: {sc.name} resident {res_med:.1f} ms, backprojection() {call_ms:.1f} ms, "
        f"images agree to {agree:.2e}")
    return {"workload": (f"{sc.name}: {grid['nu']}x{grid['nv']} pixels from {sc.npulses} pulses x "
                         f"{sc.nsamples} samples, N_fft {n_fft}, Taylor {sc.taylor}; synthetic "
                         "scene of the reference's published GPU run's sizes"),
            "pixel_pulses": pp,
            "resident_ms": round(res_med, 2), "resident_runs_ms": [round(t, 2) for t in res_ms],
            "resident_gpixel_pulses_per_s": round(pp / res_med / 1e6, 1),
            "call_ms": round(call_ms, 2), "call_runs_ms": [round(t, 2) for t in times],
            "call_gpixel_pulses_per_s": round(pp / call_ms / 1e6, 1),
            "call_rel_l2_vs_resident": agree,
            "published": PUBLISHED_GPU,
            "call_vs_published": round(pp / call_ms / 1e6 / PUBLISHED_GPU["value"], 1)}


def run_serving(args, config) -> dict:
    """Serving figures: p50/p99 latency per job and jobs/s of the continuous-batching engine and of
    the one-job-per-launch configuration, both at the same Poisson arrival rates (`stream`: 70 % of
    the engine's capacity; `stream_low`: 70 % of the one-job-per-launch capacity).

    The replay runs in a process of its own (`serving.py`), as a serving deployment would. HIP maps
    a process's streams onto its GPU_MAX_HW_QUEUES hardware queues in turn, and the streams that
    this process's other legs created earlier shifted the engine's streams onto shared queues: the
    fused launches then queued behind a large job's bands (C5 at 70 %: p50 / p99 of 16.1 / 26.7 ms
    here against 1.9 / 5.7 ms in a process of its own, on the same MI355X)."""
    if config not in ("C4", "C5"):
        raise SystemExit(f"bench.py: unknown serving stream {config!r} (C4, C5 or none)")
    cmd = [sys.executable, "-u", os.path.join(ROOT, "halide-sar-app_amd", "serving.py"),
           "--config", config, "--load", str(args.serving_load)]
    if args.serving_jobs:
        cmd += ["--jobs", str(args.serving_jobs)]
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=None, text=True, timeout=900)
    lines = [ln for ln in proc.stdout.splitlines() if ln.startswith("{")]
    if proc.returncode != 0 or not lines:
        raise SystemExit(f"bench.py: serving replay {config} failed ({proc.returncode})")
    res = json.loads(lines[-1])

    def rnd(d):
        return {k: (round(v, 3) if isinstance(v, float) else v) for k, v in d.items()}
    return {"config": config, "jobs": res["jobs"], "load": args.serving_load,
            "process": "serving.py, a process of its own",
            "arrival_rate_hz": round(res["arrival_rate_hz"], 2),
            "arrival_rate_low_hz": round(res["arrival_rate_low_hz"], 2),
            "engine": {k: rnd(v) for k, v in res["serving"].items()},
            "one_job_per_launch": {k: rnd(v) for k, v in res["one_job_per_launch"].items()}}


if __name__ == "__main__":
    main()
