"""Serving form: every job's error against the CPU FP64 oracle (test tooling; the oracle is only the
checker).

    python tests/serving_error_report.py --config C5 > serving_error_C5.jsonl

Runs a job stream (C4 or C5 as `halide-sar-app_amd/serving.py` builds them) through the serving
engine: a warm-up pass that also tunes every shape class, then one pass with all jobs submitted at
once, so that the mix is as dense as the engine makes it (small jobs fused, large jobs banded, all
sharing the device). Every job's image from that pass is kept and compared with the oracle
(range profiles and accumulation in FP64, `oracle/sarbp_oracle.c`) on sampled pixels: `--sample`
uniform random pixels plus the (2h+1)^2 window around the brightest pixel of the job's image.

Prints one JSON line per job (index, side, pulses, samples, fused, tile height, rel-L2 over the
random pixels and over the window, whether the window's peak moved) and a last line with the
summary (largest, median and 99th-percentile rel-L2, jobs above 1e-4).
"""
import argparse
import concurrent.futures
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
import serving_check  # noqa: E402


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="C5", choices=["C4", "C5"])
    ap.add_argument("--jobs", type=int, default=0, help="first N jobs of the stream only")
    ap.add_argument("--sample", type=int, default=128, help="uniform random pixels per job")
    ap.add_argument("--half", type=int, default=4, help="half-size of the window at the peak")
    ap.add_argument("--workers", type=int, default=16, help="host threads for the oracle")
    args = ap.parse_args(argv)

    sv = importlib.import_module("halide-sar-app_amd.serving")
    sv.request_hw_queues()
    sar = importlib.import_module("halide-sar-app_amd")
    import oracle

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    specs = sv.c4_specs() if args.config == "C4" else sv.c5_specs()
    if args.jobs:
        specs = specs[:args.jobs]
    n = len(specs)
    t0 = time.perf_counter()
    fac = sv.JobFactory(specs)
    log(f"serving error report: {args.config}, {n} jobs generated in "
        f"{time.perf_counter() - t0:.1f} s")
    with sv.Server() as srv:
        sv.replay(srv, fac, None)
        recs, _, kept = sv.replay(srv, fac, None, keep_outputs=range(n))
    log(f"serving error report: stream done, {sum(r['fused'] for r in recs)} of {n} jobs fused")

    def check(i):
        sp, r = specs[i], recs[i]
        res = serving_check.check_job(sar, oracle, fac, i, kept[i], sample=args.sample,
                                      half=args.half)
        return dict(job=i, side=sp["side"], pulses=sp["npulses"], samples=sp["ns"],
                    taylor=sp["taylor"], fused=int(r["fused"]), ppt=int(r["ppt"]), **res)

    t1 = time.perf_counter()
    lines = [None] * n
    with concurrent.futures.ThreadPoolExecutor(max_workers=args.workers) as ex:
        futs = {ex.submit(check, i): i for i in range(n)}
        for k, f in enumerate(concurrent.futures.as_completed(futs)):
            i = futs[f]
            lines[i] = f.result()
            kept.pop(i, None)
            if (k + 1) % 200 == 0 or k + 1 == n:
                log(f"serving error report: {k + 1} of {n} jobs checked "
                    f"({time.perf_counter() - t1:.0f} s)")
    for line in lines:
        print(json.dumps(line), flush=True)
    rel = np.array([max(x["rel_l2"], x["rel_l2_peak_window"]) for x in lines])
    print(json.dumps(dict(config=args.config, jobs=n, rel_l2_max=float(rel.max()),
                          rel_l2_median=float(np.median(rel)),
                          rel_l2_p99=float(np.percentile(rel, 99)),
                          jobs_above_1e4=int((rel > 1e-4).sum()),
                          peaks_moved=int(sum(x["peak_moved"] for x in lines)),
                          sample_pixels=args.sample, window=(2 * args.half + 1) ** 2,
                          oracle_s=round(time.perf_counter() - t1, 1))), flush=True)
    return 0 if (rel <= 1e-4).all() else 1


if __name__ == "__main__":
    sys.exit(main())
