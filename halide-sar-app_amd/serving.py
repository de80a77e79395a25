"""Serving form: bindings of the continuous-batching engine (sarbp_server_*) and the job-stream
workloads of configurations C4 and C5.

A job is one image formation: phase history, k_r and platform positions in host memory, an image
grid, and a host output buffer. The engine (csrc/serving.cpp) admits jobs under an HBM budget,
overlaps their transfers and kernels on several streams and fuses small jobs' accumulations into one
launch. This module builds job streams, replays them with Poisson arrivals and reports latency
percentiles and throughput, also for the one-job-at-a-time configuration of the same engine.
"""
from __future__ import annotations

import ctypes
import dataclasses
import importlib
import threading
import time

import numpy as np

# Hardware queues the engine's streams want: num_streams + 5 streams and the null stream, each on
# a queue of its own (HIP maps streams onto GPU_MAX_HW_QUEUES queues, 4 by default, and packets that
# share a queue run in order).
ENGINE_HW_QUEUES = 8


def request_hw_queues(n: int = ENGINE_HW_QUEUES, exact: bool = False) -> None:
    """Raise GPU_MAX_HW_QUEUES to n for this process (set it to n with exact=True); effective only
    before the HIP runtime starts (the first HIP call of the process)."""
    import os
    cur = os.environ.get("GPU_MAX_HW_QUEUES", "")
    if exact or not cur.isdigit() or int(cur) < n:
        os.environ["GPU_MAX_HW_QUEUES"] = str(n)


if __name__ == "__main__":  # run as a script: the package's parent directory goes on the path
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    # --hw-queues is applied here, before anything below can start the HIP runtime
    _hwq = [sys.argv[k + 1] for k, a in enumerate(sys.argv[:-1]) if a == "--hw-queues"]
    request_hw_queues(int(_hwq[-1]) if _hwq else ENGINE_HW_QUEUES, exact=bool(_hwq))

_pkg = importlib.import_module("halide-sar-app_amd")
_scenes = importlib.import_module("halide-sar-app_amd.scenes")
_lib = _pkg.lib()
_P, _i32, _i64, _f64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double


class ServerConfig(ctypes.Structure):
    _fields_ = [("device", _i32), ("max_inflight", _i32), ("hbm_budget_bytes", _i64),
                ("fuse_max_pixel_pulses", _i64), ("num_streams", _i32), ("autotune", _i32),
                ("fuse_launch_pixel_pulses", _i64), ("band_order", _i32), ("reserved", _i32 * 5)]


class JobDesc(ctypes.Structure):
    _fields_ = [("phs", _P), ("k_r", _P), ("pos", _P), ("nsamples", _i32), ("npulses", _i32),
                ("n_fft", _i32), ("taylor_s_l", _i32), ("delta_r", _f64), ("nu", _i32),
                ("nv", _i32), ("d_u", _f64), ("d_v", _f64), ("u_hat", _f64 * 3),
                ("v_hat", _f64 * 3), ("pixel_locs", _P), ("norm_r0_policy", _i32),
                ("taylor_den_policy", _i32), ("q_format", _i32), ("range_mode", _i32),
                ("out", _P), ("reserved", _i32 * 8)]


class JobStats(ctypes.Structure):
    _fields_ = [("t_submit_ns", _i64), ("t_admit_ns", _i64), ("t_end_ns", _i64),
                ("device_bytes", _i64), ("fused", _i32), ("tile_ppt", _i32), ("status", _i32),
                ("host_enqueue_us", _i32), ("device_us", _i32), ("row_windows", _i32),
                ("reserved", _i32 * 2), ("up_us", _i32), ("ready_us", _i32),
                ("launch_us", _i32), ("acc_us", _i32)]


_lib.sarbp_server_config_init.argtypes = [ctypes.POINTER(ServerConfig)]
_lib.sarbp_server_config_init.restype = None
_lib.sarbp_server_create.argtypes = [ctypes.POINTER(ServerConfig), ctypes.POINTER(_P)]
_lib.sarbp_server_submit.argtypes = [_P, ctypes.POINTER(JobDesc), ctypes.POINTER(ctypes.c_uint64)]
_lib.sarbp_server_wait.argtypes = [_P, ctypes.c_uint64, _i32, ctypes.POINTER(JobStats)]
_lib.sarbp_server_destroy.argtypes = [_P]


class ServingError(RuntimeError):
    def __init__(self, msg, code):
        super().__init__(msg)
        self.code = code


class Server:
    def __init__(self, device=0, max_inflight=64, hbm_budget_gb=64.0, fuse_max_pixel_pulses=2 ** 31,
                 num_streams=2, autotune=True, fuse_launch_pixel_pulses=None, band_order=0):
        c = ServerConfig()
        _lib.sarbp_server_config_init(ctypes.byref(c))
        c.device, c.max_inflight = device, max_inflight
        c.hbm_budget_bytes = int(hbm_budget_gb * (1 << 30))
        c.fuse_max_pixel_pulses = int(fuse_max_pixel_pulses)
        c.num_streams, c.autotune = num_streams, int(bool(autotune))
        if fuse_launch_pixel_pulses is not None:
            c.fuse_launch_pixel_pulses = int(fuse_launch_pixel_pulses)
        c.band_order = int(band_order)
        self.cfg = c
        self._s = _P()
        rv = _lib.sarbp_server_create(ctypes.byref(c), ctypes.byref(self._s))
        if rv != 0:
            raise ServingError(f"sarbp_server_create failed ({rv}): {_pkg.last_error()}", rv)

    def submit(self, desc: JobDesc) -> int:
        jid = ctypes.c_uint64()
        rv = _lib.sarbp_server_submit(self._s, ctypes.byref(desc), ctypes.byref(jid))
        if rv != 0:
            raise ServingError(f"sarbp_server_submit failed ({rv}): {_pkg.last_error()}", rv)
        return jid.value

    def wait(self, jid: int, timeout_ms: int = -1) -> JobStats:
        st = JobStats()
        rv = _lib.sarbp_server_wait(self._s, jid, timeout_ms, ctypes.byref(st))
        if rv == 1:
            raise TimeoutError(f"job {jid}")
        if rv != 0:
            raise ServingError(f"job {jid} failed ({rv}): {_pkg.last_error()}", rv)
        return st

    def close(self):
        if self._s is not None and self._s.value:
            _lib.sarbp_server_destroy(self._s)
        self._s = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# ---------------------------------------------------------------- jobs


@dataclasses.dataclass
class Job:
    """Host-side job: arrays must stay alive until the job completes."""
    phs: np.ndarray            # complex64 (np, ns), ideally page-locked
    k_r: np.ndarray            # float32 (ns,)
    pos: np.ndarray            # float32 (np, 3)
    n_fft: int
    taylor: int
    delta_r: float
    side: int
    d_u: float
    u_hat: np.ndarray
    v_hat: np.ndarray
    out: np.ndarray            # complex128 (side, side), ideally page-locked
    taylor_den_policy: int | None = None   # None: the process default, as for a plan
    q_format: int = _pkg.Q_C32
    range_mode: int = _pkg.RANGE_F64
    keep: tuple = ()           # owners of the memory above (e.g. pinned torch tensors)

    def desc(self) -> JobDesc:
        d = JobDesc()
        d.phs = self.phs.ctypes.data
        d.k_r = self.k_r.ctypes.data
        d.pos = self.pos.ctypes.data
        d.nsamples, d.npulses = self.phs.shape[1], self.phs.shape[0]
        d.n_fft, d.taylor_s_l, d.delta_r = self.n_fft, self.taylor, self.delta_r
        d.nu = d.nv = self.side
        d.d_u = d.d_v = self.d_u
        d.u_hat[:] = list(self.u_hat)
        d.v_hat[:] = list(self.v_hat)
        d.pixel_locs = None
        d.norm_r0_policy = _pkg.POLICY_DEFAULT
        d.taylor_den_policy = (_pkg.POLICY_DEFAULT if self.taylor_den_policy is None
                               else self.taylor_den_policy)
        d.q_format = self.q_format
        d.range_mode = self.range_mode
        d.out = self.out.ctypes.data
        return d

    @property
    def pixel_pulses(self) -> float:
        return float(self.side) ** 2 * self.phs.shape[0]


def _pinned(nbytes):
    """Page-locked host bytes (torch pin_memory) as a uint8 array; the tensor owns the memory."""
    import torch
    t = torch.empty(int(nbytes), dtype=torch.uint8, pin_memory=True)
    return t.numpy(), t


class JobFactory:
    """Seeded jobs of a spec list. Each job has its own scene (linear spotlight track, seed
    `seed_base + index`, 10-500 Rayleigh point scatterers), simulated on the GPU before the
    stream starts and kept in one page-locked slab, so transfers during the stream are
    asynchronous DMA. Outputs are page-locked images taken from per-side free lists."""

    # The streams' jobs ask for Taylor windows of 17-35 dB. Under the Int(32) reading of the
    # window's F_m denominator (the library default) every level of 31 dB or more divides by zero
    # and gives a non-finite image, as the reference's would; the streams therefore ask for the f32
    # reading, whose window is finite at every level (the readings agree up to 17 dB, n_bar 2).
    DEFAULT_TAYLOR_DEN = _pkg.TAYLOR_DEN_F32

    def __init__(self, specs, seed_base=4000, device=0, q_format=None, taylor_den_policy=None,
                 range_mode=None):
        import torch
        pipe = importlib.import_module("halide-sar-app_amd.pipeline")
        self.specs = specs
        self.scenes = []
        for i, sp in enumerate(specs):
            rng = np.random.default_rng(seed_base + i)
            nt = sp.get("targets", int(rng.integers(10, 501)))
            self.scenes.append(_scenes.linear_spotlight_scene(
                npulses=sp["npulses"], nsamples=sp["ns"], ntargets=nt, nclutter=0,
                taylor=sp["taylor"], seed=seed_base + i, name=f"job{i}"))
        sizes = [8 * sc.npulses * sc.nsamples for sc in self.scenes]
        offs = np.concatenate([[0], np.cumsum([(b + 255) // 256 * 256 for b in sizes])])
        self.slab, self._owner = _pinned(int(offs[-1]))
        dev = torch.device("cuda", device)
        self.phs = []
        for sc, o, b in zip(self.scenes, offs, sizes):
            view = self.slab[o:o + b].view(np.complex64).reshape(sc.npulses, sc.nsamples)
            view[...] = pipe.simulate_device(sc, dev).cpu().numpy()
            self.phs.append(view)
        torch.cuda.synchronize(dev)
        self.q_format = _pkg.Q_C32 if q_format is None else q_format
        self.taylor_den_policy = (self.DEFAULT_TAYLOR_DEN if taylor_den_policy is None
                                  else taylor_den_policy)
        self.range_mode = _pkg.RANGE_F64 if range_mode is None else range_mode
        self.grids = [self._grid(sc, sp["side"]) for sc, sp in zip(self.scenes, specs)]
        self.free = {}
        self.lock = threading.Lock()

    @staticmethod
    def _grid(sc, side):
        ip = _pkg.img_plane_create(sc.nsamples, sc.npulses, sc.delta_r, sc.R_c,
                                   with_pixel_locs=False)
        return dict(side=side, d_u=sc.delta_r * sc.nsamples / side, u_hat=ip["u_hat"],
                    v_hat=ip["v_hat"])

    def take_out(self, side):
        with self.lock:
            lst = self.free.setdefault(side, [])
            if lst:
                return lst.pop()
        a, t = _pinned(16 * side * side)
        return a.view(np.complex128).reshape(side, side), t

    def give_out(self, side, item):
        with self.lock:
            self.free.setdefault(side, []).append(item)

    def job(self, i, out, out_owner=None) -> Job:
        sc, g, sp = self.scenes[i], self.grids[i], self.specs[i]
        return Job(phs=self.phs[i], k_r=sc.k_r, pos=sc.pos, n_fft=_pkg.nfft(sc.nsamples, 2),
                   taylor=sp["taylor"], delta_r=sc.delta_r, side=g["side"], d_u=g["d_u"],
                   u_hat=g["u_hat"], v_hat=g["v_hat"], out=out, q_format=self.q_format,
                   taylor_den_policy=self.taylor_den_policy, range_mode=self.range_mode,
                   keep=(out_owner,))

    def pixel_pulses(self, i) -> float:
        return float(self.specs[i]["side"]) ** 2 * self.specs[i]["npulses"]


def c4_specs(njobs=512, seed=5000):
    """Configuration C4: pulses U{200..4000}, samples in {424, 1024, 2048}, square images of side
    {256, 512, 1024, 2048}, Taylor U{17..35}."""
    rng = np.random.default_rng(seed)
    return [dict(ns=int(rng.choice([424, 1024, 2048])),
                 npulses=int(rng.integers(200, 4001)), side=int(rng.choice([256, 512, 1024, 2048])),
                 taylor=int(rng.integers(17, 36))) for _ in range(njobs)]


def c5_specs(nbig=4, nsmall=2000, big_side=8192, big_pulses=6000, big_ns=4096, small_side=256,
             small_pulses=500, small_ns=424, seed=5100):
    """Configuration C5: `nbig` large jobs spread evenly through a background of small ones."""
    rng = np.random.default_rng(seed)
    specs = [dict(ns=small_ns, npulses=small_pulses,
                  side=small_side, taylor=int(rng.integers(17, 36))) for _ in range(nsmall)]
    step = max(1, nsmall // (nbig + 1))
    for b in range(nbig):
        specs.insert((b + 1) * step + b, dict(ns=big_ns, npulses=big_pulses, side=big_side,
                                              taylor=30, big=True))
    return specs


def replay(server: Server, factory: JobFactory, rate_hz: float | None, keep_outputs=(),
           seed=77):
    """Submit the factory's jobs in order with exponential inter-arrival times (rate_hz None: all
    at once), collect them on a second thread, and return per-job records (times from the
    engine's clock) plus copies of the outputs of the jobs named in `keep_outputs`."""
    n = len(factory.specs)
    rng = np.random.default_rng(seed)
    arrivals = np.zeros(n) if rate_hz is None else np.cumsum(rng.exponential(1.0 / rate_hz, n))
    ids, bufs, jobs, done = [None] * n, [None] * n, [None] * n, [None] * n
    kept, keep = {}, set(keep_outputs)
    errors = []
    submitted = threading.Semaphore(0)

    def collector():
        for i in range(n):
            submitted.acquire()
            try:
                st = server.wait(ids[i])
                done[i] = (st.t_submit_ns, st.t_admit_ns, st.t_end_ns, st.fused, st.tile_ppt,
                           st.host_enqueue_us, st.device_us, st.row_windows,
                           (st.up_us, st.ready_us, st.launch_us, st.acc_us))
            except Exception as e:  # noqa: BLE001 - reported after the stream
                errors.append((i, str(e)))
            if i in keep:
                kept[i] = bufs[i][0].copy()
            factory.give_out(factory.specs[i]["side"], bufs[i])
            jobs[i] = None

    th = threading.Thread(target=collector, daemon=True)
    th.start()
    t0 = time.perf_counter()
    for i in range(n):
        while True:
            lag = arrivals[i] - (time.perf_counter() - t0)
            if lag <= 0:
                break
            time.sleep(min(lag, 2e-4))
        bufs[i] = factory.take_out(factory.specs[i]["side"])
        jobs[i] = factory.job(i, bufs[i][0], bufs[i][1])
        ids[i] = server.submit(jobs[i].desc())
        submitted.release()
    th.join()
    wall = time.perf_counter() - t0
    if errors:
        raise RuntimeError(f"{len(errors)} jobs failed, first: {errors[0]}")
    recs = []
    for i, sp in enumerate(factory.specs):
        ts, ta, te, fused, ppt, hu, du, rw, steps = done[i]
        recs.append(dict(latency_ms=(te - ts) / 1e6, queue_ms=(ta - ts) / 1e6, t_end_ms=te / 1e6,
                         t_submit_ms=ts / 1e6, fused=fused, ppt=ppt, row_windows=rw,
                         host_enqueue_ms=hu / 1e3,
                         device_ms=du / 1e3, steps_ms=_step_ms(steps, (te - ta) / 1e6),
                         pp=factory.pixel_pulses(i), big=sp.get("big", False)))
    return recs, wall, kept


SMALL_JOB_PIXEL_PULSES = 2 ** 31  # the engine's default fusion bound: jobs at most this size are "small"
STEPS = ("upload", "profiles", "wait", "accumulate", "download")


def _step_ms(steps_us, admitted_ms):
    """Durations of a job's steps after admission from the engine's step times (microseconds after
    admission): upload, range profiles, wait for its accumulation launch, accumulation, download
    (to the end of the job). A job accumulated in bands enqueues its range profiles with its bands,
    so its profiles and wait steps are 0 and its accumulation includes the profiles."""
    up, ready, launch, acc = (t / 1e3 for t in steps_us)
    if min(up, ready, launch, acc) < 0:
        return None
    return dict(zip(STEPS, (up, ready - up, launch - ready, acc - launch, admitted_ms - acc)))


def summarize(recs, wall, alone_ms=None) -> dict:
    """Latency, queueing, device and throughput figures of one replay. Latency percentiles are
    also given per size class (small: at most SMALL_JOB_PIXEL_PULSES pixel-pulses), and, with
    alone_ms (each job's submission-to-output time when it had the device to itself), the
    slowdown latency / alone of every job: the tail of a mixed stream is otherwise dominated by
    its largest jobs' own duration."""
    lat = np.array([r["latency_ms"] for r in recs])
    span = (max(r["t_end_ms"] for r in recs) - min(r["t_submit_ms"] for r in recs)) / 1e3
    pp = sum(r["pp"] for r in recs)
    q = np.array([r["queue_ms"] for r in recs])
    hq = np.array([r["host_enqueue_ms"] for r in recs])
    dv = np.array([r["device_ms"] for r in recs])
    return dict(jobs=len(recs), p50_ms=float(np.percentile(lat, 50)),
                p99_ms=float(np.percentile(lat, 99)), mean_ms=float(lat.mean()),
                queue_p50_ms=float(np.percentile(q, 50)), queue_p99_ms=float(np.percentile(q, 99)),
                device_p50_ms=float(np.percentile(dv, 50)),
                device_p99_ms=float(np.percentile(dv, 99)),
                host_enqueue_mean_ms=float(hq.mean()), host_enqueue_max_ms=float(hq.max()),
                host_enqueue_total_s=float(hq.sum() / 1e3),
                jobs_per_s=len(recs) / span, gpixel_pulses_per_s=pp / span / 1e9,
                fused_jobs=int(sum(r["fused"] for r in recs)), span_s=span, wall_s=wall,
                **_by_class(recs, lat, alone_ms))


def _by_class(recs, lat, alone_ms) -> dict:
    out = {}
    small = np.array([r["pp"] <= SMALL_JOB_PIXEL_PULSES for r in recs])
    for name, sel in (("small", small), ("large", ~small)):
        if sel.any():
            out[f"{name}_jobs"] = int(sel.sum())
            out[f"{name}_p50_ms"] = float(np.percentile(lat[sel], 50))
            out[f"{name}_p99_ms"] = float(np.percentile(lat[sel], 99))
        recs_c = [r for r, m in zip(recs, sel) if m and r.get("steps_ms")]
        if recs_c:  # p50 and p99 of each step's duration
            out[f"{name}_steps_ms"] = {
                k: [round(float(np.percentile([r["steps_ms"][k] for r in recs_c], q)), 3)
                    for q in (50, 99)] for k in STEPS}
    if alone_ms is not None:
        slow = lat / np.maximum(np.asarray(alone_ms, np.float64), 1e-3)
        out["slowdown_p50"] = float(np.percentile(slow, 50))
        out["slowdown_p99"] = float(np.percentile(slow, 99))
        for name, sel in (("small", small), ("large", ~small)):
            if sel.any():
                out[f"{name}_slowdown_p99"] = float(np.percentile(slow[sel], 99))
    return out


def one_job_per_launch_config(device=0) -> dict:
    """The plain path on the same engine: one job at a time on one stream, no fusion, no tuning."""
    return dict(device=device, max_inflight=1, fuse_max_pixel_pulses=0, num_streams=1,
                autotune=False)


def benchmark(specs, load=0.7, seed_base=4000, device=0, log=print, **server_kw) -> dict:
    """Serving figures of a job stream for the serving engine and for the one-job-per-launch
    configuration: each one's capacity (all jobs submitted at once, after a warm-up pass that also
    tunes every shape class), then two Poisson streams run on both: `stream` at `load` x the
    engine's capacity and `stream_low` at `load` x the one-job-per-launch capacity (a rate both
    sustain)."""
    t = time.perf_counter()
    factory = JobFactory(specs, seed_base=seed_base, device=device)
    log(f"serving: {len(specs)} jobs generated in {time.perf_counter() - t:.1f} s")
    res = {}
    with Server(device=device, **server_kw) as srv, \
            Server(**one_job_per_launch_config(device)) as base:
        caps = {}
        alone = None
        for name, s in (("one_job_per_launch", base), ("serving", srv)):
            replay(s, factory, None)
            recs, wall, _ = replay(s, factory, None)
            if alone is None:
                # one job at a time: each job's own submission-to-output time on an idle device
                # (its queueing behind earlier jobs excluded), the reference for slowdowns
                alone = [r["latency_ms"] - r["queue_ms"] for r in recs]
            caps[name] = summarize(recs, wall, alone)
            log(f"{name}: capacity {caps[name]['jobs_per_s']:.1f} jobs/s")
        rate = load * caps["serving"]["jobs_per_s"]
        rate_low = load * caps["one_job_per_launch"]["jobs_per_s"]
        for name, s in (("serving", srv), ("one_job_per_launch", base)):
            res[name] = dict(capacity=caps[name])
            for key, r in (("stream", rate), ("stream_low", rate_low)):
                res[name][key] = summarize(*replay(s, factory, r)[:2], alone)
                log(f"{name}: Poisson {r:.1f}/s: {res[name][key]}")
    res["arrival_rate_hz"] = rate
    res["arrival_rate_low_hz"] = rate_low
    res["load"] = load
    return res


def main(argv=None):
    """python halide-sar-app_amd/serving.py --config C4|C5 [--jobs N] [--load F]: one JSON line."""
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="C4", choices=["C4", "C5"])
    ap.add_argument("--jobs", type=int, default=0)
    ap.add_argument("--load", type=float, default=0.7)
    ap.add_argument("--streams", type=int, default=2, help="large-job streams of the engine")
    ap.add_argument("--fuse-max", type=float, default=2 ** 31,
                    help="largest job (pixel-pulses) the engine fuses")
    ap.add_argument("--band-order", type=int, default=0, choices=[0, 1],
                    help="larger jobs sharing a stream: 0 least work left first, 1 interleaved")
    ap.add_argument("--hw-queues", type=int, default=ENGINE_HW_QUEUES,
                    help="GPU_MAX_HW_QUEUES for this process (HIP's default is 4)")
    a = ap.parse_args(argv)
    specs = c4_specs() if a.config == "C4" else c5_specs()
    if a.jobs:
        specs = specs[:a.jobs]
    res = benchmark(specs, load=a.load, log=lambda m: print(m, file=sys.stderr, flush=True),
                    num_streams=a.streams, fuse_max_pixel_pulses=a.fuse_max,
                    band_order=a.band_order)
    print(json.dumps(dict(config=a.config, jobs=len(specs), **res)), flush=True)


if __name__ == "__main__":
    main()
