"""Seeded synthetic SAR scenes (phase histories + platform data) in the reference's buffer layouts.

The reference ships no simulator; scenes here follow the pipeline's own conventions so that its
FFT-forward + expj(-k_c*dr) backprojection focuses them:

    freq_k = f_c + (k - ns/2) * df,   k_r = 4*pi*freq/c  (f32, as PlatformData loads it)
    B_IF   = (ns - 1) * df,           delta_r = c / (2 * B_IF)
    phs[p, k] = sum_s a_s * exp(+j * k_r[k] * (|pos_p| - |pos_p - x_s|)) + noise

Layouts match PlatformData (PlatformData.cpp:215-252): phs complex64 (np, ns) = Halide {2, ns, np},
pos float32 (np, 3) = Halide {3, np}, k_r float32 (ns,), R_c = pos[np // 2].

Scene parameters (scatterers, path, noise) come from numpy's seeded generator; the phase-history
sum itself is evaluated here in numpy for small scenes, or by the HIP simulator kernel
(``sarbp_sim_phs``) for large ones.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

C_LIGHT = 299792458.0


@dataclasses.dataclass
class Scene:
    name: str
    npulses: int
    nsamples: int
    f_c: float
    df: float
    pos: np.ndarray          # (np, 3) float32
    k_r: np.ndarray          # (ns,) float32
    delta_r: float
    R_c: np.ndarray          # (3,) float32
    targets: np.ndarray      # (nt, 3) float64 scatterer positions
    amps: np.ndarray         # (nt,) complex128 amplitudes
    taylor: int
    upsample: int
    db_min: float
    db_max: float
    noise_snr_db: float | None
    seed: int
    phs: np.ndarray | None = None   # (np, ns) complex64
    image: tuple | None = None      # explicit (nu, nv, d_u, d_v) image grid, else img_plane rules

    @property
    def k_r64(self) -> np.ndarray:
        freq = self.f_c + (np.arange(self.nsamples) - self.nsamples // 2) * self.df
        return 4.0 * math.pi * freq / C_LIGHT


def _base(name, npulses, nsamples, f_c, bw_or_df, pos, targets, amps, taylor, upsample, db,
          snr, seed, is_bw):
    df = bw_or_df / (nsamples - 1) if is_bw else bw_or_df
    freq = f_c + (np.arange(nsamples) - nsamples // 2) * df
    k_r = (4.0 * math.pi * freq / C_LIGHT).astype(np.float32)
    b_if = (nsamples - 1) * df
    delta_r = C_LIGHT / (2.0 * b_if)
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    return Scene(name=name, npulses=npulses, nsamples=nsamples, f_c=f_c, df=df, pos=pos, k_r=k_r,
                 delta_r=delta_r, R_c=pos[npulses // 2].copy(), targets=targets, amps=amps,
                 taylor=taylor, upsample=upsample, db_min=db[0], db_max=db[1],
                 noise_snr_db=snr, seed=seed)


def circular_arc_scene(npulses=469, nsamples=424, f_c=9.6e9, df=1.5e6, radius=10e3,
                       altitude=8e3, arc_deg=4.0, ntargets=20, patch=50.0, snr_db=20.0,
                       taylor=17, upsample=2, seed=1001, name="C1") -> Scene:
    """Circular-arc platform over `arc_deg` of azimuth; unit-amplitude point scatterers with
    uniform phase in a `patch` x `patch` m square; complex AWGN at `snr_db`."""
    rng = np.random.default_rng(seed)
    th = np.deg2rad(np.linspace(-arc_deg / 2, arc_deg / 2, npulses))
    pos = np.stack([radius * np.cos(th), radius * np.sin(th), np.full(npulses, altitude)], 1)
    xy = rng.uniform(-patch / 2, patch / 2, size=(ntargets, 2))
    targets = np.concatenate([xy, np.zeros((ntargets, 1))], 1)
    amps = np.exp(1j * rng.uniform(0, 2 * math.pi, ntargets))
    sc = _base(name, npulses, nsamples, f_c, df, pos, targets, amps, taylor, upsample,
               (-30.0, 0.0), snr_db, seed, is_bw=False)
    return sc


def linear_spotlight_scene(npulses=2000, nsamples=2048, f_c=10e9, bw=1e9, standoff=15e3,
                           altitude=5e3, ntargets=200, nclutter=10_000, clutter_sigma=0.01,
                           taylor=30, upsample=2, seed=1002, extent=None, name="C2",
                           db=(-45.0, 0.0)) -> Scene:
    """Linear track x in [-L/2, L/2] at y = -standoff, z = altitude, with L chosen for square
    resolution (cross-range = lambda R / (2 L) = delta_r). Rayleigh(1) point scatterers with uniform
    phase plus `nclutter` Rayleigh(`clutter_sigma`) clutter points over the image footprint."""
    rng = np.random.default_rng(seed)
    df = bw / (nsamples - 1)
    delta_r = C_LIGHT / (2.0 * bw)
    lam = C_LIGHT / f_c
    slant = math.hypot(standoff, altitude)
    L = lam * slant / (2.0 * delta_r)
    xs = np.linspace(-L / 2, L / 2, npulses)
    pos = np.stack([xs, np.full(npulses, -standoff), np.full(npulses, altitude)], 1)
    if extent is None:
        extent = 0.4 * delta_r * nsamples  # half-width of the populated footprint
    pts = rng.uniform(-extent, extent, size=(ntargets, 2))
    amps = rng.rayleigh(1.0, ntargets) * np.exp(1j * rng.uniform(0, 2 * math.pi, ntargets))
    cl = rng.uniform(-extent, extent, size=(nclutter, 2))
    camps = rng.rayleigh(clutter_sigma, nclutter) * np.exp(1j * rng.uniform(0, 2 * math.pi,
                                                                            nclutter))
    targets = np.concatenate([np.concatenate([pts, cl], 0),
                              np.zeros((ntargets + nclutter, 1))], 1)
    amps = np.concatenate([amps, camps])
    sc = _base(name, npulses, nsamples, f_c, bw, pos, targets, amps, taylor, upsample, db, None,
               seed, is_bw=True)
    return sc


def config_scene(name: str) -> Scene:
    """The named benchmark configurations C1, C2 and C3, and P8192_S8192: the sizes of the
    reference's published distributed-CUDA run (exp/summit/darpa-demo2/logs/SARBP-Sandia-4.763156:
    182-197: 8192 pulses x 8192 samples, an 8192 x 8192 image, Taylor 30, upsample 2, so N_fft
    32768) on a synthetic scene of the C2/C3 family."""
    if name == "C1":
        return circular_arc_scene()
    if name == "C2":
        return linear_spotlight_scene()
    if name in ("C3", "P8192_S8192"):
        c3 = name == "C3"
        sc = linear_spotlight_scene(npulses=8000 if c3 else 8192, nsamples=4096 if c3 else 8192,
                                    taylor=30, seed=1003 if c3 else 1006, name=name)
        du = sc.delta_r * sc.nsamples / 8192
        sc.image = (8192, 8192, du, du)
        return sc
    raise KeyError(name)


def simulate_phs_numpy(sc: Scene, chunk: int = 64) -> np.ndarray:
    """Evaluate the phase-history sum on the CPU (small scenes) and add the seeded noise."""
    pos = sc.pos.astype(np.float64)
    k = sc.k_r64
    out = np.zeros((sc.npulses, sc.nsamples), np.complex128)
    r0 = np.linalg.norm(pos, axis=1)
    for p0 in range(0, sc.npulses, chunk):
        p1 = min(sc.npulses, p0 + chunk)
        d = r0[p0:p1, None] - np.linalg.norm(pos[p0:p1, None, :] - sc.targets[None, :, :], axis=2)
        out[p0:p1] = np.exp(1j * k[None, :, None] * d[:, None, :]) @ sc.amps
    return add_noise(sc, out)


def add_noise(sc: Scene, clean: np.ndarray) -> np.ndarray:
    out = clean.astype(np.complex128)
    if sc.noise_snr_db is not None:
        rng = np.random.default_rng(sc.seed + 7919)
        p_sig = float(np.mean(np.abs(out) ** 2))
        sigma = math.sqrt(p_sig / (10 ** (sc.noise_snr_db / 10)) / 2)
        out = out + sigma * (rng.standard_normal(out.shape) + 1j * rng.standard_normal(out.shape))
    return out.astype(np.complex64)


def tiny_scene(npulses=16, nsamples=24, seed=7, taylor=17, ntargets=3, name="tiny") -> Scene:
    """Small seeded scene for golden fixtures and CPU tests."""
    sc = circular_arc_scene(npulses=npulses, nsamples=nsamples, f_c=9.6e9, df=15e6,
                            radius=2e3, altitude=1.5e3, arc_deg=6.0, ntargets=ntargets,
                            patch=8.0, snr_db=20.0, taylor=taylor, upsample=2, seed=seed,
                            name=name)
    sc.phs = simulate_phs_numpy(sc)
    return sc
