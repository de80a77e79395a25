"""Device-resident jobs: a scene's inputs held in HBM (torch tensors as plain device memory) and a
plan that runs the backprojection pipeline on them through the C ABI (sarbp_plan_execute).

torch is used only to own device memory and streams; every kernel is the library's HIP code.
"""
from __future__ import annotations

import importlib
import math

import numpy as np

_pkg = importlib.import_module("halide-sar-app_amd")


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("no GPU visible: the HIP backprojection path cannot run")
    return torch


class DeviceScene:
    """Phase history, k_r, platform positions and pixel locations resident on the GPU.

    A grid without host pixel_locs runs in grid mode by default: the accumulation generates the
    pixel locations itself and r stays None (grid_mode=False keeps an explicit r, generated on the
    device from the grid)."""

    def __init__(self, sc, grid: dict, device: int = 0, phs_host: np.ndarray | None = None,
                 simulate: bool = True, grid_mode: bool | None = None):
        torch = _torch()
        self.sc = sc
        self.grid = grid
        self.dev = torch.device("cuda", device)
        self.nu, self.nv = grid["nu"], grid["nv"]
        self.n_fft = _pkg.nfft(sc.nsamples, sc.upsample)
        self.k_r = torch.from_numpy(np.ascontiguousarray(sc.k_r, np.float32)).to(self.dev)
        self.pos = torch.from_numpy(np.ascontiguousarray(sc.pos, np.float32)).to(self.dev)
        self.grid_mode = grid.get("pixel_locs") is None if grid_mode is None else grid_mode
        if grid.get("pixel_locs") is not None:
            self.r = torch.from_numpy(np.ascontiguousarray(grid["pixel_locs"], np.float64)).to(self.dev)
        elif self.grid_mode:
            self.r = None
        else:
            u = torch.from_numpy(np.ascontiguousarray(grid["u"], np.float64)).to(self.dev)
            v = torch.from_numpy(np.ascontiguousarray(grid["v"], np.float64)).to(self.dev)
            self.r = torch.empty((3, self.nu * self.nv), dtype=torch.float64, device=self.dev)
            _pkg.pixel_locs_device(u, v, grid["u_hat"], grid["v_hat"], self.r)
        if phs_host is not None:
            self.phs = torch.from_numpy(np.ascontiguousarray(phs_host, np.complex64)).to(self.dev)
        elif sc.phs is not None:
            self.phs = torch.from_numpy(np.ascontiguousarray(sc.phs, np.complex64)).to(self.dev)
        elif simulate:
            self.phs = simulate_device(sc, self.dev)
        else:
            self.phs = torch.zeros((sc.npulses, sc.nsamples), dtype=torch.complex64,
                                   device=self.dev)

    def plan(self, region=None, **kw) -> "_pkg.Plan":
        sc = self.sc
        p = _pkg.Plan(sc.nsamples, sc.npulses, self.n_fft, sc.taylor, sc.delta_r, self.nu,
                      self.nv, region=region, **kw)
        if self.grid_mode:
            g = self.grid
            p.set_grid(g["d_u"], g["d_v"], g["u_hat"], g["v_hat"])
        return p

    def output(self, region=None):
        torch = _torch()
        w, h = (self.nu, self.nv) if region is None else (region[2], region[3])
        return torch.empty((h, w), dtype=torch.complex128, device=self.dev)

    def run(self, plan, out, stream=None):
        plan.execute(self.phs, self.k_r, self.pos, self.r, out, stream=stream)
        return out


def simulate_device(sc, dev):
    """Phase history of the scene's scatterers evaluated by the HIP simulator, plus the scene's
    seeded noise added on the host when it has a noise level."""
    torch = _torch()
    pos64 = torch.from_numpy(sc.pos.astype(np.float64)).to(dev)
    k64 = torch.from_numpy(sc.k_r64).to(dev)
    tg = torch.from_numpy(np.ascontiguousarray(sc.targets, np.float64)).to(dev)
    am = torch.from_numpy(np.ascontiguousarray(sc.amps, np.complex128)).to(dev)
    phs = torch.empty((sc.npulses, sc.nsamples), dtype=torch.complex64, device=dev)
    _pkg.sim_phs_device(pos64, k64, tg, am, phs)
    if sc.noise_snr_db is not None:
        import importlib as _il
        scenes = _il.import_module("halide-sar-app_amd.scenes")
        host = scenes.add_noise(sc, phs.cpu().numpy())
        phs.copy_(torch.from_numpy(host))
    torch.cuda.synchronize(dev)
    return phs


def scene_grid(sc, with_pixel_locs: bool = True) -> dict:
    """The scene's image grid: an explicit grid if the scene names one, else img_plane_create.
    Without host pixel_locs, DeviceScene generates them on the device."""
    ip = _pkg.img_plane_create(sc.nsamples, sc.npulses, sc.delta_r, sc.R_c,
                               with_pixel_locs=with_pixel_locs and getattr(sc, "image", None) is None)
    if getattr(sc, "image", None) is not None:
        nu, nv, du, dv = sc.image
        return _pkg.image_grid(nu, nv, du, dv, ip["u_hat"], ip["v_hat"],
                               with_pixel_locs=with_pixel_locs)
    return ip


def pixel_pulses(sc, grid) -> float:
    return float(grid["nu"]) * float(grid["nv"]) * float(sc.npulses)


def algorithmic_flops_per_pp() -> int:
    """Algorithmic work per pixel-pulse (sqrt, sin, cos counted as one each): r - r0 (3),
    |.|^2 (5), sqrt (1), dr (1), index (3), complex lerp (6), phase (1), sin + cos (2),
    complex multiply (6), accumulate (2)."""
    return 30


def fft_bytes(sc, n_fft: int, q_bytes: int = 8) -> int:
    """Algorithmic HBM bytes of the range-profile stage: read phs (8 B/sample), write Q."""
    return 8 * sc.npulses * sc.nsamples + q_bytes * sc.npulses * n_fft


def log2i(n: int) -> int:
    return int(round(math.log2(n)))
