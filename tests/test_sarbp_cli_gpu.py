"""The `sarbp` driver end to end on the GPU: a synthetic scene written as a platform directory, the
driver's PNG compared with the FP64 oracle's u8 image of the same inputs (image plane, N_fft rule and
dB/u8 mapping as sarbp.cpp:248-324)."""
import importlib
import os
import subprocess

import numpy as np
import pytest

from test_host_cpu import _read_png

pytestmark = pytest.mark.gpu


def _platform():
    return importlib.import_module("halide-sar-app_amd.platform_data")


def _expected_u8(orc, pd, taylor, upsample, res, aspect, lo, hi, n_hat=(0.0, 0.0, 1.0)):
    ip = orc.img_plane(pd["nsamples"], pd["npulses"], pd["delta_r"], pd["R_c"], n_hat, res,
                       aspect, upsample != 0)
    n_fft = _platform().nfft_rule(pd["nsamples"], upsample)
    img = orc.backprojection(pd["phs"], pd["k_r"], taylor, n_fft, pd["delta_r"], ip["nu"],
                             ip["nv"], pd["pos"], ip["pixel_locs"], nthreads=8)
    return orc.img_output_u8(orc.img_output_to_dB(img), lo, hi), ip


@pytest.mark.parametrize("npulses,nsamples,taylor,upsample,extra", [
    (64, 96, 17, 2, []),
    (45, 80, 19, 3, ["-s", "cuda", "-n", "2", "-r", "0.75", "-a", "1.25"]),
])
def test_sarbp_driver_matches_oracle(tmp_path, scenes, orc, npulses, nsamples, taylor, upsample,
                                     extra):
    sc = scenes.circular_arc_scene(npulses=npulses, nsamples=nsamples, ntargets=6, taylor=taylor,
                                   seed=2024 + npulses)
    phs = scenes.simulate_phs_numpy(sc)
    d = str(tmp_path / "platform")
    _platform().write_dir(d, sc, phs)
    out = str(tmp_path / "img.png")
    args = [_platform().SARBP_PATH, "-p", d, "-o", out, "-d", "-30", "-D", "0", "-t", str(taylor),
            "-u", str(upsample), *extra]
    print("running", " ".join(args), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    got = _read_png(out)
    opts = dict(zip(extra[::2], extra[1::2]))
    pd = _platform().load(d)
    want, ip = _expected_u8(orc, pd, taylor, upsample, float(opts.get("-r", 1.0)),
                            float(opts.get("-a", 1.0)), -30.0, 0.0)
    assert got.shape == (ip["nv"], ip["nu"]) == want.shape
    diff = np.abs(got.astype(int) - want.astype(int))
    # the image matches the oracle to ~1e-6 relative; a u8 value can only move by one step where
    # the dB value sits on a quantisation boundary
    assert diff.max() <= 1 and np.count_nonzero(diff) <= max(2, diff.size // 1000), \
        (diff.max(), np.count_nonzero(diff))
    assert f"X length: {ip['nu']}" in r.stdout and f"Y length: {ip['nv']}" in r.stdout
