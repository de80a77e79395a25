"""Host-side checks of the named benchmark scenes: sizes, FFT lengths and image grids."""
import importlib

import numpy as np
import pytest


@pytest.mark.parametrize("name,np_,ns,n_fft,side", [
    ("C1", 469, 424, 1024, None),
    ("C2", 2000, 2048, 8192, None),
    ("C3", 8000, 4096, 16384, 8192),
    # the sizes of the reference's published GPU run (SARBP-Sandia-4.763156:182-197)
    ("P8192_S8192", 8192, 8192, 32768, 8192),
])
def test_named_scene_sizes(sar, scenes, name, np_, ns, n_fft, side):
    sc = scenes.config_scene(name)
    assert (sc.npulses, sc.nsamples) == (np_, ns)
    assert sc.pos.shape == (np_, 3) and sc.pos.dtype == np.float32
    assert sc.k_r.shape == (ns,) and sc.k_r.dtype == np.float32
    assert sar.nfft(ns, sc.upsample) == n_fft
    if side is not None:
        assert sc.image[:2] == (side, side)
        pipe = importlib.import_module("halide-sar-app_amd.pipeline")
        grid = pipe.scene_grid(sc, with_pixel_locs=False)
        assert (grid["nu"], grid["nv"]) == (side, side)
        assert grid.get("pixel_locs") is None


def test_published_case_scene_is_seeded_apart_from_c3(scenes):
    a, b = scenes.config_scene("P8192_S8192"), scenes.config_scene("P8192_S8192")
    c3 = scenes.config_scene("C3")
    assert a.seed == b.seed and np.array_equal(a.targets, b.targets)
    assert a.seed != c3.seed and a.taylor == c3.taylor == 30
