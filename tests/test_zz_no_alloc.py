"""sarbp_img_output_device_noalloc (the allocation-free image output) against the allocating
sarbp_img_output_device on the same images, on two calls in a row, the second smaller: a large image
1000 times brighter (1024 partial maxima) and then the C1 golden 64 x 64 image (16 partial maxima),
so a partial left over from the first call would raise the second call's maximum. The outputs must
be the same bits: both reduce the same |z|^2 values and take one correctly rounded sqrt."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _both(sar, torch, img, lo, hi):
    npix = img.numel()
    out = []
    for fn in (sar.img_output_device_noalloc, sar.img_output_device):
        dB = torch.full((npix,), -7.0, dtype=torch.float64, device="cuda")
        u8 = torch.full((npix,), 7, dtype=torch.uint8, device="cuda")
        fn(img, npix, dB, u8, lo, hi)
        out.append((dB, u8))
    torch.cuda.synchronize()
    return out


def test_noalloc_matches_allocating_entry_point(sar, orc):
    import torch
    assert torch.cuda.is_available()
    rng = np.random.default_rng(2024)
    g = np.load(os.path.join(GOLDEN, "c1_64.npz"))
    small = np.ascontiguousarray(g["img"])
    big = rng.standard_normal((1000, 1000)) + 1j * rng.standard_normal((1000, 1000))
    big *= 1000.0 * np.abs(small).max() / np.abs(big).max()
    assert small.size < big.size and np.abs(small).max() < np.abs(big).max()
    for name, host in (("first, 1000 x 1000", big), ("second, 64 x 64", small)):
        print(f"no-alloc image output, {name}", flush=True)
        img = torch.from_numpy(np.ascontiguousarray(host)).cuda()
        (d0, u0), (d1, u1) = _both(sar, torch, img, -30.0, 0.0)
        assert torch.equal(d0, d1), name
        assert torch.equal(u0, u1), name
        ref = orc.img_output_to_dB(host)
        assert np.abs(d0.cpu().numpy().reshape(host.shape) - ref).max() < 1e-9, name
    # odd pixel count (a tail of 1-3 pixels for the 4-pixel dB kernel) and unaligned outputs
    img = torch.from_numpy(np.ascontiguousarray(small.ravel()[:1001])).cuda()
    npix = img.numel()
    dB = torch.zeros(npix + 1, dtype=torch.float64, device="cuda")[1:]
    u8 = torch.zeros(npix + 1, dtype=torch.uint8, device="cuda")[1:]
    sar.img_output_device_noalloc(img, npix, dB, u8, -30.0, 0.0)
    dB2 = torch.zeros(npix, dtype=torch.float64, device="cuda")
    u82 = torch.zeros(npix, dtype=torch.uint8, device="cuda")
    sar.img_output_device(img, npix, dB2, u82, -30.0, 0.0)
    torch.cuda.synchronize()
    assert torch.equal(dB, dB2) and torch.equal(u8, u82)
