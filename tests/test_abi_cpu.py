"""CPU-side checks of the C-ABI library: it loads, exports every symbol include/*.h declares, and
its host logic (argument validation, bounds queries, Taylor windows, image plane) behaves like the
reference's generated code — all without launching GPU work."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for hdr in ("sarbp_hip.h", "halide_buffer_compat.h"):
        text = open(os.path.join(ROOT, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^(?:const\s+)?(?:int|void|char)\s*\**\s*(\w+)\s*\(", text, re.M):
            names.append(m.group(1))
    return names


def test_library_exports_every_declared_symbol(sar):
    names = declared_functions()
    assert "backprojection" in names and "img_output_u8" in names and "dft_init_fftw" in names
    out = subprocess.run(["nm", "-D", "--defined-only", sar.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    lib = ctypes.CDLL(sar.LIB_PATH)
    for n in names:
        assert getattr(lib, n) is not None


def test_reference_entry_points_have_c_linkage(sar):
    out = subprocess.run(["nm", "-D", "--defined-only", sar.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    for n in ("backprojection", "backprojection_cuda", "img_output_to_dB", "img_output_u8",
              "dft_init_fftw", "dft_destroy_fftw"):
        assert re.search(rf"\bT {n}$", out, re.M), n


def test_struct_layout_matches_halide_runtime(sar):
    # halide_buffer_t: device(8) interface(8) host(8) flags(8) type(4) dimensions(4) dim(8) padding(8)
    assert ctypes.sizeof(sar.halide_buffer_t) == 56
    assert sar.halide_buffer_t.type.offset == 32
    assert sar.halide_buffer_t.dim.offset == 40
    assert ctypes.sizeof(sar.halide_dimension_t) == 16


def _valid_args(ns=24, np_=16, nu=32, nv=16):
    phs = np.zeros((np_, ns), np.complex64)
    k_r = np.ones(ns, np.float32)
    u = np.zeros(nu)
    v = np.zeros(nv)
    pos = np.ones((np_, 3), np.float32)
    r = np.zeros((3, nu * nv))
    out = np.zeros((nv, nu), np.complex128)
    return phs, k_r, u, v, pos, r, out


def test_validation_error_codes(sar):
    phs, k_r, u, v, pos, r, out = _valid_args()
    E = sar.ERR
    # wrong type: phs as complex128 (float64 parts)
    rv = sar.backprojection(phs.astype(np.complex128), k_r, 17, 64, 0.2, u, v, pos, r, out)
    assert rv == E["bad_type"]
    # wrong dimensionality: k_r 2-D
    rv = sar.backprojection(phs, k_r[None, :], 17, 64, 0.2, u, v, pos, r, out)
    assert rv == E["bad_dimensions"]
    # u float32
    rv = sar.backprojection(phs, k_r, 17, 64, 0.2, u.astype(np.float32), v, pos, r, out)
    assert rv == E["bad_type"]
    # N_fft: powers of two up to 2^17 and other lengths up to 8192 are taken (dft.cpp:20-30 plans
    # FFTW for any N); below 2, other lengths above 8192 and powers of two above 2^17 are not
    assert sar.backprojection(phs, k_r, 17, 1, 0.2, u, v, pos, r, out) == E["param_too_small"]
    assert sar.backprojection(phs, k_r, 17, 8193, 0.2, u, v, pos, r, out) == E["param_too_large"]
    assert sar.backprojection(phs, k_r, 17, 1 << 18, 0.2, u, v, pos, r, out) == E["param_too_large"]
    # k_r shorter than nsamples: access out of bounds
    assert sar.backprojection(phs, k_r[:10], 17, 64, 0.2, u, v, pos, r, out) == \
        E["access_out_of_bounds"]
    # NULL buffer pointer
    lib = sar.lib()
    b = [sar.Buffer(a) for a in (phs, k_r, u, v, pos, r, out)]
    rv = lib.backprojection(b[0].ref, None, 17, 64, 0.2, b[2].ref, b[3].ref, b[4].ref, b[5].ref,
                            b[6].ref)
    assert rv == E["buffer_argument_is_null"]
    # output with dimension 0 outside [0, 2)
    ob = sar.Buffer(np.zeros((16, 32, 3)), halide_dims=[(0, 3, 1), (0, 32, 3), (0, 16, 96)])
    assert sar.backprojection(phs, k_r, 17, 64, 0.2, u, v, pos, r, ob) == E["access_out_of_bounds"]
    # img_output_* validation
    assert sar.img_output_to_dB(out, np.zeros((16, 32), np.float32)) == E["bad_type"]
    assert sar.img_output_u8(np.zeros((16, 32)), -30, 0, np.zeros((16, 32), np.int8)) == E["bad_type"]
    assert "float" in sar.last_error() or "uint8" in sar.last_error()


def test_bounds_query_fills_required_regions(sar):
    phs, k_r, u, v, pos, r, out = _valid_args()
    F = sar.HALIDE_FLOAT
    kq = sar.Buffer.bounds_query([(0, 0, 0)], F, 32)
    pq = sar.Buffer.bounds_query([(0, 0, 0), (0, 0, 0)], F, 32)
    rq = sar.Buffer.bounds_query([(0, 0, 0), (0, 0, 0)], F, 64)
    rv = sar.backprojection(phs, kq, 17, 64, 0.2, u, v, pq, rq, out)
    assert rv == 0
    assert sar.backprojection(phs, kq, 17, 100, 0.2, u, v, pq, rq, out) == 0  # any length N_fft
    assert kq.dims() == [(0, 24, 1)]
    assert pq.dims() == [(0, 3, 1), (0, 16, 3)]
    assert rq.dims() == [(0, 32 * 16, 1), (0, 3, 32 * 16)]
    assert not out.any()  # a bounds query computes nothing


def test_no_gpu_is_reported_not_aborted(sar):
    if sar.device_count() > 0:
        pytest.skip("a GPU is visible")
    phs, k_r, u, v, pos, r, out = _valid_args()
    rv = sar.backprojection(phs, k_r, 17, 64, 0.2, u, v, pos, r, out)
    assert rv == sar.ERR["no_device_interface"]
    assert "no HIP device" in sar.last_error()
    # a length that is not a power of two passes validation and reaches the device the same way
    rv = sar.backprojection(phs, k_r, 17, 100, 0.2, u, v, pos, r, out)
    assert rv == sar.ERR["no_device_interface"]


def test_policy_defaults_and_validation(sar):
    assert sar.get_default_policies() == (sar.NORM_R0_F32_STRICT, sar.TAYLOR_DEN_INT)
    assert sar.set_default_policies(5, 0) == sar.ERR["constraint_violated"]
    assert sar.set_default_policies(1, 1) == 0
    assert sar.get_default_policies() == (1, 1)
    assert sar.set_default_policies(0, 0) == 0


@pytest.mark.parametrize("s_l", [17, 20, 25, 30, 35])
@pytest.mark.parametrize("den", [0, 1])
@pytest.mark.parametrize("num", [2, 12, 469, 2048])
def test_host_taylor_matches_oracle(sar, orc, s_l, den, num):
    a = sar.taylor_window(num, s_l, den)
    b = orc.taylor(num, s_l, int_den=(den == 0))
    assert np.array_equal(a, b, equal_nan=True)


def test_host_image_plane_matches_oracle(sar, orc):
    f = np.load(os.path.join(ROOT, "tests", "golden", "afrl_image_plane.npz"))
    a = sar.img_plane_create(int(f["nsamples"]), int(f["npulses"]), float(f["delta_r"]), f["R_c"])
    b = orc.img_plane(int(f["nsamples"]), int(f["npulses"]), float(f["delta_r"]), f["R_c"])
    for k in ("nu", "nv", "d_u", "d_v"):
        assert a[k] == b[k]
    for k in ("u", "v", "k_u", "k_v", "u_hat", "v_hat", "pixel_locs"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("u", "v", "k_u", "k_v"):
        assert np.array_equal(a[k], f[k])


def test_nfft_rule_matches_oracle(sar, orc):
    for ns in (2, 3, 24, 424, 1000, 1024, 1800, 2048, 4096):
        for up in (1, 2, 3, 6):
            assert sar.nfft(ns, up) == orc.nfft(ns, up)
