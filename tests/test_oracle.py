"""Pins the CPU FP64 oracle (oracle/) before it is trusted as the checker.

Each function of the oracle is checked against something it was not derived from:
  - known answers held by the reference's own unit tests (test/taylor_test.cpp:14-17,
    test/linspace_test.cpp:11-21, test/norm_test.cpp:15-46, test/cross_test.cpp:14-29,
    test/interp_test.cpp:25);
  - RITSAR image-plane fixtures shipped with the reference (data/AFRL/pass1/HH_ip_npy), copied as
    data into tests/golden/afrl_image_plane.npz;
  - library routines (numpy's fftshift, pocketfft, np.interp) in tests/twin.py;
  - closed forms and invariants: textbook Taylor weights, direct DFT sums, a point target at a
    known location, linearity, the even-npulses quirk identity.
Parity that none of these can pin (the Int(32)-vs-f32 reading of the Taylor F_m_den ratio for
Taylor levels >= 20 dB, and FMA contraction of the f32 ||pos||) is exposed as explicit policies.
"""
import math
import os

import numpy as np
import pytest

import twin
from conftest import rel_l2

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ------------------------------------------------------------------ helper known answers

def test_taylor_known_answer(orc):
    # test/taylor_test.cpp:14-17,27 — 12 points, 17 dB, tolerance 1e-4
    ref = np.array([0.632594, 0.662359, 0.742204, 0.846779, 0.942881, 1,
                    1, 0.942881, 0.846779, 0.742204, 0.662359, 0.632594])
    for int_den in (True, False):
        w = orc.taylor(12, 17, int_den)
        assert np.abs(w - ref).max() < 1e-6


@pytest.mark.parametrize("s_l", [17, 20, 25, 30, 33, 35, 40])
@pytest.mark.parametrize("int_den", [False, True])
@pytest.mark.parametrize("num", [12, 64, 469])
def test_taylor_textbook(orc, s_l, int_den, num):
    """The reference's F_m multiplies the sign (-1)^(m+1) in at every one of the n_bar - 1 update
    steps (signal.h:39) and omits the 1/2 of the textbook form, which the '1 +' sum absorbs as a
    factor 2 (signal.h:45): F_ref = 2 F_textbook (-1)^((m+1)(n_bar-2))."""
    A = float(np.arccosh(np.float32(10.0) ** np.float32(s_l / 20))) / math.pi
    nbar = int(2 * A * A + 0.5) + 1
    w = orc.taylor(num, s_l, int_den)
    with np.errstate(all="ignore"):
        ref, F = twin.taylor_textbook(num, s_l, nbar, int_den)
        xi = np.linspace(-0.5, 0.5, num)
        wr = 1 + 2 * sum(F[m - 1] * (-1) ** ((m + 1) * (nbar - 2)) * np.cos(2 * math.pi * m * xi)
                         for m in range(1, nbar))
        wr = wr / wr.max()
    if not np.all(np.isfinite(wr)):
        # Int(32) reading: floor(m^2/n^2) == 1 for some m != n makes a zero denominator (n_bar >= 5)
        assert int_den and nbar >= 5
        assert not np.all(np.isfinite(w))
        return
    assert np.abs(w - wr).max() < 1e-6


def test_taylor_policies_differ_only_from_20dB(orc):
    assert np.array_equal(orc.taylor(100, 17, True), orc.taylor(100, 17, False))
    assert np.abs(orc.taylor(100, 20, True) - orc.taylor(100, 20, False)).max() > 1e-2


def test_linspace_known_answer(orc):
    # test/linspace_test.cpp:11-21: 10 points on [-0.5, 0.5], last forced to stop
    out = orc.linspace(-0.5, 0.5, 10)
    ref = -0.5 + np.arange(10) / 9.0
    ref[-1] = 0.5
    assert np.abs(out - ref).max() < 1e-15
    assert out[-1] == 0.5 and out[0] == -0.5


def test_norm_known_answer(orc):
    # test/norm_test.cpp:15-46: ||(3,4)|| = 5, rows (3,4),(5,12) -> [5, 13]
    pos = np.array([[3, 4, 0], [5, 12, 0], [0, 3, 4]], np.float32)
    for pol in (orc.NORM_R0_F32_STRICT, orc.NORM_R0_F32_FMA, orc.NORM_R0_F64):
        n32, n64 = orc.norm_r0(pos, pol)
        assert list(n32) == [5.0, 13.0, 5.0]
        assert list(n64) == [5.0, 13.0, 5.0]


def test_norm_r0_policies_f32(orc):
    rng = np.random.default_rng(3)
    pos = (rng.standard_normal((4000, 3)) * 8000).astype(np.float32)
    s32, _ = orc.norm_r0(pos, orc.NORM_R0_F32_STRICT)
    f32, _ = orc.norm_r0(pos, orc.NORM_R0_F32_FMA)
    _, d64 = orc.norm_r0(pos, orc.NORM_R0_F64)
    p = pos.astype(np.float32)
    strict = np.sqrt(((p[:, 0] * p[:, 0]) + (p[:, 1] * p[:, 1])) + (p[:, 2] * p[:, 2]))
    assert np.array_equal(s32, strict.astype(np.float32))
    exact = np.sqrt((pos.astype(np.float64) ** 2).sum(1))
    assert np.abs(d64 - exact).max() <= 1e-12 * exact.max()
    assert np.abs(s32.astype(np.float64) - exact).max() < 2 ** -23 * exact.max() * 2
    assert np.count_nonzero(s32 != f32) > 0  # contraction changes some last bits


def test_cross_known_answer(orc):
    # test/cross_test.cpp:14-29: (2,3,4) x (5,6,7) = (-3, 6, -3); v_hat normalises it
    ip = orc.img_plane(8, 8, 1.0, R_c=[5, 6, 7], n_hat=[2, 3, 4], pixel_locs=False)
    c = np.array([-3.0, 6.0, -3.0])
    assert np.abs(ip["v_hat"] - c / np.linalg.norm(c)).max() < 1e-7


def test_nfft_rule(orc):
    # sarbp.cpp:258: an exact power of two doubles
    assert orc.nfft(424, 2) == 1024
    assert orc.nfft(2048, 2) == 8192
    assert orc.nfft(4096, 2) == 16384
    assert orc.nfft(1800, 2) == 4096
    assert orc.nfft(1000, 1) == 1024


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 9, 24, 1024])
def test_fftshift_matches_numpy(orc, n):
    x = np.arange(n)
    ours = np.array([x[orc.fftshift_src(i, n)] for i in range(n)])
    assert np.array_equal(ours, np.fft.fftshift(x))


@pytest.mark.parametrize("n", [1, 2, 4, 8, 64, 1024, 16384, 3, 12, 24, 100])
def test_fft_matches_library(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = np.fft.fft(x)
    got = orc.fft(x)
    assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max()) * math.log2(max(n, 2))


def test_fft_closed_forms(orc):
    n = 64
    imp = np.zeros(n, complex)
    imp[5] = 1
    assert np.allclose(orc.fft(imp), np.exp(-2j * np.pi * 5 * np.arange(n) / n), atol=1e-14)
    tone = np.exp(2j * np.pi * 7 * np.arange(n) / n)
    out = orc.fft(tone)
    assert abs(out[7] - n) < 1e-12 and np.abs(np.delete(out, 7)).max() < 1e-12
    # direct sum for a tiny non-power-of-two size
    x = np.arange(6) + 1j
    direct = np.array([sum(x[j] * np.exp(-2j * np.pi * j * k / 6) for j in range(6)) for k in range(6)])
    assert np.abs(orc.fft(x) - direct).max() < 1e-12


# ------------------------------------------------------------------ image plane vs RITSAR fixtures

def test_image_plane_matches_ritsar_fixtures(orc):
    f = np.load(os.path.join(GOLDEN, "afrl_image_plane.npz"))
    ip = orc.img_plane(int(f["nsamples"]), int(f["npulses"]), float(f["delta_r"]), f["R_c"],
                       n_hat=[0, 0, 1], pixel_locs=True)
    assert ip["nu"] == 512 and ip["nv"] == 512
    assert ip["d_u"] == pytest.approx(float(f["du"]), rel=1e-15)
    for k in ("u", "v", "k_u", "k_v"):
        assert np.array_equal(ip[k], f[k]), k
    # RITSAR computes the cross product in f64; the reference's v_hat is f32 (img_plane.cpp:31-39)
    assert np.abs(ip["v_hat"] - f["v_hat"]).max() < 1e-7
    assert np.abs(ip["u_hat"] - f["u_hat"]).max() < 1e-7
    # pixel_locs = u_hat u[x % nu] + v_hat v[x / nu] (img_plane.cpp:77-113)
    pl = ip["pixel_locs"]
    x = np.arange(512 * 512)
    for d in range(3):
        ref = ip["u_hat"][d] * ip["u"][x % 512] + ip["v_hat"][d] * ip["v"][x // 512]
        assert np.abs(pl[d] - ref).max() < 1e-12


# ------------------------------------------------------------------ pipeline vs independent twin

def _tiny(scenes, np_, seed=5, ns=24, taylor=17):
    return scenes.tiny_scene(npulses=np_, nsamples=ns, seed=seed, taylor=taylor)


@pytest.mark.parametrize("np_", [15, 16])
@pytest.mark.parametrize("quirk", [True, False])
@pytest.mark.parametrize("n_fft", [64, 16, 60, 45, 21])
def test_range_profiles_vs_twin(orc, scenes, np_, quirk, n_fft):
    """Range profiles of the oracle against numpy (np.fft.fft takes any length; its fftshift rolls
    by n // 2, which is the reference's (x + ceil(n / 2)) mod n): powers of two, an even and an odd
    other length, and an odd length shorter than the 24 samples (negative pad)."""
    sc = _tiny(scenes, np_)
    ws = orc.taylor(sc.nsamples, sc.taylor)
    wp = orc.taylor(sc.npulses, sc.taylor)
    Q = orc.range_profiles(sc.phs, sc.k_r, sc.taylor, n_fft, pulse_shift_quirk=quirk)
    T = twin.range_profiles_twin(sc.phs, sc.k_r, ws, wp, n_fft, pulse_shift_quirk=quirk)
    assert np.abs(Q - T).max() <= 1e-12 * np.abs(T).max()


def test_pulse_quirk_is_identity_for_even_npulses(orc, scenes):
    sc = _tiny(scenes, 16)
    a = orc.range_profiles(sc.phs, sc.k_r, 17, 64, pulse_shift_quirk=True)
    b = orc.range_profiles(sc.phs, sc.k_r, 17, 64, pulse_shift_quirk=False)
    assert np.array_equal(a, b)
    sc = _tiny(scenes, 15)
    a = orc.range_profiles(sc.phs, sc.k_r, 17, 64, pulse_shift_quirk=True)
    b = orc.range_profiles(sc.phs, sc.k_r, 17, 64, pulse_shift_quirk=False)
    # odd npulses: profile p holds pulse (p + 1) mod np, windowed at its own index
    assert np.array_equal(a, np.roll(b, -1, axis=0))
    assert not np.array_equal(a, b)


@pytest.mark.parametrize("np_", [15, 16])
def test_backprojection_vs_twin(orc, scenes, np_):
    sc = _tiny(scenes, np_)
    ip = orc.img_plane(sc.nsamples, sc.npulses, sc.delta_r, sc.R_c)
    nu, nv = ip["nu"], ip["nv"]
    n_fft = orc.nfft(sc.nsamples, 2)
    img = orc.backprojection(sc.phs, sc.k_r, sc.taylor, n_fft, sc.delta_r, nu, nv, sc.pos,
                             ip["pixel_locs"])
    Q = orc.range_profiles(sc.phs, sc.k_r, sc.taylor, n_fft)
    _, nr0 = orc.norm_r0(sc.pos)
    T = twin.backproject_twin(Q, sc.nsamples, sc.delta_r, sc.k_r[sc.nsamples // 2], sc.pos, nr0,
                              ip["pixel_locs"], nu, nv)
    assert rel_l2(img, T) < 1e-12


def test_region_and_points_match_full(orc, scenes):
    sc = _tiny(scenes, 16)
    ip = orc.img_plane(sc.nsamples, sc.npulses, sc.delta_r, sc.R_c)
    nu, nv = ip["nu"], ip["nv"]
    Q = orc.range_profiles(sc.phs, sc.k_r, sc.taylor, 64)
    _, nr0 = orc.norm_r0(sc.pos)
    kc = sc.k_r[sc.nsamples // 2]
    full = orc.backproject(Q, sc.nsamples, sc.delta_r, kc, sc.pos, nr0, ip["pixel_locs"], nu, nv,
                           nthreads=3)
    sub = orc.backproject(Q, sc.nsamples, sc.delta_r, kc, sc.pos, nr0, ip["pixel_locs"], nu, nv,
                          region=(3, 5, 7, 6))
    assert np.array_equal(sub, full[5:11, 3:10])
    ys, xs = np.array([0, 5, 15, 12]), np.array([1, 31, 0, 9])
    pts = orc.backproject_points(Q, sc.nsamples, sc.delta_r, kc, sc.pos, nr0, ip["pixel_locs"],
                                 nu, nv, xs, ys, nthreads=2)
    assert np.array_equal(pts, full[ys, xs])


def test_linearity(orc, scenes):
    a = _tiny(scenes, 16, seed=1)
    b = _tiny(scenes, 16, seed=2)
    ip = orc.img_plane(a.nsamples, a.npulses, a.delta_r, a.R_c)
    args = lambda phs: (phs, a.k_r, 17, 64, a.delta_r, ip["nu"], ip["nv"], a.pos, ip["pixel_locs"])  # noqa: E731
    ia = orc.backprojection(*args(a.phs))
    ib = orc.backprojection(*args(b.phs))
    iab = orc.backprojection(*args((2 * a.phs - 3j * b.phs).astype(np.complex64)))
    assert rel_l2(iab, 2 * ia - 3j * ib) < 1e-6  # f32 rounding of the combined input


def test_point_target_focuses_at_its_pixel(orc, scenes):
    """A single scatterer at a pixel location: the image peak lands on that pixel with the
    scatterer's phase (the row flip puts v index nv-1-y on output row y, backprojection.cpp:130).
    The reference interpolates on linspace(floor(-ns dr/2), floor(ns dr/2), N)
    (backprojection.cpp:120); delta_r is chosen so that ns*dr/2 is an integer (the floors are
    exact) and df so that the FFT bin spacing c/(2 N df) equals that axis' step."""
    ns, npulses, n_fft = 96, 128, 256
    delta_r = 0.25                       # ns * delta_r / 2 = 12 exactly
    lsr = 24.0 / (n_fft - 1)
    df = scenes.C_LIGHT / (2 * n_fft * lsr)
    sc = scenes.circular_arc_scene(npulses=npulses, nsamples=ns, df=df, radius=3e3, altitude=2e3,
                                   arc_deg=5.0, ntargets=1, snr_db=None, seed=11)
    ip = orc.img_plane(ns, npulses, delta_r, sc.R_c)
    nu, nv = ip["nu"], ip["nv"]
    for ix, iv in ((nu // 2 + 3, nv // 2 - 2), (nu // 2 - 9, nv // 2 + 5)):
        tgt = ip["u_hat"] * ip["u"][ix] + ip["v_hat"] * ip["v"][iv]
        sc.targets = tgt[None, :]
        sc.amps = np.array([np.exp(1j * 0.7)])
        phs = scenes.simulate_phs_numpy(sc)
        img = orc.backprojection(phs, sc.k_r, 17, n_fft, delta_r, nu, nv, sc.pos, ip["pixel_locs"])
        y, x = np.unravel_index(np.argmax(np.abs(img)), img.shape)
        assert (x, y) == (ix, nv - 1 - iv)
        # the output is re-phased by expj(k_c * dr(pixel, np/2)) (backprojection.cpp:135)
        pm = sc.pos[npulses // 2].astype(np.float64)
        _, nr0 = orc.norm_r0(sc.pos)
        k_c = float(sc.k_r[ns // 2])
        expect = 0.7 + k_c * (nr0[npulses // 2] - np.linalg.norm(tgt - pm))
        assert abs(np.angle(img[y, x] * np.exp(-1j * expect))) < 0.02


# ------------------------------------------------------------------ image output closed forms

def test_img_output_closed_forms(orc):
    img = np.array([[3 + 4j, 0.5, 0], [1j, -2.5, 5]], np.complex128)
    dB = orc.img_output_to_dB(img)
    ln10f = float(np.log(np.float32(10)))  # log(Expr(10)) is evaluated in f32 (util.h:56-58)
    ref = 10 * (np.log(np.abs(img) / 5.0) / ln10f)
    with np.errstate(divide="ignore"):
        assert np.array_equal(dB, ref)
    assert dB[0, 0] == 0.0 and dB[1, 2] == 0.0 and dB[0, 2] == -np.inf
    u8 = orc.img_output_u8(np.array([-40.0, -30.0, -29.99, -15.0, -0.0001, 0.0, 3.0, -np.inf]),
                           -30.0, 0.0)
    # 255 * (clamp(dB) - lo) / (hi - lo), truncated
    assert list(u8) == [0, 0, 0, 127, 254, 255, 255, 0]
    assert orc.img_output_u8(np.array([-15.0]), -30.0, 0.0)[0] == int(255 * 0.5)
