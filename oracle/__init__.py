"""CPU FP64 oracle for the SAR backprojection hot path — TEST INFRASTRUCTURE ONLY.

ctypes bindings over ``libsarbp_oracle.so`` (built from ``sarbp_oracle.c`` by ``oracle/Makefile``).
Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import this
package, and only as the checker (or as the timed CPU baseline); the product path never does.

Each wrapper names the reference code it restates (ISI-apex/halide-sar-app, file:line).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libsarbp_oracle.so")

NORM_R0_F32_STRICT = 0
NORM_R0_F32_FMA = 1
NORM_R0_F64 = 2

_lib = None


def build() -> str:
    """Compile the oracle with gcc (no GPU, no reference sources needed)."""
    subprocess.run(["make", "-s", "-C", _HERE, "-B"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        i32, f64, f32, i64 = ctypes.c_int, ctypes.c_double, ctypes.c_float, ctypes.c_int64
        L.orc_taylor.argtypes = [i32, f64, i32, P]
        L.orc_linspace.argtypes = [f64, f64, i32, P]
        L.orc_linspace.restype = None
        L.orc_fftshift_src.argtypes = [i32, i32]
        L.orc_nfft.argtypes = [i32, i32]
        L.orc_fft_f64.argtypes = [P, i32]
        L.orc_norm_r0.argtypes = [P, i32, i32, P, P]
        L.orc_norm_r0.restype = None
        L.orc_range_profiles.argtypes = [P, P, i32, i32, i32, i32, i32, i32, P]
        L.orc_backproject.argtypes = [P, i32, i32, i32, f64, f32, P, P, P, i64, i32, i32,
                                      i32, i32, i32, i32, i32, P]
        L.orc_backproject_points.argtypes = [P, i32, i32, i32, f64, f32, P, P, P, i64, i32, i32,
                                             P, P, i32, i32, P]
        L.orc_backprojection.argtypes = [P, P, i32, i32, i32, i32, f64, i32, i32, P, P, i32, i32,
                                         i32, i32, P]
        L.orc_img_output_to_dB.argtypes = [P, i64, P]
        L.orc_img_output_to_dB.restype = None
        L.orc_img_output_u8.argtypes = [P, i64, f64, f64, P]
        L.orc_img_output_u8.restype = None
        L.orc_img_plane.argtypes = [i32, i32, f64, P, P, f64, f64, i32, P, P, P, P, P, P, P, P, P]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def taylor(num: int, s_l: float, int_den: bool = True) -> np.ndarray:
    """Taylor window, signal.h:14-51."""
    out = np.empty(num, np.float64)
    if lib().orc_taylor(num, float(s_l), int(int_den), _p(out)) != 0:
        raise ValueError("taylor: num must be >= 2")
    return out


def linspace(start: float, stop: float, num: int) -> np.ndarray:
    """Clamped linspace, util.h:80-84."""
    out = np.empty(num, np.float64)
    lib().orc_linspace(float(start), float(stop), num, _p(out))
    return out


def fftshift_src(i: int, n: int) -> int:
    """Source index of the ceil-mid roll, signal_complex.h:24-29."""
    return lib().orc_fftshift_src(i, n)


def nfft(nsamples: int, upsample: int) -> int:
    """FFT length rule, sarbp.cpp:258."""
    return lib().orc_nfft(nsamples, upsample)


def fft(x: np.ndarray) -> np.ndarray:
    """Forward unnormalised DFT (FFTW_FORWARD semantics of dft.cpp:26,72), f64."""
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    if lib().orc_fft_f64(_p(a), a.shape[0]) != 0:
        raise ValueError("fft failed")
    return a


def norm_r0(pos: np.ndarray, policy: int = NORM_R0_F32_STRICT):
    """||pos_p|| table (backprojection.cpp:108); returns (f32 table, f64 table)."""
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    np_ = pos.shape[0]
    o32 = np.empty(np_, np.float32)
    o64 = np.empty(np_, np.float64)
    lib().orc_norm_r0(_p(pos), np_, policy, _p(o32), _p(o64))
    return o32, o64


def range_profiles(phs: np.ndarray, k_r: np.ndarray, taylor_s_l: int, n_fft: int,
                   int_den: bool = True, pulse_shift_quirk: bool = True) -> np.ndarray:
    """Q {np, N} complex128 (backprojection.cpp:80-104)."""
    phs = np.ascontiguousarray(phs, dtype=np.complex64)
    k_r = np.ascontiguousarray(k_r, dtype=np.float32)
    np_, ns = phs.shape
    Q = np.empty((np_, n_fft), np.complex128)
    rv = lib().orc_range_profiles(_p(phs), _p(k_r), ns, np_, int(taylor_s_l), int(n_fft),
                                  int(int_den), int(pulse_shift_quirk), _p(Q))
    if rv != 0:
        raise ValueError("range_profiles failed")
    return Q


def backproject(Q: np.ndarray, ns: int, delta_r: float, k_c: float, pos: np.ndarray,
                nr0: np.ndarray, r: np.ndarray, nu: int, nv: int, region=None,
                nthreads: int = 1) -> np.ndarray:
    """Accumulation over pulses for an output region (x0, y0, w, h); returns complex128 {h, w}.

    r is pixel_locs as a (3, r_extent) f64 array (Halide {r_extent, 3}, stride (1, r_extent)).
    """
    Q = np.ascontiguousarray(Q, dtype=np.complex128)
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    nr0 = np.ascontiguousarray(nr0, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    np_, N = Q.shape
    x0, y0, w, h = region if region is not None else (0, 0, nu, nv)
    out = np.empty((h, w), np.complex128)
    rv = lib().orc_backproject(_p(Q), N, ns, np_, float(delta_r), float(k_c), _p(pos), _p(nr0),
                               _p(r), r.shape[1], nu, nv, x0, y0, w, h, int(nthreads), _p(out))
    if rv != 0:
        raise ValueError("backproject failed")
    return out


def backproject_points(Q: np.ndarray, ns: int, delta_r: float, k_c: float, pos: np.ndarray,
                       nr0: np.ndarray, r: np.ndarray, nu: int, nv: int, xs: np.ndarray,
                       ys: np.ndarray, nthreads: int = 1) -> np.ndarray:
    """Accumulation at listed output coordinates (x, y); returns complex128 {count}."""
    Q = np.ascontiguousarray(Q, dtype=np.complex128)
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    nr0 = np.ascontiguousarray(nr0, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    xs = np.ascontiguousarray(xs, dtype=np.int32)
    ys = np.ascontiguousarray(ys, dtype=np.int32)
    np_, N = Q.shape
    out = np.empty(xs.shape[0], np.complex128)
    rv = lib().orc_backproject_points(_p(Q), N, ns, np_, float(delta_r), float(k_c), _p(pos),
                                      _p(nr0), _p(r), r.shape[1], nu, nv, _p(xs), _p(ys),
                                      xs.shape[0], int(nthreads), _p(out))
    if rv != 0:
        raise ValueError("backproject_points failed")
    return out


def backprojection(phs, k_r, taylor_s_l, n_fft, delta_r, nu, nv, pos, r,
                   norm_policy=NORM_R0_F32_STRICT, int_den=True, pulse_shift_quirk=True,
                   nthreads=1) -> np.ndarray:
    """The whole backprojection() pipeline (backprojection.cpp:67-139); complex128 {nv, nu}."""
    phs = np.ascontiguousarray(phs, dtype=np.complex64)
    k_r = np.ascontiguousarray(k_r, dtype=np.float32)
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    r = np.ascontiguousarray(r, dtype=np.float64)
    np_, ns = phs.shape
    out = np.empty((nv, nu), np.complex128)
    rv = lib().orc_backprojection(_p(phs), _p(k_r), ns, np_, int(taylor_s_l), int(n_fft),
                                  float(delta_r), nu, nv, _p(pos), _p(r), int(norm_policy),
                                  int(int_den), int(pulse_shift_quirk), int(nthreads), _p(out))
    if rv != 0:
        raise ValueError("backprojection failed")
    return out


def img_output_to_dB(img: np.ndarray) -> np.ndarray:
    """img_output.cpp:10-27 over a complex128 image of any shape."""
    img = np.ascontiguousarray(img, dtype=np.complex128)
    out = np.empty(img.shape, np.float64)
    lib().orc_img_output_to_dB(_p(img), img.size, _p(out))
    return out


def img_output_u8(dB: np.ndarray, dB_min: float, dB_max: float) -> np.ndarray:
    """img_output.cpp:41-52."""
    dB = np.ascontiguousarray(dB, dtype=np.float64)
    out = np.empty(dB.shape, np.uint8)
    lib().orc_img_output_u8(_p(dB), dB.size, float(dB_min), float(dB_max), _p(out))
    return out


def img_plane(nsamples: int, npulses: int, delta_r: float, R_c, n_hat=(0.0, 0.0, 1.0),
              res_factor: float = 1.0, aspect: float = 1.0, upsample: bool = True,
              pixel_locs: bool = True) -> dict:
    """img_plane_create (ImgPlane.cpp:16-63): dict of nu, nv, d_u, d_v, u, v, k_u, k_v, u_hat,
    v_hat and pixel_locs as a (3, nu*nv) array."""
    R_c = np.ascontiguousarray(R_c, dtype=np.float32)
    n_hat = np.ascontiguousarray(n_hat, dtype=np.float32)
    dims = np.zeros(2, np.int32)
    scal = np.zeros(2, np.float64)
    L = lib()
    L.orc_img_plane(nsamples, npulses, float(delta_r), _p(R_c), _p(n_hat), float(res_factor),
                    float(aspect), int(bool(upsample)), _p(dims), _p(scal), None, None, None,
                    None, None, None, None)
    nu, nv = int(dims[0]), int(dims[1])
    u = np.empty(nu)
    v = np.empty(nv)
    k_u = np.empty(nu)
    k_v = np.empty(nv)
    u_hat = np.empty(3)
    v_hat = np.empty(3)
    pl = np.empty((3, nu * nv)) if pixel_locs else None
    rv = L.orc_img_plane(nsamples, npulses, float(delta_r), _p(R_c), _p(n_hat), float(res_factor),
                         float(aspect), int(bool(upsample)), _p(dims), _p(scal), _p(u), _p(v),
                         _p(k_u), _p(k_v), _p(u_hat), _p(v_hat),
                         _p(pl) if pl is not None else None)
    if rv != 0:
        raise ValueError("img_plane failed")
    return dict(nu=nu, nv=nv, d_u=float(scal[0]), d_v=float(scal[1]), u=u, v=v, k_u=k_u, k_v=k_v,
                u_hat=u_hat, v_hat=v_hat, pixel_locs=pl)
