"""Platform directories: the reference's on-disk input format (a directory of .npy files read by
platform_load, PlatformData.cpp:21-256) — ctypes bindings of the library's loader and PNG writer
(libsarbp_host.so, include/sarbp_host.h), and a writer that stores a synthetic scene in that format
so the `sarbp` driver can run on it.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
HOST_LIB_PATH = os.path.join(_HERE, "libsarbp_host.so")
SARBP_PATH = os.path.join(_HERE, "sarbp")

_P, _i32, _i64, _f32, _f64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_float, \
    ctypes.c_double


class PlatformStruct(ctypes.Structure):
    _fields_ = [("nsamples", _i32), ("npulses", _i32), ("delta_r", _f64),
                ("has_B", _i32), ("has_B_IF", _i32), ("has_delta_t", _i32),
                ("has_chirprate", _i32), ("has_f_0", _i32), ("has_vp", _i32),
                ("B", _f64), ("B_IF", _f32), ("delta_t", _f64), ("chirprate", _f64),
                ("f_0", _f64), ("vp", _f64),
                ("freq", _P), ("k_r", _P), ("k_y", _P), ("n_hat", _P), ("R_c", _P), ("t", _P),
                ("pos", _P), ("phs", _P), ("reserved", _i32 * 8)]


_host = None


def host_lib():
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB_PATH):
            raise ImportError(f"{HOST_LIB_PATH} is missing: build it with `make -C {_HERE}`")
        lib = ctypes.CDLL(HOST_LIB_PATH)
        lib.sarbp_platform_load.argtypes = [ctypes.c_char_p,
                                            ctypes.POINTER(ctypes.POINTER(PlatformStruct))]
        lib.sarbp_platform_free.argtypes = [ctypes.POINTER(PlatformStruct)]
        lib.sarbp_platform_free.restype = None
        lib.sarbp_png_write_gray8.argtypes = [ctypes.c_char_p, _P, _i32, _i32, _i64]
        lib.sarbp_host_last_error.restype = ctypes.c_char_p
        _host = lib
    return _host


class PlatformError(RuntimeError):
    def __init__(self, msg, code):
        super().__init__(msg)
        self.code = code


def _array(ptr, ctype, shape):
    if not ptr:
        return None
    n = int(np.prod(shape))
    buf = (ctype * n).from_address(ptr)
    return np.ctypeslib.as_array(buf).reshape(shape).copy()


def load(directory: str) -> dict:
    """platform_load through the library: a dict with the PlatformData fields (absent optional
    fields are None); pos (npulses, 3) f32, phs (npulses, nsamples) complex64."""
    lib = host_lib()
    p = ctypes.POINTER(PlatformStruct)()
    rv = lib.sarbp_platform_load(os.fsencode(directory), ctypes.byref(p))
    if rv != 0:
        raise PlatformError((lib.sarbp_host_last_error() or b"").decode(), rv)
    try:
        s = p.contents
        ns, np_ = s.nsamples, s.npulses
        out = dict(nsamples=ns, npulses=np_, delta_r=s.delta_r,
                   B=s.B if s.has_B else None, B_IF=s.B_IF if s.has_B_IF else None,
                   delta_t=s.delta_t if s.has_delta_t else None,
                   chirprate=s.chirprate if s.has_chirprate else None,
                   f_0=s.f_0 if s.has_f_0 else None, vp=s.vp if s.has_vp else None,
                   freq=_array(s.freq, ctypes.c_float, (ns,)),
                   k_r=_array(s.k_r, ctypes.c_float, (ns,)),
                   k_y=_array(s.k_y, ctypes.c_double, (np_,)),
                   n_hat=_array(s.n_hat, ctypes.c_float, (3,)),
                   R_c=_array(s.R_c, ctypes.c_float, (3,)),
                   t=_array(s.t, ctypes.c_double, (ns,)),
                   pos=_array(s.pos, ctypes.c_float, (np_, 3)))
        out["phs"] = _array(s.phs, ctypes.c_float, (np_, ns, 2)).view(np.complex64)[..., 0]
        return out
    finally:
        lib.sarbp_platform_free(p)


def write_png_gray8(path: str, img: np.ndarray) -> None:
    img = np.ascontiguousarray(img, dtype=np.uint8)
    h, w = img.shape
    lib = host_lib()
    rv = lib.sarbp_png_write_gray8(os.fsencode(path), img.ctypes.data, w, h, w)
    if rv != 0:
        raise PlatformError((lib.sarbp_host_last_error() or b"").decode(), rv)


def write_dir(directory: str, sc, phs: np.ndarray | None = None, n_hat=None) -> None:
    """Store scene `sc` (halide-sar-app_amd.scenes.Scene) as a platform directory in the layout of
    the reference's example data (data/AFRL/pass1/HH_npy): scalars as 1-element arrays, counts as
    int64, frequencies, k_r, R_c and pos as f32, phs as complex64 (np, ns)."""
    os.makedirs(directory, exist_ok=True)
    phs = sc.phs if phs is None else phs
    if phs is None:
        raise ValueError("the scene has no phase history")
    ns, np_ = sc.nsamples, sc.npulses
    freq = sc.f_c + (np.arange(ns) - ns // 2) * sc.df
    b_if = (ns - 1) * sc.df

    def save(name, arr):
        np.save(os.path.join(directory, name + ".npy"), arr)

    save("B_IF", np.array([b_if], np.float32))
    save("delta_r", np.array([sc.delta_r], np.float64))
    save("f_0", np.array([freq[0]], np.float64))
    save("chirprate", np.array([b_if / 1e-6], np.float64))
    save("nsamples", np.array([ns], np.int64))
    save("npulses", np.array([np_], np.int64))
    save("freq", freq.astype(np.float32))
    save("k_r", np.ascontiguousarray(sc.k_r, np.float32))
    save("R_c", np.ascontiguousarray(sc.R_c, np.float32))
    save("pos", np.ascontiguousarray(sc.pos, np.float32))
    save("t", np.linspace(-0.5e-6, 0.5e-6, ns))
    save("phs", np.ascontiguousarray(phs, np.complex64))
    if n_hat is not None:
        save("n_hat", np.asarray(n_hat, np.float32))


def nfft_rule(nsamples: int, upsample: int) -> int:
    """N_fft of the driver (sarbp.cpp:258)."""
    return int(math.pow(2, int(math.log2(nsamples * upsample)) + 1))
