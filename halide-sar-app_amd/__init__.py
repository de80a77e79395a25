"""sarbp_hip — MI355X-native SAR backprojection behind the reference's Halide C ABI.

Python host mirror of the reference's interface for the hot path (ISI-apex/halide-sar-app): the
same entry-point names, argument meaning and error codes as the AOT-generated
``backprojection`` / ``img_output_to_dB`` / ``img_output_u8`` (backprojection.cpp:35-46,
img_output.cpp:10-52, called from sarbp.cpp:277,299,316), bound with ctypes to
``libsarbp_hip.so``, whose kernels are hand-written HIP for gfx950.

There is no CPU fallback: if the shared library is missing, importing this module raises.
Arrays use the reference's numpy layouts (PlatformData.cpp:215-252, ImgPlane.cpp:103):
  phs complex64 (np, ns) = Halide {2, ns, np};  k_r float32 (ns,);  pos float32 (np, 3) = {3, np}
  u, v float64 (nu,), (nv,);  r float64 (3, nu*nv) = {nu*nv, 3};  image complex128 (nv, nu) = {2, nu, nv}
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsarbp_hip.so")

# One HIP runtime per process: torch ships its own libamdhip64 (soname libamdhip64.so.7). When it is
# loaded first, the library's dependency on that soname binds to it; loaded the other way round the
# process would hold two runtimes and the second cannot see the GPU.
try:
    import torch  # noqa: F401
except ImportError:
    pass

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `make -C {_HERE}` (or __graft_entry__.build()); "
        "the HIP path has no fallback")

# ---------------------------------------------------------------- Halide runtime buffer ABI


class halide_dimension_t(ctypes.Structure):
    _fields_ = [("min", ctypes.c_int32), ("extent", ctypes.c_int32), ("stride", ctypes.c_int32),
                ("flags", ctypes.c_uint32)]


class halide_type_t(ctypes.Structure):
    _fields_ = [("code", ctypes.c_uint8), ("bits", ctypes.c_uint8), ("lanes", ctypes.c_uint16)]


class halide_buffer_t(ctypes.Structure):
    _fields_ = [("device", ctypes.c_uint64), ("device_interface", ctypes.c_void_p),
                ("host", ctypes.c_void_p), ("flags", ctypes.c_uint64), ("type", halide_type_t),
                ("dimensions", ctypes.c_int32), ("dim", ctypes.POINTER(halide_dimension_t)),
                ("padding", ctypes.c_void_p)]


HALIDE_INT, HALIDE_UINT, HALIDE_FLOAT = 0, 1, 2

# halide_error_code_t values returned by the entry points
ERR = dict(success=0, generic_error=-1, bad_type=-3, access_out_of_bounds=-4,
           buffer_allocation_too_large=-5,
           buffer_extents_too_large=-6, constraint_violated=-8, param_too_small=-9,
           param_too_large=-10, out_of_memory=-11, buffer_argument_is_null=-12,
           copy_to_host_failed=-14, copy_to_device_failed=-15, device_malloc_failed=-16,
           no_device_interface=-19, device_run_failed=-23, buffer_extents_negative=-28,
           bad_dimensions=-43)

NORM_R0_F32_STRICT, NORM_R0_F32_FMA, NORM_R0_F64 = 0, 1, 2
TAYLOR_DEN_INT, TAYLOR_DEN_F32 = 0, 1
POLICY_DEFAULT = -1   # serving jobs: the process default at submission
Q_C32, Q_C16 = 0, 1
RANGE_F64, RANGE_DF32 = 0, 1
GRID_FORM_NONE, GRID_FORM_PQ, GRID_FORM_CUBIC = 0, 1, 2


class Buffer:
    """A halide_buffer_t over a numpy array (or a bounds-query buffer when array is None).

    `halide_dims` lists (min, extent, stride) per Halide dimension, innermost first; by default it
    is derived from the array's strides, reversed to Halide's order, with complex arrays gaining a
    leading {0, 2, 1} dimension.
    """

    def __init__(self, array, halide_dims=None, type_code=None, bits=None):
        self.array = array
        if array is not None:
            dt = array.dtype
            if halide_dims is None:
                itemsize = dt.itemsize
                if dt.kind == "c":
                    base = itemsize // 2
                    dims = [(0, 2, 1)]
                    for ext, st in zip(reversed(array.shape), reversed(array.strides)):
                        dims.append((0, ext, st // base))
                else:
                    dims = [(0, e, s // itemsize) for e, s in zip(reversed(array.shape),
                                                                  reversed(array.strides))]
                halide_dims = dims
            if type_code is None:
                type_code = {"f": HALIDE_FLOAT, "c": HALIDE_FLOAT, "u": HALIDE_UINT,
                             "i": HALIDE_INT}[dt.kind]
                bits = dt.itemsize * 8 // (2 if dt.kind == "c" else 1)
        n = len(halide_dims)
        self._dims = (halide_dimension_t * n)(*[halide_dimension_t(m, e, s, 0)
                                                 for (m, e, s) in halide_dims])
        self.raw = halide_buffer_t()
        self.raw.device = 0
        self.raw.host = array.ctypes.data if array is not None else None
        self.raw.flags = 0
        self.raw.type = halide_type_t(type_code, bits, 1)
        self.raw.dimensions = n
        self.raw.dim = ctypes.cast(self._dims, ctypes.POINTER(halide_dimension_t))

    @classmethod
    def bounds_query(cls, halide_dims, type_code, bits):
        return cls(None, halide_dims, type_code, bits)

    def dims(self):
        return [(self._dims[i].min, self._dims[i].extent, self._dims[i].stride)
                for i in range(self.raw.dimensions)]

    @property
    def ref(self):
        return ctypes.byref(self.raw)


# ---------------------------------------------------------------- library binding

_lib = ctypes.CDLL(LIB_PATH)
_P = ctypes.c_void_p
_HB = ctypes.POINTER(halide_buffer_t)
_i32, _i64, _f64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_double

for _name in ("backprojection", "backprojection_cuda"):
    getattr(_lib, _name).argtypes = [_HB, _HB, _i32, _i32, _f64, _HB, _HB, _HB, _HB, _HB]
    getattr(_lib, _name).restype = ctypes.c_int
_lib.img_output_to_dB.argtypes = [_HB, _HB]
_lib.img_output_u8.argtypes = [_HB, _f64, _f64, _HB]
_lib.dft_init_fftw.argtypes = [ctypes.c_size_t]
_lib.dft_init_fftw.restype = None
_lib.dft_destroy_fftw.argtypes = []
_lib.dft_destroy_fftw.restype = None
_lib.sarbp_last_error.restype = ctypes.c_char_p
_lib.sarbp_device_count.restype = ctypes.c_int
_lib.sarbp_set_default_policies.argtypes = [ctypes.c_int, ctypes.c_int]
_lib.sarbp_get_default_policies.argtypes = [_P, _P]
_lib.sarbp_taylor_window.argtypes = [_i32, _i32, _i32, _P]
_lib.sarbp_img_plane_create.argtypes = [_i32, _i32, _f64, _P, _P, _f64, _f64, _i32, _P, _P, _P,
                                        _P, _P, _P, _P, _P, _P]


class PlanDesc(ctypes.Structure):
    _fields_ = [("nsamples", _i32), ("npulses", _i32), ("n_fft", _i32), ("taylor_s_l", _i32),
                ("delta_r", _f64), ("nu", _i32), ("nv", _i32), ("x0", _i32), ("y0", _i32),
                ("width", _i32), ("height", _i32), ("r_min", _i64), ("r_extent", _i64),
                ("norm_r0_policy", _i32), ("taylor_den_policy", _i32), ("q_format", _i32),
                ("range_mode", _i32), ("device", _i32), ("tile_ppt", _i32),
                ("fft_blocks", _i32), ("row_windows", _i32), ("reserved", _i32 * 4)]


_lib.sarbp_plan_desc_init.argtypes = [ctypes.POINTER(PlanDesc)]
_lib.sarbp_plan_desc_init.restype = None
_lib.sarbp_plan_create.argtypes = [ctypes.POINTER(PlanDesc), ctypes.POINTER(_P)]
_lib.sarbp_plan_destroy.argtypes = [_P]
_lib.sarbp_plan_execute.argtypes = [_P, _P, _P, _P, _P, _P, _P]
_lib.sarbp_plan_range_profiles.argtypes = [_P, _P, _P, _P, _P]
_lib.sarbp_plan_accumulate.argtypes = [_P, _P, _P, _P, _P, _P, _P]
_lib.sarbp_plan_set_grid.argtypes = [_P, _f64, _f64, _P, _P]
_lib.sarbp_plan_q_buffer.argtypes = [_P]
_lib.sarbp_plan_q_buffer.restype = _P
_lib.sarbp_plan_q_bytes.argtypes = [_P]
_lib.sarbp_plan_q_bytes.restype = ctypes.c_size_t
_lib.sarbp_plan_tile_ppt.argtypes = [_P]
_lib.sarbp_plan_tile_ppt.restype = _i32
_lib.sarbp_plan_grid_form.argtypes = [_P]
_lib.sarbp_plan_grid_form.restype = _i32
_lib.sarbp_plan_img_output.argtypes = [_P, _P, _P, _P, _f64, _f64, _P]
_lib.sarbp_plan_tile_paths.argtypes = [_P, _P]
_lib.sarbp_plan_read_pulse_norms.argtypes = [_P, _P, _i32]
_lib.sarbp_plan_read_profiles.argtypes = [_P, _P]
_lib.sarbp_img_output_device.argtypes = [_P, _i64, _P, _P, _f64, _f64, _P]
_lib.sarbp_img_output_device_noalloc.argtypes = [_P, _i64, _P, _P, _f64, _f64, _P]
_lib.sarbp_sim_phs_device.argtypes = [_P, _i32, _P, _i32, _P, _P, _i32, _P, _P]
_lib.sarbp_pixel_locs_device.argtypes = [_P, _i32, _P, _i32, _P, _P, _P, _P]


def lib():
    return _lib


def last_error() -> str:
    return (_lib.sarbp_last_error() or b"").decode()


def abi_last_call_bands() -> tuple[int, int]:
    """(bands on the image plane, bands) of this thread's last backprojection() call."""
    g, n = ctypes.c_int32(0), ctypes.c_int32(0)
    _lib.sarbp_abi_last_call_bands(ctypes.byref(g), ctypes.byref(n))
    return g.value, n.value


def device_count() -> int:
    return int(_lib.sarbp_device_count())


def set_default_policies(norm_r0_policy: int, taylor_den_policy: int) -> int:
    return _lib.sarbp_set_default_policies(norm_r0_policy, taylor_den_policy)


def get_default_policies():
    a, b = ctypes.c_int(), ctypes.c_int()
    _lib.sarbp_get_default_policies(ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


# ---------------------------------------------------------------- the reference's entry points


def _as_buf(x):
    return x if isinstance(x, Buffer) else Buffer(x)


def backprojection(phs, k_r, taylor_s_l: int, N_fft: int, delta_r: float, u, v, pos, r,
                   output_img) -> int:
    """int backprojection(phs, k_r, taylor_s_l, N_fft, delta_r, u, v, pos, r, output_img)
    (backprojection.cpp:35-46). Arguments are Buffers or numpy arrays in the reference layouts
    (r as a (3, nu*nv) float64 array, output_img as complex128 (nv, nu))."""
    bufs = [_as_buf(b) for b in (phs, k_r, u, v, pos, r, output_img)]
    return _lib.backprojection(bufs[0].ref, bufs[1].ref, int(taylor_s_l), int(N_fft),
                               float(delta_r), bufs[2].ref, bufs[3].ref, bufs[4].ref,
                               bufs[5].ref, bufs[6].ref)


def backprojection_cuda(phs, k_r, taylor_s_l, N_fft, delta_r, u, v, pos, r, output_img) -> int:
    bufs = [_as_buf(b) for b in (phs, k_r, u, v, pos, r, output_img)]
    return _lib.backprojection_cuda(bufs[0].ref, bufs[1].ref, int(taylor_s_l), int(N_fft),
                                    float(delta_r), bufs[2].ref, bufs[3].ref, bufs[4].ref,
                                    bufs[5].ref, bufs[6].ref)


def img_output_to_dB(img, out) -> int:
    """int img_output_to_dB(img, out) (img_output.cpp:10-35)."""
    a, b = _as_buf(img), _as_buf(out)
    return _lib.img_output_to_dB(a.ref, b.ref)


def img_output_u8(dB, dB_min: float, dB_max: float, out) -> int:
    """int img_output_u8(dB, dB_min, dB_max, out) (img_output.cpp:41-52)."""
    a, b = _as_buf(dB), _as_buf(out)
    return _lib.img_output_u8(a.ref, float(dB_min), float(dB_max), b.ref)


def dft_init_fftw(N_fft: int) -> None:
    _lib.dft_init_fftw(int(N_fft))


def dft_destroy_fftw() -> None:
    _lib.dft_destroy_fftw()


def nfft(nsamples: int, upsample: int) -> int:
    """FFT length rule of sarbp.cpp:258: 2^(floor(log2(nsamples * upsample)) + 1)."""
    n = nsamples * upsample
    return 1 << (int(np.floor(np.log2(n))) + 1)


# ---------------------------------------------------------------- host-side helpers


def taylor_window(num: int, taylor_s_l: int, den_policy: int = TAYLOR_DEN_INT) -> np.ndarray:
    out = np.empty(num, np.float64)
    rv = _lib.sarbp_taylor_window(num, int(taylor_s_l), int(den_policy),
                                  out.ctypes.data_as(_P))
    if rv != 0:
        raise ValueError(last_error())
    return out


def img_plane_create(nsamples: int, npulses: int, delta_r: float, R_c, n_hat=(0.0, 0.0, 1.0),
                     res_factor: float = 1.0, aspect: float = 1.0, upsample: bool = True,
                     with_pixel_locs: bool = True) -> dict:
    """img_plane_create (ImgPlane.cpp:16-63): nu, nv, d_u, d_v, u, v, k_u, k_v, u_hat, v_hat and
    pixel_locs as a (3, nu*nv) float64 array (Halide {nu*nv, 3})."""
    R_c = np.ascontiguousarray(R_c, dtype=np.float32)
    n_hat = np.ascontiguousarray(n_hat, dtype=np.float32)
    dims = np.zeros(2, np.int32)
    scal = np.zeros(2, np.float64)
    p = lambda a: a.ctypes.data_as(_P) if a is not None else None  # noqa: E731
    rv = _lib.sarbp_img_plane_create(nsamples, npulses, float(delta_r), p(R_c), p(n_hat),
                                     float(res_factor), float(aspect), int(bool(upsample)),
                                     p(dims), p(scal), None, None, None, None, None, None, None)
    if rv != 0:
        raise ValueError(last_error())
    nu, nv = int(dims[0]), int(dims[1])
    out = dict(nu=nu, nv=nv, d_u=float(scal[0]), d_v=float(scal[1]), u=np.empty(nu),
               v=np.empty(nv), k_u=np.empty(nu), k_v=np.empty(nv), u_hat=np.empty(3),
               v_hat=np.empty(3))
    out["pixel_locs"] = np.empty((3, nu * nv)) if with_pixel_locs else None
    rv = _lib.sarbp_img_plane_create(nsamples, npulses, float(delta_r), p(R_c), p(n_hat),
                                     float(res_factor), float(aspect), int(bool(upsample)),
                                     p(dims), p(scal), p(out["u"]), p(out["v"]), p(out["k_u"]),
                                     p(out["k_v"]), p(out["u_hat"]), p(out["v_hat"]),
                                     p(out["pixel_locs"]))
    if rv != 0:
        raise ValueError(last_error())
    return out


def image_grid(nu: int, nv: int, d_u: float, d_v: float, u_hat, v_hat,
               with_pixel_locs: bool = True) -> dict:
    """An explicit nu x nv grid in the img_plane_create conventions (for configs whose image size
    is not the img_plane rule's): u = (x + floor(-nu/2)) d_u, likewise v."""
    u = (np.arange(nu) + (-(nu // 2) if nu % 2 == 0 else -(nu // 2) - 1)) * d_u
    v = (np.arange(nv) + (-(nv // 2) if nv % 2 == 0 else -(nv // 2) - 1)) * d_v
    u_hat = np.asarray(u_hat, np.float64)
    v_hat = np.asarray(v_hat, np.float64)
    pl = None
    if with_pixel_locs:
        pl = np.empty((3, nu * nv))
        for d in range(3):
            pl[d] = ((0.0 + u_hat[d] * u)[None, :] + (v_hat[d] * v)[:, None]).ravel()
    return dict(nu=nu, nv=nv, d_u=d_u, d_v=d_v, u=u, v=v, u_hat=u_hat, v_hat=v_hat, pixel_locs=pl)


def grid_points(grid: dict, xs, ys) -> np.ndarray:
    """pixel_locs of output pixels (xs, ys) as a (3, count) array: output row y holds v index
    nv - 1 - y (backprojection.cpp:130)."""
    xs = np.asarray(xs)
    vy = grid["nv"] - 1 - np.asarray(ys)
    u, v = grid["u"][xs], grid["v"][vy]
    return np.stack([(0.0 + grid["u_hat"][d] * u) + grid["v_hat"][d] * v for d in range(3)])


# ---------------------------------------------------------------- device-resident plans


def _ptr(t) -> int | None:
    if t is None:
        return None
    return int(t.data_ptr()) if hasattr(t, "data_ptr") else int(t)


def _stream_handle(stream) -> int | None:
    if stream is None:
        import torch
        return int(torch.cuda.current_stream().cuda_stream)
    return int(getattr(stream, "cuda_stream", stream))


class Plan:
    """A device-resident plan (sarbp_plan_create): windows, range-profile buffer, pulse table.

    Tensors passed to execute() must be device tensors in the layouts of sarbp_plan_execute:
    phs complex64 (np, ns), k_r float32 (ns,), pos float32 (np, 3), r float64 (3, r_extent),
    out complex128 (height, width).
    """

    def __init__(self, nsamples, npulses, n_fft, taylor_s_l, delta_r, nu, nv, region=None,
                 r_min=0, r_extent=None, norm_r0_policy=None, taylor_den_policy=None,
                 q_format=Q_C32, range_mode=RANGE_F64, device=0, tile_ppt=0, fft_blocks=0,
                 row_windows=0):
        d = PlanDesc()
        _lib.sarbp_plan_desc_init(ctypes.byref(d))
        d.nsamples, d.npulses, d.n_fft, d.taylor_s_l = nsamples, npulses, n_fft, taylor_s_l
        d.delta_r = delta_r
        d.nu, d.nv = nu, nv
        x0, y0, w, h = region if region is not None else (0, 0, nu, nv)
        d.x0, d.y0, d.width, d.height = x0, y0, w, h
        d.r_min = r_min
        d.r_extent = nu * nv if r_extent is None else r_extent
        if norm_r0_policy is not None:
            d.norm_r0_policy = norm_r0_policy
        if taylor_den_policy is not None:
            d.taylor_den_policy = taylor_den_policy
        d.q_format, d.range_mode, d.device = q_format, range_mode, device
        d.tile_ppt = tile_ppt
        d.fft_blocks = fft_blocks
        d.row_windows = row_windows
        self.desc = d
        self._p = _P()
        rv = _lib.sarbp_plan_create(ctypes.byref(d), ctypes.byref(self._p))
        if rv != 0:
            raise RuntimeError(f"sarbp_plan_create failed ({rv}): {last_error()}")

    def __del__(self, _destroy=_lib.sarbp_plan_destroy):
        # the library function is bound at definition: at interpreter exit the module's globals may
        # already be cleared when the last plans are collected
        p = getattr(self, "_p", None)
        if p is not None and p.value:
            _destroy(p)
            self._p = None

    def set_grid(self, d_u, d_v, u_hat, v_hat):
        """Grid mode (sarbp_plan_set_grid): pixel locations generated in the accumulation kernel
        from the image grid; execute/accumulate then take r=None. u_hat=None switches back."""
        if u_hat is None:
            rv = _lib.sarbp_plan_set_grid(self._p, 0.0, 0.0, None, None)
        else:
            uh = np.ascontiguousarray(u_hat, np.float64)
            vh = np.ascontiguousarray(v_hat, np.float64)
            rv = _lib.sarbp_plan_set_grid(self._p, float(d_u), float(d_v), uh.ctypes.data,
                                          vh.ctypes.data)
        if rv != 0:
            raise RuntimeError(f"sarbp_plan_set_grid failed ({rv}): {last_error()}")

    def execute(self, phs, k_r, pos, r, out, stream=None):
        rv = _lib.sarbp_plan_execute(self._p, _ptr(phs), _ptr(k_r), _ptr(pos), _ptr(r), _ptr(out),
                                     _stream_handle(stream))
        if rv != 0:
            raise RuntimeError(f"sarbp_plan_execute failed ({rv}): {last_error()}")

    def range_profiles(self, phs, k_r, q=None, stream=None):
        rv = _lib.sarbp_plan_range_profiles(self._p, _ptr(phs), _ptr(k_r), _ptr(q),
                                            _stream_handle(stream))
        if rv != 0:
            raise RuntimeError(f"sarbp_plan_range_profiles failed ({rv}): {last_error()}")

    def accumulate(self, k_r, pos, r, out, q=None, stream=None):
        rv = _lib.sarbp_plan_accumulate(self._p, _ptr(q), _ptr(k_r), _ptr(pos), _ptr(r), _ptr(out),
                                        _stream_handle(stream))
        if rv != 0:
            raise RuntimeError(f"sarbp_plan_accumulate failed ({rv}): {last_error()}")

    def read_profiles(self) -> np.ndarray:
        """The plan's range profiles Q as complex64 (npulses, n_fft) (synchronises the device)."""
        out = np.empty((self.desc.npulses, self.desc.n_fft), np.complex64)
        rv = _lib.sarbp_plan_read_profiles(self._p, out.ctypes.data_as(_P))
        if rv != 0:
            raise RuntimeError(f"sarbp_plan_read_profiles failed ({rv}): {last_error()}")
        return out

    def pulse_norms(self, n=None) -> np.ndarray:
        """The ||pos_p|| table of the last accumulate (synchronous)."""
        n = self.desc.npulses if n is None else n
        out = np.empty(n, np.float64)
        rv = _lib.sarbp_plan_read_pulse_norms(self._p, out.ctypes.data_as(_P), n)
        if rv != 0:
            raise RuntimeError(f"sarbp_plan_read_pulse_norms failed ({rv}): {last_error()}")
        return out

    @property
    def q_ptr(self) -> int:
        return int(_lib.sarbp_plan_q_buffer(self._p))

    @property
    def q_bytes(self) -> int:
        """Size of the stored range-profile rows (the buffer range_profiles(q=...) fills)."""
        return int(_lib.sarbp_plan_q_bytes(self._p))

    @property
    def tile_ppt(self) -> int:
        """Pixels per thread of the accumulation tiles (the default resolved for the region)."""
        return int(_lib.sarbp_plan_tile_ppt(self._p))

    @property
    def grid_form(self) -> int:
        """GRID_FORM_NONE / _PQ / _CUBIC: the expansion form of the plan's grid-mode tiles."""
        return int(_lib.sarbp_plan_grid_form(self._p))

    TILE_PATHS = ("exact", "direct", "pq", "cubic_gather", "cubic_window", "cubic_df")

    def tile_paths(self) -> dict:
        """Tiles of the last accumulation per loop (sarbp_plan_tile_paths; synchronises)."""
        out = np.zeros(len(self.TILE_PATHS), np.int64)
        rv = _lib.sarbp_plan_tile_paths(self._p, out.ctypes.data_as(_P))
        if rv != 0:
            raise RuntimeError(f"sarbp_plan_tile_paths failed ({rv}): {last_error()}")
        return dict(zip(self.TILE_PATHS, (int(x) for x in out)))

    def img_output(self, img, dB=None, u8=None, dB_min=0.0, dB_max=1.0, stream=None):
        """dB / u8 of the image the last accumulate wrote, in one pass (sarbp_plan_img_output)."""
        rv = _lib.sarbp_plan_img_output(self._p, _ptr(img), _ptr(dB), _ptr(u8), float(dB_min),
                                        float(dB_max), _stream_handle(stream))
        if rv != 0:
            raise RuntimeError(f"sarbp_plan_img_output failed ({rv}): {last_error()}")


def img_output_device(img, npix: int, dB=None, u8=None, dB_min=0.0, dB_max=1.0, stream=None):
    rv = _lib.sarbp_img_output_device(_ptr(img), int(npix), _ptr(dB), _ptr(u8), float(dB_min),
                                      float(dB_max), _stream_handle(stream))
    if rv != 0:
        raise RuntimeError(f"sarbp_img_output_device failed ({rv}): {last_error()}")


def img_output_device_noalloc(img, npix: int, dB=None, u8=None, dB_min=0.0, dB_max=1.0,
                              stream=None):
    """sarbp_img_output_device_noalloc: the same output without allocating device memory."""
    rv = _lib.sarbp_img_output_device_noalloc(_ptr(img), int(npix), _ptr(dB), _ptr(u8),
                                              float(dB_min), float(dB_max), _stream_handle(stream))
    if rv != 0:
        raise RuntimeError(f"sarbp_img_output_device_noalloc failed ({rv}): {last_error()}")


def sim_phs_device(pos64, k_r64, targets, amps, phs, stream=None):
    """Evaluate the point-scatterer phase-history sum on the device (tensors: pos64 f64 (np, 3),
    k_r64 f64 (ns,), targets f64 (nt, 3), amps complex128 (nt,), phs complex64 (np, ns))."""
    np_, ns, nt = pos64.shape[0], k_r64.shape[0], targets.shape[0]
    rv = _lib.sarbp_sim_phs_device(_ptr(pos64), np_, _ptr(k_r64), ns, _ptr(targets), _ptr(amps),
                                   nt, _ptr(phs), _stream_handle(stream))
    if rv != 0:
        raise RuntimeError(f"sarbp_sim_phs_device failed ({rv}): {last_error()}")


def pixel_locs_device(u, v, u_hat, v_hat, out, stream=None):
    """pixel_locs (img_plane.cpp:77-113) on the device: u, v device float64 tensors, u_hat/v_hat
    3-vectors on the host, out a device float64 tensor (3, len(u) * len(v))."""
    uh = np.ascontiguousarray(u_hat, np.float64)
    vh = np.ascontiguousarray(v_hat, np.float64)
    rv = _lib.sarbp_pixel_locs_device(_ptr(u), int(u.shape[0]), _ptr(v), int(v.shape[0]),
                                      uh.ctypes.data_as(_P), vh.ctypes.data_as(_P), _ptr(out),
                                      _stream_handle(stream))
    if rv != 0:
        raise RuntimeError(f"sarbp_pixel_locs_device failed ({rv}): {last_error()}")
