"""Python front end of libtiltgpu (ctypes over include/tiltgpu.h).

Mirrors the interfaces of the `tilt` reconstruction core of imod-rs:

* ``TiltOptions`` -- the option surface of apps/tilt/src/main.rs:18-120, same
  names and defaults;
* ``reconstruct`` -- the pipeline of main.rs:601-1093 (after file reading) on
  the GPU: returns the output volume [ny][out_nz][out_nx] and the header
  statistics (amin, amax, amean);
* ``GpuSession`` -- the GpuSession seam of apps/tilt/src/gpu.rs:26-269
  (``new`` returns None when no GPU can be used, ``reconstruct_row``), fed with
  what the CPU path also needs so its rows equal main.rs:1011-1073.

There is no CPU fallback: when libtiltgpu.so is missing every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libtiltgpu.so")

OK, ERR_ARG, ERR_HIP, ERR_NODEV, ERR_BUDGET, ERR_STATE, ERR_IO, ERR_KERNEL = 0, -1, -2, -3, -4, -5, -6, -7
FILL = {"mean": 0, "zero": 1, "edge": 2}
STORAGE = {"f32": 0, "f16": 1, "bf16": 2}
GEOMETRY = {"f32": 0, "f64": 1}


class TiltGpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"tiltgpu error {code}: {msg}")
        self.code = code


class Dims(C.Structure):
    _fields_ = [("in_nx", C.c_int32), ("in_ny", C.c_int32), ("nz", C.c_int32),
                ("out_nx", C.c_int32), ("out_nz", C.c_int32)]


class Options(C.Structure):
    _fields_ = [("radial_cutoff", C.c_float), ("radial_falloff", C.c_float),
                ("no_weight", C.c_int32), ("exact_filter_size", C.c_int32),
                ("hamming", C.c_int32), ("no_cosine_stretch", C.c_int32),
                ("cosine_cubic", C.c_int32), ("max_cosine_tilt", C.c_float),
                ("log", C.c_int32), ("fill_mode", C.c_int32),
                ("densweight", C.c_float), ("outscale", C.c_float), ("outadd", C.c_float),
                ("sirt", C.c_int32), ("sirt_relax", C.c_float), ("fake_sirt", C.c_int32),
                ("storage", C.c_int32), ("geometry", C.c_int32)]


class ViewParams(C.Structure):
    _fields_ = [("cos_t", C.c_float), ("sin_t", C.c_float), ("cos_xt", C.c_float),
                ("sin_xt", C.c_float), ("zfactor", C.c_float)]


class Timing(C.Structure):
    _fields_ = [("prep_ms", C.c_double), ("filter_ms", C.c_double), ("bp_ms", C.c_double),
                ("fp_ms", C.c_double), ("other_ms", C.c_double),
                ("bp_launches", C.c_int64), ("filter_launches", C.c_int64),
                ("fp_launches", C.c_int64)]


class Phantom(C.Structure):
    _fields_ = [("vol_nx", C.c_int32), ("vol_ny", C.c_int32), ("vol_nz", C.c_int32),
                ("n_ellipsoids", C.c_int32), ("n_beads", C.c_int32),
                ("axis_min", C.c_float), ("axis_max", C.c_float),
                ("bead_radius", C.c_float), ("bead_density", C.c_float),
                ("noise_rel", C.c_float), ("seed", C.c_uint64)]


class MrcInfo(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("mode", C.c_int32),
                ("mx", C.c_int32), ("my", C.c_int32), ("mz", C.c_int32), ("next", C.c_int32),
                ("nlabl", C.c_int32), ("swapped", C.c_int32), ("is_imod", C.c_int32),
                ("xlen", C.c_float), ("ylen", C.c_float), ("zlen", C.c_float),
                ("amin", C.c_float), ("amax", C.c_float), ("amean", C.c_float),
                ("label0", C.c_char * 81)]


class ServerStats(C.Structure):
    _fields_ = [("completed", C.c_int64), ("failed", C.c_int64), ("rounds", C.c_int64),
                ("bp_launches", C.c_int64), ("bp_batched_requests", C.c_int64),
                ("max_active", C.c_int32), ("active", C.c_int32), ("queued", C.c_int32),
                ("pad", C.c_int32), ("bytes_in_use", C.c_int64), ("budget_bytes", C.c_int64),
                ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double)]


class ViewGeometry(C.Structure):
    _fields_ = [("xtilt", C.POINTER(C.c_float)), ("n_xtilt", C.c_int32),
                ("zfactor", C.POINTER(C.c_float)), ("n_zfactor", C.c_int32),
                ("lnx", C.POINTER(C.c_int32)), ("lny", C.POINTER(C.c_int32)),
                ("lxs", C.POINTER(C.c_float)), ("lys", C.POINTER(C.c_float)),
                ("ldx", C.POINTER(C.c_float)), ("ldy", C.POINTER(C.c_float))]


class AlignView(C.Structure):
    _fields_ = [("dx", C.c_double), ("dy", C.c_double), ("rotation", C.c_double),
                ("mag", C.c_double), ("tilt", C.c_double)]


class CtfParams(C.Structure):
    _fields_ = [("voltage_kv", C.c_float), ("cs_mm", C.c_float), ("pixel_a", C.c_float),
                ("amplitude_contrast", C.c_float), ("cuton_freq", C.c_float),
                ("plate_phase_deg", C.c_float), ("tilt_axis_deg", C.c_float),
                ("defocus_nm", C.c_float), ("strip_width", C.c_int32),
                ("max_strip_width", C.c_int32)]


_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_vp = C.c_void_p

# (name, restype, argtypes) of every symbol include/tiltgpu.h declares.
SIGNATURES = [
    ("tiltgpu_last_error", C.c_char_p, []),
    ("tiltgpu_version", C.c_char_p, []),
    ("tiltgpu_device_count", C.c_int, []),
    ("tiltgpu_default_options", None, [C.POINTER(Options)]),
    ("tiltgpu_proj_params", C.c_int, [_fp, C.c_int32, _fp, C.c_int32, _fp, C.c_int32,
                                      C.POINTER(ViewParams)]),
    ("tiltgpu_geometry_tables", C.c_int, [_vp, C.POINTER(ViewParams), _fp, _fp]),
    ("tiltgpu_radial_weights", C.c_int, [C.c_int32, C.c_int32, C.c_float, C.c_float,
                                         C.c_int32, _fp]),
    ("tiltgpu_create", C.c_int, [C.POINTER(Dims), C.POINTER(Options), C.POINTER(_vp)]),
    ("tiltgpu_set_angles", C.c_int, [_vp, _fp, C.c_int32]),
    ("tiltgpu_set_xtilt", C.c_int, [_vp, _fp, C.c_int32]),
    ("tiltgpu_set_zfactor", C.c_int, [_vp, _fp, C.c_int32]),
    ("tiltgpu_set_local_alignment", C.c_int, [_vp, _ip, _ip, _fp, _fp, _fp, _fp]),
    ("tiltgpu_upload_projections", C.c_int, [_vp, _fp, C.c_int64]),
    ("tiltgpu_upload_views", C.c_int, [_vp, C.c_int32, C.c_int32, _fp]),
    ("tiltgpu_host_alloc", C.c_void_p, [C.c_int64]),
    ("tiltgpu_device_alloc", C.c_void_p, [C.c_int64]),
    ("tiltgpu_device_free", None, [C.c_void_p]),
    ("tiltgpu_copy_to_device", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64]),
    ("tiltgpu_copy_to_host", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64]),
    ("tiltgpu_host_free", None, [C.c_void_p]),
    ("tiltgpu_run", C.c_int, [_vp]),
    ("tiltgpu_run_rows", C.c_int, [_vp, C.c_int32, C.c_int32]),
    ("tiltgpu_read_rows", C.c_int, [_vp, C.c_int32, C.c_int32, _fp, _fp]),
    ("tiltgpu_sirt_residuals", C.c_int, [_vp, _dp, C.c_int32]),
    ("tiltgpu_reproject", C.c_int, [_vp, _fp, _fp]),
    ("tiltgpu_set_tuning", C.c_int, [_vp, C.c_int32, C.c_int32]),
    ("tiltgpu_get_tuning", C.c_int, [_vp, _ip]),
    ("tiltgpu_set_bp_shape", C.c_int, [_vp, C.c_int32, C.c_int32]),
    ("tiltgpu_get_bp_shape", C.c_int, [_vp, _ip]),
    ("tiltgpu_set_sirt_trace", C.c_int, [_vp, C.c_int32, C.c_int32]),
    ("tiltgpu_read_sirt_trace", C.c_int, [_vp, C.c_int32, _fp]),
    ("tiltgpu_read_filtered", C.c_int, [_vp, _fp]),
    ("tiltgpu_set_section_transforms", C.c_int, [_vp, _fp, C.c_int32, C.c_float]),
    ("tiltgpu_read_projections", C.c_int, [_vp, _fp]),
    ("tiltgpu_read_xf_file", C.c_int, [C.c_char_p, _fp, C.c_int32]),
    ("tiltgpu_read_means", C.c_int, [_vp, _fp]),
    ("tiltgpu_synchronize", C.c_int, [_vp]),
    ("tiltgpu_get_timing", C.c_int, [_vp, C.POINTER(Timing)]),
    ("tiltgpu_reset_timing", C.c_int, [_vp]),
    ("tiltgpu_device_bytes", C.c_int64, [_vp]),
    ("tiltgpu_destroy", None, [_vp]),
    ("tiltgpu_bp_session_new", C.c_int, [_fp, C.POINTER(ViewParams), _fp, C.c_int32,
                                         C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                         C.c_int32, C.c_int32, C.POINTER(_vp)]),
    ("tiltgpu_reconstruct_row", C.c_int, [_vp, C.c_int32, _fp]),
    ("tiltgpu_reconstruct_rows", C.c_int, [_vp, C.c_int32, C.c_int32, _fp]),
    ("tiltgpu_reconstruct_rows_noalloc", C.c_int, [_vp, C.POINTER(ViewParams), _fp, C.c_int32,
                                                   C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                                   C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                                   _vp, _vp]),
    ("tiltgpu_server_create", C.c_int, [C.c_int64, C.POINTER(_vp)]),
    ("tiltgpu_server_submit", C.c_int, [_vp, C.POINTER(Dims), C.POINTER(Options), _fp,
                                        C.c_int32, _fp, _fp, C.POINTER(C.c_int64)]),
    ("tiltgpu_server_submit_geometry", C.c_int, [_vp, C.POINTER(Dims), C.POINTER(Options), _fp,
                                                 C.c_int32, C.POINTER(ViewGeometry), _fp, _fp,
                                                 C.POINTER(C.c_int64)]),
    ("tiltgpu_server_add_series", C.c_int, [_vp, C.POINTER(Dims), C.POINTER(Options), _fp,
                                            C.c_int32, _fp, C.c_int32, _ip, _ip, _fp, _fp,
                                            _fp, _fp, _fp, _ip]),
    ("tiltgpu_server_submit_slab", C.c_int, [_vp, C.c_int32, C.c_int32, C.c_int32, _fp,
                                             C.POINTER(C.c_int64)]),
    ("tiltgpu_server_poll", C.c_int, [_vp, C.c_int64]),
    ("tiltgpu_server_wait", C.c_int, [_vp, C.c_int64, _dp]),
    ("tiltgpu_server_get_stats", C.c_int, [_vp, C.POINTER(ServerStats)]),
    ("tiltgpu_server_reset_stats", C.c_int, [_vp]),
    ("tiltgpu_server_destroy", None, [_vp]),
    ("tiltgpu_copy_calibration", C.c_int, [C.c_int64, C.c_int32, _dp]),
    ("tiltgpu_fma_calibration", C.c_int, [C.c_int32, _dp]),
    ("tiltgpu_issue_calibration", C.c_int, [C.c_int32, C.c_int32, _dp]),
    ("tiltgpu_ctf_default_params", None, [C.POINTER(CtfParams)]),
    ("tiltgpu_electron_wavelength", C.c_double, [C.c_float]),
    ("tiltgpu_ctf_strip_width", C.c_int, [C.POINTER(CtfParams), C.c_int32, _fp, C.c_int32]),
    ("tiltgpu_ctf_strip_layout", C.c_int, [C.c_int32, C.c_int32, C.c_float, C.c_int32, _ip, _dp,
                                           C.c_int32]),
    ("tiltgpu_ctf_create", C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.POINTER(CtfParams), _fp,
                                     C.c_int32, _fp, C.c_int32, C.POINTER(_vp)]),
    ("tiltgpu_ctf_strip_width_used", C.c_int, [_vp]),
    ("tiltgpu_ctf_upload", C.c_int, [_vp, C.c_int32, C.c_int32, _fp]),
    ("tiltgpu_ctf_download", C.c_int, [_vp, C.c_int32, C.c_int32, _fp]),
    ("tiltgpu_ctf_run", C.c_int, [_vp]),
    ("tiltgpu_ctf_last_run", C.c_int, [_vp, _dp, C.POINTER(C.c_int64)]),
    ("tiltgpu_ctf_destroy", None, [_vp]),
    ("tiltgpu_ctf_phaseflip", C.c_int, [_fp, _fp, C.c_int32, C.c_int32, C.c_int32, _fp,
                                        C.c_int32, _fp, C.c_int32, C.POINTER(CtfParams)]),
    ("tiltgpu_read_defocus_file", C.c_int, [C.c_char_p, _ip, _fp, C.c_int32]),
    ("tiltgpu_interpolate_defocus", C.c_int, [_ip, _fp, C.c_int32, C.c_int32, _fp]),
    ("tiltgpu_read_tilt_file", C.c_int, [C.c_char_p, _fp, C.c_int32]),
    ("tiltgpu_read_per_view_file", C.c_int, [C.c_char_p, _fp, C.c_int32]),
    ("tiltgpu_local_corrections", C.c_int, [_dp, _dp, C.POINTER(AlignView), C.c_int32,
                                            C.c_int32, C.c_int32, C.c_float, C.c_double,
                                            C.c_double, C.c_int32, C.c_double, C.c_double,
                                            C.c_int32, _ip, _ip, _ip, _dp, _dp]),
    ("tiltgpu_write_local_corrections", C.c_int, [C.c_char_p, C.c_int32, _ip, _ip, _ip, _dp,
                                                  _dp]),
    ("tiltgpu_parse_local_file", C.c_int, [C.c_char_p, C.c_int32, _ip, _ip, _ip, _fp, _fp,
                                           _fp, _fp]),
    ("tiltgpu_mrc_info_read", C.c_int, [C.c_char_p, C.POINTER(MrcInfo)]),
    ("tiltgpu_mrc_read_sections", C.c_int, [C.c_char_p, C.c_int32, C.c_int32, _fp, C.c_int64]),
    ("tiltgpu_mrc_write_float", C.c_int, [C.c_char_p, _fp, C.c_int32, C.c_int32, C.c_int32,
                                          C.c_float, C.c_float, C.c_float, C.c_char_p]),
    ("tiltgpu_synth_series", C.c_int, [C.POINTER(Phantom), C.c_int32, C.c_int32, C.c_int32,
                                       _fp, _fp, _fp, _fp, C.c_int32, _fp]),
]

_lib = None


def lib():
    """Load libtiltgpu.so (built by imod-rs_amd/build.py); raises if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TiltGpuError(ERR_STATE, f"{LIB_PATH} is missing: run imod-rs_amd/build.py")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc < 0:
        raise TiltGpuError(rc, lib().tiltgpu_last_error().decode(errors="replace"))
    return rc


def _f(a):
    return None if a is None else a.ctypes.data_as(_fp)


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


def device_count() -> int:
    return lib().tiltgpu_device_count()


@dataclass
class TiltOptions:
    """apps/tilt/src/main.rs:18-120 (file names aside) plus ``storage``."""
    thickness: int = 0
    width: int | None = None
    radial: tuple = (0.35, 0.05)
    no_weight: bool = False
    sirt: int = 0
    sirt_relax: float = 1.0
    no_cosine_stretch: bool = False
    log: bool = False
    densweight: float = 1.0
    outscale: float = 1.0
    outadd: float = 0.0
    exact_filter_size: int | None = None
    fill_mode: str = "mean"
    cosine_interp: str = "linear"
    max_cosine_tilt: float = 80.0
    hamming: bool = False
    fake_sirt: int = 0
    storage: str = "f32"
    geometry: str = "f32"  # per-view tables: "f32" (the reference's) or "f64" (GPU, rounded once)

    def to_c(self) -> Options:
        o = Options()
        lib().tiltgpu_default_options(C.byref(o))
        o.radial_cutoff, o.radial_falloff = float(self.radial[0]), float(self.radial[1])
        o.no_weight = int(self.no_weight)
        o.exact_filter_size = int(self.exact_filter_size or 0)
        o.hamming = int(self.hamming)
        o.no_cosine_stretch = int(self.no_cosine_stretch)
        o.cosine_cubic = int(self.cosine_interp.lower() == "cubic")
        o.max_cosine_tilt = float(self.max_cosine_tilt)
        o.log = int(self.log)
        o.fill_mode = FILL.get(self.fill_mode.lower(), 0)  # unknown -> mean (main.rs:593)
        o.densweight, o.outscale, o.outadd = self.densweight, self.outscale, self.outadd
        o.sirt, o.sirt_relax, o.fake_sirt = int(self.sirt), float(self.sirt_relax), int(self.fake_sirt)
        o.storage = STORAGE[self.storage]
        o.geometry = GEOMETRY.get(self.geometry.lower(), -1)  # unknown: refused by the library
        return o


def proj_params(angles, xtilt=None, zfactor=None):
    ang = _f32(angles)
    xt, zf = _f32(xtilt), _f32(zfactor)
    out = (ViewParams * max(1, ang.size))()
    _check(lib().tiltgpu_proj_params(_f(ang), ang.size, _f(xt), 0 if xt is None else xt.size,
                                     _f(zf), 0 if zf is None else zf.size, out))
    return out


def radial_weights(n, cutoff=0.35, falloff=0.05, exact_filter_size=0, hamming=False):
    out = np.zeros(n, np.float32)
    _check(lib().tiltgpu_radial_weights(n, exact_filter_size or 0, cutoff, falloff,
                                        int(hamming), _f(out)))
    return out


class PinnedArray:
    """A float32 numpy array in page-locked host memory (tiltgpu_host_alloc)."""

    def __init__(self, shape):
        n = int(np.prod(shape))
        self.ptr = lib().tiltgpu_host_alloc(max(4, 4 * n))
        if not self.ptr:
            raise TiltGpuError(-2, lib().tiltgpu_last_error().decode())
        self.array = np.ctypeslib.as_array((C.c_float * n).from_address(self.ptr)).reshape(shape)

    def close(self):
        if getattr(self, "ptr", None):
            self.array = None
            lib().tiltgpu_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Session:
    """One reconstruction resident on the GPU (tiltgpu_create ... tiltgpu_destroy)."""

    def __init__(self, in_nx, in_ny, nz, opts: TiltOptions | None = None):
        self.opts = opts or TiltOptions()
        self.out_nx = self.opts.width or in_nx
        self.out_nz = self.opts.thickness if self.opts.thickness > 0 else in_nx
        self.dims = Dims(in_nx, in_ny, nz, self.out_nx, self.out_nz)
        self._trace_rows = (0, 0)
        self._n_angles = 0
        self._copts = self.opts.to_c()
        h = _vp()
        _check(lib().tiltgpu_create(C.byref(self.dims), C.byref(self._copts), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().tiltgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_angles(self, a):
        a = _f32(a)
        _check(lib().tiltgpu_set_angles(self.h, _f(a), a.size))
        self._n_angles = a.size

    def set_xtilt(self, a):
        a = _f32(a)
        _check(lib().tiltgpu_set_xtilt(self.h, _f(a), a.size))

    def set_zfactor(self, a):
        a = _f32(a)
        _check(lib().tiltgpu_set_zfactor(self.h, _f(a), a.size))

    def set_local(self, nx, ny, xs, ys, dx, dy):
        nx = np.ascontiguousarray(nx, np.int32)
        ny = np.ascontiguousarray(ny, np.int32)
        self._local = [_f32(v) if np.size(v) else np.zeros(1, np.float32) for v in (xs, ys, dx, dy)]
        _check(lib().tiltgpu_set_local_alignment(
            self.h, nx.ctypes.data_as(_ip), ny.ctypes.data_as(_ip), *[_f(v) for v in self._local]))

    def upload(self, P):
        P = _f32(P)
        _check(lib().tiltgpu_upload_projections(self.h, _f(P), P.size))

    def upload_views(self, v0, P):
        """Raw projections of sections [v0, v0 + len(P)) (a stack uploaded in chunks)."""
        P = _f32(P)
        _check(lib().tiltgpu_upload_views(self.h, v0, P.shape[0], _f(P)))

    def run(self):
        _check(lib().tiltgpu_run(self.h))

    def run_rows(self, y0, y1):
        """WBP of output rows [y0, y1) (filtering once per upload); read them with read_rows."""
        _check(lib().tiltgpu_run_rows(self.h, y0, y1))

    def read_rows(self, y0=0, y1=None, stats=True):
        y1 = self.dims.in_ny if y1 is None else y1
        out = np.empty((y1 - y0, self.out_nz, self.out_nx), np.float32)
        st = np.empty((y1 - y0, 3), np.float32) if stats else None
        _check(lib().tiltgpu_read_rows(self.h, y0, y1, _f(out), _f(st)))
        return out, st

    def residuals(self):
        out = np.zeros(max(1, self.opts.sirt), np.float64)
        k = _check(lib().tiltgpu_sirt_residuals(self.h, out.ctypes.data_as(_dp), out.size))
        return out[:k]

    def geometry_tables(self):
        """(views [nviews][5] cos_t, sin_t, cos_xt, sin_xt, zfactor; stretch [nz];
        zc [nviews][out_nz]) the session reconstructs with."""
        nv = min(self.dims.nz, self._n_angles)
        vp = (ViewParams * max(1, nv))()
        st = np.zeros(self.dims.nz, np.float32)
        zc = np.zeros((max(1, nv), self.out_nz), np.float32)
        _check(lib().tiltgpu_geometry_tables(self.h, vp, _f(st), _f(zc)))
        views = np.array([[v.cos_t, v.sin_t, v.cos_xt, v.sin_xt, v.zfactor] for v in vp[:nv]],
                         np.float32).reshape(nv, 5)
        return views, st, zc[:nv]

    def set_tuning(self, bp_tile_z=0, fp_bins_per_lane=0):
        """Force launch shapes (0 = tuned table / analytic rule)."""
        _check(lib().tiltgpu_set_tuning(self.h, bp_tile_z, fp_bins_per_lane))

    def set_bp_shape(self, views_per_stage=0, x_per_lane=0):
        """Force the backprojection launch shape beyond its depth (0 = tuned table)."""
        _check(lib().tiltgpu_set_bp_shape(self.h, views_per_stage, x_per_lane))

    def bp_shape(self):
        """{'tile_z', 'views_per_stage', 'x_per_lane'} of the next run."""
        out = np.zeros(3, np.int32)
        _check(lib().tiltgpu_get_bp_shape(self.h, out.ctypes.data_as(_ip)))
        return dict(zip(("tile_z", "views_per_stage", "x_per_lane"), map(int, out)))

    def tuning(self):
        """{'tile_z', 'tile_source', 'fp_bins', 'fp_source'}; source 0 rule, 1 table, 2 forced."""
        out = np.zeros(4, np.int32)
        _check(lib().tiltgpu_get_tuning(self.h, out.ctypes.data_as(_ip)))
        return dict(zip(("tile_z", "tile_source", "fp_bins", "fp_source"), map(int, out)))

    def set_sirt_trace(self, y0, y1):
        """Record rows [y0, y1) after the WBP and every SIRT iteration of run()."""
        _check(lib().tiltgpu_set_sirt_trace(self.h, y0, y1))
        self._trace_rows = (y0, y1)

    def sirt_trace(self):
        """[sirt_done+1][rows][out_nz][out_nx] recorded by the last run()."""
        y0, y1 = self._trace_rows
        k = self.residuals().size
        out = np.empty((k + 1, y1 - y0, self.out_nz, self.out_nx), np.float32)
        for i in range(k + 1):
            _check(lib().tiltgpu_read_sirt_trace(self.h, i, _f(out[i])))
        return out

    def reproject(self, vol):
        """Forward projection of vol [in_ny][out_nz][out_nx] (main.rs:830-896)."""
        vol = np.ascontiguousarray(vol, np.float32)
        if vol.shape != (self.dims.in_ny, self.out_nz, self.out_nx):
            raise ValueError(f"volume shape {vol.shape} != "
                             f"{(self.dims.in_ny, self.out_nz, self.out_nx)}")
        out = np.empty((self.dims.nz, self.dims.in_ny, self.dims.in_nx), np.float32)
        _check(lib().tiltgpu_reproject(self.h, _f(vol), _f(out)))
        return out

    def set_section_transforms(self, xf, fill=0.0):
        """newstack -x transforms [n][6] applied to the uploaded stack (n = 0 removes them)."""
        x = np.ascontiguousarray(np.asarray(xf, np.float32).reshape(-1, 6))
        _check(lib().tiltgpu_set_section_transforms(self.h, _f(x), x.shape[0], fill))

    def read_projections(self):
        """The stack the reconstruction reads (after the section transforms)."""
        out = np.empty((self.dims.nz, self.dims.in_ny, self.dims.in_nx), np.float32)
        _check(lib().tiltgpu_read_projections(self.h, _f(out)))
        return out

    def read_filtered(self):
        out = np.empty((self.dims.nz, self.dims.in_ny, self.dims.in_nx), np.float32)
        _check(lib().tiltgpu_read_filtered(self.h, _f(out)))
        return out

    def read_means(self):
        out = np.empty(self.dims.nz, np.float32)
        _check(lib().tiltgpu_read_means(self.h, _f(out)))
        return out

    def synchronize(self):
        _check(lib().tiltgpu_synchronize(self.h))

    def timing(self):
        t = Timing()
        _check(lib().tiltgpu_get_timing(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in Timing._fields_}

    def reset_timing(self):
        _check(lib().tiltgpu_reset_timing(self.h))

    def device_bytes(self):
        return lib().tiltgpu_device_bytes(self.h)


def header_stats(row_stats, per_row):
    """Header amin/amax/amean from per-section (min, max, mean) as main.rs:1080-1092
    combines them: f64 running sum of smean * (out_nx*out_nz)."""
    gmin = float(np.min(row_stats[:, 0])) if len(row_stats) else float(np.finfo(np.float32).max)
    gmax = float(np.max(row_stats[:, 1])) if len(row_stats) else float(np.finfo(np.float32).min)
    gsum = 0.0
    for m in row_stats[:, 2]:
        gsum += float(m) * float(per_row)
    gmean = float(np.float32(gsum / (per_row * len(row_stats)))) if len(row_stats) else 0.0
    return gmin, gmax, gmean


def reconstruct(P, angles, opts: TiltOptions | None = None, xtilt=None, zfactor=None,
                local=None, return_session=False, trace_rows=None):
    """The whole reconstruction on the GPU.  P: raw projections [nz][ny][nx];
    local: (nx, ny, xs, ys, dx, dy) flat per-view grids or an object with .flat();
    trace_rows=(y0, y1) records those rows after every SIRT iteration
    (Session.sirt_trace, with return_session=True)."""
    P = _f32(P)
    nz, ny, nx = P.shape
    s = Session(nx, ny, nz, opts)
    try:
        s.set_angles(angles)
        if xtilt is not None:
            s.set_xtilt(xtilt)
        if zfactor is not None:
            s.set_zfactor(zfactor)
        if local is not None:
            s.set_local(*(local.flat() if hasattr(local, "flat") else local))
        if trace_rows is not None:
            s.set_sirt_trace(*trace_rows)
        s.upload(P)
        s.run()
        vol, st = s.read_rows()
    except BaseException:
        s.close()  # release the device buffers now, not when the traceback goes
        raise
    hs = header_stats(st, s.out_nx * s.out_nz)
    if return_session:
        return vol, hs, s
    s.close()
    return vol, hs


class GpuSession:
    """Shape of GpuSession (apps/tilt/src/gpu.rs:26-269) over tiltgpu_bp_session_new."""

    def __init__(self, h, out_nx, out_nz):
        self.h, self.out_nx, self.out_nz = h, out_nx, out_nz

    @classmethod
    def new(cls, projections, params, proj_means, in_nx, in_ny, n_projs, out_nx, out_nz,
            fill_mode="mean"):
        """Returns None when no session can be created (gpu.rs:43-74 -> Option)."""
        P = _f32(projections)
        m = _f32(proj_means)
        if not isinstance(params, C.Array):
            params = proj_params(*params)
        h = _vp()
        rc = lib().tiltgpu_bp_session_new(_f(P), params, _f(m), len(params), in_nx, in_ny,
                                          n_projs, out_nx, out_nz, FILL[fill_mode], C.byref(h))
        if rc != OK:
            return None
        return cls(h, out_nx, out_nz)

    def reconstruct_row(self, iy):
        out = np.empty(self.out_nx * self.out_nz, np.float32)
        _check(lib().tiltgpu_reconstruct_row(self.h, iy, _f(out)))
        return out

    def reconstruct_rows(self, y0, y1):
        out = np.empty((y1 - y0, self.out_nz, self.out_nx), np.float32)
        _check(lib().tiltgpu_reconstruct_rows(self.h, y0, y1, _f(out)))
        return out

    def close(self):
        if self.h:
            lib().tiltgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """A float32 device buffer from tiltgpu_device_alloc (for the device-pointer
    entry points): upload from / download into numpy arrays."""

    def __init__(self, shape):
        self.shape = tuple(int(x) for x in shape)
        self.nbytes = int(np.prod(self.shape)) * 4
        self.ptr = lib().tiltgpu_device_alloc(max(4, self.nbytes))
        if not self.ptr:
            raise TiltGpuError(ERR_HIP, lib().tiltgpu_last_error().decode(errors="replace"))

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        assert a.nbytes == self.nbytes
        _check(lib().tiltgpu_copy_to_device(self.ptr, a.ctypes.data, a.nbytes))

    def download(self):
        out = np.empty(self.shape, np.float32)
        _check(lib().tiltgpu_copy_to_host(out.ctypes.data, self.ptr, self.nbytes))
        return out

    def close(self):
        if self.ptr:
            lib().tiltgpu_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def reconstruct_rows_noalloc(d_proj, params, proj_means, in_nx, in_ny, n_projs, out_nx, out_nz,
                             y0, y1, d_out, fill_mode="mean"):
    """tiltgpu_reconstruct_rows_noalloc on DeviceArrays: d_proj [n_projs][in_ny][in_nx],
    d_out [y1-y0][out_nz][out_nx]; params a ViewParams array (proj_params)."""
    m = _f32(proj_means)
    _check(lib().tiltgpu_reconstruct_rows_noalloc(
        d_proj.ptr, params, _f(m), len(params), in_nx, in_ny, n_projs, out_nx, out_nz,
        FILL[fill_mode], y0, y1, d_out.ptr, None))


def local_corrections(obs, beads, views, patch_size=500, overlap=0.5, cx=0.0, cy=0.0,
                      skip_edge=0, img_w=0.0, img_h=0.0):
    """tiltalign's --output-local corrections (apps/tiltalign/src/main.rs:864-940,
    1374-1402).  obs [n_beads][n_views][2] centred observations (NaN x: none),
    beads [n_beads][3] centred, views: sequence of (dx, dy, rotation, mag, tilt).
    Returns a list of (view, patch_x, patch_y, dx, dy)."""
    obs = np.ascontiguousarray(obs, np.float64)
    beads = np.ascontiguousarray(beads, np.float64)
    nb, nv = obs.shape[0], obs.shape[1]
    av = (AlignView * max(1, nv))(*[AlignView(*map(float, v)) for v in views])
    cap = 1024
    while True:
        ov, opx, opy = (np.zeros(cap, np.int32) for _ in range(3))
        odx, ody = np.zeros(cap), np.zeros(cap)
        n = _check(lib().tiltgpu_local_corrections(
            obs.ctypes.data_as(_dp), beads.ctypes.data_as(_dp), av, nb, nv, patch_size, overlap,
            cx, cy, skip_edge, img_w, img_h, cap, ov.ctypes.data_as(_ip), opx.ctypes.data_as(_ip),
            opy.ctypes.data_as(_ip), odx.ctypes.data_as(_dp), ody.ctypes.data_as(_dp)))
        if n <= cap:
            return [(int(ov[i]), int(opx[i]), int(opy[i]), float(odx[i]), float(ody[i]))
                    for i in range(n)]
        cap = n


def write_local_corrections(path, corrections):
    """The --output-local file of tiltalign (main.rs:1404-1409)."""
    n = len(corrections)
    cols = list(zip(*corrections)) if n else [[]] * 5
    v, px, py = (np.ascontiguousarray(c, np.int32) if n else np.zeros(1, np.int32) for c in cols[:3])
    dx, dy = (np.ascontiguousarray(c, np.float64) if n else np.zeros(1) for c in cols[3:])
    _check(lib().tiltgpu_write_local_corrections(
        str(path).encode(), n, v.ctypes.data_as(_ip), px.ctypes.data_as(_ip),
        py.ctypes.data_as(_ip), dx.ctypes.data_as(_dp), dy.ctypes.data_as(_dp)))


def read_tilt_file(path, cap=100000):
    out = np.zeros(cap, np.float32)
    n = _check(lib().tiltgpu_read_tilt_file(path.encode(), _f(out), cap))
    return out[:min(n, cap)]


def read_per_view_file(path, cap=100000):
    out = np.zeros(cap, np.float32)
    n = _check(lib().tiltgpu_read_per_view_file(path.encode(), _f(out), cap))
    return out[:min(n, cap)]


def synth_series(nx, ny, angles, vol_nz, n_ellipsoids=200, n_beads=30, seed=0, noise_rel=0.1,
                 xtilt=None, shift_x=None, shift_y=None, vol_nx=None, axis=(5.0, 40.0),
                 bead_radius=5.0, bead_density=10.0, nthreads=0):
    """Seeded ellipsoid + bead phantom projected analytically with the tilt
    geometry; returns [n_views][ny][nx] float32."""
    ang = _f32(angles)
    ph = Phantom(vol_nx or nx, ny, vol_nz, n_ellipsoids, n_beads, axis[0], axis[1],
                 bead_radius, bead_density, noise_rel, seed)
    out = np.empty((ang.size, ny, nx), np.float32)
    xt, sx, sy = _f32(xtilt), _f32(shift_x), _f32(shift_y)
    _check(lib().tiltgpu_synth_series(C.byref(ph), nx, ny, ang.size, _f(ang), _f(xt), _f(sx),
                                      _f(sy), nthreads, _f(out)))
    return out


def parse_local_file(path, nz):
    """Per-view local-alignment grids (nx, ny, xs, ys, dx, dy) of a --local-file."""
    sizes = np.zeros(3, np.int32)
    _check(lib().tiltgpu_parse_local_file(path.encode(), nz, sizes.ctypes.data_as(_ip), None,
                                          None, None, None, None, None))
    nx = np.zeros(nz, np.int32)
    ny = np.zeros(nz, np.int32)
    xs, ys = np.zeros(max(1, sizes[0]), np.float32), np.zeros(max(1, sizes[1]), np.float32)
    dx, dy = np.zeros(max(1, sizes[2]), np.float32), np.zeros(max(1, sizes[2]), np.float32)
    _check(lib().tiltgpu_parse_local_file(path.encode(), nz, sizes.ctypes.data_as(_ip),
                                          nx.ctypes.data_as(_ip), ny.ctypes.data_as(_ip),
                                          _f(xs), _f(ys), _f(dx), _f(dy)))
    return nx, ny, xs[:sizes[0]], ys[:sizes[1]], dx[:sizes[2]], dy[:sizes[2]]


def mrc_info(path):
    info = MrcInfo()
    _check(lib().tiltgpu_mrc_info_read(path.encode(), C.byref(info)))
    return info


def mrc_read(path):
    """All sections of an MRC file as float32 [nz][ny][nx] (modes with one value
    per pixel)."""
    info = mrc_info(path)
    out = np.empty((info.nz, info.ny, info.nx), np.float32)
    _check(lib().tiltgpu_mrc_read_sections(path.encode(), 0, info.nz, _f(out), out.size))
    return out


def read_xf_file(path):
    """Transforms of an .xf file as [n][6] (crates/imod-transforms/src/xf.rs:147-174)."""
    n = _check(lib().tiltgpu_read_xf_file(path.encode(), None, 0))
    out = np.zeros((max(n, 1), 6), np.float32)
    _check(lib().tiltgpu_read_xf_file(path.encode(), _f(out), n))
    return out[:n]


def mrc_write(path, vol, xlen=0.0, ylen=0.0, zlen=0.0, label=None):
    v = _f32(vol)
    nz, ny, nx = v.shape
    _check(lib().tiltgpu_mrc_write_float(path.encode(), _f(v), nx, ny, nz, xlen, ylen, zlen,
                                         None if label is None else label.encode()))


def copy_calibration(nbytes, iters=1):
    """ms per launch of the library's plain device copy kernel (diagnostic)."""
    ms = C.c_double()
    _check(lib().tiltgpu_copy_calibration(int(nbytes), int(iters), C.byref(ms)))
    return ms.value


def fma_calibration(iters=1 << 14):
    """Attainable non-packed FP32 FMA rate of the device (VGPR operands, both
    VALU pipes), TFLOP/s."""
    t = C.c_double()
    _check(lib().tiltgpu_fma_calibration(iters, C.byref(t)))
    return t.value


ISSUE_MIXES = {0: "fp32_fma", 1: "fract_alone", 2: "fract_with_fp32_fma",
               3: "fract_with_int_add", 4: "bp_two_row_voxel_step"}


def issue_calibration(iters=1 << 13):
    """ns per wave instruction per SIMD of each calibration mix (8 waves/SIMD)."""
    out = {}
    for mix, name in ISSUE_MIXES.items():
        t = C.c_double()
        n = iters if mix != 4 else max(1, iters // 8)
        _check(lib().tiltgpu_issue_calibration(mix, n, C.byref(t)))
        out[name] = t.value
    return out


def read_defocus_file(path):
    """(entries [n][3] f32, views [n] int) of a ctfphaseflip defocus file."""
    n = lib().tiltgpu_read_defocus_file(str(path).encode(), None, None, 0)
    if n <= -100:
        raise TiltGpuError(ERR_IO, f"{path}: malformed line {-100 - n}")
    _check(n)
    views = np.zeros(max(n, 1), np.int32)
    ent = np.zeros((max(n, 1), 3), np.float32)
    lib().tiltgpu_read_defocus_file(str(path).encode(), views.ctypes.data_as(_ip), _f(ent), n)
    return ent[:n], views[:n]


def interpolate_defocus(entries, views, n_views):
    ent = np.ascontiguousarray(np.asarray(entries, np.float32).reshape(-1, 3))
    vw = np.ascontiguousarray(np.asarray(views, np.int32))
    out = np.zeros((max(n_views, 1), 3), np.float32)
    n = _check(lib().tiltgpu_interpolate_defocus(vw.ctypes.data_as(_ip), _f(ent), vw.size,
                                                  int(n_views), _f(out)))
    return out[:n]


def ctf_params(pixel_a=1.0, voltage_kv=300.0, cs_mm=2.7, defocus_nm=-3000.0, strip_width=256,
               amp=0.07, cuton=0.0, plate_deg=0.0, tilt_axis_deg=0.0, max_strip_width=0):
    """tiltgpu_ctf_params with ctfphaseflip's option names (main.rs:20-89);
    max_strip_width 0 selects the automatic limit."""
    p = CtfParams()
    lib().tiltgpu_ctf_default_params(C.byref(p))
    p.pixel_a, p.voltage_kv, p.cs_mm, p.defocus_nm = pixel_a, voltage_kv, cs_mm, defocus_nm
    p.strip_width, p.amplitude_contrast, p.cuton_freq = int(strip_width), amp, cuton
    p.plate_phase_deg, p.tilt_axis_deg, p.max_strip_width = plate_deg, tilt_axis_deg, int(max_strip_width)
    return p


def electron_wavelength(voltage_kv):
    return float(lib().tiltgpu_electron_wavelength(float(voltage_kv)))


def _vdef(vdef):
    if vdef is None:
        return None, 0
    v = np.ascontiguousarray(np.asarray(vdef, np.float32).reshape(-1, 3))
    return v, v.shape[0]


def ctf_strip_width(params: CtfParams, nx, vdef=None):
    v, n = _vdef(vdef)
    return _check(lib().tiltgpu_ctf_strip_width(C.byref(params), int(nx), _f(v), n))


def ctf_strip_layout(nx, ny, tilt_axis_deg, strip_width):
    """(boxes [n][4] = bx0, bx1, by0, by1; centres [n]) of the strips."""
    n = _check(lib().tiltgpu_ctf_strip_layout(int(nx), int(ny), float(tilt_axis_deg),
                                              int(strip_width), None, None, 0))
    boxes = np.zeros((max(n, 1), 4), np.int32)
    cen = np.zeros(max(n, 1), np.float64)
    _check(lib().tiltgpu_ctf_strip_layout(int(nx), int(ny), float(tilt_axis_deg), int(strip_width),
                                          boxes.ctypes.data_as(_ip), cen.ctypes.data_as(_dp), n))
    return boxes[:n], cen[:n]


class CtfCorrector:
    """A device-resident stack for the CTF phase flip (tiltgpu_ctf_*)."""

    def __init__(self, nx, ny, nz, params: CtfParams, tilt_deg=None, vdef=None):
        self._h = _vp()
        t = _f32(tilt_deg)
        v, n = _vdef(vdef)
        self.shape = (int(nz), int(ny), int(nx))
        _check(lib().tiltgpu_ctf_create(int(nx), int(ny), int(nz), C.byref(params), _f(t),
                                        0 if t is None else t.size, _f(v), n, C.byref(self._h)))

    @property
    def strip_width(self):
        return _check(lib().tiltgpu_ctf_strip_width_used(self._h))

    def upload(self, P, z0=0):
        a = _f32(P)
        nsec = a.shape[0] if a.ndim == 3 else 1
        _check(lib().tiltgpu_ctf_upload(self._h, int(z0), nsec, _f(a)))

    def run(self):
        _check(lib().tiltgpu_ctf_run(self._h))

    def download(self, z0=0, nsec=None):
        nz, ny, nx = self.shape
        nsec = nz - z0 if nsec is None else nsec
        out = np.empty((nsec, ny, nx), np.float32)
        _check(lib().tiltgpu_ctf_download(self._h, int(z0), int(nsec), _f(out)))
        return out

    def last_run(self):
        ms = C.c_double()
        n = C.c_int64()
        _check(lib().tiltgpu_ctf_last_run(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def close(self):
        if self._h:
            lib().tiltgpu_ctf_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ctf_phaseflip(P, tilt_deg, pixel_a, voltage_kv=300.0, cs_mm=2.7, defocus_nm=-3000.0,
                  vdef=None, strip_width=256, amp=0.07, cuton=0.0, plate_deg=0.0,
                  tilt_axis_deg=0.0, max_strip_width=0):
    """CTF phase flip of a host stack [nz][ny][nx] on the GPU (one call)."""
    a = _f32(P)
    nz, ny, nx = a.shape
    p = ctf_params(pixel_a, voltage_kv, cs_mm, defocus_nm, strip_width, amp, cuton, plate_deg,
                   tilt_axis_deg, max_strip_width)
    t = _f32(tilt_deg)
    v, n = _vdef(vdef)
    out = np.empty_like(a)
    _check(lib().tiltgpu_ctf_phaseflip(_f(a), _f(out), nx, ny, nz, _f(t),
                                       0 if t is None else t.size, _f(v), n, C.byref(p)))
    return out


def _view_geometry(xtilt, zfactor, local):
    """A ViewGeometry over numpy copies (returned too: they must outlive the call)."""
    xt, zf = _f32(xtilt), _f32(zfactor)
    g = ViewGeometry()
    keep = [xt, zf]
    if xt is not None:
        g.xtilt, g.n_xtilt = _f(xt), xt.size
    if zf is not None:
        g.zfactor, g.n_zfactor = _f(zf), zf.size
    if local is not None:
        parts = local.flat() if hasattr(local, "flat") else local
        lnx = np.ascontiguousarray(parts[0], np.int32)
        lny = np.ascontiguousarray(parts[1], np.int32)
        fl = [_f32(v) if np.size(v) else np.zeros(1, np.float32) for v in parts[2:]]
        keep += [lnx, lny] + fl
        g.lnx, g.lny = lnx.ctypes.data_as(_ip), lny.ctypes.data_as(_ip)
        g.lxs, g.lys, g.ldx, g.ldy = [_f(v) for v in fl]
    return g, keep


class Server:
    """Continuous-batching server (tiltgpu_server_*): requests from any thread are
    admitted by device-memory budget and batched into shared launches."""

    def __init__(self, budget_bytes):
        h = _vp()
        _check(lib().tiltgpu_server_create(int(budget_bytes), C.byref(h)))
        self.h = h
        self._pending = {}
        self._series_shape = {}
        self._series_keep = []

    def submit(self, P, angles, opts: TiltOptions | None = None, xtilt=None, zfactor=None,
               local=None):
        """Queue a whole reconstruction (with optional x-axis tilts, z factors and
        local grids); returns the request id."""
        P = _f32(P)
        ang = _f32(angles)
        opts = opts or TiltOptions()
        nz, ny, nx = P.shape
        out_nx = opts.width or nx
        out_nz = opts.thickness if opts.thickness > 0 else nx
        d = Dims(nx, ny, nz, out_nx, out_nz)
        o = opts.to_c()
        out = np.empty((ny, out_nz, out_nx), np.float32)
        rid = C.c_int64()
        if xtilt is None and zfactor is None and local is None:
            _check(lib().tiltgpu_server_submit(self.h, C.byref(d), C.byref(o), _f(ang),
                                               ang.size, _f(P), _f(out), C.byref(rid)))
        else:
            g, keep = _view_geometry(xtilt, zfactor, local)
            _check(lib().tiltgpu_server_submit_geometry(self.h, C.byref(d), C.byref(o), _f(ang),
                                                        ang.size, C.byref(g), _f(P), _f(out),
                                                        C.byref(rid)))
            del keep  # copied at submit
        self._pending[rid.value] = (out, P)  # the server reads P until the request completes
        return rid.value

    def add_series(self, P, angles, opts: TiltOptions | None = None, xtilt=None, local=None):
        P = _f32(P)
        ang = _f32(angles)
        opts = opts or TiltOptions()
        nz, ny, nx = P.shape
        d = Dims(nx, ny, nz, opts.width or nx, opts.thickness if opts.thickness > 0 else nx)
        o = opts.to_c()
        xt = _f32(xtilt)
        loc = [None] * 6
        if local is not None:
            parts = local.flat() if hasattr(local, "flat") else local
            loc = [np.ascontiguousarray(parts[0], np.int32), np.ascontiguousarray(parts[1], np.int32)] + \
                  [_f32(v) if np.size(v) else np.zeros(1, np.float32) for v in parts[2:]]
        self._series_keep.append((loc, xt))
        sid = C.c_int32()
        _check(lib().tiltgpu_server_add_series(
            self.h, C.byref(d), C.byref(o), _f(ang), ang.size, _f(xt),
            0 if xt is None else xt.size,
            None if loc[0] is None else loc[0].ctypes.data_as(_ip),
            None if loc[1] is None else loc[1].ctypes.data_as(_ip),
            *[None if v is None else _f(v) for v in loc[2:]], _f(P), C.byref(sid)))
        self._series_shape[sid.value] = (d.out_nz, d.out_nx)
        return sid.value

    def submit_slab(self, series_id, y0, y1):
        if series_id not in self._series_shape:
            raise TiltGpuError(ERR_ARG, f"unknown series {series_id}")
        onz, onx = self._series_shape[series_id]
        out = np.empty((y1 - y0, onz, onx), np.float32)
        rid = C.c_int64()
        _check(lib().tiltgpu_server_submit_slab(self.h, series_id, y0, y1, _f(out), C.byref(rid)))
        self._pending[rid.value] = (out, None)
        return rid.value

    def poll(self, rid):
        return _check(lib().tiltgpu_server_poll(self.h, rid))

    def wait(self, rid):
        """Blocks until done; returns (output array, latency in ms)."""
        lat = C.c_double()
        out, _ = self._pending[rid]
        try:
            _check(lib().tiltgpu_server_wait(self.h, rid, C.byref(lat)))
        finally:
            self._pending.pop(rid, None)
        return out, lat.value

    def stats(self):
        st = ServerStats()
        _check(lib().tiltgpu_server_get_stats(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_ if k != "pad"}

    def reset_stats(self):
        _check(lib().tiltgpu_server_reset_stats(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().tiltgpu_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
