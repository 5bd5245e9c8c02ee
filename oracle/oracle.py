"""TEST INFRASTRUCTURE ONLY: ctypes front end of the oracle (liboracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may import
this module.  It restates the `tilt` reconstruction core of imod-rs
(apps/tilt/src/main.rs) on the CPU, in two forms:

* ``f32`` -- the reference's FP32 operation order (C, -ffp-contract=off);
* ``f64`` -- the same algorithm in double precision (the FP64 oracle).

It also restates the text parsers the path uses: ``read_tilt_file``
(crates/imod-transforms/src/tilt.rs:6-23), ``load_per_view_file``
(apps/tilt/src/main.rs:310-326) and ``load_local_alignment`` (main.rs:232-307).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

FILL = {"mean": 0, "zero": 1, "edge": 2}


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OrOpts(C.Structure):
    _fields_ = [
        ("in_nx", C.c_int), ("in_ny", C.c_int), ("nz", C.c_int),
        ("out_nx", C.c_int), ("out_nz", C.c_int),
        ("cutoff", C.c_float), ("falloff", C.c_float),
        ("no_weight", C.c_int), ("exact_filter_size", C.c_int), ("hamming", C.c_int),
        ("no_cosine_stretch", C.c_int), ("cosine_cubic", C.c_int),
        ("max_cosine_tilt", C.c_float),
        ("log_", C.c_int), ("fill_mode", C.c_int),
        ("densweight", C.c_float), ("outscale", C.c_float), ("outadd", C.c_float),
        ("sirt", C.c_int), ("sirt_relax", C.c_float), ("fake_sirt", C.c_int),
        ("nthreads", C.c_int),
    ]


def _declare(L):
    L.or_min_max_mean.argtypes = [_fp, C.c_long, _fp]
    for suf in ("f32", "f64"):
        getattr(L, f"or_tilt_run_{suf}").argtypes = [
            C.POINTER(OrOpts), _fp, _fp, C.c_int, _fp, C.c_int, _fp, C.c_int,
            _ip, _ip, _fp, _fp, _fp, _fp, _fp, _fp]
        getattr(L, f"or_tilt_run_rows_{suf}").argtypes = [
            C.POINTER(OrOpts), _fp, _fp, C.c_int, _fp, C.c_int, _fp, C.c_int,
            _ip, _ip, _fp, _fp, _fp, _fp, C.c_int, C.c_int, _fp, _fp, _dp]
        getattr(L, f"or_sirt_trace_{suf}").argtypes = [
            C.POINTER(OrOpts), _fp, _fp, C.c_int, _fp, C.c_int, C.c_int, C.c_int, _fp, _dp]
        getattr(L, f"or_backproject_simple_{suf}").argtypes = [
            C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
            _fp, _fp, _fp, C.c_int, C.c_int, _fp, C.c_int]
        getattr(L, f"or_forward_project_simple_{suf}").argtypes = [
            C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _fp, _fp, _fp, C.c_int]
        getattr(L, f"or_forward_project_full_{suf}").argtypes = [
            C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _fp, C.c_int, _fp, C.c_int, _fp,
            C.c_int, _ip, _ip, _fp, _fp, _fp, _fp, _fp, _fp, C.c_int]
        getattr(L, f"or_rweight_float_{suf}").argtypes = [
            _fp, C.c_long, C.c_int, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int]
        getattr(L, f"or_cosine_stretch_float_{suf}").argtypes = [
            _fp, C.c_int, C.c_int, C.c_int, _fp, C.c_int, C.c_float]
        getattr(L, f"or_dft_float_{suf}").argtypes = [_fp, _fp, C.c_int, C.c_int]
        getattr(L, f"or_local_interp_float_{suf}").argtypes = [
            _fp, C.c_int, _fp, C.c_int, _fp, _fp, C.c_float, C.c_float, _fp]
        getattr(L, f"or_apply_xf_stack_{suf}").argtypes = [
            _fp, C.c_int, C.c_int, C.c_int, _fp, C.c_int, C.c_float, _fp, C.c_int]
        getattr(L, f"or_ctf_phaseflip_{suf}").argtypes = [
            _fp, C.c_int, C.c_int, C.c_int, _fp, C.c_int, _fp, C.c_float, _dp, C.c_int, _fp,
            C.c_int]
    L.or_xf_inverse_f32.argtypes = [_fp, _fp]
    L.or_studio_reconstruct.argtypes = [_fp, C.c_int, C.c_int, C.c_int, _fp, C.c_int, _fp]
    L.or_xf_inverse_f64.argtypes = [_fp, _dp]
    L.or_radial_weights_f32.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_int, _fp]
    L.or_radial_weights_f64.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_int, _dp]
    L.or_proj_params_f32.argtypes = [_fp, C.c_int, _fp, C.c_int, _fp, C.c_int,
                                     _fp, _fp, _fp, _fp, _fp]
    L.or_proj_params_f64.argtypes = [_fp, C.c_int, _fp, C.c_int, _fp, C.c_int,
                                     _dp, _dp, _dp, _dp, _dp]


def _f(a):
    return a.ctypes.data_as(_fp) if a is not None else None


def _f32c(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


# ------------------------------------------------------------------ parsers --

def read_tilt_file(path):
    """read_tilt_file, crates/imod-transforms/src/tilt.rs:6-23: one f32 per line,
    blank lines skipped, unparsable lines silently skipped."""
    out = []
    with open(path, "r") as fh:
        for line in fh:
            t = line.strip()
            if not t:
                continue
            v = _rust_parse_f32(t)
            if v is not None:
                out.append(v)
    return np.array(out, dtype=np.float32)


def read_xf_file(path):
    """read_xf_file, crates/imod-transforms/src/xf.rs:147-174: blank lines
    skipped; the whitespace-separated tokens that parse as f32 are kept
    (filter_map), and a line with at least six of them gives
    (a11, a12, a21, a22, dx, dy)."""
    out = []
    with open(path, "r") as fh:
        for line in fh:
            t = line.strip()
            if not t:
                continue
            vals = [v for v in (_rust_parse_f32(tok) for tok in t.split()) if v is not None]
            if len(vals) >= 6:
                out.append(vals[:6])
    return np.array(out, dtype=np.float32).reshape(-1, 6)


def xf_inverse(xf6, prec="f32"):
    """LinearTransform::inverse (xf.rs:95-109)."""
    a = _f32c(np.asarray(xf6).reshape(6))
    dt = np.float32 if prec == "f32" else np.float64
    out = np.zeros(6, dt)
    getattr(lib(), f"or_xf_inverse_{prec}")(_f(a), out.ctypes.data_as(_fp if prec == "f32" else _dp))
    return out


def apply_xf_stack(P, xf, fill=0.0, prec="f32", nthreads=8):
    """newstack -x (apps/newstack/src/main.rs:1405-1462) over a stack [nz][ny][nx]."""
    a = _f32c(P)
    nz, ny, nx = a.shape
    x = _f32c(np.asarray(xf).reshape(-1, 6))
    out = np.empty_like(a)
    getattr(lib(), f"or_apply_xf_stack_{prec}")(_f(a), nx, ny, nz, _f(x), x.shape[0], fill,
                                                _f(out), nthreads)
    return out


def studio_reconstruct(P, angles):
    """imod-studio step_reconstruction (apps/imod-studio/src/main.rs:725-770):
    [in_ny][in_nx (z)][in_nx]."""
    a = _f32c(P)
    nz, ny, nx = a.shape
    ang = _f32c(angles)
    out = np.empty((ny, nx, nx), np.float32)
    lib().or_studio_reconstruct(_f(a), nx, ny, nz, _f(ang), ang.size, _f(out))
    return out


def electron_wavelength(voltage_kv):
    """Relativistic electron wavelength in Angstroms, apps/ctfphaseflip/src/main.rs:707-717."""
    v = float(np.float32(voltage_kv)) * 1000.0
    m0, e, c, h = 9.10938e-31, 1.60218e-19, 2.99792e8, 6.62607e-34
    lam = h / np.sqrt(2.0 * m0 * e * v * (1.0 + e * v / (2.0 * m0 * c * c)))
    return lam * 1e10


def ctf_strip_width(strip_width, nx, pixel_a, wavelength, defocus_nm, vdef=None,
                    max_strip_width=None):
    """The strip width of apps/ctfphaseflip/src/main.rs:272-298: --max-strip-width,
    else pixel / (wavelength * max|defocus| / nx) (at least 16) when the defocus is
    meaningful; the requested width clamped to it."""
    if max_strip_width is not None:
        max_w = int(max_strip_width)
    else:
        if vdef is not None:  # a defocus file, possibly without entries (fold from 0)
            mx = max([max(abs(float(np.float32(a))), abs(float(np.float32(b)))) * 10.0
                      for a, b, _ in vdef], default=0.0)
        else:
            mx = abs(float(np.float32(defocus_nm))) * 10.0
        if mx > 1.0 and wavelength > 0.0:
            per_px = mx / float(nx)
            max_w = int(max(float(np.float32(pixel_a)) / (wavelength * per_px), 16.0))
        else:
            max_w = int(strip_width)
    return min(int(strip_width), max_w)


def read_defocus_file(path):
    """read_defocus_file, apps/ctfphaseflip/src/main.rs:103-157: '#' and blank lines
    skipped; 'view d1 [d2 [astig]]' (nm, nm, degrees); a malformed line raises."""
    entries, views = [], []
    with open(path) as fh:
        for ln, line in enumerate(fh, 1):
            t = line.strip()
            if not t or t.startswith("#"):
                continue
            f = t.split()
            if len(f) < 2:
                raise ValueError(f"line {ln}: expected at least 'view defocus1'")
            if _rust_parse_usize(f[0]) is None:
                raise ValueError(f"line {ln}: invalid view number {f[0]!r}")
            d1 = _rust_parse_f32(f[1])
            if d1 is None:
                raise ValueError(f"line {ln}: invalid defocus1 {f[1]!r}")
            d2, aa = d1, 0.0
            if len(f) >= 3:
                d2 = _rust_parse_f32(f[2])
                if d2 is None:
                    raise ValueError(f"line {ln}: invalid defocus2 {f[2]!r}")
            if len(f) >= 4:
                aa = _rust_parse_f32(f[3])
                if aa is None:
                    raise ValueError(f"line {ln}: invalid astig_angle {f[3]!r}")
            views.append(_rust_parse_usize(f[0]))
            entries.append((np.float32(d1), np.float32(d2), np.float32(aa)))
    return entries, views


def interpolate_defocus(entries, views, n_views):
    """interpolate_defocus, apps/ctfphaseflip/src/main.rs:164-198 (f32 lerp
    between the nearest listed views, the nearest one beyond the ends)."""
    if not entries or n_views == 0:
        return np.zeros((0, 3), np.float32)
    res = np.zeros((n_views, 3), np.float32)
    have = np.zeros(n_views, bool)
    for v, e in zip(views, entries):
        if v < n_views:
            res[v] = e
            have[v] = True
    for v in range(n_views):
        if have[v]:
            continue
        b = next((i for i in range(v - 1, -1, -1) if have[i]), None)
        a = next((i for i in range(v + 1, n_views) if have[i]), None)
        if b is not None and a is not None:
            t = np.float32(v - b) / np.float32(a - b)
            res[v] = res[b] + t * (res[a] - res[b])
        elif b is not None:
            res[v] = res[b]
        elif a is not None:
            res[v] = res[a]
    return res


def ctf_phaseflip(P, tilt_deg, pixel_a, voltage_kv=300.0, cs_mm=2.7, defocus_nm=-3000.0,
                  vdef=None, strip_width=256, amp=0.07, cuton=0.0, plate_deg=0.0,
                  tilt_axis_deg=0.0, prec="f32", nthreads=8):
    """ctfphaseflip's CPU path (apps/ctfphaseflip/src/main.rs:369-712) over a stack;
    strip_width is the width actually used (ctf_strip_width)."""
    a = _f32c(P)
    nz, ny, nx = a.shape
    t = _f32c(tilt_deg)
    f32pi = float(np.float32(np.pi))
    phys = np.array([electron_wavelength(voltage_kv), float(np.float32(cs_mm)) * 1e7,
                     float(np.float32(amp)), float(np.float32(cuton)),
                     float(np.float32(plate_deg)) * f32pi / 180.0,
                     float(np.float32(tilt_axis_deg)) * f32pi / 180.0,
                     float(np.float32(pixel_a))], np.float64)
    vd = None if vdef is None or len(vdef) == 0 else _f32c(np.asarray(vdef).reshape(-1, 3))
    if vd is not None:
        assert vd.shape[0] >= nz
    out = np.empty_like(a)
    getattr(lib(), f"or_ctf_phaseflip_{prec}")(_f(a), nx, ny, nz, _f(t), t.size, _f(vd),
                                               float(defocus_nm), phys.ctypes.data_as(_dp),
                                               int(strip_width), _f(out), nthreads)
    return out


def load_per_view_file(path):
    """load_per_view_file, main.rs:310-326: '#' and blank lines skipped, a parse
    error raises (the reference exits with status 1)."""
    out = []
    with open(path, "r") as fh:
        for line in fh:
            t = line.strip()
            if not t or t.startswith("#"):
                continue
            v = _rust_parse_f32(t)
            if v is None:
                raise ValueError(f"parse error: {t!r}")
            out.append(v)
    return np.array(out, dtype=np.float32)


def _rust_parse_f32(t):
    """str::parse::<f32>: decimal / exponent / inf / nan, no hex, no spaces."""
    s = t
    if not s or s != s.strip() or "_" in s:
        return None
    low = s.lower().lstrip("+-")
    if low in ("inf", "infinity", "nan"):
        return float(np.float32(float(s)))
    if low.startswith("0x"):
        return None
    try:
        return float(np.float32(float(s)))
    except ValueError:
        return None


def _rust_parse_usize(t):
    """str::parse::<usize>: ASCII digits after at most one leading '+'."""
    d = t[1:] if t.startswith("+") else t
    if not d or not all("0" <= ch <= "9" for ch in d):
        return None
    return int(d)


@dataclass
class LocalGrids:
    """Per-view patch grids (main.rs:153-168): xs, ys sorted unique, dx/dy
    row-major [iy*nx + ix]."""
    nz: int
    xs: list = field(default_factory=list)
    ys: list = field(default_factory=list)
    dx: list = field(default_factory=list)
    dy: list = field(default_factory=list)

    def flat(self):
        nx = np.array([len(x) for x in self.xs], dtype=np.int32)
        ny = np.array([len(y) for y in self.ys], dtype=np.int32)
        cat = lambda L: (np.concatenate([np.asarray(a, np.float32) for a in L])
                         if sum(len(a) for a in L) else np.zeros(1, np.float32))
        return nx, ny, cat(self.xs), cat(self.ys), cat(self.dx), cat(self.dy)


def _dedup_sorted(vals):
    """Vec::dedup_by(|a, b| (a - b).abs() < 0.01) on a sorted list: an element is
    dropped when it is within 0.01 of the last element kept."""
    out = []
    for v in vals:
        if out and abs(np.float32(v) - np.float32(out[-1])) < np.float32(0.01):
            continue
        out.append(v)
    return out


def build_local_grids(patches, nz):
    """load_local_alignment grid building, main.rs:255-306.  patches = list of
    (view, x, y, dx, dy)."""
    g = LocalGrids(nz, [[] for _ in range(nz)], [[] for _ in range(nz)],
                   [[] for _ in range(nz)], [[] for _ in range(nz)])
    by_view = [[] for _ in range(nz)]
    for p in patches:
        if p[0] < nz:
            by_view[p[0]].append(p)
    for vi, vp in enumerate(by_view):
        if not vp:
            continue
        xs = _dedup_sorted(sorted(np.float32(p[1]) for p in vp))
        ys = _dedup_sorted(sorted(np.float32(p[2]) for p in vp))
        nx, ny = len(xs), len(ys)
        gdx = np.zeros(nx * ny, np.float32)
        gdy = np.zeros(nx * ny, np.float32)
        for p in vp:
            ix = next((i for i, xv in enumerate(xs)
                       if abs(np.float32(xv) - np.float32(p[1])) < np.float32(0.01)), 0)
            iy = next((i for i, yv in enumerate(ys)
                       if abs(np.float32(yv) - np.float32(p[2])) < np.float32(0.01)), 0)
            gdx[iy * nx + ix] = p[3]
            gdy[iy * nx + ix] = p[4]
        g.xs[vi], g.ys[vi] = np.array(xs, np.float32), np.array(ys, np.float32)
        g.dx[vi], g.dy[vi] = gdx, gdy
    return g


def parse_local_file(path):
    """Line parsing of load_local_alignment, main.rs:237-253: '#'/blank skipped,
    fewer than 5 fields skipped, each field unwrap_or(0)."""
    patches = []
    with open(path) as fh:
        for line in fh:
            t = line.strip()
            if not t or t.startswith("#"):
                continue
            parts = t.split()
            if len(parts) < 5:
                continue
            view = _rust_parse_usize(parts[0])
            vals = [_rust_parse_f32(x) for x in parts[1:5]]
            patches.append((view if view is not None else 0,
                            *[v if v is not None else 0.0 for v in vals]))
    return patches


# ----------------------------------------------------------------- numerics --

def align_project(bead, view):
    """project, apps/tiltalign/src/main.rs:169-184 (f64)."""
    dx, dy, rot, mag, tilt = view
    ct, st, cr, sr = math.cos(tilt), math.sin(tilt), math.cos(rot), math.sin(rot)
    xp = bead[0] * ct + bead[2] * st
    yp = bead[1]
    return mag * (cr * xp - sr * yp) + dx, mag * (sr * xp + cr * yp) + dy


def local_corrections(obs, beads, views, patch_size=500, overlap=0.5, cx=0.0, cy=0.0,
                      skip_edge=0, img_w=0.0, img_h=0.0):
    """compute_local_alignment (apps/tiltalign/src/main.rs:864-940) and the
    --skip-edge-patches filter (:1374-1402), restated with Python floats (f64).
    obs[b][v] = (x, y) or None; returns [(view, px, py, dx, dy)]."""
    if not beads:
        return []
    xs = [b[0] + cx for b in beads]
    ys = [b[1] + cy for b in beads]
    min_x, max_x, min_y, max_y = min(xs), max(xs), min(ys), max(ys)
    ps = float(patch_size)
    step = ps * (1.0 - float(np.float32(overlap)))
    if step < 1.0:
        return []
    npx = max(0, math.ceil((max_x - min_x) / step)) + 1
    npy = max(0, math.ceil((max_y - min_y) / step)) + 1
    out = []
    for py in range(npy):
        for px in range(npx):
            pcx = min_x + px * step + ps / 2.0
            pcy = min_y + py * step + ps / 2.0
            half = ps / 2.0
            for vi, view in enumerate(views):
                sx = sy = 0.0
                count = 0
                for bi, track in enumerate(obs):
                    o = track[vi]
                    if o is None:
                        continue
                    if abs(o[0] + cx - pcx) <= half and abs(o[1] + cy - pcy) <= half:
                        qx, qy = align_project(beads[bi], view)
                        sx += o[0] - qx
                        sy += o[1] - qy
                        count += 1
                if count >= 2:
                    out.append((vi, px, py, sx / count, sy / count))
    if skip_edge > 0:
        w = img_w if img_w > 0 else cx * 2.0
        h = img_h if img_h > 0 else cy * 2.0
        m = float(skip_edge)

        def keep(c):
            ex, ey = c[1] * step + ps / 2.0, c[2] * step + ps / 2.0
            return ex >= m and ex <= w - m and ey >= m and ey <= h - m
        out = [c for c in out if keep(c)]
    return out


def min_max_mean(a):
    a = np.ascontiguousarray(a, dtype=np.float32).ravel()
    out = np.zeros(3, np.float32)
    lib().or_min_max_mean(_f(a), a.size, _f(out))
    return float(out[0]), float(out[1]), float(out[2])


def radial_weights(n, cutoff=0.35, falloff=0.05, exact_filter_size=0, hamming=False,
                   prec="f32"):
    dt = np.float32 if prec == "f32" else np.float64
    out = np.zeros(n, dt)
    ptr = out.ctypes.data_as(_fp if prec == "f32" else _dp)
    getattr(lib(), f"or_radial_weights_{prec}")(n, exact_filter_size, cutoff, falloff,
                                                 int(hamming), ptr)
    return out


def rweight(rows2d, cutoff=0.35, falloff=0.05, exact_filter_size=0, hamming=False,
            prec="f32", nthreads=8):
    a = np.array(rows2d, dtype=np.float32, copy=True)
    n = a.shape[-1]
    getattr(lib(), f"or_rweight_float_{prec}")(_f(a), a.size // n, n, cutoff, falloff,
                                                exact_filter_size, int(hamming), nthreads)
    return a


def dft(x, sign=-1, prec="f64"):
    x = np.asarray(x, np.complex128)
    re = np.ascontiguousarray(x.real, np.float32)
    im = np.ascontiguousarray(x.imag, np.float32)
    getattr(lib(), f"or_dft_float_{prec}")(_f(re), _f(im), x.size, sign)
    return re.astype(np.float64) + 1j * im.astype(np.float64)


def cosine_stretch(P, angles, cubic=False, max_tilt=80.0, prec="f32"):
    a = np.array(P, dtype=np.float32, copy=True)
    nz, ny, nx = a.shape
    ang = _f32c(angles)
    assert ang.size >= nz
    getattr(lib(), f"or_cosine_stretch_float_{prec}")(_f(a), nx, ny, nz, _f(ang),
                                                       int(cubic), max_tilt)
    return a


def proj_params(angles, xtilt=None, zfactor=None, prec="f32"):
    ang = _f32c(angles)
    n = ang.size
    dt = np.float32 if prec == "f32" else np.float64
    tab = np.zeros((5, n), dt)
    xt = _f32c(xtilt)
    zf = _f32c(zfactor)
    ptr = (lambda k: tab[k].ctypes.data_as(_fp if prec == "f32" else _dp))
    getattr(lib(), f"or_proj_params_{prec}")(
        _f(ang), n, _f(xt), 0 if xt is None else xt.size, _f(zf),
        0 if zf is None else zf.size, ptr(0), ptr(1), ptr(2), ptr(3), ptr(4))
    return tab


def backproject(P, ptab, means, out_nx, out_nz, fill="mean", y0=0, y1=None,
                prec="f32", nthreads=8):
    """Plain BP (no local alignment): out[(iy-y0)][oz][ox], times 1/nz."""
    P = _f32c(P)
    nz, ny, nx = P.shape
    ptab = np.ascontiguousarray(ptab, np.float32)
    nviews = ptab.shape[1]
    y1 = ny if y1 is None else y1
    out = np.zeros((y1 - y0, out_nz, out_nx), np.float32)
    getattr(lib(), f"or_backproject_simple_{prec}")(
        nx, ny, nz, nviews, out_nx, out_nz, FILL[fill] if isinstance(fill, str) else fill,
        _f(ptab), _f(_f32c(means)), _f(P), y0, y1, _f(out), nthreads)
    return out


def forward_project(V, ptab, in_nx, nz, prec="f32", nthreads=8):
    V = _f32c(V)
    ny, out_nz, out_nx = V.shape
    ptab = np.ascontiguousarray(ptab, np.float32)
    nviews = ptab.shape[1]
    R = np.zeros((nz, ny, in_nx), np.float32)
    getattr(lib(), f"or_forward_project_simple_{prec}")(
        in_nx, ny, nz, nviews, out_nx, out_nz, _f(ptab), _f(V), _f(R), nthreads)
    return R


def forward_project_full(V, angles, in_nx, nz, xtilt=None, zfactor=None, local=None,
                         prec="f32", nthreads=8):
    """forward_project (main.rs:830-896) of V [in_ny][out_nz][out_nx] with the whole
    geometry: tilt angles, x-axis tilts, z factors and local grids (each optional).
    Returns R [nz][in_ny][in_nx]."""
    V = _f32c(V)
    ny, out_nz, out_nx = V.shape
    ang = _f32c(angles)
    xt, zf = _f32c(xtilt), _f32c(zfactor)
    if local is not None:
        lnx, lny, lxs, lys, ldx, ldy = local.flat()
        lp = (lnx.ctypes.data_as(_ip), lny.ctypes.data_as(_ip), _f(lxs), _f(lys),
              _f(ldx), _f(ldy))
    else:
        lp = (None,) * 6
    R = np.zeros((nz, ny, in_nx), np.float32)
    getattr(lib(), f"or_forward_project_full_{prec}")(
        in_nx, ny, nz, out_nx, out_nz, _f(ang), ang.size, _f(xt), 0 if xt is None else xt.size,
        _f(zf), 0 if zf is None else zf.size, *lp, _f(V), _f(R), nthreads)
    return R


def lowpass_3d(V, prec="f32"):
    """lowpass_3d_volume (main.rs:493-538): one clamped 3x3x3 box pass over
    V [ny][nz][nx], computed in the given precision."""
    dt = np.float32 if prec == "f32" else np.float64
    a = np.array(V, dtype=dt, copy=True, order="C")
    ny, nz, nx = a.shape
    fn = getattr(lib(), f"or_lowpass_3d_{prec}")
    fn.argtypes = [_fp if prec == "f32" else _dp, C.c_int, C.c_int, C.c_int]
    fn(a.ctypes.data_as(_fp if prec == "f32" else _dp), ny, nz, nx)
    return a.astype(np.float64)


def local_interp(grids, view, x, y, prec="f32"):
    out = np.zeros(2, np.float32)
    xs = _f32c(grids.xs[view]) if len(grids.xs[view]) else np.zeros(1, np.float32)
    ys = _f32c(grids.ys[view]) if len(grids.ys[view]) else np.zeros(1, np.float32)
    gdx = _f32c(grids.dx[view]) if len(grids.dx[view]) else np.zeros(1, np.float32)
    gdy = _f32c(grids.dy[view]) if len(grids.dy[view]) else np.zeros(1, np.float32)
    getattr(lib(), f"or_local_interp_float_{prec}")(
        _f(xs), len(grids.xs[view]), _f(ys), len(grids.ys[view]), _f(gdx), _f(gdy),
        x, y, _f(out))
    return float(out[0]), float(out[1])


@dataclass
class TiltOptions:
    """The `tilt` option surface (apps/tilt/src/main.rs:18-120) with its defaults."""
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
    exact_filter_size: int = 0
    fill_mode: str = "mean"
    cosine_interp: str = "linear"
    max_cosine_tilt: float = 80.0
    hamming: bool = False
    fake_sirt: int = 0


def _or_opts(opts, nx, ny, nz, out_nx, out_nz, nthreads):
    return OrOpts(nx, ny, nz, out_nx, out_nz, opts.radial[0], opts.radial[1],
                  int(opts.no_weight), int(opts.exact_filter_size or 0), int(opts.hamming),
                  int(opts.no_cosine_stretch), int(opts.cosine_interp.lower() == "cubic"),
                  opts.max_cosine_tilt, int(opts.log), FILL.get(opts.fill_mode.lower(), 0),
                  opts.densweight, opts.outscale, opts.outadd, int(opts.sirt),
                  opts.sirt_relax, int(opts.fake_sirt), nthreads)


def sirt_trace(P, angles, opts: TiltOptions, rows, zfactor=None, prec="f64", nthreads=8):
    """SIRT on output rows (y0, y1) of the single-row geometry (no x-axis tilt, no
    local alignment), with the volume after the initial WBP and after each of the
    opts.sirt iterations: returns (trace [sirt+1][rows][out_nz][out_nx] float32,
    resid [sirt] float64 = |O - FP(V)| over those rows)."""
    P = np.array(P, dtype=np.float32, copy=True)
    nz, ny, nx = P.shape
    ang = _f32c(angles)
    out_nx = opts.width if opts.width else nx
    out_nz = opts.thickness if opts.thickness > 0 else nx
    if not opts.no_cosine_stretch and ang.size < nz:
        raise IndexError("fewer tilt angles than sections (main.rs:450 would panic)")
    y0, y1 = rows
    o = _or_opts(opts, nx, ny, nz, out_nx, out_nz, nthreads)
    trace = np.zeros((opts.sirt + 1, y1 - y0, out_nz, out_nx), np.float32)
    resid = np.zeros(max(opts.sirt, 1), np.float64)
    zf = _f32c(zfactor)
    getattr(lib(), f"or_sirt_trace_{prec}")(
        C.byref(o), _f(P), _f(ang), ang.size, _f(zf), 0 if zf is None else zf.size, y0, y1,
        _f(trace), resid.ctypes.data_as(_dp))
    return trace, resid[:opts.sirt]


def tilt_run(P, angles, opts: TiltOptions, xtilt=None, zfactor=None, local=None,
             prec="f32", nthreads=8, rows=None, means=None):
    """The whole reconstruction (main.rs:601-1093 after reading).  P: [nz][ny][nx]
    raw projections.  Returns (volume [ny][out_nz][out_nx] float32, (amin, amax, amean)).
    rows=(y0, y1) reconstructs (and returns) only those output rows.  means: the
    per-view means (f64) of the whole images when P is a band of their rows."""
    P = np.array(P, dtype=np.float32, copy=True)
    nz, ny, nx = P.shape
    ang = _f32c(angles)
    out_nx = opts.width if opts.width else nx
    out_nz = opts.thickness if opts.thickness > 0 else nx
    if not opts.no_cosine_stretch and ang.size < nz:
        raise IndexError("fewer tilt angles than sections (main.rs:450 would panic)")
    o = _or_opts(opts, nx, ny, nz, out_nx, out_nz, nthreads)
    y0, y1 = (0, ny) if rows is None else rows
    vol = np.zeros((y1 - y0, out_nz, out_nx), np.float32)
    stats = np.zeros(3, np.float32)
    xt = _f32c(xtilt)
    zf = _f32c(zfactor)
    if local is not None:
        lnx, lny, lxs, lys, ldx, ldy = local.flat()
        lp = (lnx.ctypes.data_as(_ip), lny.ctypes.data_as(_ip), _f(lxs), _f(lys),
              _f(ldx), _f(ldy))
    else:
        lp = (None,) * 6
    getattr(lib(), f"or_tilt_run_rows_{prec}")(
        C.byref(o), _f(P), _f(ang), ang.size, _f(xt), 0 if xt is None else xt.size,
        _f(zf), 0 if zf is None else zf.size, *lp, y0, y1, _f(vol), _f(stats),
        None if means is None else np.ascontiguousarray(means, np.float64).ctypes.data_as(_dp))
    return vol, (float(stats[0]), float(stats[1]), float(stats[2]))
