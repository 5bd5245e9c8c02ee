"""Benchmark of the tilt reconstruction core on one MI355X.

Workload (the configuration the headline metric is quoted on): weighted
backprojection of 2048x2048 float32 projections, 121 views from -60 to +60 deg
in 1 deg steps, x-axis tilt 0.5 deg, per-view in-plane shifts N(0, 2 px)
(passed as one local-alignment patch per view), into a 2048 x 2048 x 512 volume,
radial filter 0.35, 0.05.  Input: the seeded ellipsoid/bead phantom (800
ellipsoids, 120 beads) projected analytically, noise sigma = 0.1 x stack std.

A step = one pass of the hot path over the resident series: per-view means,
R-weighting + cosine stretch, backprojection (output transform fused), with
inputs already in HBM.  Metric: Gvoxel*view/s = out_nx*out_ny*thickness*views/t.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-fp16]
"""
from __future__ import annotations

import argparse
import json
import re
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "imod-rs_amd"))

METRIC = "Gvoxel·view/s WBP 2k²×512×121; served p50/p99 latency vs plain FP32"
# BP algorithmic work per voxel-view: proj_x (2 adds), floor/frac (1), lerp of
# two rows (2 FMA = 4 flop counted as the reference's 2 mul + 2 add), y-blend and
# accumulate: 7 flop per voxel-view in all.
BP_FLOP_PER_VV = 7.0
PEAK_FP32_TFLOPS = 157.3
# VALU lane-instructions per voxel*view of the 2048^2 x 121 -> z512 backprojection launch
# (SQ_INSTS_VALU x 64 / voxel*views, profiles/bp_c1_sq_counters.md) and the
# SIMDs of an MI355X (256 CUs x 4)
BP_VALU_PER_VV = 11.66
N_SIMDS = 1024
PEAK_HBM_GBS = 8000.0


def c1_inputs(tg, nx=2048, ny=2048, thickness=512, n_views=121, seed=1, nthreads=16):
    ang = np.linspace(-60.0, 60.0, n_views).astype(np.float32)
    rng = np.random.default_rng(seed)
    sx = rng.normal(0.0, 2.0, n_views).astype(np.float32)
    sy = rng.normal(0.0, 2.0, n_views).astype(np.float32)
    xt = np.full(n_views, 0.5, np.float32)
    P = tg.synth_series(nx, ny, ang, vol_nz=thickness, n_ellipsoids=800, n_beads=120,
                        seed=seed, noise_rel=0.1, xtilt=xt, shift_x=sx, shift_y=sy,
                        nthreads=nthreads)
    # one patch per view at the centre = a constant shift (tiltalign local file)
    nxg = np.ones(n_views, np.int32)
    local = (nxg, nxg.copy(), np.full(n_views, nx / 2, np.float32),
             np.full(n_views, ny / 2, np.float32), sx, sy)
    return P, ang, xt, local


def load_traffic():
    """HBM bytes per BP launch measured with rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, scaled by a copy-kernel calibration), if a summary exists."""
    path = os.path.join(REPO, "profiles", "bp_c1_traffic.json")
    try:
        with open(path) as fh:
            return json.load(fh).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


ERR_ROWS = (992, 1056)  # the 64-slice subset the FP64-oracle error is measured on


def run_gpu(tg, P, ang, xt, local, storage, steps, warmup, keep_rows=None, geometry="f32"):
    opts = tg.TiltOptions(thickness=512, radial=(0.35, 0.05), storage=storage, geometry=geometry)
    s = tg.Session(P.shape[2], P.shape[1], P.shape[0], opts)
    s.set_angles(ang)
    s.set_xtilt(xt)
    s.set_local(*local)
    s.upload(P)
    for _ in range(warmup):
        s.run()
    s.synchronize()
    s.reset_timing()
    t0 = time.perf_counter()
    for i in range(steps):
        s.run()  # synchronous: returns when the step's kernels are done
        print(f"# step {i + 1}/{steps} done", file=sys.stderr, flush=True)
    s.synchronize()
    dt = time.perf_counter() - t0
    tm = s.timing()
    rows = s.read_rows(*keep_rows, stats=False)[0] if keep_rows else None
    s.close()
    if keep_rows:
        return dt, tm, rows
    return dt, tm


def oracle_rows_f64(P, ang, xt, local, rows=ERR_ROWS, halo=24):
    """The FP64 oracle's output rows [rows) of the WBP workload (the checker for
    the error figures), from the projection rows they can touch."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O  # noqa: E402 -- checker only
    y0, y1 = rows
    p0, p1 = y0 - halo, y1 + halo
    sub = np.ascontiguousarray(P[:, p0:p1, :])
    nv = P.shape[0]
    patches = [(v, float(local[2][v]), float(local[3][v] - p0), float(local[4][v]),
                float(local[5][v])) for v in range(nv)]
    grids = O.build_local_grids(patches, nv)
    opts = O.TiltOptions(thickness=512, radial=(0.35, 0.05))
    vol, _ = O.tilt_run(sub, ang, opts, xtilt=xt, local=grids, prec="f64",
                        nthreads=host_threads(), rows=(halo, halo + (y1 - y0)),
                        means=view_means(P))
    return vol


def oracle_rows_f32(P, ang, xt, local, rows=ERR_ROWS):
    """The FP32 oracle's output rows [rows) of the WBP workload, run on the whole
    stack so that every coordinate (iy + dx*sin(xt) + shift) rounds as in the
    whole reconstruction: the parity checker at this size."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O  # noqa: E402 -- checker only
    nv = P.shape[0]
    patches = [(v, float(local[2][v]), float(local[3][v]), float(local[4][v]),
                float(local[5][v])) for v in range(nv)]
    grids = O.build_local_grids(patches, nv)
    opts = O.TiltOptions(thickness=512, radial=(0.35, 0.05))
    vol, _ = O.tilt_run(P, ang, opts, xtilt=xt, local=grids, prec="f32",
                        nthreads=host_threads(), rows=rows)
    return vol


def host_threads():
    """Threads for the host-side oracle work: the CPUs this process may run on,
    capped by OMP_NUM_THREADS when it is set (the machine's share for one GPU)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(cap))) if cap.isdigit() and int(cap) > 0 else max(1, n)


def host_description():
    """nproc, the usable CPU count and the CPU model of this host."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None
    return {"nproc": os.cpu_count(), "usable_cpus": aff,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "cpu_model": model}


def view_means(P):
    """Per-view means of the whole images (the fill values), in f64."""
    return np.array([np.mean(P[v], dtype=np.float64) for v in range(P.shape[0])])


def err_vs(x, ref):
    ref = ref.astype(np.float64)
    d = x.astype(np.float64) - ref
    rms = float(np.sqrt(np.mean(ref ** 2)))
    return {"max_abs_over_rms": float(f"{np.max(np.abs(d)) / rms:.3g}"),
            "rms_over_rms": float(f"{np.sqrt(np.mean(d ** 2)) / rms:.3g}")}


def evict_from_page_cache(path):
    """Write the file back and drop its pages from the page cache, so the next
    read comes from the device (fsync + POSIX_FADV_DONTNEED on the file; no
    privileges needed).  Returns whether the advice was taken."""
    fd = os.open(path, os.O_RDONLY)
    try:
        os.fsync(fd)
        if hasattr(os, "posix_fadvise"):
            os.posix_fadvise(fd, 0, 0, os.POSIX_FADV_DONTNEED)
            return True
        return False
    finally:
        os.close(fd)


PAGE_CACHE_NOTE = ("input written, synced and evicted from the page cache before the run "
                   "(a cold read); output written through the page cache, not synced")


def end_to_end_section(tg, P, ang, xt, local):
    """The `tilt` command line on the 2048^2 x 121 -> z512 workload from files: MRC read in chunks
    overlapped with upload, WBP, and read-back overlapped with the MRC write
    (page-locked buffers both ways).  Wall times as the command reports them;
    the files live in a temporary directory that is removed afterwards."""
    import shutil
    import subprocess
    import tempfile
    tilt = os.path.join(REPO, "imod-rs_amd", "_build", "tilt")
    if not os.path.exists(tilt):
        return {"skipped": "tilt executable not built"}
    d = tempfile.mkdtemp(prefix="tilt_e2e_")
    try:
        need = P.nbytes + P.shape[1] * 512 * P.shape[2] * 4 + (1 << 30)
        free = shutil.disk_usage(d).free
        if free < need:
            return {"skipped": f"{free / 1e9:.1f} GB free in {d}, {need / 1e9:.1f} GB needed"}
        inp, out = os.path.join(d, "c1.mrc"), os.path.join(d, "c1.rec")
        tg.mrc_write(inp, P)
        nv = P.shape[0]
        with open(os.path.join(d, "c1.tlt"), "w") as fh:
            fh.write("".join(f"{a:.9g}\n" for a in ang))
        with open(os.path.join(d, "xtilt.txt"), "w") as fh:
            fh.write("".join(f"{v:.9g}\n" for v in xt))
        with open(os.path.join(d, "local.txt"), "w") as fh:
            fh.write("".join(f"{v} {local[2][v]:.9g} {local[3][v]:.9g} {local[4][v]:.9g} "
                             f"{local[5][v]:.9g}\n" for v in range(nv)))
        cold = evict_from_page_cache(inp)
        cmd = [tilt, "-i", inp, "-o", out, "-t", os.path.join(d, "c1.tlt"), "-z", "512",
               "--radial", "0.35,0.05", "--xtilt-file", os.path.join(d, "xtilt.txt"),
               "--local-file", os.path.join(d, "local.txt")]
        t0 = time.perf_counter()
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            return {"error": r.stderr[-400:]}
        line = [l for l in r.stderr.splitlines() if l.startswith("tilt: wall time")]
        vals = [float(x) for x in re.findall(r"([0-9.]+) s", line[0])] if line else []
        res = {"command": "tilt -i c1.mrc -o c1.rec -t c1.tlt -z 512 --radial 0.35,0.05 "
                          "--xtilt-file xtilt.txt --local-file local.txt",
               "process_wall_s": round(wall, 3),
               "page_cache": PAGE_CACHE_NOTE if cold else "hot (eviction advice unavailable)",
               "input_bytes": int(os.path.getsize(inp)), "output_bytes": int(os.path.getsize(out))}
        if len(vals) == 4:
            res.update({"read_upload_s": vals[0], "reconstruction_s": vals[1],
                        "readback_write_s": vals[2], "total_s": vals[3],
                        "gvoxel_view_per_s": round(float(P.shape[2]) * P.shape[1] * 512 * nv
                                                   / vals[3] / 1e9, 1)})
        return res
    finally:
        shutil.rmtree(d, ignore_errors=True)


# The `tilt` WBP rows the reference publishes whole-process times for
# (BENCHMARKS.md:42-43, Apple Silicon): (width = rows, views, thickness,
# seconds on 1 thread, seconds on 8 threads), and the CPU-runnable
# 512^2 x 61 -> z128 workload, which has no published time.
PUBLISHED_TILT = [(256, 41, 128, 2.8, 0.9), (512, 61, 128, None, None),
                  (512, 61, 256, 33.3, 9.1)]


def published_sizes_section(tg, cpu=True, reps=3):
    """The reference's two published `tilt` sizes: the `tilt` command on files
    (process wall including MRC read and write, median of `reps`, the files just
    written and so in the page cache), the reconstruction alone on the resident
    series (HIP events of the session), and the FP32 oracle on this host's
    threads (the whole job) -- beside the reference's published seconds."""
    import shutil
    import subprocess
    import tempfile
    tilt = os.path.join(REPO, "imod-rs_amd", "_build", "tilt")
    if not os.path.exists(tilt):
        return {"skipped": "tilt executable not built"}
    if cpu:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle as O  # noqa: E402 -- CPU-baseline leg only
    out = []
    for nx, nv, thick, ref1, ref8 in PUBLISHED_TILT:
        ang = np.linspace(-60, 60, nv).astype(np.float32)
        P = tg.synth_series(nx, nx, ang, vol_nz=thick, n_ellipsoids=60, n_beads=10, seed=nx)
        vv = float(nx) * nx * thick * nv
        res = {"workload": f"tilt WBP {nx}x{nx}x{nv} -> z={thick}, default options"}
        if ref8 is not None:
            res["reference_published_s"] = {"1_thread": ref1, "8_threads": ref8,
                                            "host": "Apple Silicon, 8 cores (BENCHMARKS.md:3,42-43)"}
        d = tempfile.mkdtemp(prefix="tilt_pub_")
        try:
            inp, rec, tlt = (os.path.join(d, f) for f in ("ts.mrc", "ts.rec", "ts.tlt"))
            tg.mrc_write(inp, P)
            with open(tlt, "w") as fh:
                fh.write("".join(f"{a:.9g}\n" for a in ang))
            walls = []
            for _ in range(reps):
                t0 = time.perf_counter()
                r = subprocess.run([tilt, "-i", inp, "-o", rec, "-t", tlt, "-z", str(thick)],
                                   capture_output=True, text=True, timeout=300)
                walls.append(time.perf_counter() - t0)
                if r.returncode != 0:
                    res["error"] = r.stderr[-300:]
                    break
            if "error" not in res:
                w = float(np.median(walls))
                res["tilt_process_wall_s"] = round(w, 3)
                if ref8 is not None:
                    res["speedup_vs_reference_8_threads"] = round(ref8 / w, 1)
        finally:
            shutil.rmtree(d, ignore_errors=True)
        with tg.Session(nx, nx, nv, tg.TiltOptions(thickness=thick)) as s:
            s.set_angles(ang)
            s.upload(P)
            s.run()
            s.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                s.run()
            s.synchronize()
            ms = (time.perf_counter() - t0) / reps * 1e3
        res["gpu_resident_ms"] = round(ms, 3)
        res["gpu_resident_gvoxel_view_per_s"] = round(vv / ms / 1e6, 1)
        if cpu:
            nthreads = host_threads()
            t0 = time.perf_counter()
            O.tilt_run(P, ang, O.TiltOptions(thickness=thick), prec="f32", nthreads=nthreads)
            dt = time.perf_counter() - t0
            res["cpu_oracle"] = {"seconds": round(dt, 3), "cores": nthreads,
                                 "gvoxel_view_per_s": round(vv / dt / 1e9, 3),
                                 "note": "compute only (no file I/O)"}
        out.append(res)
        print(f"# published size: {res}", file=sys.stderr, flush=True)
    return out


def section_transforms_section(tg, P, cpu=True, reps=3):
    """newstack -x on the 2048x2048x121 stack: one seeded small rotation +
    shift per section applied on the device (xf_apply_kernel: 4 B read + 4 B
    written per pixel), timed with the session's HIP events; the oracle on two
    sections, single-threaded as the reference's loop, is the CPU figure and the
    bit-for-bit check."""
    nz, ny, nx = P.shape
    rng = np.random.default_rng(11)
    a = np.deg2rad(rng.uniform(-3, 3, nz))
    m = rng.uniform(0.99, 1.01, nz)
    xf = np.stack([m * np.cos(a), -m * np.sin(a), m * np.sin(a), m * np.cos(a),
                   rng.uniform(-8, 8, nz), rng.uniform(-8, 8, nz)], 1).astype(np.float32)
    s = tg.Session(nx, ny, nz, tg.TiltOptions(thickness=512))
    try:
        s.set_angles(np.linspace(-60, 60, nz).astype(np.float32))
        s.upload(P)
        ms = []
        got = None
        for _ in range(reps + 1):
            s.set_section_transforms(xf)
            s.reset_timing()
            got = s.read_projections()
            ms.append(s.timing()["prep_ms"])
        kms = float(np.median(ms[1:]))
    finally:
        s.close()
    px = float(nx) * ny * nz
    res = {"workload": f"newstack -x {nx}x{ny}x{nz}, per-section rotation +-3 deg, scale 0.99-1.01, "
                       "shift +-8 px, bilinear, fill 0",
           "kernel_ms": round(kms, 3), "gpixel_per_s": round(px / kms / 1e6, 2),
           "hbm_gb_per_s": round(8.0 * px / kms / 1e6, 1),
           "frac_of_hbm_peak": round(8.0 * px / kms / 1e6 / 8000.0, 3)}
    if cpu:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle as O  # noqa: E402 -- checker and CPU figure only
        t0 = time.perf_counter()
        ref = O.apply_xf_stack(P[:2], xf[:2], nthreads=1)
        dt = time.perf_counter() - t0
        res["cpu_oracle"] = {"mpixel_per_s": round(2.0 * nx * ny / dt / 1e6, 1), "cores": 1,
                             "sample": "2 sections, the reference's serial per-pixel loop restated"}
        res["bitwise_equal_to_oracle_on_sample"] = bool(
            np.array_equal(got[:2].view(np.uint32), ref.view(np.uint32)))
    return res


def ctf_end_to_end(P, tilt, pixel=4.0, axis=85.0):
    """The `ctfphaseflip` command line on the stack from files (chunks of
    sections read while the previous chunk uploads, read-back overlapped with
    the MRC write).  Wall times as the command reports them; the files live in
    a temporary directory that is removed afterwards."""
    import shutil
    import subprocess
    import tempfile
    sys.path.insert(0, os.path.join(REPO, "imod-rs_amd"))
    import tiltgpu as tg  # noqa: E402
    exe = os.path.join(REPO, "imod-rs_amd", "_build", "ctfphaseflip")
    if not os.path.exists(exe):
        return {"skipped": "ctfphaseflip executable not built"}
    d = tempfile.mkdtemp(prefix="ctf_e2e_")
    try:
        need = 2 * P.nbytes + (1 << 30)
        free = shutil.disk_usage(d).free
        if free < need:
            return {"skipped": f"{free / 1e9:.1f} GB free in {d}, {need / 1e9:.1f} GB needed"}
        nz, ny, nx = P.shape
        inp, out, tlt = (os.path.join(d, f) for f in ("ts.mrc", "ts_ctf.mrc", "ts.tlt"))
        tg.mrc_write(inp, P, xlen=nx * pixel, ylen=ny * pixel, zlen=float(nz))
        with open(tlt, "w") as fh:
            fh.write("".join(f"{a:.9g}\n" for a in tilt))
        cold = evict_from_page_cache(inp)
        cmd = [exe, "-i", inp, "-o", out, "-t", tlt, "--tilt-axis-angle", f"{axis:g}"]
        t0 = time.perf_counter()
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            return {"error": r.stderr[-400:]}
        line = [l for l in r.stderr.splitlines() if l.startswith("ctfphaseflip: wall time")]
        vals = [float(x) for x in re.findall(r"([0-9.]+) s", line[0])] if line else []
        res = {"command": f"ctfphaseflip -i ts.mrc -o ts_ctf.mrc -t ts.tlt --tilt-axis-angle {axis:g}"
                          f" ({nx}x{ny}x{nz}, {pixel} A pixels from the header)",
               "process_wall_s": round(wall, 3),
               "page_cache": PAGE_CACHE_NOTE if cold else "hot (eviction advice unavailable)",
               "bytes_each_way": int(os.path.getsize(inp))}
        if len(vals) == 4:
            res.update({"read_upload_s": vals[0], "correction_s": vals[1],
                        "device_s": vals[2], "readback_write_s": vals[3],
                        "mpixel_per_s_file_to_file": round(float(nx) * ny * nz / wall / 1e6, 1)})
        return res
    finally:
        shutil.rmtree(d, ignore_errors=True)


def ctf_section(tg, P, cpu=True, reps=3, pixel=4.0, e2e=True):
    """ctfphaseflip on the 2048x2048x121 stack: 300 kV, Cs 2.7 mm, defocus
    -3000 nm, 4 A pixels (automatic strip width: 16 px), tilt axis at 0 deg (the
    reference's default: strips are column bands) and at 85 deg (bounding boxes
    the full width and ~195 rows).  Device time of tiltgpu_ctf_run (HIP events,
    the whole stack resident); the oracle (FP32, the reference's CPU path
    restated) on a few sections is the CPU figure and the parity check."""
    nz, ny, nx = P.shape
    tilt = np.linspace(-60.0, 60.0, nz).astype(np.float32)
    out = {}
    if cpu:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle as O  # noqa: E402 -- checker and CPU figure only
    for axis, ncpu in ((0.0, 8), (85.0, 1)):
        prm = tg.ctf_params(pixel_a=pixel, tilt_axis_deg=axis)
        c = tg.CtfCorrector(nx, ny, nz, prm, tilt)
        try:
            c.upload(P)
            ms = []
            for _ in range(reps + 1):
                c.run()
                ms.append(c.last_run())
            g = c.download(0, ncpu) if cpu else None
            width = c.strip_width
        finally:
            c.close()
        kms = float(np.median([m for m, _ in ms[1:]]))
        boxes, _ = tg.ctf_strip_layout(nx, ny, axis, width)
        area = sum(int(b[1] - b[0]) * int(b[3] - b[2]) for b in boxes if b[1] > b[0] and b[3] > b[2])
        r = {"workload": f"ctfphaseflip {nx}x{ny}x{nz}, pixel {pixel} A, defocus -3000 nm, "
                         f"tilt axis {axis:g} deg, strip width {width}",
             "strips": int(len(boxes)), "transformed_pixels_per_section": area,
             "device_ms": round(kms, 2), "launches": int(ms[-1][1]),
             "mpixel_per_s": round(float(nx) * ny * nz / kms / 1e3, 1)}
        if cpu:
            t0 = time.perf_counter()
            ref = O.ctf_phaseflip(P[:ncpu], tilt[:ncpu], pixel, strip_width=width,
                                  tilt_axis_deg=axis, nthreads=ncpu)
            dt = time.perf_counter() - t0
            r["cpu_oracle"] = {"mpixel_per_s": round(ncpu * float(nx) * ny / dt / 1e6, 2),
                               "cores": ncpu, "sample": f"{ncpu} section(s), FP32 restatement"}
            r["max_abs_err_over_rms_vs_oracle"] = float(f"{err_vs(g, ref)['max_abs_over_rms']:.3e}")
        out[f"axis_{axis:g}"] = r
        print(f"# ctfphaseflip axis {axis:g}: {r}", file=sys.stderr, flush=True)
    if e2e:
        try:
            out["end_to_end"] = ctf_end_to_end(P, tilt, pixel)
        except Exception as e:  # noqa: BLE001 -- reported, the device figures stand
            out["end_to_end"] = {"error": str(e)[:300]}
        print(f"# ctfphaseflip end to end: {out['end_to_end']}", file=sys.stderr, flush=True)
    return out


def _cpu_band(O, P, ang, xt, local, rows, nthreads):
    """One timed FP32-oracle run: output rows [ny/2 - rows/2, +rows) of the WBP
    workload from the projection rows they can touch (+-24 halo), with the whole
    images' means.  Returns Gvoxel*view/s and seconds."""
    ny = P.shape[1]
    y0 = ny // 2 - rows // 2
    halo = 24  # |dx*sin(0.5 deg)| <= 9 rows plus |shift| of a few rows
    p0, p1 = y0 - halo, y0 + rows + halo
    sub = np.ascontiguousarray(P[:, p0:p1, :])
    nv = P.shape[0]
    patches = [(v, float(local[2][v]), float(local[3][v] - p0), float(local[4][v]),
                float(local[5][v])) for v in range(nv)]
    grids = O.build_local_grids(patches, nv)
    opts = O.TiltOptions(thickness=512, radial=(0.35, 0.05))
    means = view_means(P)
    t0 = time.perf_counter()
    O.tilt_run(sub, ang, opts, xtilt=xt, local=grids, prec="f32", nthreads=nthreads,
               rows=(halo, halo + rows), means=means)
    dt = time.perf_counter() - t0
    return rows * 512 * P.shape[2] * nv / dt / 1e9, dt


def cpu_baseline(P, ang, xt, local, rows=32, rows_1t=4):
    """The FP32 oracle (the reference's algorithm restated in C, its threading
    model: rows of a chunk spread over the threads) on bounded slabs of the same
    workload: `rows` output rows on every thread this process may use, and
    `rows_1t` rows on one thread."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O  # noqa: E402 -- CPU-baseline leg only
    nthreads = host_threads()
    v, dt = _cpu_band(O, P, ang, xt, local, rows, nthreads)
    v1, dt1 = _cpu_band(O, P, ang, xt, local, rows_1t, 1)
    return {"value": v, "unit": "Gvoxel·view/s", "cores": nthreads,
            "kind": "port",
            "sample": f"{rows} output rows (of 2048) of the same WBP workload, "
                      f"projection rows +-24 halo, filter+stretch+BP, {dt:.1f} s on "
                      f"{nthreads} threads",
            "one_thread": {"value": v1, "cores": 1,
                           "sample": f"{rows_1t} output rows, {dt1:.1f} s"},
            "host": host_description()}


# --------------------------------------------------------------------- SIRT

def c2_inputs(tg, nx=1024, thickness=300, n_views=61, seed=2, nthreads=16):
    """SIRT configuration: 1024x1024 projections, 61 views -60..60 in 2 deg
    steps, thickness 300, the phantom of the 512x512x128 case scaled to the size
    (400 ellipsoids, 60 beads), noise sigma 0.1 x stack std."""
    ang = np.linspace(-60.0, 60.0, n_views).astype(np.float32)
    P = tg.synth_series(nx, nx, ang, vol_nz=thickness, n_ellipsoids=400, n_beads=60,
                        seed=seed, noise_rel=0.1, nthreads=nthreads)
    return P, ang


def rel_errors(x, ref):
    ref = ref.astype(np.float64)
    d = x.astype(np.float64) - ref
    rms = float(np.sqrt(np.mean(ref ** 2)))
    return float(np.max(np.abs(d))) / rms, float(np.sqrt(np.mean(d ** 2))) / rms


def sirt_section(tg, iters=20, rows=(504, 520), oracle=True):
    """SIRT-20 of the 1024^2 x 61 series: FP32 and BF16 projection storage (per-view
    scale, FP32 accumulation); FP64 residual norms from the GPU; every iterate of
    output rows [504, 520) against the FP64 oracle's iterates of those rows."""
    t0 = time.perf_counter()
    P, ang = c2_inputs(tg)
    print(f"# SIRT inputs generated in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    nv, ny, nx = P.shape
    th = 300
    vv = float(nx) * ny * th * nv
    out = {"config": {"workload": f"SIRT-{iters} 1024x1024 x61 views (-60..60, 2 deg) -> "
                                  f"1024x1024x300, relax 1.0", "traced_rows": list(rows)}}
    ref = None
    if oracle:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle as O  # noqa: E402 -- checker for the error figures only
        t1 = time.perf_counter()
        ref, ref_res = O.sirt_trace(P, ang, O.TiltOptions(thickness=th, sirt=iters), rows,
                                    prec="f64", nthreads=host_threads())
        out["oracle_f64_seconds"] = round(time.perf_counter() - t1, 2)
        print(f"# FP64 oracle trace: {out['oracle_f64_seconds']} s", file=sys.stderr, flush=True)
    for storage in ("f32", "bf16"):
        opts = tg.TiltOptions(thickness=th, sirt=iters, sirt_relax=1.0, storage=storage)
        s = tg.Session(nx, ny, nv, opts)
        s.set_angles(ang)
        s.set_sirt_trace(*rows)
        s.upload(P)
        s.run()  # warm-up
        s.synchronize()
        s.reset_timing()
        t1 = time.perf_counter()
        s.run()
        dt = time.perf_counter() - t1
        tm = s.timing()
        res = s.residuals()
        tr = s.sirt_trace()
        s.close()
        # work: the initial WBP plus one FP and one BP per iteration
        work = vv * (1 + 2 * iters)
        entry = {"seconds": round(dt, 4), "ms_per_iteration": round((dt * 1e3 - tm["prep_ms"]) / (iters + 0.5), 3),
                 "gvoxel_view_per_s": round(work / dt / 1e9, 1),
                 "breakdown_ms": {"fp": round(tm["fp_ms"], 2), "bp": round(tm["bp_ms"], 2),
                                  "filter": round(tm["filter_ms"], 2), "other": round(tm["other_ms"] + tm["prep_ms"], 2)},
                 "fp_launch_ms": round(tm["fp_ms"] / max(1, tm["fp_launches"]), 3),
                 "bp_launch_ms": round(tm["bp_ms"] / max(1, tm["bp_launches"]), 3),
                 "residual_norm_f64": [float(f"{r:.9g}") for r in res]}
        if ref is not None:
            errs = [rel_errors(tr[k], ref[k]) for k in range(iters + 1)]
            entry["error_vs_f64_oracle"] = {
                "per_iterate_max_abs_over_rms": [float(f"{e[0]:.3g}") for e in errs],
                "per_iterate_rms_over_rms": [float(f"{e[1]:.3g}") for e in errs]}
        out[storage] = entry
        print(f"# SIRT {storage}: {dt:.3f} s, {entry['gvoxel_view_per_s']} Gvv/s", file=sys.stderr, flush=True)
    return out


# ------------------------------------------------------------------ serving

def c3_series(tg, n_distinct=64, seed=3000, nthreads=16):
    """Series of the batched configuration: width in {256, 512, 768, 1024}
    (square), views uniform 41..121 over +-60 deg, thickness uniform 64..300,
    the seeded phantom (ellipsoid count scaled with the area), noise 0.1."""
    out = []
    for i in range(n_distinct):
        r = np.random.default_rng(seed + i)
        w = int(r.choice([256, 512, 768, 1024]))
        nv = int(r.integers(41, 122))
        th = int(r.integers(64, 301))
        ang = np.linspace(-60.0, 60.0, nv).astype(np.float32)
        n_ell = int(min(400, max(20, round(200 * (w / 512.0) ** 2))))
        P = tg.synth_series(w, w, ang, vol_nz=th, n_ellipsoids=n_ell, n_beads=30,
                            seed=seed + i, noise_rel=0.1, nthreads=nthreads)
        out.append((P, ang, th))
    return out


def c3_requests(series, n_req, storage):
    """Request i uses series i % len(series); even i WBP, odd i 10-iteration SIRT."""
    import tiltgpu as tg
    reqs = []
    for i in range(n_req):
        P, ang, th = series[i % len(series)]
        sirt = 10 if i % 2 else 0
        vv = float(P.shape[2]) * P.shape[1] * th * P.shape[0]
        passes = 1 + 2 * sirt  # backprojections + forward projections
        reqs.append((P, ang, tg.TiltOptions(thickness=th, sirt=sirt, storage=storage), vv * passes))
    return reqs


def plain_run(tg, P, ang, opts):
    vol, _ = tg.reconstruct(P, ang, opts)
    return vol


def serve_batch(tg, reqs, budget):
    with tg.Server(budget) as srv:
        t0 = time.perf_counter()
        ids = [srv.submit(P, a, o) for P, a, o, _ in reqs]
        for i in ids:
            srv.wait(i)
        dt = time.perf_counter() - t0
        st = srv.stats()
    return dt, st


def poisson_times(rate, n, seed):
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, n))


def serve_stream(tg, jobs, arrivals, budget, slab_series=None):
    """Server: submit each job at its arrival time; latency = submit -> result.
    jobs: ("recon", request) or ("slab", y0) on the resident series."""
    lat = [None] * len(jobs)
    with tg.Server(budget) as srv:
        sid = None
        if slab_series is not None:
            P1, ang1, xt1, loc1 = slab_series
            sid = srv.add_series(P1, ang1, tg.TiltOptions(thickness=512), xtilt=xt1, local=loc1)
        ids = []
        t0 = time.perf_counter()
        for job, ta in zip(jobs, arrivals):
            now = time.perf_counter() - t0
            if ta > now:
                time.sleep(ta - now)
            if job[0] == "slab":
                ids.append(srv.submit_slab(sid, job[1], job[1] + 64))
            else:
                P, a, o, _ = job[1]
                ids.append(srv.submit(P, a, o))
        for k, i in enumerate(ids):
            lat[k] = srv.wait(i)[1]
        st = srv.stats()
    return np.array(lat), st


def plain_stream(tg, jobs, arrivals, slab_session=None):
    """Plain per-request FP32 launches: one request at a time in arrival order,
    a session of its own per reconstruction, slabs as row-range WBP launches of a
    session holding the resident series; latency = arrival -> result."""
    lat = []
    t0 = time.perf_counter()
    for job, ta in zip(jobs, arrivals):
        now = time.perf_counter() - t0
        if ta > now:
            time.sleep(ta - now)
        if job[0] == "slab":
            slab_session.run_rows(job[1], job[1] + 64)
            slab_session.read_rows(job[1], job[1] + 64, stats=False)
        else:
            P, a, o, _ = job[1]
            plain_run(tg, P, a, o)
        lat.append((time.perf_counter() - t0 - ta) * 1e3)
    return np.array(lat)


def pct(x, q):
    return float(np.percentile(x, q)) if len(x) else None


def stream_jobs(reqs, with_slabs, seed):
    """Reconstructions in order, with a slab request (64 output rows of the resident
    2048^2 x 121 series, at a random 64-row offset) after every third one."""
    rng = np.random.default_rng(seed)
    jobs = []
    for i, r in enumerate(reqs):
        jobs.append(("recon", r))
        if with_slabs and i % 3 == 2:
            jobs.append(("slab", int(rng.integers(0, 32)) * 64))
    return jobs


def serving_section(tg, n_req, n_stream, c1=None, budget=96 << 30, streams=True):
    t0 = time.perf_counter()
    series = c3_series(tg)
    print(f"# serving inputs generated in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    out = {"config": {"requests": n_req, "distinct_series": len(series),
                      "mix": "even WBP / odd SIRT-10; widths 256-1024, 41-121 views, z 64-300",
                      "budget_bytes": budget,
                      "scheduler": "one round in flight; arrivals during it join the next round; "
                                   "requests of <= 5e8 voxel-views join their variant's "
                                   "shallowest backprojection launch",
                      "slowdown": "latency / the same job's time alone on the device "
                                  "(plain FP32 session; slabs: run_rows + read-back)"}}
    # batched capacity (all requests at once), FP32 and FP16 storage; FP32 also with
    # one backprojection launch per tile depth (the launches before merging)
    runs = [("f32", None), ("f16", None), ("f32", "0")]
    for storage, merge in runs:
        reqs = c3_requests(series, n_req, storage)
        work = sum(r[3] for r in reqs)
        if merge is not None:
            os.environ["TILTGPU_SERVER_MERGE_TILES"] = merge
        try:
            dt, st = serve_batch(tg, reqs, budget)
        finally:
            os.environ.pop("TILTGPU_SERVER_MERGE_TILES", None)
        key = f"batched_{storage}" + ("_launch_per_tile_depth" if merge == "0" else "")
        out[key] = {
            "seconds": round(dt, 3), "requests_per_s": round(n_req / dt, 3),
            "gvoxel_view_per_s": round(work / dt / 1e9, 2),
            "rounds": st["rounds"], "bp_launches": st["bp_launches"],
            "mean_requests_per_bp_launch": round(st["bp_batched_requests"] / max(1, st["bp_launches"]), 2),
            "max_active": st["max_active"]}
        print(f"# {key}: {dt:.2f} s", file=sys.stderr, flush=True)
    reqs = c3_requests(series, n_req, "f32")
    work = sum(r[3] for r in reqs)
    alone = []  # each request's time alone on the device, for the slowdowns below
    t1 = time.perf_counter()
    for P, a, o, _ in reqs:
        t2 = time.perf_counter()
        plain_run(tg, P, a, o)
        alone.append((time.perf_counter() - t2) * 1e3)
    dtp = time.perf_counter() - t1
    plain_cap = n_req / dtp
    out["plain_f32_sequential"] = {"seconds": round(dtp, 3), "requests_per_s": round(plain_cap, 3),
                                   "gvoxel_view_per_s": round(work / dtp / 1e9, 2)}
    print(f"# plain sequential: {dtp:.2f} s", file=sys.stderr, flush=True)
    if not streams:
        return out
    cap = out["batched_f32"]["requests_per_s"]
    # served request streams at 20 / 60 / 90 % of the batched capacity, with slab
    # requests of the resident 2048^2 x 121 series mixed in when it is given; the
    # plain baseline at the 20 % rate (below its own capacity) and at 90 % of its
    # own capacity, where its queue is still stable
    slab_session = None
    if c1 is not None:
        P1, ang1, xt1, loc1 = c1
        slab_session = tg.Session(P1.shape[2], P1.shape[1], P1.shape[0],
                                  tg.TiltOptions(thickness=512))
        slab_session.set_angles(ang1)
        slab_session.set_xtilt(xt1)
        slab_session.set_local(*loc1)
        slab_session.upload(P1)
        slab_session.run_rows(0, 64)  # filters the resident series once
    slab_alone = None
    if slab_session is not None:
        ts = []
        for y0 in (0, 640, 1280, 1920, 960):
            t2 = time.perf_counter()
            slab_session.run_rows(y0, y0 + 64)
            slab_session.read_rows(y0, y0 + 64, stats=False)
            ts.append((time.perf_counter() - t2) * 1e3)
        slab_alone = float(np.median(ts))
        out["slab_alone_ms"] = round(slab_alone, 2)

    def own_times(kinds):
        # each job's own time alone: stream request k repeats request k % n_req
        # of the batched mix (same series, same SIRT count: n_req is even)
        own, k = [], 0
        for kind in kinds:
            if kind == "recon":
                own.append(alone[k % n_req])
                k += 1
            else:
                own.append(slab_alone)
        return np.array(own, np.float64)

    def slowdowns(lat, kinds):
        return lat / own_times(kinds)

    def summary(lat, kinds, st=None):
        e = {"p50_ms": round(pct(lat, 50), 2), "p99_ms": round(pct(lat, 99), 2),
             "p99_over_p50": round(pct(lat, 99) / pct(lat, 50), 2)}
        if n_req % len(series) == 0 and len(series) % 2 == 0 and (
                slab_alone is not None or not np.any(kinds == "slab")):
            sd = slowdowns(lat, kinds)
            e["slowdown_p50"] = round(pct(sd, 50), 2)
            e["slowdown_p99"] = round(pct(sd, 99), 2)
            # the same jobs' own times alone: the spread the mix itself has,
            # which no scheduler removes from the latency percentiles
            own = own_times(kinds)
            e["alone_p50_ms"] = round(pct(own, 50), 2)
            e["alone_p99_ms"] = round(pct(own, 99), 2)
            e["alone_p99_over_p50"] = round(pct(own, 99) / pct(own, 50), 2)
        rec = kinds == "recon"
        e["reconstruction_p50_ms"] = round(pct(lat[rec], 50), 2)
        e["reconstruction_p99_ms"] = round(pct(lat[rec], 99), 2)
        if st is not None:
            e["mean_requests_per_bp_launch"] = round(st["bp_batched_requests"] / max(1, st["bp_launches"]), 2)
            e["rounds"] = st["rounds"]
        if c1 is not None:
            sl = kinds == "slab"
            e["slab_p50_ms"] = round(pct(lat[sl], 50), 2)
            e["slab_p99_ms"] = round(pct(lat[sl], 99), 2)
        return e

    def stream(rate, seed):
        jobs = stream_jobs(c3_requests(series, n_stream, "f32"), c1 is not None, seed=seed)
        n_recon = sum(1 for j in jobs if j[0] == "recon")
        # offered rate counts reconstructions; slabs ride along
        arr = poisson_times(rate * len(jobs) / n_recon, len(jobs), seed=seed)
        return jobs, arr, np.array([j[0] for j in jobs])

    served = []
    for load in (0.2, 0.6, 0.9):  # the configured loads: 20, 60 and 90% of batched capacity
        jobs, arr, kinds = stream(load * cap, int(load * 100))
        lat_s, st = serve_stream(tg, jobs, arr, budget, slab_series=c1)
        entry = {"load": load, "offered_requests_per_s": round(load * cap, 3),
                 "jobs": {"reconstructions": int(np.sum(kinds == "recon")),
                          "slabs": int(np.sum(kinds == "slab"))},
                 "served": summary(lat_s, kinds, st)}
        served.append(entry)
        print(f"# served stream load {load}: {entry['served']}", file=sys.stderr, flush=True)
    out["stream"] = served
    plain = []
    for name, rate, seed in (("at_served_20pct_rate", 0.2 * cap, 20),
                             ("at_90pct_of_plain_capacity", 0.9 * plain_cap, 91)):
        if rate >= plain_cap:
            continue
        jobs, arr, kinds = stream(rate, seed)
        lat_p = plain_stream(tg, jobs, arr, slab_session=slab_session)
        entry = {"name": name, "offered_requests_per_s": round(rate, 3),
                 "fraction_of_plain_capacity": round(rate / plain_cap, 3),
                 "plain_f32": summary(lat_p, kinds)}
        if name == "at_served_20pct_rate":
            entry["served_same_jobs"] = served[0]["served"]
        plain.append(entry)
        print(f"# plain stream {name}: {entry['plain_f32']}", file=sys.stderr, flush=True)
    out["plain_stream"] = plain
    if slab_session is not None:
        slab_session.close()
    out["matched_accuracy"] = "served results are bit-identical to plain per-request FP32 runs"
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp16", action="store_true")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving section")
    ap.add_argument("--serve-only", action="store_true", help="only the serving section")
    ap.add_argument("--serve-batch-only", action="store_true",
                    help="only the batched and plain runs of the serving section (for profiles)")
    ap.add_argument("--serve-requests", type=int, default=64)
    ap.add_argument("--stream-requests", type=int, default=225,
                    help="reconstructions per stream (a slab request rides along every third)")
    ap.add_argument("--no-sirt", action="store_true", help="skip the SIRT section")
    ap.add_argument("--no-xf", action="store_true",
                    help="skip the newstack -x section-transform measurement")
    ap.add_argument("--no-e2e", action="store_true",
                    help="skip the end-to-end run of the tilt command on files")
    ap.add_argument("--no-ctf", action="store_true", help="skip the ctfphaseflip section")
    ap.add_argument("--ctf-only", action="store_true", help="only the ctfphaseflip section")
    ap.add_argument("--sirt-only", action="store_true", help="only the SIRT section")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="one plain 2 GiB device copy, then one step; for rocprofv3 --pmc runs")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU only: the multi-GPU path (WBP Y-slab "
              "sharding) is not built; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import tiltgpu as tg
    if tg.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if a.pmc_probe:
        copy_bytes = 2 << 30
        ms = tg.copy_calibration(copy_bytes, 1)
        print(json.dumps({"copy_bytes": copy_bytes, "copy_ms": ms,
                          "copy_gbs": 2 * copy_bytes / ms / 1e6}), flush=True)
        P, ang, xt, local = c1_inputs(tg)
        dt, tm = run_gpu(tg, P, ang, xt, local, "f32", 1, 0)
        print(json.dumps({"step_s": dt, **tm}), flush=True)
        return
    if a.serve_batch_only:
        print(json.dumps({"serving": serving_section(tg, a.serve_requests, 0, streams=False)}),
              flush=True)
        return
    if a.serve_only:
        c1 = c1_inputs(tg)
        print(json.dumps({"serving": serving_section(tg, a.serve_requests, a.stream_requests,
                                                     c1=c1)}), flush=True)
        return
    if a.sirt_only:
        print(json.dumps({"sirt": sirt_section(tg, oracle=not a.no_cpu)}), flush=True)
        return
    t0 = time.perf_counter()
    P, ang, xt, local = c1_inputs(tg)
    print(f"# inputs generated in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    if a.ctf_only:
        print(json.dumps({"ctfphaseflip": ctf_section(tg, P, cpu=not a.no_cpu, e2e=not a.no_e2e)}),
              flush=True)
        return
    nx, ny, nv = P.shape[2], P.shape[1], P.shape[0]
    vv = float(nx) * ny * 512 * nv

    dt, tm, rows32 = run_gpu(tg, P, ang, xt, local, "f32", a.steps, a.warmup, keep_rows=ERR_ROWS)
    ms_step = dt / a.steps * 1e3
    value = vv / (dt / a.steps) / 1e9
    bp_ms = tm["bp_ms"] / max(1, tm["bp_launches"])
    achieved = BP_FLOP_PER_VV * vv / (bp_ms * 1e-3) / 1e12
    fma_tflops = tg.fma_calibration()  # attainable non-packed FP32 rate (both VALU pipes), measured
    issue = tg.issue_calibration()     # ns per wave instruction per SIMD of the calibration mixes
    traffic = load_traffic()
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gvoxel·view/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded analytic ellipsoid/bead phantom projections)",
        "config": {
            "workload": "WBP 2048x2048 x121 views (-60..60, 1 deg) -> 2048x2048x512, "
                        "xtilt 0.5 deg, per-view shifts N(0,2px) via local file, "
                        "radial 0.35,0.05, FP32 storage",
            "in": [nx, ny, nv], "out": [nx, ny, 512],
        },
        "breakdown_ms_per_step": {
            "means": round(tm["prep_ms"] / a.steps, 3),
            "filter_stretch": round(tm["filter_ms"] / a.steps, 3),
            "backprojection": round(tm["bp_ms"] / a.steps, 3),
        },
        "roofline": {
            "kernel": "bp_kernel (two-row, local)",
            "bound": "fp32_valu",
            "pipe": "FP32 vector ALU: gather/interpolation, no matrix work, no MFMA; "
                    "peak = MI355X FP32 vector 157.3 TFLOP/s",
            "achieved": round(achieved, 3),
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            "traffic": traffic,
            "flop_per_voxel_view": BP_FLOP_PER_VV,
            "launch_ms": round(bp_ms, 3),
            "fp32_fma_attainable_tflops": round(fma_tflops, 1),
            "frac_of_attainable": round(achieved / fma_tflops, 4),
            "valu_issue_ns_per_wave_instr": {k: round(v, 4) for k, v in issue.items()},
            "bp_launch_if_issue_bound_ms": round(
                vv / 64.0 * BP_VALU_PER_VV * issue["bp_two_row_voxel_step"] / N_SIMDS / 1e6, 2),
        },
    }
    rows16 = None
    if not a.no_fp16:
        dt16, tm16, rows16 = run_gpu(tg, P, ang, xt, local, "f16", a.steps, a.warmup,
                                     keep_rows=ERR_ROWS)
        line["fp16_storage"] = {
            "value": round(vv / (dt16 / a.steps) / 1e9, 3),
            "ms_per_step": round(dt16 / a.steps * 1e3, 3),
            "bp_launch_ms": round(tm16["bp_ms"] / max(1, tm16["bp_launches"]), 3),
        }
    if not a.no_cpu:
        t1 = time.perf_counter()
        ref64 = oracle_rows_f64(P, ang, xt, local)
        err = {"subset": f"output rows [{ERR_ROWS[0]}, {ERR_ROWS[1]}) of 2048 (64 slices), "
                         f"FP64 oracle {time.perf_counter() - t1:.1f} s",
               "f32_storage": err_vs(rows32, ref64)}
        if rows16 is not None:
            err["f16_storage"] = err_vs(rows16, ref64)
            err["f16_vs_f32_gpu"] = err_vs(rows16, rows32)
        # the per-view tables computed in FP64 on the GPU (one step each)
        for st in ("f32",) + (() if a.no_fp16 else ("f16",)):
            r = run_gpu(tg, P, ang, xt, local, st, 1, 0, keep_rows=ERR_ROWS, geometry="f64")[2]
            err[f"{st}_storage_f64_geometry"] = err_vs(r, ref64)
        line["error_vs_f64_oracle"] = err
        print(f"# error vs FP64 oracle: {err}", file=sys.stderr, flush=True)
        t1 = time.perf_counter()
        ref32 = oracle_rows_f32(P, ang, xt, local)
        err32 = {"subset": f"output rows [{ERR_ROWS[0]}, {ERR_ROWS[1]}) of 2048, FP32 oracle "
                           f"on the whole stack {time.perf_counter() - t1:.1f} s",
                 "bar": "max_abs_over_rms <= 1e-4 for FP32 storage",
                 "f32_storage": err_vs(rows32, ref32)}
        if rows16 is not None:
            err32["f16_storage"] = err_vs(rows16, ref32)
        line["error_vs_f32_oracle"] = err32
        print(f"# error vs FP32 oracle: {err32}", file=sys.stderr, flush=True)
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(P, ang, xt, local)
    if not a.no_xf:
        try:
            line["section_transforms"] = section_transforms_section(tg, P, cpu=not a.no_cpu)
        except Exception as e:  # noqa: BLE001 -- reported, the device figures stand
            line["section_transforms"] = {"error": str(e)[:300]}
        print(f"# section transforms: {line['section_transforms']}", file=sys.stderr, flush=True)
    if not a.no_e2e:
        try:
            line["end_to_end"] = end_to_end_section(tg, P, ang, xt, local)
        except Exception as e:  # noqa: BLE001 -- reported, the device figures stand
            line["end_to_end"] = {"error": str(e)[:300]}
        print(f"# end to end: {line['end_to_end']}", file=sys.stderr, flush=True)
    if not a.no_e2e:
        try:
            line["published_sizes"] = published_sizes_section(tg, cpu=not a.no_cpu)
        except Exception as e:  # noqa: BLE001 -- reported, the device figures stand
            line["published_sizes"] = {"error": str(e)[:300]}
    if not a.no_ctf:
        try:
            line["ctfphaseflip"] = ctf_section(tg, P, cpu=not a.no_cpu, e2e=not a.no_e2e)
        except Exception as e:  # noqa: BLE001 -- reported, the device figures stand
            line["ctfphaseflip"] = {"error": str(e)[:300]}
    if not a.no_sirt:
        line["sirt"] = sirt_section(tg, oracle=not a.no_cpu)
    if not a.no_serve:
        line["serving"] = serving_section(tg, a.serve_requests, a.stream_requests,
                                          c1=(P, ang, xt, local))
    line["child_processes_at_exit"] = child_processes()
    print(json.dumps(line), flush=True)


def child_processes():
    """Processes this run started that are still alive (should be none)."""
    try:
        import psutil
        return [f"{c.pid} {c.name()}" for c in psutil.Process().children(recursive=True)]
    except Exception as e:  # noqa: BLE001 -- diagnostic only
        return [f"unknown: {e}"[:80]]


if __name__ == "__main__":
    main()
