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
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "imod-rs_amd"))

METRIC = "Gvoxel·view/s WBP 2k²×512×121; served p50/p99 latency vs plain FP32"
# BP algorithmic work per voxel-view: proj_x (2 adds), floor/frac (1), lerp of
# two rows (2 FMA = 4 flop counted as the reference's 2 mul + 2 add), y-blend and
# accumulate; counted as in the survey's roofline: 7 flop per voxel-view.
BP_FLOP_PER_VV = 7.0
PEAK_FP32_TFLOPS = 157.3
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


def run_gpu(tg, P, ang, xt, local, storage, steps, warmup):
    opts = tg.TiltOptions(thickness=512, radial=(0.35, 0.05), storage=storage)
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
    s.close()
    return dt, tm


def cpu_baseline(P, ang, xt, local, rows=32):
    """The FP32 oracle (the reference's algorithm restated in C, its threading
    model) on a bounded slab: output rows [1008, 1040) of the same workload,
    from the projection rows they can touch."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O  # noqa: E402 -- CPU-baseline leg only
    nthreads = min(16, os.cpu_count() or 1)
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
    t0 = time.perf_counter()
    O.tilt_run(sub, ang, opts, xtilt=xt, local=grids, prec="f32", nthreads=nthreads,
               rows=(halo, halo + rows))
    dt = time.perf_counter() - t0
    vv = rows * 512 * P.shape[2] * nv
    return {"value": vv / dt / 1e9, "unit": "Gvoxel·view/s", "cores": nthreads,
            "kind": "port",
            "sample": f"{rows} output rows (of 2048) of the same WBP workload, "
                      f"projection rows +-{halo} halo, filter+stretch+BP, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp16", action="store_true")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="one plain 2 GiB device copy, then one step; for rocprofv3 --pmc runs")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the WBP Y-slab sharding is described in DESIGN.md); run with --gpus 1",
              file=sys.stderr)
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
    t0 = time.perf_counter()
    P, ang, xt, local = c1_inputs(tg)
    print(f"# inputs generated in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    nx, ny, nv = P.shape[2], P.shape[1], P.shape[0]
    vv = float(nx) * ny * 512 * nv

    dt, tm = run_gpu(tg, P, ang, xt, local, "f32", a.steps, a.warmup)
    ms_step = dt / a.steps * 1e3
    value = vv / (dt / a.steps) / 1e9
    bp_ms = tm["bp_ms"] / max(1, tm["bp_launches"])
    achieved = BP_FLOP_PER_VV * vv / (bp_ms * 1e-3) / 1e12
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
            "bound": "mfma",
            "pipe": "FP32 vector ALU (gather/interpolation, no MFMA); peak = MI355X FP32 157.3 TFLOP/s",
            "achieved": round(achieved, 3),
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            "traffic": traffic,
            "flop_per_voxel_view": BP_FLOP_PER_VV,
            "launch_ms": round(bp_ms, 3),
        },
    }
    if not a.no_fp16:
        dt16, tm16 = run_gpu(tg, P, ang, xt, local, "f16", a.steps, a.warmup)
        line["fp16_storage"] = {
            "value": round(vv / (dt16 / a.steps) / 1e9, 3),
            "ms_per_step": round(dt16 / a.steps * 1e3, 3),
            "bp_launch_ms": round(tm16["bp_ms"] / max(1, tm16["bp_launches"]), 3),
        }
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(P, ang, xt, local)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
