"""Autotuner for the launch shapes of libtiltgpu on the device it runs on.

Measures, for every backprojection kernel variant (single-row; two-row with
x-axis tilt; two-row with local shifts) x projection storage (f32, f16) x width
x thickness class, the backprojection time at each tile depth (32, 64, 128, 256),
and, per width, the forward-projection time at 4, 8 and 16 bins per lane.  The
measurements are written as DIR/autotune.json and turned into the C++ table
csrc/tile_table.inc (compiled into the library; sessions and the server's
batched launches choose their shapes from it) and a markdown summary.

    python imod-rs_amd/autotune.py --out DIR            # measure (GPU), write DIR/*
    python imod-rs_amd/autotune.py --out DIR --from-json DIR/autotune.json   # re-emit

Launch shapes change only how the work is split, never the results.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import tiltgpu as tg  # noqa: E402

VARIANTS = [  # (name, two_row, local)
    ("single-row", 0, 0),
    ("two-row (x-axis tilt)", 1, 0),
    ("two-row + local shifts", 1, 1),
]
WIDTHS = [256, 512, 1024, 2048]
THICK = [64, 128, 256, 512]
TILES = [32, 64, 128, 256]
FP_WIDTHS = [256, 384, 512, 768, 1024, 1536, 2048]
FP_BINS = [4, 8, 16]


def bp_time(P, th, two_row, local, storage, tz, reps=3):
    """Median backprojection ms of a square series (as many rows as columns, so
    that the launch fills the device as a real one does) at tile depth tz."""
    nviews, rows, w = P.shape
    ang = np.linspace(-60.0, 60.0, nviews).astype(np.float32)
    rng = np.random.default_rng(w + th)
    s = tg.Session(w, rows, nviews, tg.TiltOptions(thickness=th, storage=storage))
    try:
        s.set_angles(ang)
        if two_row:
            s.set_xtilt(np.full(nviews, 0.5, np.float32))
        if local:
            one = np.ones(nviews, np.int32)
            s.set_local(one, one, np.full(nviews, w / 2, np.float32), np.full(nviews, rows / 2, np.float32),
                        rng.normal(0, 2, nviews).astype(np.float32), rng.normal(0, 2, nviews).astype(np.float32))
        s.set_tuning(tz, 0)
        s.upload(P)
        s.run()
        ts = []
        for _ in range(reps):
            s.reset_timing()
            s.run()
            t = s.timing()
            ts.append(t["bp_ms"] / max(1, t["bp_launches"]))
        return float(np.median(ts))
    except tg.TiltGpuError:
        return None  # the depth does not fit the geometry's LDS need
    finally:
        s.close()


def fp_time(w, bins, th=128, nviews=61, reps=3):
    ang = np.linspace(-60.0, 60.0, nviews).astype(np.float32)
    V = np.ones((w, th, w), np.float32)
    s = tg.Session(w, w, nviews, tg.TiltOptions(thickness=th, no_cosine_stretch=True))
    try:
        s.set_angles(ang)
        s.set_tuning(0, bins)
        s.reproject(V)
        ts = []
        for _ in range(reps):
            s.reset_timing()
            s.reproject(V)
            ts.append(s.timing()["fp_ms"])
        return float(np.median(ts))
    finally:
        s.close()


def pad(th, tz):
    return (th + tz - 1) // tz * tz


def emit(meas, out):
    """Write tile_table.inc and autotune.md from the measurements.

    Backprojection time is proportional to the padded depth: t = pad(th, tz) *
    k(tz) for a variant, storage and width.  The table stores k(tz) (the median
    over the measured thicknesses of t / pad, relative to k(128) = 1000); the
    library picks the depth that minimises pad(out_nz, tz) * k(tz) at the
    nearest measured width."""
    md = ["# Autotuned launch shapes (gfx950)", "",
          "Written by `python imod-rs_amd/autotune.py --out DIR`.  Backprojection:",
          "median ms per launch of a w x w series at 61 views (w output rows), and the",
          "per-padded-slice cost k(tz) derived from them (relative, k(128) = 1000);",
          "forward projection: median ms of one reprojection of a w x w x 128 volume",
          "at 61 views.", "",
          "## Backprojection tile depth", "",
          "| variant | storage | width | thickness | tz=32 | tz=64 | tz=128 | tz=256 | fastest |",
          "|---|---|---|---|---|---|---|---|---|"]
    rows = []
    groups = {}
    for m in meas["bp"]:
        groups.setdefault((m["two_row"], m["local"], m["half"], m["width"]), []).append(m)
        ms = m["ms"]
        ok = {int(k): v for k, v in ms.items() if v is not None}
        best = min(ok, key=ok.get)
        cells = [f"{ms[str(tz)]:.3f}" if ms.get(str(tz)) is not None else "-" for tz in TILES]
        md.append(f"| {m['variant']} | {'f16' if m['half'] else 'f32'} | {m['width']} | {m['thickness']} | "
                  + " | ".join(cells) + f" | {best} |")
    md += ["", "Per-padded-slice cost k(tz), relative to k(128) = 1000:", "",
           "| two_row | local | half | width | k32 | k64 | k128 | k256 |", "|---|---|---|---|---|---|---|---|"]
    for key, ms in sorted(groups.items()):
        k = {}
        for tz in TILES:
            vals = [m["ms"][str(tz)] / pad(m["thickness"], tz) for m in ms if m["ms"].get(str(tz)) is not None]
            k[tz] = float(np.median(vals)) if vals else None
        base = k[128] if k[128] else min(v for v in k.values() if v)
        kk = [int(round(1000 * k[tz] / base)) if k[tz] else 0 for tz in TILES]
        rows.append((*key, *kk))
        md.append("| %d | %d | %d | %d | %d | %d | %d | %d |" % (*key, *kk))
    md += ["", "## Forward-projection bins per lane", "",
           "| width | 4 | 8 | 16 | chosen |", "|---|---|---|---|---|"]
    fp_rows = []
    for m in meas["fp"]:
        ms = {int(b): v for b, v in m["ms"].items()}
        best = min(ms, key=ms.get)
        fp_rows.append((m["width"], best))
        md.append(f"| {m['width']} | " + " | ".join(f"{ms[b]:.3f}" for b in FP_BINS) + f" | {best} |")
    with open(os.path.join(out, "tile_table.inc"), "w") as fh:
        fh.write("// Measured launch shapes for gfx950, written by imod-rs_amd/autotune.py.\n")
        fh.write("// Backprojection: per-padded-slice cost of tile depth 32, 64, 128, 256 (relative,\n")
        fh.write("// k128 = 1000; 0 = not measured) per kernel variant, storage and width.\n")
        fh.write("// {two_row, local, half, in_nx, k32, k64, k128, k256}\n")
        fh.write("static const BpTune kBpTable[] = {\n")
        for r in rows:
            fh.write("    {%d, %d, %d, %d, %d, %d, %d, %d},\n" % r)
        fh.write("};\n// Forward projection: fastest bins per lane per width.  {in_nx, bins_per_lane}\n")
        fh.write("static const FpTune kFpTable[] = {\n")
        for r in fp_rows:
            fh.write("    {%d, %d},\n" % r)
        fh.write("};\n")
    with open(os.path.join(out, "autotune.md"), "w") as fh:
        fh.write("\n".join(md) + "\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--quick", action="store_true", help="two widths and thicknesses only")
    ap.add_argument("--from-json", help="re-emit the table from a saved autotune.json")
    a = ap.parse_args()
    os.makedirs(a.out, exist_ok=True)
    if a.from_json:
        with open(a.from_json) as fh:
            emit(json.load(fh), a.out)
        return
    widths, thick = (WIDTHS[1::2], THICK[1::2]) if a.quick else (WIDTHS, THICK)
    t0 = time.time()
    meas = {"bp": [], "fp": []}
    for name, two, loc in VARIANTS:
        for storage in ("f32", "f16"):
            for w in widths:
                # the backprojection's time does not depend on the values
                P = np.ones((61, w, w), np.float32)
                for th in thick:
                    ms = {str(tz): bp_time(P, th, two, loc, storage, tz) for tz in TILES}
                    meas["bp"].append({"variant": name, "two_row": two, "local": loc,
                                       "half": int(storage != "f32"), "width": w, "thickness": th,
                                       "ms": ms})
                    print(f"# bp {name} {storage} w={w} z={th}: {ms}", flush=True)
    for w in FP_WIDTHS if not a.quick else FP_WIDTHS[::3]:
        ms = {str(b): fp_time(w, b) for b in FP_BINS}
        meas["fp"].append({"width": w, "ms": ms})
        print(f"# fp w={w}: {ms}", flush=True)
    meas["seconds"] = round(time.time() - t0, 1)
    with open(os.path.join(a.out, "autotune.json"), "w") as fh:
        json.dump(meas, fh, indent=1)
    emit(meas, a.out)
    print(f"# wrote {a.out}/autotune.json, tile_table.inc and autotune.md", flush=True)


if __name__ == "__main__":
    main()
