"""Autotuner for the launch shapes of libtiltgpu on the device it runs on.

Measures, for every backprojection kernel variant (single-row; two-row with
x-axis tilt; two-row with local shifts) x projection storage (f32, f16) x width
x thickness class, the backprojection time at each tile depth (32, 64, 128, 256)
and at the other built launch shapes (two views staged per barrier at depths
32-128; two x positions per lane, tile x 128, at depths 32 and 64), and, per
width, the forward-projection time at 4, 8 and 16 bins per lane.  The
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
# (tile depth, views per stage, x per lane) beyond shape (1, 1)
SHAPES = [(32, 2, 1), (64, 2, 1), (128, 2, 1), (32, 1, 2), (64, 1, 2)]
FP_WIDTHS = [256, 384, 512, 768, 1024, 1536, 2048]
FP_BINS = [4, 8, 16]
FP_ROWS2_MAX = 2048  # widest volume row the two-rows-per-workgroup walk takes


def bp_time(P, th, two_row, local, storage, tz, reps=3, vps=1, xpl=1):
    """Median backprojection ms of a square series (as many rows as columns, so
    that the launch fills the device as a real one does) at tile depth tz and
    launch shape (vps views per stage, xpl x positions per lane)."""
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
        s.set_bp_shape(vps, xpl)
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


def fp_time(w, bins, th=256, nviews=61, reps=2, rows=1, slide=0):
    """Median ms per forward-projection launch inside SIRT-4 runs of a w x w
    series of seeded normal projections, thickness th: the launches then run as
    they do in a reconstruction, between backprojections and filters, on a
    volume of realistic values (a constant volume draws less power and ran some
    shapes faster than real volumes do).  rows: volume rows per workgroup,
    slide: lane window map, 0 fixed or 1 sliding (the library reads
    TILTGPU_FP_ROWS and TILTGPU_FP_SLIDE when it builds each launch)."""
    ang = np.linspace(-60.0, 60.0, nviews).astype(np.float32)
    P = np.random.default_rng(w).standard_normal((nviews, w, w)).astype(np.float32)
    os.environ["TILTGPU_FP_ROWS"] = str(rows)
    os.environ["TILTGPU_FP_SLIDE"] = str(slide)
    s = tg.Session(w, w, nviews, tg.TiltOptions(thickness=th, sirt=4))
    try:
        s.set_angles(ang)
        s.set_tuning(0, bins)
        s.upload(P)
        s.run()
        ts = []
        for _ in range(reps):
            s.reset_timing()
            s.run()
            t = s.timing()
            ts.append(t["fp_ms"] / max(1, t["fp_launches"]))
        return float(np.median(ts))
    finally:
        s.close()
        del os.environ["TILTGPU_FP_ROWS"]
        del os.environ["TILTGPU_FP_SLIDE"]


def pad(th, tz):
    return (th + tz - 1) // tz * tz


def shape_key(tz, vps, xpl):
    return f"{tz}/{vps}/{xpl}"


def emit(meas, out):
    """Write tile_table.inc and autotune.md from the measurements.

    Per kernel variant, storage, width and thickness, each launch shape's time is
    stored as a per-padded-slice cost k = t / pad(th, tz), relative to shape
    (128, 1, 1) of the same measurement = 1000.  The library takes the
    measurement nearest in width, then thickness, and picks the shape that
    minimises pad(out_nz, tz) * k."""
    md = ["# Autotuned launch shapes (gfx950)", "",
          "Written by `python imod-rs_amd/autotune.py --out DIR`.  Backprojection: median",
          "ms per launch of a w x w series at 61 views (w output rows) for every launch",
          "shape -- tile depth tz, views staged per barrier vps, x positions per lane xpl",
          "(tile x 64 * xpl); forward projection: median ms per launch, reprojecting a",
          "w x w x 256 volume inside SIRT-4 runs of seeded normal projections at 61 views.", "",
          "## Backprojection tile depth (1 view per stage, 1 x per lane)", "",
          "| variant | storage | width | thickness | tz=32 | tz=64 | tz=128 | tz=256 | fastest |",
          "|---|---|---|---|---|---|---|---|---|"]
    cost_rows = []
    for m in meas["bp"]:
        ms = m["ms"]
        ok = {int(k): v for k, v in ms.items() if v is not None}
        best = min(ok, key=ok.get)
        cells = [f"{ms[str(tz)]:.3f}" if ms.get(str(tz)) is not None else "-" for tz in TILES]
        md.append(f"| {m['variant']} | {'f16' if m['half'] else 'f32'} | {m['width']} | {m['thickness']} | "
                  + " | ".join(cells) + f" | {best} |")
    md += ["", "## Other launch shapes", "",
           "| variant | storage | width | thickness | " + " | ".join(
               f"tz={tz} vps={v} xpl={x}" for tz, v, x in SHAPES) + " |",
           "|---|---|---|---|" + "---|" * len(SHAPES)]
    for m in meas["bp"]:
        sh = m.get("shapes", {})
        md.append(f"| {m['variant']} | {'f16' if m['half'] else 'f32'} | {m['width']} | "
                  f"{m['thickness']} | " + " | ".join(
                      f"{sh[shape_key(*c)]:.3f}" if sh.get(shape_key(*c)) is not None else "-"
                      for c in SHAPES) + " |")
    md += ["", "## The library's choice at the measured thickness", "",
           "The shape minimising pad(thickness, tz) * k, and its measured ms against the",
           "fastest tile depth of shape (1, 1).", "",
           "| variant | storage | width | thickness | chosen (tz, vps, xpl) | ms | best (1, 1) ms |",
           "|---|---|---|---|---|---|---|"]
    for m in meas["bp"]:
        th = m["thickness"]
        cand = [((tz, 1, 1), m["ms"].get(str(tz))) for tz in TILES]
        cand += [(c, m.get("shapes", {}).get(shape_key(*c))) for c in SHAPES]
        cand = [(c, t) for c, t in cand if t is not None]
        base_t = m["ms"].get("128") or min(t / pad(th, c[0]) * pad(th, 128) for c, t in cand)
        base = base_t / pad(th, 128)
        for (tz, v, x), t in cand:
            cost_rows.append((m["two_row"], m["local"], m["half"], m["width"], th, tz, v, x,
                              int(round(1000 * (t / pad(th, tz)) / base))))
        ch = min(cand, key=lambda ct: ct[1])  # pad(th, tz) * k is proportional to t
        b11 = min(t for c, t in cand if c[1:] == (1, 1))
        md.append(f"| {m['variant']} | {'f16' if m['half'] else 'f32'} | {m['width']} | {th} | "
                  f"({ch[0][0]}, {ch[0][1]}, {ch[0][2]}) | {ch[1]:.3f} | {b11:.3f} |")
    md += ["", "## Forward-projection bins per lane, volume rows per workgroup, window map", "",
           FP_MD_HEAD, "|---|" + "---|" * (len(FP_FORMS) * len(FP_BINS) + 1)]
    fp_rows, fp_md = fp_table_rows(meas["fp"])
    md += fp_md
    with open(os.path.join(out, "tile_table.inc"), "w") as fh:
        fh.write("// Measured launch shapes for gfx950, written by imod-rs_amd/autotune.py.\n")
        fh.write("// Backprojection: per-padded-slice cost k of each launch shape (tile depth,\n")
        fh.write("// views per LDS stage, x positions per lane) per kernel variant, storage,\n")
        fh.write("// width and thickness, relative to shape (128, 1, 1) = 1000.\n")
        fh.write("// {two_row, local, half, in_nx, thickness, tile_z, views_per_stage, x_per_lane, k}\n")
        fh.write("static const BpCost kBpCostTable[] = {\n")
        for r in cost_rows:
            fh.write("    {%d, %d, %d, %d, %d, %d, %d, %d, %d},\n" % r)
        if not cost_rows:
            fh.write("    {-1, 0, 0, 0, 0, 0, 0, 0, 0},\n")
        fh.write("};\n// Forward projection: fastest bins per lane, volume rows per workgroup and lane\n")
        fh.write("// window map (0 fixed, 1 sliding) per width.\n")
        fh.write("// {in_nx, bins_per_lane, rows_per_workgroup, sliding_windows}\n")
        fh.write("static const FpTune kFpTable[] = {\n")
        for r in fp_rows:
            fh.write("    {%d, %d, %d, %d},\n" % r)
        fh.write("};\n")
    with open(os.path.join(out, "autotune.md"), "w") as fh:
        fh.write("\n".join(md) + "\n")


def fp_measure(w):
    """Every forward-projection form at width w (two rows only up to FP_ROWS2_MAX)."""
    entry = {"width": w}
    for key, r, sl in FP_FORMS:
        if r == 1 or w <= FP_ROWS2_MAX:
            entry[key] = {str(b): fp_time(w, b, rows=r, slide=sl) for b in FP_BINS}
    print(f"# fp w={w}: {entry}", flush=True)
    return entry


# measurement keys of the forward-projection forms: (rows per workgroup, window map)
FP_FORMS = [("ms", 1, 0), ("ms_rows2", 2, 0), ("ms_slide", 1, 1), ("ms_rows2_slide", 2, 1)]
FP_MD_HEAD = ("| width | " + " | ".join(
    f"{b}{', two rows' if r == 2 else ''}{', sliding' if sl else ''}"
    for _, r, sl in FP_FORMS for b in FP_BINS) + " | chosen (bins, rows, sliding) |")


def fp_table_rows(meas_fp):
    """(width, bins, rows, sliding) of the fastest forward projection per measured
    width, and the markdown rows of the measurements."""
    rows, md = [], []
    for m in meas_fp:
        cand = {}
        for key, r, sl in FP_FORMS:
            cand.update({(int(b), r, sl): v for b, v in m.get(key, {}).items()})
        best = min(cand, key=cand.get)
        rows.append((m["width"], best[0], best[1], best[2]))
        cells = [f"{cand[(b, r, sl)]:.3f}" if (b, r, sl) in cand else "-"
                 for _, r, sl in FP_FORMS for b in FP_BINS]
        md.append(f"| {m['width']} | " + " | ".join(cells)
                  + f" | ({best[0]}, {best[1]}, {best[2]}) |")
    return rows, md


def replace_fp_table(inc_text, fp_rows):
    """The table source with its kFpTable block replaced by fp_rows."""
    head, sep, rest = inc_text.partition("static const FpTune kFpTable[] = {\n")
    if not sep:
        raise ValueError("no kFpTable block in the table source")
    _, sep2, tail = rest.partition("};\n")
    body = "".join("    {%d, %d, %d, %d},\n" % r for r in fp_rows)
    return head + sep + body + sep2 + tail


def measure_fp_only(out, quick=False):
    """Measure the forward projection only and write DIR/tile_table.inc (the
    compiled table's backprojection rows, new FP rows) and DIR/autotune_fp.md."""
    t0 = time.time()
    meas = []
    for w in FP_WIDTHS if not quick else FP_WIDTHS[::3]:
        meas.append(fp_measure(w))
    rows, md = fp_table_rows(meas)
    src = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "tile_table.inc")
    with open(src) as fh:
        text = replace_fp_table(fh.read(), rows)
    with open(os.path.join(out, "tile_table.inc"), "w") as fh:
        fh.write(text)
    with open(os.path.join(out, "autotune_fp.json"), "w") as fh:
        json.dump({"fp": meas, "seconds": round(time.time() - t0, 1)}, fh, indent=1)
    with open(os.path.join(out, "autotune_fp.md"), "w") as fh:
        fh.write("\n".join([FP_MD_HEAD, "|---|" + "---|" * (len(FP_FORMS) * len(FP_BINS) + 1)]
                            + md) + "\n")
    print(f"# wrote {out}/tile_table.inc, autotune_fp.json and autotune_fp.md", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--quick", action="store_true", help="two widths and thicknesses only")
    ap.add_argument("--from-json", help="re-emit the table from a saved autotune.json")
    ap.add_argument("--fp-only", action="store_true",
                    help="measure FP again; with --from-json keep its backprojection "
                         "measurements, without it keep the compiled table's backprojection rows")
    a = ap.parse_args()
    os.makedirs(a.out, exist_ok=True)
    if a.fp_only and not a.from_json:
        measure_fp_only(a.out, a.quick)
        return
    if a.from_json and not a.fp_only:
        with open(a.from_json) as fh:
            emit(json.load(fh), a.out)
        return
    widths, thick = (WIDTHS[1::2], THICK[1::2]) if a.quick else (WIDTHS, THICK)
    t0 = time.time()
    meas = {"bp": [], "fp": []}
    if a.fp_only:
        with open(a.from_json) as fh:
            meas = json.load(fh)
        meas["fp"] = []
    for name, two, loc in ([] if a.fp_only else VARIANTS):
        for storage in ("f32", "f16"):
            for w in widths:
                # the backprojection's time does not depend on the values
                P = np.ones((61, w, w), np.float32)
                for th in thick:
                    ms = {str(tz): bp_time(P, th, two, loc, storage, tz) for tz in TILES}
                    shapes = {shape_key(tz, v, x): bp_time(P, th, two, loc, storage, tz, vps=v, xpl=x)
                              for tz, v, x in SHAPES}
                    meas["bp"].append({"variant": name, "two_row": two, "local": loc,
                                       "half": int(storage != "f32"), "width": w, "thickness": th,
                                       "ms": ms, "shapes": shapes})
                    print(f"# bp {name} {storage} w={w} z={th}: {ms} {shapes}", flush=True)
    for w in FP_WIDTHS if not a.quick else FP_WIDTHS[::3]:
        meas["fp"].append(fp_measure(w))
    meas["seconds"] = round(time.time() - t0, 1)
    with open(os.path.join(a.out, "autotune.json"), "w") as fh:
        json.dump(meas, fh, indent=1)
    emit(meas, a.out)
    print(f"# wrote {a.out}/autotune.json, tile_table.inc and autotune.md", flush=True)


if __name__ == "__main__":
    main()
