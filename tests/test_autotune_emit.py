"""The autotuner's table writer (imod-rs_amd/autotune.py emit) on synthetic
measurements: per-thickness costs relative to shape (128, 1, 1), every measured
shape present, unmeasured ones absent, and the FP rows.  CPU only."""
import os
import re
import sys

import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "imod-rs_amd"))
import autotune as at  # noqa: E402


def _meas():
    bp = []
    for th, ms128 in ((64, 2.0), (512, 10.0)):
        ms = {"32": 3.0 * th / 64, "64": 2.5 * th / 64, "128": ms128, "256": None}
        shapes = {at.shape_key(*c): None for c in at.SHAPES}
        shapes[at.shape_key(64, 1, 2)] = 1.5 * th / 64
        bp.append({"variant": "single-row", "two_row": 0, "local": 0, "half": 0, "width": 1024,
                   "thickness": th, "ms": ms, "shapes": shapes})
    fp = [{"width": 1024, "ms": {"4": 3.0, "8": 2.0, "16": 2.5},
           "ms_slide": {"4": 3.1, "8": 2.1, "16": 2.4}},
          {"width": 2048, "ms": {"4": 9.0, "8": 7.0, "16": 6.0},
           "ms_rows2": {"4": 8.0, "8": 5.5, "16": 6.5},
           "ms_rows2_slide": {"4": 7.0, "8": 5.4, "16": 6.6}}]
    return {"bp": bp, "fp": fp}


def test_emit_writes_per_thickness_costs(tmp_path):
    at.emit(_meas(), str(tmp_path))
    inc = (tmp_path / "tile_table.inc").read_text()
    rows = [tuple(int(v) for v in m.groups()) for m in re.finditer(
        r"\{(-?\d+), (\d+), (\d+), (\d+), (\d+), (\d+), (\d+), (\d+), (\d+)\}", inc)]
    # two thicknesses x (3 measured depths + 1 measured shape); depth 256 unmeasured
    assert len(rows) == 8
    for r in rows:
        assert r[:4] == (0, 0, 0, 1024) and r[4] in (64, 512)
    by = {(r[4], r[5], r[6], r[7]): r[8] for r in rows}
    assert by[(64, 128, 1, 1)] == 1000 and by[(512, 128, 1, 1)] == 1000
    # thickness 64: t / pad relative to 2.0 / 128; depth 64 pads 64 -> 2.5/64 vs 2/128
    assert by[(64, 64, 1, 1)] == round(1000 * (2.5 / 64) / (2.0 / 128))
    assert by[(64, 64, 1, 2)] == round(1000 * (1.5 / 64) / (2.0 / 128))
    assert (64, 256, 1, 1) not in by and (64, 128, 2, 1) not in by
    # bins and rows per workgroup: one row where only one was measured, two where
    # the two-row walk was fastest
    assert ("static const FpTune kFpTable[] = {\n    {1024, 8, 1, 0},\n    {2048, 8, 2, 1},\n};"
            in inc)
    md = (tmp_path / "autotune.md").read_text()
    # the choice at the measured thickness is the fastest measured launch
    assert "| single-row | f32 | 1024 | 64 | (64, 1, 2) | 1.500 |" in md


def test_emit_without_measurements_writes_a_sentinel(tmp_path):
    at.emit({"bp": [], "fp": []}, str(tmp_path))
    inc = (tmp_path / "tile_table.inc").read_text()
    assert "{-1, 0, 0, 0, 0, 0, 0, 0, 0}" in inc


def test_fp_only_rewrite_keeps_backprojection_rows():
    """--fp-only without a saved measurement: the compiled table's backprojection
    rows stay as they are and only the FP block is replaced."""
    src = os.path.join(os.path.dirname(at.__file__), "csrc", "tile_table.inc")
    with open(src) as fh:
        text = fh.read()
    rows, md = at.fp_table_rows(_meas()["fp"])
    assert rows == [(1024, 8, 1, 0), (2048, 8, 2, 1)]
    assert md[1].endswith("| (8, 2, 1) |")
    # unmeasured forms are dashes: 1024 wide has no two-row columns
    assert md[0].count(" - |") == 6
    new = at.replace_fp_table(text, rows)
    before, _, fp_old = text.partition("static const FpTune kFpTable[] = {")
    after, _, fp_new = new.partition("static const FpTune kFpTable[] = {")
    assert before == after
    assert fp_new.startswith("\n    {1024, 8, 1, 0},\n    {2048, 8, 2, 1},\n};\n")
    assert fp_new.split("};\n", 1)[1] == fp_old.split("};\n", 1)[1]
    with pytest.raises(ValueError):
        at.replace_fp_table("no table here", rows)
