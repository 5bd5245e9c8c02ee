"""tiltalign's local-alignment output (apps/tiltalign/src/main.rs:864-940,
1362-1409): the library's tiltgpu_local_corrections against a hand-computed
patch grid, against the oracle's restatement on a random series, and the file
it writes as `tilt --local-file` reads it.  CPU only."""
import math

import numpy as np
import pytest

import oracle as O
import tiltgpu as tg

CX = CY = 50.0
# centred bead positions, observed positions per view (None: not observed) and
# the two views' solved parameters (dx, dy, rotation, mag, tilt)
BEADS = [(-5.0, -5.0, 0.0), (-2.0, -3.0, 0.0), (20.0, 15.0, 0.0), (22.0, 13.0, 0.0)]
VIEWS = [(0.0, 0.0, 0.0, 1.0, 0.0), (1.0, 1.0, 0.0, 1.0, 0.0)]
OBS = [[(-4.5, -4.8), (-3.7, -3.8)],
       [(-1.7, -2.9), None],
       [(20.4, 15.0), (21.1, 16.3)],
       [(22.2, 13.4), (22.9, 14.5)]]
# patch 20 px at overlap 0.5: step 10 over the beads' box x 45..72, y 45..65, so
# 4 x 3 patches centred at (55 + 10 px, 55 + 10 py), half width 10 (inclusive)
EXPECTED = [(0, 0, 0, 0.4, 0.15),   # beads 0, 1 in view 0
            (0, 1, 0, 0.3, 0.2),    # beads 2 (y = 65.0 on the edge), 3
            (0, 2, 0, 0.3, 0.2),
            (0, 1, 1, 0.3, 0.2),
            (1, 1, 1, 0.0, 0.4),    # view 1: bead 1 unobserved; beads 2, 3
            (0, 2, 1, 0.3, 0.2),
            (1, 2, 1, 0.0, 0.4)]


def _obs_array(obs):
    a = np.full((len(obs), len(obs[0]), 2), np.nan)
    for b, track in enumerate(obs):
        for v, o in enumerate(track):
            if o is not None:
                a[b, v] = o
    return a


def _close(got, exp):
    assert len(got) == len(exp)
    for g, e in zip(got, exp):
        assert g[:3] == e[:3]
        assert g[3] == pytest.approx(e[3], abs=1e-12) and g[4] == pytest.approx(e[4], abs=1e-12)


def test_hand_computed_patch_grid():
    for got in (tg.local_corrections(_obs_array(OBS), BEADS, VIEWS, 20, 0.5, CX, CY),
                O.local_corrections(OBS, BEADS, VIEWS, 20, 0.5, CX, CY)):
        _close(got, EXPECTED)


def test_skip_edge_patches_uses_indices_times_step():
    # --skip-edge-patches 20 on a 100 x 80 image keeps patches whose index-based
    # centre (10 * index + 10) lies in [20, 80] x [20, 60]: px >= 1 and py >= 1
    exp = [e for e in EXPECTED if e[1] >= 1 and e[2] >= 1]
    got = tg.local_corrections(_obs_array(OBS), BEADS, VIEWS, 20, 0.5, CX, CY, skip_edge=20,
                               img_w=100.0, img_h=80.0)
    _close(got, exp)
    _close(O.local_corrections(OBS, BEADS, VIEWS, 20, 0.5, CX, CY, skip_edge=20, img_w=100.0,
                               img_h=80.0), exp)


def test_degenerate_inputs():
    assert tg.local_corrections(np.zeros((0, 2, 2)), np.zeros((0, 3)), VIEWS, 20, 0.5, CX, CY) == []
    # a step below one pixel gives no grid (main.rs:897-899)
    assert tg.local_corrections(_obs_array(OBS), BEADS, VIEWS, 20, 0.97, CX, CY) == []
    # a single observed bead per patch never reaches the two-bead minimum
    one = [[OBS[0][0], OBS[0][1]]]
    assert tg.local_corrections(_obs_array(one), BEADS[:1], VIEWS, 20, 0.5, CX, CY) == []


def test_library_equals_oracle_on_a_random_series():
    rng = np.random.default_rng(5)
    nb, nv = 40, 31
    beads = [(rng.uniform(-400, 400), rng.uniform(-400, 400), rng.uniform(-60, 60))
             for _ in range(nb)]
    views = [(rng.normal(0, 3), rng.normal(0, 3), rng.normal(0, 0.01), rng.uniform(0.99, 1.01),
              math.radians(a)) for a in np.linspace(-60, 60, nv)]
    obs = []
    for b in beads:
        track = []
        for v in views:
            if rng.random() < 0.1:
                track.append(None)
                continue
            px, py = O.align_project(b, v)
            track.append((px + rng.normal(0, 0.7), py + rng.normal(0, 0.7)))
        obs.append(track)
    for kw in (dict(), dict(skip_edge=150, img_w=1024.0, img_h=1024.0)):
        got = tg.local_corrections(_obs_array(obs), beads, views, 300, 0.5, 512.0, 512.0, **kw)
        ref = O.local_corrections(obs, beads, views, 300, 0.5, 512.0, 512.0, **kw)
        assert len(got) > 50
        _close(got, ref)


def test_written_file_is_read_by_tilt_as_coordinates(tmp_path):
    """The file has tiltalign's header and "view px py dx dy" lines (dx, dy to 4
    decimals); tilt's --local-file parser reads the patch INDICES as the patch
    positions in pixels (apps/tilt/src/main.rs:232-307) -- the reference's
    behaviour, kept."""
    path = tmp_path / "local.txt"
    tg.write_local_corrections(path, EXPECTED)
    lines = path.read_text().splitlines()
    assert lines[0] == "# Local alignment corrections: view patch_x patch_y dx dy"
    assert lines[1] == "0 0 0 0.4000 0.1500"
    assert lines[5] == "1 1 1 0.0000 0.4000"
    assert O.parse_local_file(str(path)) == [tuple([e[0], float(e[1]), float(e[2]),
                                                    float(np.float32(round(e[3], 4))),
                                                    float(np.float32(round(e[4], 4)))])
                                             for e in EXPECTED]
    nx, ny, xs, ys, dx, dy = tg.parse_local_file(str(path), 2)
    # view 0: patch positions 0, 1, 2 (x) and 0, 1 (y) -- indices, not pixels
    assert list(xs[:nx[0]]) == [0.0, 1.0, 2.0] and list(ys[:ny[0]]) == [0.0, 1.0]
