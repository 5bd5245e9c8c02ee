"""Pins the oracle (CPU restatement of apps/tilt/src/main.rs) before it is
trusted as the GPU path's checker: the reference's own known-answer tests,
closed forms, invariants and brute force on tiny inputs.  CPU only."""
import json
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---------------------------------------------------------- reference KATs --

def test_min_max_mean_known_answer():
    # crates/imod-math/src/tests.rs:17-23
    g = json.load(open(os.path.join(GOLDEN, "min_max_mean.json")))
    mn, mx, mean = O.min_max_mean(np.array(g["input"], np.float32))
    assert mn == g["min"] and mx == g["max"]
    assert abs(mean - g["mean"]) < g["mean_tol"]


def test_min_max_mean_f64_accumulation():
    # the sum is kept in f64 (stats.rs:851-858): an f32 sum would lose the 1
    a = np.array([1e8, 1.0, -1e8], np.float32)
    assert O.min_max_mean(a)[2] == pytest.approx(1.0 / 3.0, rel=1e-6)
    assert O.min_max_mean(np.zeros(0, np.float32)) == (0.0, 0.0, 0.0)


def test_tilt_file_roundtrip(tmp_path):
    # crates/imod-transforms/src/tests.rs:118-131: -60..60 step 3, tolerance 0.1
    ang = O.read_tilt_file(os.path.join(GOLDEN, "tilt_roundtrip.tlt"))
    ref = np.arange(-60, 61, 3, dtype=np.float32)
    assert ang.shape == ref.shape
    assert np.all(np.abs(ang - ref) < 0.1)


def test_tilt_file_skips_blank_and_garbage(tmp_path):
    p = tmp_path / "x.tlt"
    p.write_text("  -60.0 \n\nabc\n1e1\n0x10\n 3,5\n+2.5\ninf\n")
    ang = O.read_tilt_file(str(p))
    assert list(ang[:3]) == [-60.0, 10.0, 2.5] and np.isinf(ang[3]) and len(ang) == 4


def test_generate_tilt_angles():
    # crates/imod-transforms/src/tests.rs:133-140
    g = json.load(open(os.path.join(GOLDEN, "generate_tilt_angles.json")))
    a = np.array([np.float32(g["start"]) + np.float32(i) * np.float32(g["increment"])
                  for i in range(g["count"])], np.float32)
    assert len(a) == g["count"]
    for i, v, tol in g["checks"]:
        assert abs(a[i] - v) < tol


def test_per_view_file_parse_error(tmp_path):
    p = tmp_path / "xt.txt"
    p.write_text("# x tilt\n0.5\n\n0.25\n")
    assert list(O.load_per_view_file(str(p))) == [0.5, 0.25]
    p.write_text("0.5\nnope\n")
    with pytest.raises(ValueError):
        O.load_per_view_file(str(p))


# ------------------------------------------------------------------ the DFT --

@pytest.mark.parametrize("n", [2, 7, 12, 30, 97, 256, 768, 1000])
def test_dft_matches_numpy(n):
    x = np.random.default_rng(n).normal(size=n) + 1j * np.random.default_rng(n + 1).normal(size=n)
    x32 = x.astype(np.complex64)
    ref = np.fft.fft(x32.astype(np.complex128))
    scale = np.max(np.abs(ref))
    # the wrapper returns float32, so the f64 form is checked to f32 rounding
    assert np.max(np.abs(O.dft(x32, -1, "f64") - ref)) < 1.2e-7 * scale
    assert np.max(np.abs(O.dft(x32, -1, "f32") - ref)) < 2e-6 * scale
    inv = np.fft.ifft(x32.astype(np.complex128)) * n
    assert np.max(np.abs(O.dft(x32, +1, "f64") - inv)) < 1.2e-7 * scale


# -------------------------------------------------------- radial weighting --

@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_radial_weight_shape(prec):
    n, c, f = 256, 0.35, 0.05
    W = O.radial_weights(n, c, f, prec=prec)
    nxc = n // 2 + 1
    k = np.arange(n)
    fb = np.where(k <= n // 2, k, n - k)
    ramp = fb / nxc
    below = ramp <= c
    # a pure ramp up to the cutoff (the Ram-Lak filter), DC removed
    assert W[0] == 0
    assert np.allclose(W[below], ramp[below], rtol=1e-7, atol=0)
    # symmetric: W(k) = W(n-k)
    assert np.array_equal(W[1:], W[1:][::-1])
    # Gaussian roll-off: at one falloff past the cutoff the weight is w*exp(-1/2)
    w1 = (c + f)
    Wc = O.radial_weights(n, c, f, exact_filter_size=100, prec=prec)  # w = k/100
    k1 = int(round(w1 * 100))
    assert Wc[k1] == pytest.approx(w1 * np.exp(-0.5), rel=1e-5)
    assert np.all(np.diff(W[: n // 2 + 1][~below[: n // 2 + 1]]) < 0)


def test_radial_weight_hamming():
    n, c = 200, 0.3
    W = O.radial_weights(n, c, 0.05, exact_filter_size=100, hamming=True, prec="f64")
    w = np.arange(n) / 100.0
    w = np.minimum(w, (n - np.arange(n)) / 100.0)
    # window 0.54 + 0.46 cos(pi w / c): 1 at w=0, 0.08 at w=c, 0 beyond
    c32 = float(np.float32(c))  # the cutoff arrives as an f32 (main.rs:582)
    assert W[10] == pytest.approx(0.1 * (0.54 + 0.46 * np.cos(np.pi * 0.1 / c32)), rel=1e-12)
    assert W[30] == pytest.approx(0.3 * 0.08, rel=1e-9)
    assert np.all(W[31:n - 30] == 0)


@pytest.mark.parametrize("prec,tol", [("f32", 2e-6), ("f64", 2e-7)])  # f64 returned as f32
@pytest.mark.parametrize("n", [64, 96, 127])
def test_rweight_cosine_rows_are_eigenvectors(prec, tol, n):
    """A pure cosine at bin k is multiplied by W(k): the filter is diagonal in
    the Fourier basis (the definition of apply_rweighting, main.rs:333-381)."""
    W = O.radial_weights(n, 0.35, 0.05, prec="f64")
    x = np.arange(n)
    for k in (0, 1, 5, n // 4, n // 2 - 1):
        row = np.cos(2 * np.pi * k * x / n + 0.3).astype(np.float32)
        out = O.rweight(row[None, :], prec=prec)[0]
        assert np.max(np.abs(out - W[k] * row)) < tol * max(1.0, n / 64)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_rweight_linear_and_shift_equivariant(prec):
    rng = np.random.default_rng(3)
    a = rng.normal(size=(4, 80)).astype(np.float32)
    b = rng.normal(size=(4, 80)).astype(np.float32)
    fa, fb = O.rweight(a, prec=prec), O.rweight(b, prec=prec)
    fab = O.rweight((a + 2 * b).astype(np.float32), prec=prec)
    tol = 1e-5 if prec == "f32" else 1e-6
    assert np.max(np.abs(fab - (fa + 2 * fb))) < tol * np.max(np.abs(fab))
    sh = O.rweight(np.roll(a, 7, axis=1), prec=prec)
    assert np.max(np.abs(sh - np.roll(fa, 7, axis=1))) < tol * np.max(np.abs(fa))
    # constant rows vanish (DC weight 0)
    assert np.max(np.abs(O.rweight(np.full((2, 80), 3.0, np.float32), prec=prec))) < 1e-5


def test_rweight_matches_numpy_fft():
    rng = np.random.default_rng(4)
    rows = rng.normal(size=(6, 120)).astype(np.float32)
    W = O.radial_weights(120, 0.35, 0.05, prec="f64")
    ref = np.real(np.fft.ifft(np.fft.fft(rows.astype(np.float64), axis=1) * W, axis=1))
    assert np.max(np.abs(O.rweight(rows, prec="f64") - ref)) < 1e-6
    assert np.max(np.abs(O.rweight(rows, prec="f32") - ref)) < 2e-5


# ----------------------------------------------------------- cosine stretch --

@pytest.mark.parametrize("cubic", [False, True])
def test_cosine_stretch_identity_and_limits(cubic):
    rng = np.random.default_rng(5)
    P = rng.normal(size=(4, 3, 50)).astype(np.float32)
    ang = np.array([0.0, 85.0, -85.0, 40.0], np.float32)
    S = O.cosine_stretch(P, ang, cubic=cubic)
    assert np.array_equal(S[0], P[0])          # theta = 0: identity
    assert np.array_equal(S[1:3], P[1:3])      # |theta| > 80: untouched
    assert not np.array_equal(S[3], P[3])


@pytest.mark.parametrize("cubic", [False, True])
def test_cosine_stretch_linear_ramp(cubic):
    n = 64
    ramp = (0.25 * np.arange(n) + 1.0).astype(np.float32)
    P = np.tile(ramp, (1, 2, 1)).astype(np.float32)
    deg = 40.0
    S = O.cosine_stretch(P, np.array([deg], np.float32), cubic=cubic, prec="f64")[0, 0]
    s = 1.0 / np.cos(np.deg2rad(deg))
    src = (np.arange(n) - n / 2) * s + n / 2
    ok = (src >= 0) & (src <= n - 1)
    # a linear ramp stays exactly linear where the source is inside the row ...
    assert np.allclose(S[ok], 0.25 * src[ok] + 1.0, atol=1e-9)
    # ... and is zero past the row (main.rs:481-483)
    assert np.all(S[src >= n] == 0) and np.all(S[src < 0] == 0)


# ---------------------------------------------------------- backprojection --

def _tab(angles, xtilt=None, zf=None):
    return O.proj_params(angles, xtilt, zf, prec="f32")


def test_bp_bruteforce_fixture():
    g = np.load(os.path.join(GOLDEN, "bp_bruteforce.npz"))
    P, ang, means, V = g["P"], g["angles"], g["means"], g["V"]
    nx, nz = int(g["out_nx"]), int(g["out_nz"])
    o64 = O.backproject(P, O.proj_params(ang, prec="f32"), means, nx, nz, fill="mean", prec="f64")
    o32 = O.backproject(P, _tab(ang), means, nx, nz, fill="mean", prec="f32")
    rms = np.sqrt(np.mean(V ** 2))
    # f64 oracle: geometry in f64 from the f32 angles, as the brute force
    assert np.max(np.abs(o64 - V)) < 1e-6 * rms
    assert np.max(np.abs(o32 - V)) < 1e-5 * rms


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_bp_constant_projections(prec):
    """All projections c, Mean fill: the volume is c (weights sum to 1, fill = c)."""
    P = np.full((9, 4, 40), 2.5, np.float32)
    ang = np.linspace(-60, 60, 9).astype(np.float32)
    V = O.backproject(P, _tab(ang), np.full(9, 2.5, np.float32), 56, 30, prec=prec)
    assert np.max(np.abs(V - 2.5)) < 1e-5


def test_bp_single_view_identity():
    """One view at 0 deg, out_nx = in_nx: V[iy][oz][ox] = P[iy][ox] for
    ox <= n-2 and the view mean at ox = n-1 (floor(px)+1 = n is out of range)."""
    rng = np.random.default_rng(6)
    P = rng.normal(size=(1, 5, 16)).astype(np.float32)
    m = np.float32(0.123)
    V = O.backproject(P, _tab(np.zeros(1, np.float32)), np.array([m]), 16, 3)
    for oz in range(3):
        assert np.array_equal(V[:, oz, :15], P[0, :, :15])
        assert np.all(V[:, oz, 15] == m)


def test_bp_linear_in_projections():
    rng = np.random.default_rng(7)
    a = rng.normal(size=(7, 3, 30)).astype(np.float32)
    b = rng.normal(size=(7, 3, 30)).astype(np.float32)
    ang = np.linspace(-45, 45, 7).astype(np.float32)
    z = np.zeros(7, np.float32)
    A = O.backproject(a, _tab(ang), z, 30, 12, fill="zero", prec="f64")
    B = O.backproject(b, _tab(ang), z, 30, 12, fill="zero", prec="f64")
    AB = O.backproject((a + 3 * b).astype(np.float32), _tab(ang), z, 30, 12, fill="zero", prec="f64")
    assert np.max(np.abs(AB - (A + 3 * B))) < 1e-5


def test_bp_f32_f64_agree():
    rng = np.random.default_rng(8)
    P = rng.normal(size=(11, 4, 48)).astype(np.float32)
    ang = np.linspace(-60, 60, 11).astype(np.float32)
    means = P.reshape(11, -1).mean(axis=1).astype(np.float32)
    a = O.backproject(P, _tab(ang), means, 48, 20, prec="f32")
    b = O.backproject(P, _tab(ang), means, 48, 20, prec="f64")
    assert np.sqrt(np.mean((a - b) ** 2)) < 1e-5 * np.sqrt(np.mean(b ** 2))


# ---------------------------------------------------- forward projection --

def test_fp_is_adjoint_of_zero_fill_bp():
    """<FP(V), S> = nz * <V, BP_zero-fill(S)> -- forward_project (main.rs:831-896)
    is the transpose of the backprojection interpolation."""
    rng = np.random.default_rng(9)
    nz, ny, nx, onx, onz = 7, 3, 26, 22, 10
    ang = np.linspace(-50, 50, nz).astype(np.float32)
    V = rng.normal(size=(ny, onz, onx)).astype(np.float32)
    S = rng.normal(size=(nz, ny, nx)).astype(np.float32)
    tab = _tab(ang)
    R = O.forward_project(V, tab, nx, nz, prec="f64")
    B = O.backproject(S, tab, np.zeros(nz, np.float32), onx, onz, fill="zero", prec="f64")
    lhs = float(np.sum(R.astype(np.float64) * S))
    rhs = nz * float(np.sum(V.astype(np.float64) * B))
    assert lhs == pytest.approx(rhs, rel=1e-5)


def test_fp_conserves_mass():
    nz, ny, nx = 5, 1, 40
    V = np.zeros((ny, 8, 30), np.float32)
    V[0, 4, 15] = 2.0
    ang = np.linspace(-30, 30, nz).astype(np.float32)
    R = O.forward_project(V, _tab(ang), nx, nz, prec="f64")
    assert np.allclose(R.reshape(nz, -1).sum(axis=1), 2.0, atol=1e-6)


# --------------------------------------------------------- local alignment --

def test_local_grid_bilinear_exact():
    """interpolate (main.rs:172-198) is bilinear on the grid cell and clamps
    outside; a bilinear field is reproduced exactly inside."""
    xs, ys = [0.0, 10.0, 30.0], [0.0, 20.0]
    f = lambda x, y: 0.5 * x - 0.25 * y + 0.01 * x * y + 1.0
    patches = [(0, x, y, f(x, y), -f(x, y)) for x in xs for y in ys]
    g = O.build_local_grids(patches, 2)
    for (x, y) in [(5.0, 5.0), (12.5, 7.0), (29.0, 19.0)]:
        dx, dy = O.local_interp(g, 0, x, y, prec="f64")
        assert dx == pytest.approx(f(x, y), abs=1e-5) and dy == pytest.approx(-f(x, y), abs=1e-5)
    assert O.local_interp(g, 0, -5.0, -5.0, prec="f64")[0] == pytest.approx(f(0, 0))
    assert O.local_interp(g, 0, 50.0, 50.0, prec="f64")[0] == pytest.approx(f(30, 20))
    assert O.local_interp(g, 1, 3.0, 3.0) == (0.0, 0.0)  # no patches


def test_local_grid_dedup():
    # positions within 0.01 of the last kept one merge (Vec::dedup_by, main.rs:281-283)
    g = O.build_local_grids([(0, 1.0, 0.0, 1, 1), (0, 1.005, 0.0, 2, 2), (0, 2.0, 0.0, 3, 3)], 1)
    assert list(g.xs[0]) == [1.0, 2.0]


# ------------------------------------------------------------------- SIRT --

def test_sirt_relax_zero_is_wbp():
    rng = np.random.default_rng(10)
    P = rng.normal(size=(9, 3, 32)).astype(np.float32)
    ang = np.linspace(-40, 40, 9).astype(np.float32)
    w, _ = O.tilt_run(P, ang, O.TiltOptions(thickness=16))
    s, _ = O.tilt_run(P, ang, O.TiltOptions(thickness=16, sirt=2, sirt_relax=0.0))
    assert np.array_equal(w, s)


def test_sirt_one_iteration_by_composition():
    """V1 = V0 + relax * BP(RW(O - FP(V0))) with the residual unstretched and
    the fill taken from the original means (main.rs:919-953)."""
    rng = np.random.default_rng(11)
    P = rng.normal(size=(7, 2, 24)).astype(np.float32)
    ang = np.linspace(-30, 30, 7).astype(np.float32)
    means = np.array([O.min_max_mean(P[i])[2] for i in range(7)], np.float32)
    tab = _tab(ang)
    Pf = O.cosine_stretch(O.rweight(P.reshape(-1, 24)).reshape(P.shape), ang)
    V0 = O.backproject(Pf, tab, means, 24, 10)
    D = (P - O.forward_project(V0, tab, 24, 7)).astype(np.float32)
    Df = O.rweight(D.reshape(-1, 24)).reshape(D.shape)
    V1 = V0 + np.float32(0.7) * O.backproject(Df, tab, means, 24, 10)
    got, _ = O.tilt_run(P, ang, O.TiltOptions(thickness=10, sirt=1, sirt_relax=0.7))
    assert np.max(np.abs(got - V1)) < 1e-5 * np.sqrt(np.mean(V1 ** 2))


def test_sirt_trace_rows_equal_whole_run():
    """SIRT on a row subset (single-row geometry, full-image means) gives those
    rows of the whole run at every iterate, in FP32 and FP64: the iterates the
    per-iteration error is measured against are the whole run's."""
    rng = np.random.default_rng(14)
    P = (rng.random((11, 9, 40)) + 1.0).astype(np.float32)
    ang = np.linspace(-50, 50, 11).astype(np.float32)
    for prec in ("f32", "f64"):
        tr, res = O.sirt_trace(P, ang, O.TiltOptions(thickness=14, sirt=3), (2, 7), prec=prec)
        assert tr.shape == (4, 5, 14, 40) and res.shape == (3,)
        for k in range(4):
            whole, _ = O.tilt_run(P, ang, O.TiltOptions(thickness=14, sirt=k), prec=prec)
            assert np.array_equal(tr[k], whole[2:7]), (prec, k)
        assert np.all(np.diff(res) < 0)


def test_row_band_with_whole_image_means_equals_whole_run():
    """A band of projection rows (with a halo for x-axis tilt) and the whole
    images' means reconstructs those output rows as the whole run: exactly in
    FP64 (the form the FP64-oracle error of the large benchmark is computed in);
    in FP32 to the rounding of iy + dx*sin(xt) at a shifted iy."""
    rng = np.random.default_rng(16)
    nz, ny, nx = 11, 30, 48
    P = (rng.random((nz, ny, nx)) + 0.5).astype(np.float32)
    ang = np.linspace(-50, 50, nz).astype(np.float32)
    xt = np.full(nz, 2.0, np.float32)
    sx, sy = rng.normal(0, 1, nz), rng.normal(0, 1, nz)
    opts = O.TiltOptions(thickness=16)
    full_grids = O.build_local_grids([(v, 24.0, 15.0, sx[v], sy[v]) for v in range(nz)], nz)
    whole, _ = O.tilt_run(P, ang, opts, xtilt=xt, local=full_grids)
    p0, p1 = 4, 22
    band_grids = O.build_local_grids([(v, 24.0, 15.0 - p0, sx[v], sy[v]) for v in range(nz)], nz)
    means = np.array([np.mean(P[v], dtype=np.float64) for v in range(nz)])
    sub = np.ascontiguousarray(P[:, p0:p1])
    whole64, _ = O.tilt_run(P, ang, opts, xtilt=xt, local=full_grids, prec="f64")
    band64, _ = O.tilt_run(sub, ang, opts, xtilt=xt, local=band_grids, rows=(6, 12),
                           means=means, prec="f64")
    assert np.array_equal(band64, whole64[10:16])
    band, _ = O.tilt_run(sub, ang, opts, xtilt=xt, local=band_grids, rows=(6, 12), means=means)
    ref = whole[10:16].astype(np.float64)
    assert np.max(np.abs(band - ref)) <= 1e-5 * np.sqrt(np.mean(ref ** 2))
    # the band's own means (fill values) would be wrong by percents
    own, _ = O.tilt_run(sub, ang, opts, xtilt=xt, local=band_grids, rows=(6, 12))
    assert np.max(np.abs(own - ref)) > 1e-3 * np.sqrt(np.mean(ref ** 2))


def test_fp_threads_do_not_change_bits():
    rng = np.random.default_rng(15)
    V = rng.normal(size=(3, 12, 30)).astype(np.float32)
    ang = np.linspace(-60, 60, 9).astype(np.float32)
    tab = O.proj_params(ang, xtilt=np.full(9, 0.8, np.float32))
    a = O.forward_project(V, tab, 34, 9, nthreads=1)
    b = O.forward_project(V, tab, 34, 9, nthreads=8)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_fake_sirt_constant_volume_unchanged():
    P = np.full((5, 4, 20), 1.5, np.float32)
    ang = np.linspace(-30, 30, 5).astype(np.float32)
    v, _ = O.tilt_run(P, ang, O.TiltOptions(thickness=8, no_weight=True, no_cosine_stretch=True,
                                            fake_sirt=3))
    assert np.max(np.abs(v - 1.5)) < 1e-5


def test_output_transform_and_header_stats():
    rng = np.random.default_rng(12)
    P = rng.normal(size=(5, 3, 20)).astype(np.float32)
    ang = np.linspace(-30, 30, 5).astype(np.float32)
    v0, _ = O.tilt_run(P, ang, O.TiltOptions(thickness=8))
    v1, (mn, mx, mean) = O.tilt_run(P, ang, O.TiltOptions(thickness=8, densweight=2.0,
                                                          outscale=0.5, outadd=3.0))
    assert np.array_equal(v1, (v0 * np.float32(2.0)) * np.float32(0.5) + np.float32(3.0))
    assert mn == v1.min() and mx == v1.max()
    assert mean == pytest.approx(float(v1.astype(np.float64).mean()), rel=1e-6)


# ------------------------------------------------- newstack -x transforms --

def _then(a, b):
    """LinearTransform::then, crates/imod-transforms/src/xf.rs:83-92 (a, then b)."""
    a11, a12, a21, a22, adx, ady = [np.float32(v) for v in a]
    b11, b12, b21, b22, bdx, bdy = [np.float32(v) for v in b]
    return [b11 * a11 + b12 * a21, b11 * a12 + b12 * a22, b21 * a11 + b22 * a21,
            b21 * a12 + b22 * a22, b11 * adx + b12 * ady + bdx, b21 * adx + b22 * ady + bdy]


def test_xf_inverse_roundtrip_known_answer():
    # crates/imod-transforms/src/tests.rs:40-57
    xf = [0.95, -0.31, 0.31, 0.95, 5.0, -3.0]
    r = _then(xf, O.xf_inverse(xf))
    assert abs(r[0] - 1.0) < 1e-5 and abs(r[1]) < 1e-5 and abs(r[2]) < 1e-5
    assert abs(r[3] - 1.0) < 1e-5 and abs(r[4]) < 1e-4 and abs(r[5]) < 1e-4
    # the FP64 form inverts exactly to rounding
    inv64 = O.xf_inverse(xf, prec="f64")
    A = np.array([[0.95, -0.31], [0.31, 0.95]], np.float64)
    assert np.allclose(np.linalg.inv(np.array(A, np.float32).astype(np.float64)),
                       inv64[:4].reshape(2, 2), rtol=1e-12)


def test_xf_file_roundtrip_known_answer(tmp_path):
    # crates/imod-transforms/src/tests.rs:81-115, written in the 4f12.7 2f12.3 format
    # of write_xf_file (xf.rs:176-191); both parsers (oracle and library host) read it
    import tiltgpu
    xforms = [[1.0, 0.0, 0.0, 1.0, 0.0, 0.0],
              [0.9848077, -0.1736482, 0.1736482, 0.9848077, 5.123, -3.456]]
    p = tmp_path / "test.xf"
    p.write_text("".join("%12.7f%12.7f%12.7f%12.7f%12.3f%12.3f\n" % tuple(x) for x in xforms)
                 + "\n  bad line 1 2\n")
    for got in (O.read_xf_file(str(p)), tiltgpu.read_xf_file(str(p))):
        assert got.shape == (2, 6)
        for a, b in zip(got, xforms):
            assert np.all(np.abs(a[:4] - np.float32(b[:4])) < 1e-5)
            assert np.all(np.abs(a[4:] - np.float32(b[4:])) < 0.01)
    assert np.array_equal(O.read_xf_file(str(p)), tiltgpu.read_xf_file(str(p)))


def test_xf_apply_identity_translation_rotation():
    rng = np.random.default_rng(5)
    P = rng.normal(size=(3, 20, 30)).astype(np.float32)
    ident = np.tile(np.array([1, 0, 0, 1, 0, 0], np.float32), (3, 1))
    assert np.array_equal(O.apply_xf_stack(P, ident), P)
    # an integer translation moves the image and fills the uncovered band
    Q = O.apply_xf_stack(P, [1, 0, 0, 1, 3, -2], fill=7.0)
    assert np.array_equal(Q[:, :18, 3:], P[:, 2:, :27])
    assert np.all(Q[:, 18:, :] == 7.0) and np.all(Q[:, :, :3] == 7.0)
    # +90 degrees about the centre of an even square image: a permutation,
    # out[y][x] = in[n - x][y], with the x = 0 column outside
    S = P[:, :, :20]
    Q = O.apply_xf_stack(S, [0, -1, 1, 0, 0, 0], fill=-5.0)
    assert np.all(Q[:, :, 0] == -5.0)
    for x in range(1, 20):
        assert np.array_equal(Q[:, :, x], S[:, 20 - x, :])


def test_xf_apply_linear_ramp_is_reproduced():
    # bilinear interpolation is exact for an affine image: inside the image the
    # output is the ramp at the inverse-transformed position
    ny, nx = 40, 50
    yy, xx = np.mgrid[0:ny, 0:nx].astype(np.float64)
    img = (0.3 * xx - 0.7 * yy + 2.0).astype(np.float32)[None]
    xf = np.array([0.98, -0.12, 0.11, 1.02, 1.7, -2.3], np.float32)
    Q = O.apply_xf_stack(img, xf, fill=np.nan)
    inv = O.xf_inverse(xf, prec="f64")
    sx = inv[0] * (xx - nx / 2) + inv[1] * (yy - ny / 2) + inv[4] + nx / 2
    sy = inv[2] * (xx - nx / 2) + inv[3] * (yy - ny / 2) + inv[5] + ny / 2
    inside = (sx >= 0) & (sx <= nx - 1.001) & (sy >= 0) & (sy <= ny - 1.001)
    ref = 0.3 * sx - 0.7 * sy + 2.0
    assert inside.sum() > 0.8 * nx * ny
    assert np.max(np.abs(Q[0][inside] - ref[inside])) < 1e-4


def _xf_numpy_f32(P, xf, fill):
    """An independent vectorised restatement in numpy float32, same operation order."""
    nz, ny, nx = P.shape
    out = np.empty_like(P)
    f32 = np.float32
    xcen, ycen = f32(nx) / f32(2), f32(ny) / f32(2)
    yy, xx = np.mgrid[0:ny, 0:nx]
    xadj = xx.astype(f32) - xcen
    yadj = yy.astype(f32) - ycen
    for z in range(nz):
        t = np.asarray(xf, f32).reshape(-1, 6)[min(z, len(np.asarray(xf).reshape(-1, 6)) - 1)]
        a11, a12, a21, a22, dx, dy = [f32(v) for v in t]
        det = a11 * a22 - a12 * a21
        i11, i12, i21, i22 = a22 / det, -a12 / det, -a21 / det, a11 / det
        idx, idy = -(i11 * dx + i12 * dy), -(i21 * dx + i22 * dy)
        sx = i11 * xadj + i12 * yadj + idx + xcen
        sy = i21 * xadj + i22 * yadj + idy + ycen
        fx0, fy0 = np.floor(sx), np.floor(sy)
        fx, fy = sx - fx0, sy - fy0

        def tap(fxx, fyy):
            ok = (fxx >= 0) & (fxx < nx) & (fyy >= 0) & (fyy < ny)
            v = np.full(sx.shape, f32(fill), f32)
            v[ok] = P[z][fyy[ok].astype(np.int64), fxx[ok].astype(np.int64)]
            return v
        v00, v10 = tap(fx0, fy0), tap(fx0 + 1, fy0)
        v01, v11 = tap(fx0, fy0 + 1), tap(fx0 + 1, fy0 + 1)
        ofx, ofy = f32(1) - fx, f32(1) - fy
        out[z] = v00 * ofx * ofy + v10 * fx * ofy + v01 * ofx * fy + v11 * fx * fy
    return out


def test_xf_apply_equals_independent_numpy_restatement():
    rng = np.random.default_rng(9)
    P = rng.normal(size=(5, 33, 47)).astype(np.float32)
    xf = []
    for _ in range(3):  # fewer transforms than sections: the last one repeats
        a = np.deg2rad(rng.uniform(-20, 20))
        m = rng.uniform(0.9, 1.1)
        xf.append([m * np.cos(a), -m * np.sin(a), m * np.sin(a), m * np.cos(a),
                   rng.uniform(-6, 6), rng.uniform(-6, 6)])
    xf = np.array(xf, np.float32)
    Q = O.apply_xf_stack(P, xf, fill=0.25)
    assert np.array_equal(Q.view(np.uint32), _xf_numpy_f32(P, xf, 0.25).view(np.uint32))
    Q64 = O.apply_xf_stack(P, xf, fill=0.25, prec="f64")
    assert np.max(np.abs(Q - Q64)) < 1e-4 * np.sqrt(np.mean(Q64.astype(np.float64) ** 2)) * 50


def test_studio_reconstruction_equals_tilt_without_weighting():
    """imod-studio's reconstruction step (apps/imod-studio/src/main.rs:725-770)
    restated on its own is bit for bit the tilt oracle with --no-weight
    --no-cosine-stretch --fill-mode zero and thickness = width: two independent
    restatements of the same loop (cos_xt = zfactor = 1 multiply exactly)."""
    rng = np.random.default_rng(13)
    nz, ny, nx = 9, 4, 40
    P = rng.normal(size=(nz, ny, nx)).astype(np.float32)
    ang = np.linspace(-55, 55, 8).astype(np.float32)  # fewer angles than sections
    studio = O.studio_reconstruct(P, ang)
    vol, _ = O.tilt_run(P, ang, O.TiltOptions(thickness=nx, no_weight=True,
                                              no_cosine_stretch=True, fill_mode="zero"))
    assert np.array_equal(studio.view(np.uint32), vol.view(np.uint32))


# ------------------------------------------------------------ ctfphaseflip --

def test_electron_wavelength_known_values():
    # relativistic wavelength (textbook values: 0.01969 A at 300 kV, 0.02508 A at 200 kV)
    assert abs(O.electron_wavelength(300.0) - 0.019687) < 2e-6
    assert abs(O.electron_wavelength(200.0) - 0.025079) < 2e-6


@pytest.mark.parametrize("defocus_nm", [-3000.0, 3000.0, 12000.0])
def test_ctf_flip_of_a_single_frequency(defocus_nm):
    """One strip covering a cosine of 5 cycles: its two Fourier coefficients are
    negated exactly when CTF(s) = -sin(pi*lambda*s^2*def) (Cs = 0, w = 0) is
    negative, so the output is +-input; column 0 sits on the strip edge where the
    triangular weight is 0 and keeps the input."""
    nx, ny, pix = 64, 48, 2.0
    x = np.arange(nx)
    img = np.repeat(np.cos(2 * np.pi * 5 * x / nx)[None, :], ny, 0).astype(np.float32)
    P = np.stack([img, 2 * img])
    s = 5.0 / (nx * pix)
    ctf = -np.sin(np.pi * O.electron_wavelength(300.0) * s * s * defocus_nm * 10.0)
    Q = O.ctf_phaseflip(P, [0.0, 0.0], pix, cs_mm=0.0, amp=0.0, defocus_nm=defocus_nm,
                        strip_width=nx)
    sign = np.sign(ctf)
    assert np.max(np.abs(Q[:, :, 1:] - sign * P[:, :, 1:])) < 1e-5
    assert np.array_equal(Q[:, :, 0], P[:, :, 0])


def test_ctf_no_flip_is_identity_and_precisions_agree():
    """Zero defocus, Cs 0, full amplitude contrast: CTF = cos(0) = 1 everywhere,
    nothing is flipped, and the strips give back the image to transform
    rounding; on a real case the f32 and f64 paths agree."""
    rng = np.random.default_rng(17)
    P = rng.normal(size=(3, 40, 56)).astype(np.float32)
    Q = O.ctf_phaseflip(P, [0.0, 20.0, -40.0], 2.5, cs_mm=0.0, amp=1.0, defocus_nm=0.0,
                        strip_width=16)
    assert np.max(np.abs(Q - P)) < 1e-5
    kw = dict(defocus_nm=-4000.0, strip_width=16, tilt_axis_deg=7.0, plate_deg=30.0,
              vdef=[(-4000, -4300, 20), (-4100, -4100, 0), (-3900, -4200, 35)])
    a = O.ctf_phaseflip(P, [0.0, 20.0, -40.0], 2.5, **kw)
    b = O.ctf_phaseflip(P, [0.0, 20.0, -40.0], 2.5, prec="f64", **kw)
    assert np.sqrt(np.mean((a.astype(np.float64) - b) ** 2)) < 1e-5 * np.sqrt(np.mean(b ** 2))


def test_ctf_defocus_file_and_interpolation(tmp_path):
    # read_defocus_file / interpolate_defocus, apps/ctfphaseflip/src/main.rs:103-198
    p = tmp_path / "d.txt"
    p.write_text("# view d1 d2 astig\n1 -3000\n\n4 -3600 -3900 30\n6 -4000 -4100\n")
    e, v = O.read_defocus_file(str(p))
    assert v == [1, 4, 6]
    assert e[0] == (np.float32(-3000), np.float32(-3000), np.float32(0))
    assert e[2] == (np.float32(-4000), np.float32(-4100), np.float32(0))
    d = O.interpolate_defocus(e, v, 8)
    assert np.array_equal(d[0], d[1]) and np.array_equal(d[7], d[6])
    t = np.float32(1) / np.float32(3)
    assert d[2][0] == np.float32(-3000) + t * (np.float32(-3600) - np.float32(-3000))
    assert d[5][1] == np.float32(-3900) + np.float32(0.5) * (np.float32(-4100) - np.float32(-3900))
    bad = tmp_path / "bad.txt"
    bad.write_text("2\n")
    with pytest.raises(ValueError):
        O.read_defocus_file(str(bad))


def test_ctf_strip_width_rule():
    # apps/ctfphaseflip/src/main.rs:272-298
    lam = O.electron_wavelength(300.0)
    # pixel / (lambda * max|defocus| / nx), at least 16: 6.9 -> 16 here
    assert O.ctf_strip_width(256, 2048, 2.0, lam, -3000.0) == 16
    assert O.ctf_strip_width(256, 2048, 20.0, lam, -3000.0) == int(20.0 / (lam * 30000 / 2048))
    assert O.ctf_strip_width(256, 2048, 2.0, lam, -3000.0, max_strip_width=100) == 100
    assert O.ctf_strip_width(256, 2048, 2.0, lam, 0.0) == 256


# ------------------------------- two-row geometry, local grids, fill modes --

def _geometry_fixture():
    g = np.load(os.path.join(GOLDEN, "geometry_bruteforce.npz"))
    return g, [k.split("/")[0] for k in g.files if k.endswith("/angles")]


def _case_inputs(g, name):
    nz = int(g["nz"])
    pts = g[f"{name}/patches"]
    grids = None
    if len(pts):
        grids = O.build_local_grids([(int(q[0]), q[1], q[2], q[3], q[4]) for q in pts], nz)
    return g[f"{name}/angles"], g[f"{name}/xtilt"], g[f"{name}/zf"], grids


@pytest.mark.parametrize("fill", ["mean", "zero", "edge"])
def test_two_row_geometry_bp_bruteforce_fixture(fill):
    """The oracle's backprojection with x-axis tilt, z factors and local grids,
    under each fill mode, equals an FP64 sparse-operator brute force that shares
    no code with it (tests/golden/make_golden.py: tent weights, numpy.interp grid
    interpolation and Edge clamp): the row selection with the last-row fallback
    and the y-blend (main.rs:1051-1066), the fills (:386-406), the local shifts
    (:171-228)."""
    g, cases = _geometry_fixture()
    assert sorted(cases) == ["grid3x3", "xtilt", "xtilt_zf_local1"]
    P = g["P"]
    opts = O.TiltOptions(thickness=int(g["out_nz"]), width=int(g["out_nx"]), no_weight=True,
                         no_cosine_stretch=True, fill_mode=fill)
    for name in cases:
        ang, xt, zf, grids = _case_inputs(g, name)
        ref = g[f"{name}/bp_{fill}"]
        rms = np.sqrt(np.mean(ref ** 2))
        v64, _ = O.tilt_run(P, ang, opts, xtilt=xt, zfactor=zf, local=grids, prec="f64")
        v32, _ = O.tilt_run(P, ang, opts, xtilt=xt, zfactor=zf, local=grids, prec="f32")
        assert np.max(np.abs(v64 - ref)) <= 1e-6 * rms, (name, fill)
        assert np.max(np.abs(v32 - ref)) <= 1e-5 * rms, (name, fill)


def test_two_row_geometry_fp_is_the_transposed_operator():
    """The oracle's forward projection with x-axis tilt, z factors and local grids
    is the transpose of the zero-fill sampling operator (without 1/nz), entry by
    entry: the four two-row scatter weights and the one-row fallback of
    main.rs:866-889."""
    g, cases = _geometry_fixture()
    nx, nz = int(g["nx"]), int(g["nz"])
    for name in cases:
        ang, xt, zf, grids = _case_inputs(g, name)
        ref = g[f"{name}/fp"]
        scale = np.max(np.abs(ref))
        r64 = O.forward_project_full(g["Vin"], ang, nx, nz, xtilt=xt, zfactor=zf, local=grids,
                                     prec="f64")
        r32 = O.forward_project_full(g["Vin"], ang, nx, nz, xtilt=xt, zfactor=zf, local=grids,
                                     prec="f32")
        assert np.max(np.abs(r64 - ref)) <= 1e-6 * scale, name
        assert np.max(np.abs(r32 - ref)) <= 1e-5 * scale, name


def test_two_row_fixture_hits_every_rule():
    """The fixture's cases reach each branch the rules name: both rows, the
    one-row fallback at the last row, a first row with nothing added, samples
    left and right of the row, a view with y offset 0 on the x-tilt case and a
    view with no patches on the grid case."""
    g, _ = _geometry_fixture()
    ny, nx, out_nx, out_nz = int(g["ny"]), int(g["nx"]), int(g["out_nx"]), int(g["out_nz"])
    ang = g["xtilt/angles"].astype(np.float64)
    xt = np.deg2rad(g["xtilt/xtilt"].astype(np.float64))
    dx = np.arange(out_nx) - out_nx / 2.0
    dz = np.arange(out_nz) - out_nz / 2.0
    sy_last = (ny - 1) + dx * np.sin(xt[0])
    assert np.any(np.floor(sy_last) == ny - 1) and np.any(np.floor(sy_last) == ny - 2)
    assert np.any(np.floor(0 + dx * np.sin(xt[0])) == -1)
    px = (dx[None, :] * np.cos(np.deg2rad(ang[0])) + dz[:, None] * np.sin(np.deg2rad(ang[0]))
          + nx / 2.0)
    assert np.any(px < 0) and np.any(np.floor(px) >= nx - 1)
    assert g["xtilt/xtilt"][3] == 0.0
    assert not np.any(g["grid3x3/patches"][:, 0] == 2)
    # the fills differ where samples fall outside the row, so each fill is seen
    a, b, c = (g[f"xtilt/bp_{f}"] for f in ("mean", "zero", "edge"))
    assert not np.allclose(a, b) and not np.allclose(a, c) and not np.allclose(b, c)


def test_two_row_adjoint_identity_with_local_grid():
    """<FP V, S> = nz <V, BP_zero-fill S> for the two-row geometry with a local
    grid, in FP64 (main.rs:866-889 against :1040-1066)."""
    g, _ = _geometry_fixture()
    ang, xt, zf, grids = _case_inputs(g, "xtilt_zf_local1")
    nz, ny, nx = int(g["nz"]), int(g["ny"]), int(g["nx"])
    rng = np.random.default_rng(77)
    S = rng.normal(size=(nz, ny, nx)).astype(np.float32)
    V = g["Vin"]
    R = O.forward_project_full(V, ang, nx, nz, xtilt=xt, zfactor=zf, local=grids, prec="f64")
    opts = O.TiltOptions(thickness=int(g["out_nz"]), width=int(g["out_nx"]), no_weight=True,
                         no_cosine_stretch=True, fill_mode="zero")
    B, _ = O.tilt_run(S, ang, opts, xtilt=xt, zfactor=zf, local=grids, prec="f64")
    lhs = float(np.sum(R.astype(np.float64) * S))
    rhs = nz * float(np.sum(V.astype(np.float64) * B))
    assert lhs == pytest.approx(rhs, rel=1e-6)


# ------------------------------------------------------------- fake-SIRT --

def _box3_clamped(V):
    """One separable 3x3x3 box pass with clamped (replicated) edges, the textbook
    form: scipy.ndimage.uniform_filter with mode 'nearest'."""
    from scipy.ndimage import uniform_filter
    return uniform_filter(np.asarray(V, np.float64), size=3, mode="nearest")


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 2e-6)])
def test_lowpass_impulse_response_and_edges(prec, tol):
    """lowpass_3d_volume (main.rs:493-538) on unit impulses: 1/27 on the 3x3x3
    cube around an interior voxel; at a corner the clamped edge gives weights
    (2/3, 1/3) per axis.  A ramp along each axis is kept inside and pulled in by
    1/3 of a step at both ends."""
    ny, nz, nx = 5, 6, 7
    V = np.zeros((ny, nz, nx))
    V[2, 3, 3] = 1.0
    out = O.lowpass_3d(V, prec=prec)
    exp = np.zeros_like(V)
    exp[1:4, 2:5, 2:5] = 1.0 / 27.0
    assert np.max(np.abs(out - exp)) < tol
    V = np.zeros((ny, nz, nx))
    V[0, 0, 0] = 1.0
    w = np.array([2.0 / 3.0, 1.0 / 3.0])
    out = O.lowpass_3d(V, prec=prec)
    exp = np.zeros_like(V)
    exp[:2, :2, :2] = w[:, None, None] * w[None, :, None] * w[None, None, :]
    assert np.max(np.abs(out - exp)) < tol
    for axis, n in ((0, ny), (1, nz), (2, nx)):
        shape = [1, 1, 1]
        shape[axis] = n
        ramp = np.broadcast_to(np.arange(n, dtype=np.float64).reshape(shape), (ny, nz, nx))
        out = O.lowpass_3d(np.array(ramp), prec=prec)
        exp = np.array(ramp, np.float64)
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[axis] = 0
        hi[axis] = n - 1
        exp[tuple(lo)] += 1.0 / 3.0
        exp[tuple(hi)] -= 1.0 / 3.0
        assert np.max(np.abs(out - exp)) < tol * n, axis


@pytest.mark.parametrize("passes", [1, 2])
def test_fake_sirt_equals_box_filtered_wbp(passes):
    """--fake-sirt N (main.rs:976-1001) is N clamped 3x3x3 box passes of the WBP
    volume, before the output transform."""
    rng = np.random.default_rng(19)
    P = rng.normal(size=(7, 5, 24)).astype(np.float32)
    ang = np.linspace(-45, 45, 7).astype(np.float32)
    wbp, _ = O.tilt_run(P, ang, O.TiltOptions(thickness=9), prec="f64")
    ref = wbp.astype(np.float64)
    for _ in range(passes):
        ref = _box3_clamped(ref)
    got64, _ = O.tilt_run(P, ang, O.TiltOptions(thickness=9, fake_sirt=passes), prec="f64")
    got32, _ = O.tilt_run(P, ang, O.TiltOptions(thickness=9, fake_sirt=passes), prec="f32")
    rms = np.sqrt(np.mean(ref ** 2))
    assert np.max(np.abs(got64 - ref)) < 1e-6 * rms
    assert np.max(np.abs(got32 - ref)) < 1e-5 * rms
