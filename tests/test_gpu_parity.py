"""GPU parity: libtiltgpu (through its C ABI) against the FP32 oracle, which
restates apps/tilt/src/main.rs in the reference's operation order.

Bar: max |gpu - oracle| <= 1e-4 * RMS(oracle volume) for every FP32
result.  FP16/BF16 storage is compared with the FP64 oracle at the tolerance
stated in its test.
"""
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-4


def case(tg, nx=128, ny=20, nviews=31, thick=48, seed=0, span=60.0, xtilt=None,
         shifts=None, n_ell=30, n_beads=5):
    angles = np.linspace(-span, span, nviews).astype(np.float32)
    sx = sy = None
    if shifts is not None:
        sx, sy = shifts
    P = tg.synth_series(nx, ny, angles, vol_nz=thick, n_ellipsoids=n_ell, n_beads=n_beads,
                        seed=seed, xtilt=xtilt, shift_x=sx, shift_y=sy,
                        axis=(2.0, 10.0), bead_radius=2.0)
    return P, angles


def both(tg, P, angles, gopts, xtilt=None, zfactor=None, local=None, prec="f32"):
    print(f"case: shape={P.shape} opts={gopts}", flush=True)
    vol_g, st_g = tg.reconstruct(P, angles, gopts, xtilt=xtilt, zfactor=zfactor, local=local)
    oopts = O.TiltOptions(**{k: v for k, v in vars(gopts).items() if k not in ("storage", "geometry")})
    if oopts.exact_filter_size is None:
        oopts.exact_filter_size = 0
    vol_o, st_o = O.tilt_run(P, angles, oopts, xtilt=xtilt, zfactor=zfactor, local=local,
                             prec=prec)
    return vol_g, st_g, vol_o, st_o


def test_wbp_defaults(gpu, rel):
    P, ang = case(gpu)
    g, sg, o, so = both(gpu, P, ang, gpu.TiltOptions(thickness=48))
    assert g.shape == o.shape == (20, 48, 128)
    assert rel(g, o) <= TOL
    assert np.allclose(sg, so, rtol=1e-4, atol=1e-4 * np.sqrt(np.mean(o.astype(np.float64) ** 2)))


@pytest.mark.parametrize("opts", [
    dict(no_weight=True, no_cosine_stretch=True),
    dict(no_weight=True),
    dict(no_cosine_stretch=True),
    dict(fill_mode="zero"),
    dict(fill_mode="edge"),
    dict(hamming=True, radial=(0.3, 0.05)),
    dict(exact_filter_size=40, radial=(0.4, 0.1)),
    dict(cosine_interp="cubic"),
    dict(max_cosine_tilt=30.0),
    dict(densweight=2.0, outscale=0.5, outadd=3.0),
    dict(width=96),
    dict(width=200, thickness=100),
    dict(log=True),
])
def test_wbp_options(gpu, rel, opts):
    P, ang = case(gpu, seed=1)
    if opts.get("log"):
        P = np.abs(P) + 0.5
    kw = dict(thickness=48)
    kw.update(opts)
    g, sg, o, so = both(gpu, P, ang, gpu.TiltOptions(**kw))
    assert g.shape == o.shape
    assert rel(g, o) <= TOL, opts


@pytest.mark.parametrize("nx", [96, 120, 160, 250, 768, 131, 194, 259, 1958, 182, 195, 1001, 154])
def test_wbp_mixed_radix_widths(gpu, rel, nx):
    P, ang = case(gpu, nx=nx, ny=6, nviews=21, thick=32, seed=nx)
    g, _, o, _ = both(gpu, P, ang, gpu.TiltOptions(thickness=32))
    assert rel(g, o) <= TOL


@pytest.mark.parametrize("nx", [4100, 5760, 10200])
def test_wbp_wide_rows(gpu, rel, nx):
    """Rows wider than 4096 take the mixed-radix filter with more than 64 KiB of
    LDS per workgroup (159.4 KiB at 10200)."""
    P, ang = case(gpu, nx=nx, ny=3, nviews=7, thick=16, seed=nx, n_ell=40)
    g, _, o, _ = both(gpu, P, ang, gpu.TiltOptions(thickness=16))
    assert rel(g, o) <= TOL


def test_wbp_width_beyond_lds_is_refused(gpu):
    P = np.zeros((2, 1, 10240), np.float32)  # 160 KiB of dynamic LDS: over the limit
    with pytest.raises(gpu.TiltGpuError) as ei:
        gpu.reconstruct(P, np.zeros(2, np.float32), gpu.TiltOptions(thickness=4))
    assert ei.value.code == gpu.ERR_ARG and "too large" in str(ei.value)


def test_wbp_odd_rows_and_fewer_angles(gpu, rel):
    # in_ny odd: FFT row pairs straddle views; 25 angles for 27 sections
    P, ang = case(gpu, ny=7, nviews=27, thick=40, seed=3)
    g, _, o, _ = both(gpu, P, ang[:25], gpu.TiltOptions(thickness=40, no_cosine_stretch=True))
    assert rel(g, o) <= TOL


def test_wbp_xtilt(gpu, rel):
    P, ang = case(gpu, seed=4, xtilt=np.full(31, 0.5, np.float32))
    xt = np.full(31, 0.5, np.float32)
    g, _, o, _ = both(gpu, P, ang, gpu.TiltOptions(thickness=48), xtilt=xt)
    assert rel(g, o) <= TOL


def test_wbp_xtilt_large_and_zfactor(gpu, rel):
    P, ang = case(gpu, seed=5, ny=40)
    xt = np.linspace(-3, 3, 31).astype(np.float32)
    zf = np.linspace(0.9, 1.1, 31).astype(np.float32)
    g, _, o, _ = both(gpu, P, ang, gpu.TiltOptions(thickness=48), xtilt=xt, zfactor=zf)
    assert rel(g, o) <= TOL


def test_wbp_local_constant_shifts(gpu, rel):
    rng = np.random.default_rng(6)
    sx = rng.normal(0, 2, 31).astype(np.float32)
    sy = rng.normal(0, 2, 31).astype(np.float32)
    P, ang = case(gpu, seed=6, shifts=(sx, sy), xtilt=np.full(31, 0.5, np.float32))
    patches = [(v, 64.0, 10.0, float(sx[v]), float(sy[v])) for v in range(31)]
    grids = O.build_local_grids(patches, 31)
    xt = np.full(31, 0.5, np.float32)
    g, _, o, _ = both(gpu, P, ang, gpu.TiltOptions(thickness=48), xtilt=xt, local=grids)
    assert rel(g, o) <= TOL


def test_wbp_local_grid(gpu, rel):
    rng = np.random.default_rng(7)
    P, ang = case(gpu, seed=7, ny=32)
    patches = []
    for v in range(31):
        if v == 5:
            continue  # a view without patches: no shift for it
        for x in (20.0, 64.0, 110.0):
            for y in (4.0, 16.0, 28.0):
                patches.append((v, x, y, float(rng.normal(0, 1.5)), float(rng.normal(0, 1.5))))
    grids = O.build_local_grids(patches, 31)
    g, _, o, _ = both(gpu, P, ang, gpu.TiltOptions(thickness=48), local=grids)
    assert rel(g, o) <= TOL


def test_filter_stage_and_means(gpu, rel):
    P, ang = case(gpu, seed=8, ny=9)
    opts = gpu.TiltOptions(thickness=48)
    vol, hs, s = gpu.reconstruct(P, ang, opts, return_session=True)
    filt = s.read_filtered()
    means = s.read_means()
    s.close()
    ref_means = np.array([O.min_max_mean(P[v])[2] for v in range(P.shape[0])], np.float32)
    assert np.array_equal(means, ref_means)
    rw = O.rweight(P.reshape(-1, P.shape[2]), prec="f32").reshape(P.shape)
    st = O.cosine_stretch(rw, ang, prec="f32")
    assert rel(filt, st) <= 1e-5


@pytest.mark.parametrize("nx", [64, 256, 512, 1024, 2048, 4096])
def test_filter_power_of_two_widths(gpu, rel, nx):
    """The power-of-two filter (register radix-8 passes, 64 <= n <= 2048; 4096
    takes the mixed-radix passes) against the oracle's R-weighting + stretch,
    with an odd row count so a row pair straddles two views and the last
    workgroup holds fewer pairs than it can.  Widths 256-2048 also take the
    per-view means from the row sums of the filter's load; every width's means
    equal the oracle's (min_max_mean) bit for bit."""
    ny, nv = 5, 7
    ang = np.linspace(-60.0, 60.0, nv).astype(np.float32)
    rng = np.random.default_rng(nx)
    P = rng.normal(1.0, 0.5, size=(nv, ny, nx)).astype(np.float32)
    print(f"case: filter width {nx}", flush=True)
    vol, hs, s = gpu.reconstruct(P, ang, gpu.TiltOptions(thickness=8), return_session=True)
    filt = s.read_filtered()
    means = s.read_means()
    s.close()
    ref_means = np.array([O.min_max_mean(P[v])[2] for v in range(nv)], np.float32)
    assert np.array_equal(means, ref_means)
    rw = O.rweight(P.reshape(-1, nx), prec="f32").reshape(P.shape)
    st = O.cosine_stretch(rw, ang, prec="f32")
    print(f"width {nx}: vs FP32 oracle {rel(filt, st):.2e}")
    assert rel(filt, st) <= 1e-5


def test_sirt_matches_oracle(gpu, rel):
    P, ang = case(gpu, nx=96, ny=6, nviews=21, thick=32, seed=9)
    opts = gpu.TiltOptions(thickness=32, sirt=4, sirt_relax=0.8)
    vol, hs, s = gpu.reconstruct(P, ang, opts, return_session=True)
    res = s.residuals()
    s.close()
    o, _ = O.tilt_run(P, ang, O.TiltOptions(thickness=32, sirt=4, sirt_relax=0.8))
    assert rel(vol, o) <= TOL
    assert res.shape == (4,) and np.all(np.isfinite(res)) and np.all(res > 0)


@pytest.mark.parametrize("in_nx,ny,out_nx,thick,nviews", [
    (200, 3, 200, 40, 23),    # table: 4 bins per lane
    (400, 2, 333, 57, 21),    # 8 bins per lane, odd output width
    (700, 2, 900, 33, 18),    # 4 bins per lane, wider output than input
    (1040, 1, 1024, 24, 9),   # 8 bins per lane, two lane windows per row
    (2300, 3, 2300, 12, 7),   # 16 bins per lane; wider than the two-row walk stages
])
@pytest.mark.parametrize("rows", [1, 2])
@pytest.mark.parametrize("slide", [0, 1])
def test_reproject_bitwise(gpu, monkeypatch, in_nx, ny, out_nx, thick, nviews, rows, slide):
    """Forward projection (SIRT reprojection) equals the reference's serial
    scatter (main.rs:830-896) bit for bit in the single-row geometry, including
    views at 0 and +-60 deg, z factors (one of 3.5, whose z rows move the
    sliding windows by more than one window), and views steeper than 87 deg and
    past 90 deg (the per-bin scan path), with one or two volume rows per
    workgroup (TILTGPU_FP_ROWS; an odd row count leaves the last row alone, and
    rows wider than the two-row walk stages walk one row), with fixed and with
    sliding lane windows (TILTGPU_FP_SLIDE)."""
    monkeypatch.setenv("TILTGPU_FP_ROWS", str(rows))
    monkeypatch.setenv("TILTGPU_FP_SLIDE", str(slide))
    rng = np.random.default_rng(in_nx + thick)
    ang = np.linspace(-62, 62, nviews).astype(np.float32)
    ang[1], ang[2], ang[-1] = 0.0, 87.5, 95.0
    zf = np.where(np.arange(nviews) % 3 == 0, 1.3, 1.0).astype(np.float32)
    zf[4] = 3.5
    V = rng.standard_normal((ny, thick, out_nx)).astype(np.float32)
    print(f"case: reproject in_nx={in_nx} out_nx={out_nx} thick={thick} views={nviews} "
          f"rows={rows} slide={slide}", flush=True)
    with gpu.Session(in_nx, ny, nviews, gpu.TiltOptions(thickness=thick, width=out_nx,
                                                        no_cosine_stretch=True)) as s:
        s.set_angles(ang)
        s.set_zfactor(zf)
        R = s.reproject(V)
        R2 = s.reproject(V)
    ref = O.forward_project(V, O.proj_params(ang, zfactor=zf), in_nx, nviews)
    assert np.array_equal(R.view(np.uint32), ref.view(np.uint32))  # bit patterns
    assert np.array_equal(R.view(np.uint32), R2.view(np.uint32))


@pytest.mark.parametrize("local", [False, True])
def test_reproject_xtilt_scatter(gpu, rel, local):
    """Two-row reprojection (x-axis tilt, optionally a 3x2 local grid): the same
    bits on repeated runs (fixed-point sums), within FP32 summation rounding of
    the FP32 oracle's serial scatter and closer still to the FP64 oracle's."""
    rng = np.random.default_rng(3)
    ang = np.linspace(-60, 60, 15).astype(np.float32)
    xt = np.linspace(-0.7, 0.9, 15).astype(np.float32)
    V = rng.standard_normal((8, 30, 96)).astype(np.float32)
    grids = None
    if local:
        pts = [(v, x, y, float(rng.normal(0, 1.2)), float(rng.normal(0, 0.8)))
               for v in range(15) for x in (10.0, 48.0, 90.0) for y in (1.0, 6.0)]
        grids = O.build_local_grids(pts, 15)
    with gpu.Session(96, 8, 15, gpu.TiltOptions(thickness=30, no_cosine_stretch=True)) as s:
        s.set_angles(ang)
        s.set_xtilt(xt)
        if grids is not None:
            s.set_local(*grids.flat())
        R = s.reproject(V)
        R2 = s.reproject(V)
    assert np.array_equal(R.view(np.uint32), R2.view(np.uint32))
    ref = O.forward_project_full(V, ang, 96, 15, xtilt=xt, local=grids)
    ref64 = O.forward_project_full(V, ang, 96, 15, xtilt=xt, local=grids, prec="f64")
    assert rel(R, ref) <= 1e-5
    assert rel(R, ref64) <= rel(ref, ref64) + 1e-7


def test_sirt_trace_per_iteration(gpu, rel):
    """Every SIRT iterate of rows [3, 9) against the FP32 oracle's iterates of the
    same rows (the oracle runs only those rows: in the single-row geometry rows
    evolve independently); the last iterate equals those rows of the output."""
    P, ang = case(gpu, nx=128, ny=12, nviews=25, thick=40, seed=21)
    opts = gpu.TiltOptions(thickness=40, sirt=5, sirt_relax=0.9)
    print("case: SIRT-5 trace rows 3..9 of 128x12x25", flush=True)
    vol, _, s = gpu.reconstruct(P, ang, opts, return_session=True, trace_rows=(3, 9))
    tr, res = s.sirt_trace(), s.residuals()
    s.close()
    ref, ref_res = O.sirt_trace(P, ang, O.TiltOptions(thickness=40, sirt=5, sirt_relax=0.9),
                                (3, 9), prec="f32")
    assert tr.shape == ref.shape == (6, 6, 40, 128)
    for k in range(6):
        assert rel(tr[k], ref[k]) <= TOL, k
    assert np.array_equal(tr[-1], vol[3:9])
    assert res.shape == (5,) and np.all(np.diff(res) < 0)  # the residual norm falls


def test_run_rows_equals_full_run(gpu):
    """Row ranges (tiltgpu_run_rows) equal those rows of the whole WBP, bit for bit,
    with x-axis tilt (two-row geometry); the filtering is reused across ranges and
    redone after new angles."""
    xt = np.full(21, 0.5, np.float32)
    P, ang = case(gpu, nx=128, ny=24, nviews=21, thick=40, seed=31, xtilt=xt)
    opts = gpu.TiltOptions(thickness=40)
    print("case: run_rows 128x24x21 xtilt 0.5", flush=True)
    full, _ = gpu.reconstruct(P, ang, opts, xtilt=xt)
    with gpu.Session(128, 24, 21, opts) as s:
        s.set_angles(ang)
        s.set_xtilt(xt)
        s.upload(P)
        s.run_rows(5, 13)
        a, _ = s.read_rows(5, 13)
        s.run_rows(0, 4)
        b, _ = s.read_rows(0, 4)
        s.set_angles(ang[::-1].copy())
        s.run_rows(0, 4)
        c, _ = s.read_rows(0, 4)
    assert np.array_equal(a, full[5:13]) and np.array_equal(b, full[0:4])
    flipped, _ = gpu.reconstruct(P, ang[::-1].copy(), opts, xtilt=xt)
    assert np.array_equal(c, flipped[0:4])
    with gpu.Session(128, 24, 21, gpu.TiltOptions(thickness=40, sirt=2)) as s:
        s.set_angles(ang)
        s.upload(P)
        with pytest.raises(gpu.TiltGpuError):
            s.run_rows(0, 4)


@pytest.mark.parametrize("geometry", ["plain", "xtilt", "local"])
def test_launch_shapes_do_not_change_results(gpu, geometry):
    """Every backprojection launch shape (tile depth, views staged per barrier,
    x positions per lane) and FP bins-per-lane choice gives the same bits (WBP
    and SIRT); the default comes from the tuned table."""
    xt = None if geometry == "plain" else np.full(21, 0.5, np.float32)
    P, ang = case(gpu, nx=160, ny=10, nviews=21, thick=72, seed=41, xtilt=xt)
    grids = None
    if geometry == "local":
        rng = np.random.default_rng(43)
        grids = O.build_local_grids([(v, x, y, float(rng.normal(0, 1.5)), float(rng.normal(0, 0.7)))
                                     for v in range(21) for x in (20.0, 140.0) for y in (1.0, 8.0)],
                                    21)
    outs = []
    for tz, bins, vps, xpl in [(0, 0, 0, 0), (32, 4, 1, 1), (64, 8, 1, 1), (128, 16, 1, 1),
                               (256, 8, 1, 1), (32, 8, 2, 1), (64, 4, 2, 1), (128, 8, 2, 1),
                               (32, 16, 1, 2), (64, 8, 1, 2)]:
        with gpu.Session(160, 10, 21, gpu.TiltOptions(thickness=72, sirt=2)) as s:
            s.set_angles(ang)
            if xt is not None:
                s.set_xtilt(xt)
            if grids is not None:
                s.set_local(*grids.flat())
            s.set_tuning(tz, bins)
            s.set_bp_shape(vps, xpl)
            s.upload(P)
            s.run()
            outs.append(s.read_rows()[0])
            t = s.tuning()
            if tz:
                assert t["tile_z"] == tz and t["tile_source"] == 2 and t["fp_bins"] == bins
                assert s.bp_shape() == {"tile_z": tz, "views_per_stage": vps, "x_per_lane": xpl}
            else:
                assert t["tile_source"] == 1 and t["fp_source"] == 1  # the tuned table
    for o in outs[1:]:
        # the two-row reprojection sums in fixed point: order-independent too
        assert np.array_equal(o, outs[0])


def test_unbuilt_launch_shape_is_refused(gpu):
    P, ang = case(gpu, nx=96, ny=4, nviews=9, thick=40, seed=44)
    with gpu.Session(96, 4, 9, gpu.TiltOptions(thickness=40)) as s:
        s.set_angles(ang)
        s.set_tuning(256, 0)
        s.set_bp_shape(2, 1)  # two views per stage is not built at depth 256
        s.upload(P)
        with pytest.raises(gpu.TiltGpuError) as e:
            s.run()
        assert e.value.code == gpu.ERR_ARG
    with pytest.raises(gpu.TiltGpuError):
        with gpu.Session(96, 4, 9, gpu.TiltOptions(thickness=40)) as s:
            s.set_bp_shape(3, 1)


def test_sirt_xtilt_scatter(gpu, rel):
    P, ang = case(gpu, nx=96, ny=10, nviews=21, thick=32, seed=10)
    xt = np.full(21, 1.0, np.float32)
    opts = gpu.TiltOptions(thickness=32, sirt=2)
    g, _, o, _ = both(gpu, P, ang, opts, xtilt=xt)
    assert rel(g, o) <= TOL


def test_fake_sirt(gpu, rel):
    P, ang = case(gpu, seed=11, ny=8)
    g, _, o, _ = both(gpu, P, ang, gpu.TiltOptions(thickness=48, fake_sirt=2,
                                                   densweight=1.5))
    assert rel(g, o) <= TOL


def test_seam_session_rows(gpu, rel):
    P, ang = case(gpu, seed=12, ny=10)
    rw = O.rweight(P.reshape(-1, P.shape[2]), prec="f32").reshape(P.shape)
    st = O.cosine_stretch(rw, ang, prec="f32")
    means = np.array([O.min_max_mean(P[v])[2] for v in range(P.shape[0])], np.float32)
    tab = O.proj_params(ang, prec="f32")
    s = gpu.GpuSession.new(st, (ang,), means, 128, 10, 31, 128, 48)
    assert s is not None
    row3 = s.reconstruct_row(3).reshape(48, 128)
    rows = s.reconstruct_rows(0, 10)
    s.close()
    ref = O.backproject(st, tab, means, 128, 48, fill="mean", prec="f32")
    assert rel(row3, ref[3]) <= TOL
    assert rel(rows, ref) <= TOL


def test_deterministic(gpu):
    P, ang = case(gpu, seed=13, ny=8, xtilt=None)
    a, _ = gpu.reconstruct(P, ang, gpu.TiltOptions(thickness=48))
    b, _ = gpu.reconstruct(P, ang, gpu.TiltOptions(thickness=48))
    assert np.array_equal(a, b)


@pytest.mark.parametrize("storage,tol_max,tol_rms", [("f16", 5e-3, 5e-4), ("bf16", 4e-2, 4e-3)])
def test_half_storage(gpu, rel, storage, tol_max, tol_rms):
    """FP16/BF16 projection storage (per-view scale, FP32 accumulation).  Its
    error is measured against the FP32 GPU path, which makes the same
    in-range/fill decisions, and stated against the FP64 oracle.  (The FP32
    reference itself differs from FP64 by up to a few 1e-2 x RMS at voxels whose
    fill decision flips between FP32 and FP64 coordinates, so max|d| vs FP64 is
    not a storage-precision measure; RMS is.)"""
    P, ang = case(gpu, seed=14, ny=8)
    g32, _ = gpu.reconstruct(P, ang, gpu.TiltOptions(thickness=48))
    g, _ = gpu.reconstruct(P, ang, gpu.TiltOptions(thickness=48, storage=storage))
    o64, _ = O.tilt_run(P, ang, O.TiltOptions(thickness=48), prec="f64")
    rms64 = np.sqrt(np.mean(o64.astype(np.float64) ** 2))
    rms_err = lambda a, b: float(np.sqrt(np.mean((a.astype(np.float64) - b) ** 2)) / rms64)
    print(f"{storage}: vs FP32 max {rel(g, g32):.3e} rms {rms_err(g, g32):.3e}; "
          f"vs FP64 max {rel(g, o64):.3e} rms {rms_err(g, o64):.3e}; "
          f"FP32 vs FP64 max {rel(g32, o64):.3e} rms {rms_err(g32, o64):.3e}")
    assert rel(g, g32) <= tol_max
    assert rms_err(g, g32) <= tol_rms
    assert rms_err(g, o64) <= rms_err(g32, o64) + tol_rms


def _ulps(a, b):
    """Distance in FP32 units in the last place (values of one sign or zero)."""
    a = np.asarray(a, np.float32).ravel()
    b = np.asarray(b, np.float32).ravel()
    return np.abs(a.view(np.int32).astype(np.int64) - b.view(np.int32).astype(np.int64))


def _geometry_case(gpu, geometry):
    rng = np.random.default_rng(21)
    nv, out_nz = 41, 37
    ang = np.sort(rng.uniform(-70.0, 70.0, nv)).astype(np.float32)
    ang[3] = -84.0  # beyond --max-cosine-tilt 80: stretch factor 0
    xt = rng.uniform(-3.0, 3.0, nv).astype(np.float32)
    zf = rng.uniform(0.9, 1.1, nv).astype(np.float32)
    s = gpu.Session(64, 4, nv, gpu.TiltOptions(thickness=out_nz, geometry=geometry))
    try:
        s.set_angles(ang)
        s.set_xtilt(xt)
        s.set_zfactor(zf)
        views, st, zc = s.geometry_tables()
    finally:
        s.close()
    return ang, xt, zf, out_nz, views, st, zc


def test_geometry_tables_f32_are_the_references():
    """The default (FP32) tables: cos/sin equal the oracle's FP32 ProjParams
    (main.rs:688-697) bit for bit, and the z terms are (dz * sin t) * zfactor in
    FP32 (main.rs:1015-1016), checked by recomputing those two IEEE products."""
    import tiltgpu as gpu
    ang, xt, zf, out_nz, views, st, zc = _geometry_case(gpu, "f32")
    tab = np.ascontiguousarray(O.proj_params(ang, xtilt=xt, zfactor=zf, prec="f32").T)
    assert np.array_equal(views.view(np.uint32), tab.view(np.uint32))
    dz = (np.arange(out_nz, dtype=np.float32) - np.float32(out_nz / 2.0))
    exp = (dz[None, :] * views[:, 1:2]) * views[:, 4:5]
    assert np.array_equal(zc.view(np.uint32), exp.astype(np.float32).view(np.uint32))
    assert st[3] == 0.0 and np.all(st[np.abs(ang) <= 80.0] > 1.0 - 1e-7)


def test_geometry_tables_f64_closed_form():
    """TILTGPU_GEOMETRY_F64: cos/sin of the tilt and x-axis tilt, the stretch
    factors 1/cos t and the z terms dz * sin t * zfactor computed in FP64 on the
    GPU and rounded once to FP32, against the same closed forms evaluated by
    numpy in float64 and rounded once.  Within 1 ulp: the device's double cos/sin
    and the host libm's may differ in the last double bit, which moves the FP32
    rounding only at a near-tie.  A table computed in FP32 (the default) sits
    several ulps away for some views, which this bound catches."""
    import tiltgpu as gpu
    ang, xt, zf, out_nz, views, st, zc = _geometry_case(gpu, "f64")
    d = np.deg2rad(ang.astype(np.float64))
    dx = np.deg2rad(xt.astype(np.float64))
    exp_views = np.stack([np.cos(d), np.sin(d), np.cos(dx), np.sin(dx), zf.astype(np.float64)], 1)
    assert _ulps(views, exp_views).max() <= 1
    exp_st = np.where(np.abs(ang) <= 80.0, 1.0 / np.cos(d), 0.0)
    assert st[3] == 0.0 and _ulps(st, exp_st).max() <= 1
    dz = np.arange(out_nz, dtype=np.float64) - out_nz / 2.0
    exp_zc = dz[None, :] * np.sin(d)[:, None] * zf.astype(np.float64)[:, None]
    assert _ulps(zc, exp_zc).max() <= 1
    _, _, _, _, v32, _, zc32 = _geometry_case(gpu, "f32")
    print(f"FP32 tables vs FP64 closed form: views {_ulps(v32, exp_views).max()} ulp, "
          f"z terms {_ulps(zc32, exp_zc).max()} ulp")
    assert _ulps(v32, exp_views).max() > 1


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_geometry_f64_reconstruction(gpu, rel, storage):
    """Reconstruction with FP64 geometry tables (x-axis tilt and z factors
    included).  Against the FP32-geometry path most voxels move only by the
    tables' rounding; the rest are voxels whose fill decision (a coordinate at
    the edge of the image) flips with a last-bit change of the tables, as it
    does between FP32 and FP64.  The error against the FP64 oracle is stated and
    is no larger than the FP32-geometry path's in RMS."""
    nv = 31
    xt = np.full(nv, 1.5, np.float32)
    P, ang = case(gpu, seed=23, ny=8, nviews=nv, xtilt=xt)
    zf = np.linspace(0.95, 1.05, nv).astype(np.float32)
    print(f"case: geometry f64, storage {storage}", flush=True)
    g32, _ = gpu.reconstruct(P, ang, gpu.TiltOptions(thickness=48, storage=storage),
                             xtilt=xt, zfactor=zf)
    g64, _ = gpu.reconstruct(P, ang, gpu.TiltOptions(thickness=48, storage=storage, geometry="f64"),
                             xtilt=xt, zfactor=zf)
    o64, _ = O.tilt_run(P, ang, O.TiltOptions(thickness=48), xtilt=xt, zfactor=zf, prec="f64")
    rms64 = np.sqrt(np.mean(o64.astype(np.float64) ** 2))
    rms_err = lambda a, b: float(np.sqrt(np.mean((a.astype(np.float64) - b) ** 2)) / rms64)
    print(f"{storage} geometry f64: vs FP64 oracle max {rel(g64, o64):.3e} rms {rms_err(g64, o64):.3e}; "
          f"geometry f32: max {rel(g32, o64):.3e} rms {rms_err(g32, o64):.3e}; "
          f"f64 vs f32 geometry rms {rms_err(g64, g32):.3e}")
    d = np.abs(g64.astype(np.float64) - g32) / rms64
    print(f"voxels moved by more than 1e-4 x RMS: {np.mean(d > 1e-4):.2%}")
    assert np.median(d) <= 1e-5
    assert rms_err(g64, o64) <= rms_err(g32, o64) * 1.05 + 1e-6


def test_chunked_upload_equals_whole_upload(gpu):
    """Sections uploaded in chunks, out of order and from page-locked memory
    (tiltgpu_upload_views) reconstruct bit for bit as one whole upload; a
    session missing a chunk refuses to run."""
    P, ang = case(gpu, seed=24, ny=10, nviews=31)
    opts = gpu.TiltOptions(thickness=48)
    whole, _ = gpu.reconstruct(P, ang, opts)
    print("case: chunked upload 31 views in chunks of 7", flush=True)
    s = gpu.Session(P.shape[2], P.shape[1], P.shape[0], opts)
    buf = gpu.PinnedArray((7,) + P.shape[1:])
    try:
        s.set_angles(ang)
        starts = list(range(0, 31, 7))
        for v0 in starts[1:]:
            n = min(7, 31 - v0)
            buf.array[:n] = P[v0:v0 + n]
            s.upload_views(v0, buf.array[:n])
        with pytest.raises(gpu.TiltGpuError):
            s.run()  # sections 0..6 missing
        buf.array[:7] = P[:7]
        s.upload_views(0, buf.array[:7])
        s.run()
        vol, _ = s.read_rows(0, P.shape[1])
    finally:
        buf.close()
        s.close()
    assert np.array_equal(vol.view(np.uint32), whole.view(np.uint32))


def _small_xf(rng, n):
    out = []
    for _ in range(n):
        a = np.deg2rad(rng.uniform(-15, 15))
        m = rng.uniform(0.95, 1.05)
        out.append([m * np.cos(a), -m * np.sin(a), m * np.sin(a), m * np.cos(a),
                    rng.uniform(-5, 5), rng.uniform(-5, 5)])
    return np.array(out, np.float32)


def test_section_transforms_bitwise(gpu, rel):
    """newstack -x on the device: the transformed stack equals the oracle's bit
    for bit (fewer transforms than sections: the last repeats), the
    reconstruction of it equals the oracle's within 1e-4 x RMS, and removing the
    transforms gives the uploaded stack back."""
    P, ang = case(gpu, seed=25, ny=12, nviews=9)
    xf = _small_xf(np.random.default_rng(25), 5)
    print("case: section transforms 9 sections, 5 transforms", flush=True)
    opts = gpu.TiltOptions(thickness=48)
    s = gpu.Session(P.shape[2], P.shape[1], P.shape[0], opts)
    try:
        s.set_angles(ang)
        s.set_section_transforms(xf, fill=0.5)
        s.upload(P)
        got = s.read_projections()
        exp = O.apply_xf_stack(P, xf, fill=0.5)
        assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
        s.run()
        vol, _ = s.read_rows()
        s.set_section_transforms(np.zeros((0, 6), np.float32))
        assert np.array_equal(s.read_projections(), P)
    finally:
        s.close()
    ref, _ = O.tilt_run(exp, ang, O.TiltOptions(thickness=48))
    assert rel(vol, ref) <= TOL


def test_studio_reconstruction_route(gpu, rel):
    """imod-studio's step_reconstruction (apps/imod-studio/src/main.rs:725-770)
    routed through the library: no weighting, no stretch, zero fill, thickness
    = width, against the restatement of the studio's own loop."""
    P, ang = case(gpu, nx=96, ny=6, nviews=21, thick=32, seed=26)
    print("case: studio route 96x6x21", flush=True)
    g, _ = gpu.reconstruct(P, ang[:19], gpu.TiltOptions(thickness=96, no_weight=True,
                                                         no_cosine_stretch=True,
                                                         fill_mode="zero"))
    ref = O.studio_reconstruct(P, ang[:19])
    assert g.shape == ref.shape == (6, 96, 96)
    assert rel(g, ref) <= TOL


def test_new_entry_points_reject_bad_arguments(gpu):
    """Chunked uploads outside the sections, a run before every section has
    arrived, negative transform counts and bad geometry modes are refused with
    an error code and leave the session usable."""
    P, ang = case(gpu, seed=27, ny=4, nviews=5)
    s = gpu.Session(P.shape[2], P.shape[1], P.shape[0], gpu.TiltOptions(thickness=16))
    try:
        s.set_angles(ang)
        with pytest.raises(gpu.TiltGpuError):
            s.upload_views(4, P[:2])        # sections 4..5 of 5
        with pytest.raises(gpu.TiltGpuError):
            s.upload_views(-1, P[:1])
        s.upload_views(0, P[:3])
        with pytest.raises(gpu.TiltGpuError):
            s.run()                         # sections 3, 4 missing
        with pytest.raises(gpu.TiltGpuError):
            s.read_projections()
        assert gpu.lib().tiltgpu_set_section_transforms(s.h, None, -1, 0.0) < 0
        s.upload_views(3, P[3:])
        s.run()
        vol, _ = s.read_rows()
        assert np.isfinite(vol).all()
    finally:
        s.close()
    with pytest.raises(gpu.TiltGpuError):
        gpu.Session(32, 4, 5, gpu.TiltOptions(thickness=16, geometry="f128"))


def test_c0_config(gpu, rel):
    """Config 0: 512x512 projections, 61 views -60..60 step 2, thickness 128,
    radial 0.35,0.05, phantom of 200 ellipsoids + 30 beads, noise 0.1."""
    ang = np.arange(-60, 61, 2, dtype=np.float32)
    P = gpu.synth_series(512, 512, ang, vol_nz=128, n_ellipsoids=200, n_beads=30, seed=0,
                         noise_rel=0.1)
    opts = gpu.TiltOptions(thickness=128, radial=(0.35, 0.05))
    print("case: WBP 512x512x61 -> z128", flush=True)
    g, sg = gpu.reconstruct(P, ang, opts)
    o, so = O.tilt_run(P, ang, O.TiltOptions(thickness=128), nthreads=16)
    err = rel(g, o)
    print(f"512x512x61 -> z128 max|d|/rms = {err:.3e}")
    assert err <= TOL


@pytest.mark.parametrize("tile_z", [32, 64, 128, 256])
@pytest.mark.parametrize("geom", ["plain", "xtilt", "local"])
def test_tile_depths(gpu, rel, monkeypatch, tile_z, geom):
    """Every backprojection tile depth gives the oracle's result (one-row,
    two-row and local-shift kernels)."""
    monkeypatch.setenv("TILTGPU_TILE_Z", str(tile_z))
    rng = np.random.default_rng(20 + tile_z)
    sx = rng.normal(0, 2, 31).astype(np.float32)
    sy = rng.normal(0, 2, 31).astype(np.float32)
    xt = np.full(31, 0.5, np.float32) if geom != "plain" else None
    P, ang = case(gpu, seed=21, ny=12, thick=150, xtilt=xt,
                  shifts=(sx, sy) if geom == "local" else None)
    local = None
    if geom == "local":
        local = O.build_local_grids([(v, 64.0, 6.0, float(sx[v]), float(sy[v])) for v in range(31)], 31)
    g, _, o, _ = both(gpu, P, ang, gpu.TiltOptions(thickness=150), xtilt=xt, local=local)
    assert rel(g, o) <= TOL


def test_forced_allocation_failures_fail_cleanly(gpu, monkeypatch):
    """TILTGPU_FAIL_ALLOC_AT=k refuses the k-th device allocation: every call
    that needs it returns TILTGPU_ERR_BUDGET with its message, a run that needs
    fewer allocations is unaffected, and the device is usable afterwards."""
    P, ang = case(gpu, nx=64, ny=4, nviews=9, thick=16, seed=5)
    opts = gpu.TiltOptions(thickness=16, sirt=2)
    ref, _ = gpu.reconstruct(P, ang, opts)
    failed = 0
    for k in range(1, 41):
        monkeypatch.setenv("TILTGPU_FAIL_ALLOC_AT", str(k))
        print(f"case: allocation {k} refused", flush=True)
        try:
            v, _ = gpu.reconstruct(P, ang, opts)
        except gpu.TiltGpuError as e:
            assert e.code == gpu.ERR_BUDGET and "TILTGPU_FAIL_ALLOC_AT" in str(e)
            failed += 1
            continue
        assert np.array_equal(v, ref)
    monkeypatch.delenv("TILTGPU_FAIL_ALLOC_AT")
    assert failed >= 5
    v, _ = gpu.reconstruct(P, ang, opts)
    assert np.array_equal(v, ref)



@pytest.mark.parametrize("nv,ny,nx,thick,kw", [
    (1, 3, 16, 8, {}),                          # one view
    (5, 1, 32, 4, {"xtilt": 1.0}),              # one row, two-row geometry at the last row
    (7, 2, 2, 3, {}),                           # the narrowest width the filter takes
    (4, 3, 48, 1, {}),                          # thickness 1
    (6, 3, 48, 16, {"width": 1}),               # one output column
    (6, 3, 48, 16, {"width": 80}),              # output wider than the projections
    (5, 1, 24, 6, {"sirt": 2}),                 # SIRT on a one-row series
])
def test_degenerate_shapes(gpu, rel, nv, ny, nx, thick, kw):
    kw = dict(kw)
    xt = None
    if "xtilt" in kw:
        xt = np.full(nv, kw.pop("xtilt"), np.float32)
    print(f"case: views {nv} rows {ny} width {nx} thickness {thick} {kw}", flush=True)
    rng = np.random.default_rng(nv * 100 + ny * 10 + nx)
    P = rng.normal(1.0, 0.3, (nv, ny, nx)).astype(np.float32)
    ang = np.linspace(-50.0, 50.0, nv).astype(np.float32) if nv > 1 else np.array([10.0], np.float32)
    g, sg, o, so = both(gpu, P, ang, gpu.TiltOptions(thickness=thick, **kw), xtilt=xt)
    assert g.shape == o.shape
    assert rel(g, o) <= TOL


def test_section_transforms_odd_sizes_bitwise(gpu):
    """newstack -x on sections whose size is not a multiple of the kernel's
    64 x 32 pixel blocks (333 x 77, and 65 x 33), bit for bit the oracle's."""
    rng = np.random.default_rng(26)
    for nx, ny, nz in ((333, 77, 4), (65, 33, 3)):
        P = rng.standard_normal((nz, ny, nx)).astype(np.float32)
        xf = _small_xf(rng, nz)
        print(f"case: section transforms {nx}x{ny}x{nz}", flush=True)
        with gpu.Session(nx, ny, nz, gpu.TiltOptions(thickness=16)) as s:
            s.set_angles(np.linspace(-30, 30, nz).astype(np.float32))
            s.set_section_transforms(xf, fill=-1.25)
            s.upload(P)
            got = s.read_projections()
        exp = O.apply_xf_stack(P, xf, fill=-1.25)
        assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
