"""ctfphaseflip through libtiltgpu's C ABI (tiltgpu_ctf_*), checked against the
oracle's restatement of apps/ctfphaseflip/src/main.rs:353-705.

CPU tests: the strip layout (bisection per row) against the reference's full
scan restated in numpy, the strip-width rule and the wavelength against the
oracle.  GPU tests: the corrected stack against the oracle (FP32 and FP64) at
max |gpu - oracle| <= 1e-4 * RMS(oracle).
"""
import numpy as np
import pytest

import oracle as O

TOL = 1e-4


def scan_layout(nx, ny, axis_deg, width):
    """main.rs:389-487: corner range, strip count, and each strip's bounding box
    by testing every pixel (numpy float64, the reference's expression)."""
    f32pi = float(np.float32(np.pi))
    ax = float(np.float32(axis_deg)) * f32pi / 180.0
    px, py = np.cos(ax), np.sin(ax)
    cx, cy = nx / 2.0, ny / 2.0
    corners = [(0.0, 0.0), (float(nx), 0.0), (0.0, float(ny)), (float(nx), float(ny))]
    ds = [(x - cx) * px + (y - cy) * py for x, y in corners]
    mn, mx = min(ds), max(ds)
    n = int(np.ceil((mx - mn) / width))
    X, Y = np.meshgrid(np.arange(nx, dtype=np.float64), np.arange(ny, dtype=np.float64))
    D = (X - cx) * px + (Y - cy) * py
    boxes, cens = [], []
    for si in range(n):
        scp = mn + (si + 0.5) * float(width)
        lo, hi = scp - width / 2.0, scp + width / 2.0
        m = (D >= lo) & (D < hi)
        if m.any():
            ys, xs = np.nonzero(m)
            boxes.append((xs.min(), xs.max() + 1, ys.min(), ys.max() + 1))
        else:
            boxes.append(None)
        cens.append(scp)
    return boxes, cens


@pytest.mark.parametrize("nx,ny,axis,width", [
    (64, 48, 0.0, 16), (64, 48, 30.0, 16), (67, 53, -85.0, 16), (40, 90, 90.0, 20),
    (100, 100, 45.0, 32), (33, 17, -135.0, 7), (128, 96, 5.5, 256), (50, 60, 180.0, 16)])
def test_strip_layout_equals_full_scan(nx, ny, axis, width):
    import tiltgpu as tg
    boxes, cen = tg.ctf_strip_layout(nx, ny, axis, width)
    rb, rc = scan_layout(nx, ny, axis, width)
    assert len(boxes) == len(rb)
    assert np.array_equal(cen, np.array(rc))
    for b, r in zip(boxes, rb):
        if r is None:
            assert b[1] <= b[0] or b[3] <= b[2]
        else:
            assert tuple(int(v) for v in b) == tuple(int(v) for v in r)


def test_wavelength_equals_oracle():
    import tiltgpu as tg
    for kv in (60.0, 80.0, 120.0, 200.0, 300.0, 1000.0):
        assert tg.electron_wavelength(kv) == O.electron_wavelength(kv)


@pytest.mark.parametrize("kw", [
    dict(pixel_a=2.0), dict(pixel_a=10.0), dict(pixel_a=10.0, defocus_nm=-500.0),
    dict(pixel_a=5.0, max_strip_width=40), dict(pixel_a=5.0, strip_width=8),
    dict(pixel_a=1.0, defocus_nm=0.0), dict(pixel_a=8.0, vdef=[(-2000, -2500, 10), (-4000, -3000, 0)]),
    dict(pixel_a=8.0, vdef=[]), dict(pixel_a=3.0, voltage_kv=120.0)])
def test_strip_width_rule_equals_oracle(kw):
    import tiltgpu as tg
    kw = dict(kw)
    vdef = kw.pop("vdef", None)
    p = tg.ctf_params(**kw)
    got = tg.ctf_strip_width(p, 1024, vdef)
    want = O.ctf_strip_width(p.strip_width, 1024, p.pixel_a, O.electron_wavelength(p.voltage_kv),
                             p.defocus_nm, vdef=vdef,
                             max_strip_width=p.max_strip_width or None)
    assert got == want


def test_ctf_rejects_bad_arguments():
    import tiltgpu as tg
    with pytest.raises(tg.TiltGpuError):
        tg.ctf_strip_width(tg.ctf_params(pixel_a=0.0), 64)
    with pytest.raises(tg.TiltGpuError):
        tg.ctf_strip_width(tg.ctf_params(strip_width=0), 64)
    with pytest.raises(tg.TiltGpuError):
        tg.ctf_strip_layout(0, 10, 0.0, 16)


def _stack(nz, ny, nx, seed):
    rng = np.random.default_rng(seed)
    X, Y = np.meshgrid(np.arange(nx), np.arange(ny))
    base = np.sin(X / 3.1) * np.cos(Y / 4.3) + 0.3 * np.sin((X + Y) / 1.7)
    return (base[None] + rng.normal(0, 0.5, (nz, ny, nx))).astype(np.float32)


CASES = [
    # (nz, ny, nx, options)
    (4, 48, 64, dict(pixel_a=3.0, defocus_nm=-2500.0)),
    (3, 53, 67, dict(pixel_a=4.0, tilt_axis_deg=-85.0, defocus_nm=-4000.0)),
    (3, 40, 60, dict(pixel_a=2.5, tilt_axis_deg=30.0, plate_deg=90.0, amp=0.1, cuton=0.01,
                     max_strip_width=24, strip_width=24)),
    (4, 64, 64, dict(pixel_a=3.0, max_strip_width=32, strip_width=32,
                     vdef=[(-2000.0, -2600.0, 25.0), (-2100.0, -2700.0, 30.0),
                           (-2200.0, -2200.0, 0.0), (-2300.0, -2900.0, -40.0)])),
    (2, 96, 128, dict(pixel_a=6.0, voltage_kv=200.0, cs_mm=2.0, tilt_axis_deg=5.5)),
    (2, 256, 256, dict(pixel_a=2.0, max_strip_width=64, strip_width=64, tilt_axis_deg=-12.0)),
    (2, 5000, 20, dict(pixel_a=3.0)),  # columns of 5000: 80 KB of LDS per workgroup
    # power-of-two rows / columns (register radix-8 kernels), odd row counts
    (2, 256, 192, dict(pixel_a=3.0, max_strip_width=64, strip_width=64)),
    (3, 45, 128, dict(pixel_a=3.0, max_strip_width=64, strip_width=64)),
    (3, 128, 128, dict(pixel_a=2.0, tilt_axis_deg=90.0, max_strip_width=64, strip_width=64,
                       vdef=[(-2000.0, -2500.0, 30.0)] * 3)),
    (1, 2048, 32, dict(pixel_a=4.0)),
    (1, 40, 2048, dict(pixel_a=4.0, tilt_axis_deg=90.0)),
    (2, 96, 1024, dict(pixel_a=4.0, tilt_axis_deg=85.0)),
    # rotated axis / astigmatism with power-of-two rows: the Nyquist column of
    # the flipped spectrum is not Hermitian, the inverse keeps the real part
    (3, 128, 128, dict(pixel_a=2.0, tilt_axis_deg=89.0, max_strip_width=64, strip_width=64)),
    (3, 128, 128, dict(pixel_a=2.0, max_strip_width=64, strip_width=64,
                       vdef=[(-2000.0, -2500.0, 30.0)] * 3)),
    # Bluestein (a prime factor above 13): 68 = 4*17 and 97 over 256 points,
    # 1843 = 19*97 over 4096 points (512 threads); 4096 directly
    (2, 97, 68, dict(pixel_a=3.0, max_strip_width=68, strip_width=68, tilt_axis_deg=0.0)),
    (1, 34, 1843, dict(pixel_a=4.0, tilt_axis_deg=90.0, plate_deg=30.0)),
    (1, 4096, 24, dict(pixel_a=4.0, max_strip_width=24, strip_width=24)),
]


@pytest.mark.gpu
@pytest.mark.parametrize("ci", range(len(CASES)))
def test_ctf_phaseflip_matches_oracle(gpu, rel, ci):
    nz, ny, nx, kw = CASES[ci]
    kw = dict(kw)
    vdef = kw.pop("vdef", None)
    P = _stack(nz, ny, nx, ci)
    tilt = np.linspace(-50.0, 50.0, nz).astype(np.float32)
    p = gpu.ctf_params(**kw)
    width = gpu.ctf_strip_width(p, nx, vdef)
    print(f"case {ci}: {nz}x{ny}x{nx} {kw} width={width}", flush=True)
    g = gpu.ctf_phaseflip(P, tilt, vdef=vdef, **kw)
    okw = {k: v for k, v in kw.items() if k not in ("max_strip_width", "strip_width")}
    for prec in ("f32", "f64"):
        o = O.ctf_phaseflip(P, tilt, vdef=vdef, strip_width=width, prec=prec, **okw)
        err = rel(g, o)
        print(f"  {prec}: max|gpu-oracle|/RMS = {err:.3e}", flush=True)
        assert err <= TOL
    # pixels outside every strip's weight keep their input value exactly
    o32 = O.ctf_phaseflip(P, tilt, vdef=vdef, strip_width=width, prec="f32", **okw)
    same = o32 == P
    assert np.array_equal(g[same], P[same])
    assert not np.allclose(g, P)


@pytest.mark.gpu
def test_ctf_bluestein_for_every_length(gpu, rel, monkeypatch):
    # the transform choice changes rounding only
    monkeypatch.setenv("TILTGPU_CTF_BLUESTEIN", "all")
    nz, ny, nx = 2, 45, 195
    P = _stack(nz, ny, nx, 5)
    tilt = np.array([-30.0, 40.0], np.float32)
    kw = dict(pixel_a=3.0, tilt_axis_deg=-20.0, amp=0.1)
    g = gpu.ctf_phaseflip(P, tilt, **kw)
    width = gpu.ctf_strip_width(gpu.ctf_params(**kw), nx)
    o = O.ctf_phaseflip(P, tilt, strip_width=width, prec="f64", **kw)
    assert rel(g, o) <= TOL


@pytest.mark.gpu
def test_ctf_corrector_reuse_and_sections(gpu, rel):
    nz, ny, nx = 5, 48, 80
    P = _stack(nz, ny, nx, 7)
    tilt = np.linspace(-60.0, 60.0, nz).astype(np.float32)
    p = gpu.ctf_params(pixel_a=3.0)
    c = gpu.CtfCorrector(nx, ny, nz, p, tilt)
    try:
        c.upload(P)
        c.run()
        a = c.download()
        c.run()  # a second run over the same input gives the same output
        b = c.download()
        assert np.array_equal(a, b)
        c.upload(P[2:3] * 2.0, z0=2)  # linear in the input section by section
        c.run()
        s = c.download(2, 1)
        assert rel(s[0], 2.0 * a[2]) <= 1e-5
        ms, launches = c.last_run()
        assert ms > 0 and launches > 0
    finally:
        c.close()
    one = gpu.ctf_phaseflip(P, tilt, pixel_a=3.0)
    assert np.array_equal(one, a)


@pytest.mark.gpu
def test_ctf_gpu_rejects_bad_arguments(gpu):
    p = gpu.ctf_params(pixel_a=3.0)
    with pytest.raises(gpu.TiltGpuError):
        gpu.CtfCorrector(0, 10, 1, p)
    c = gpu.CtfCorrector(16, 16, 2, p)
    try:
        with pytest.raises(gpu.TiltGpuError):
            c.upload(np.zeros((2, 16, 16), np.float32), z0=1)
        with pytest.raises(gpu.TiltGpuError) as e:
            c.run()  # nothing uploaded yet
        assert e.value.code == gpu.ERR_STATE
        c.upload(np.ones((1, 16, 16), np.float32), z0=0)
        with pytest.raises(gpu.TiltGpuError) as e:
            c.run()  # section 1 still missing
        assert e.value.code == gpu.ERR_STATE
        c.upload(np.ones((1, 16, 16), np.float32), z0=1)
        with pytest.raises(gpu.TiltGpuError) as e:
            c.download()  # no run since the upload
        assert e.value.code == gpu.ERR_STATE
        c.run()
        with pytest.raises(gpu.TiltGpuError):
            c.download(1, 2)
        assert c.download().shape == (2, 16, 16)
    finally:
        c.close()


DEFOCUS_FILE = """# view defocus1 [defocus2 [astig]]

0 -2500
+3 -2600.5 -2900 12.5
  7\t-2800 -3100
9 -3000 -3000 0 extra
"""


def test_defocus_file_and_interpolation_equal_oracle(tmp_path):
    import tiltgpu as tg
    p = tmp_path / "def.txt"
    p.write_text(DEFOCUS_FILE)
    ent, views = tg.read_defocus_file(p)
    oent, oviews = O.read_defocus_file(str(p))
    assert list(views) == oviews == [0, 3, 7, 9]
    assert np.array_equal(ent, np.array(oent, np.float32))
    for nv in (1, 5, 8, 12):
        got = tg.interpolate_defocus(ent, views, nv)
        want = O.interpolate_defocus(oent, oviews, nv)
        assert got.shape == want.shape and np.array_equal(got, want)
    assert tg.interpolate_defocus(np.zeros((0, 3)), np.zeros(0), 5).shape == (0, 3)
    # views listed twice: the later line wins; views beyond the stack are ignored
    e2 = np.array([[-1, -1, 0], [-2, -2, 0], [-9, -9, 0]], np.float32)
    v2 = np.array([2, 2, 40], np.int32)
    assert np.array_equal(tg.interpolate_defocus(e2, v2, 4),
                          O.interpolate_defocus([tuple(r) for r in e2], list(v2), 4))


@pytest.mark.parametrize("bad", ["0\n", "x -2500\n", "1 abc\n", "1 -2500 q\n", "1 -2 -3 z\n",
                                 "-1 -2500\n"])
def test_defocus_file_malformed_lines(tmp_path, bad):
    import tiltgpu as tg
    p = tmp_path / "bad.txt"
    p.write_text("# ok\n0 -2000\n" + bad)
    with pytest.raises(tg.TiltGpuError, match="line 3"):
        tg.read_defocus_file(p)
    with pytest.raises(ValueError):
        O.read_defocus_file(str(p))


def _cli():
    import os
    return os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                        "imod-rs_amd", "_build", "ctfphaseflip")


def test_ctf_cli_usage_errors(tmp_path):
    import subprocess
    r = subprocess.run([_cli()], capture_output=True, text=True)
    assert r.returncode == 2 and "required" in r.stderr
    r = subprocess.run([_cli(), "-i", "a", "-o", "b", "-t", "c", "--bogus"], capture_output=True,
                       text=True)
    assert r.returncode == 2
    r = subprocess.run([_cli(), "-i", "a", "-o", "b", "-t", "c", "-w", "x"], capture_output=True,
                       text=True)
    assert r.returncode == 2
    r = subprocess.run([_cli(), "-i", str(tmp_path / "none.mrc"), "-o", str(tmp_path / "o.mrc"),
                        "-t", str(tmp_path / "none.tlt")], capture_output=True, text=True)
    assert r.returncode == 1 and "tilt file" in r.stderr


@pytest.mark.gpu
def test_ctf_cli_end_to_end_matches_oracle(gpu, tmp_path, rel):
    import subprocess
    nz, ny, nx = 6, 72, 96
    P = _stack(nz, ny, nx, 11)
    tilt = np.linspace(-45.0, 45.0, nz).astype(np.float32)
    src, out = str(tmp_path / "in.mrc"), str(tmp_path / "out.mrc")
    gpu.mrc_write(src, P, xlen=nx * 3.5, ylen=ny * 3.5, zlen=float(nz))
    (tmp_path / "a.tlt").write_text("".join(f"{t:.2f}\n" for t in tilt))
    (tmp_path / "d.txt").write_text("0 -2500 -2800 20\n5 -3000 -3300 35\n")
    cmd = [_cli(), "-i", src, "-o", out, "-t", str(tmp_path / "a.tlt"), "--defocus-file",
           str(tmp_path / "d.txt"), "--tilt-axis-angle", "10", "--plate-phase", "45", "--gpu",
           "--upload-sections", "4"]  # two chunks of 4 and 2 sections each way
    print("cmd:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    print(r.stderr, flush=True)
    assert r.returncode == 0
    assert "interpolating missing values" in r.stderr and "strips=16px" in r.stderr
    g = gpu.mrc_read(out)
    tl = np.array(O.read_tilt_file(str(tmp_path / "a.tlt")), np.float32)
    ent, views = O.read_defocus_file(str(tmp_path / "d.txt"))
    vdef = O.interpolate_defocus(ent, views, nz)
    lam = O.electron_wavelength(300.0)
    width = O.ctf_strip_width(256, nx, 3.5, lam, -3000.0, vdef=vdef)
    o = O.ctf_phaseflip(P, tl, 3.5, vdef=vdef, strip_width=width, tilt_axis_deg=10.0,
                        plate_deg=45.0)
    assert rel(g, o) <= TOL
    info = gpu.mrc_info(out)
    assert info.label0.decode() == "ctfphaseflip: CTF phase correction (2D strips)"
    assert info.amin == g.min() and info.amax == g.max()
    assert info.xlen == pytest.approx(nx * 3.5)


@pytest.mark.gpu
def test_ctf_forced_allocation_failures(gpu, monkeypatch):
    P = _stack(3, 40, 60, 3)
    tilt = np.array([-20.0, 0.0, 20.0], np.float32)
    ref = gpu.ctf_phaseflip(P, tilt, pixel_a=3.0, tilt_axis_deg=20.0)
    failed = 0
    for k in range(1, 30):
        monkeypatch.setenv("TILTGPU_FAIL_ALLOC_AT", str(k))
        try:
            g = gpu.ctf_phaseflip(P, tilt, pixel_a=3.0, tilt_axis_deg=20.0)
        except gpu.TiltGpuError as e:
            assert e.code == gpu.ERR_BUDGET
            failed += 1
            continue
        assert np.array_equal(g, ref)
    monkeypatch.delenv("TILTGPU_FAIL_ALLOC_AT")
    assert failed >= 5
    assert np.array_equal(gpu.ctf_phaseflip(P, tilt, pixel_a=3.0, tilt_axis_deg=20.0), ref)



@pytest.mark.gpu
@pytest.mark.parametrize("axis", [0.0, 85.0])
def test_ctf_full_size_properties(gpu, axis):
    """At the benchmark's image size (2048 x 2048, 4 A pixels, strip width 16):
    scaling the input by 2 scales the output by 2 bit for bit (every step is
    linear and a power of two is exact), and the correction is linear to
    rounding, ctf(P + Q) = ctf(P) + ctf(Q)."""
    nz, ny, nx = 4, 2048, 2048
    rng = np.random.default_rng(int(axis))
    P = rng.normal(size=(nz, ny, nx)).astype(np.float32)
    Q = rng.normal(size=(nz, ny, nx)).astype(np.float32)
    tilt = np.array([-60.0, -20.0, 20.0, 60.0], np.float32)
    p = gpu.ctf_params(pixel_a=4.0, tilt_axis_deg=axis)
    c = gpu.CtfCorrector(nx, ny, nz, p, tilt)
    try:
        assert c.strip_width == 16
        out = []
        for X in (P, 2.0 * P, Q, P + Q):
            c.upload(X)
            c.run()
            out.append(c.download())
    finally:
        c.close()
    a, a2, b, ab = out
    assert np.array_equal(a2, 2.0 * a)
    rms = np.sqrt(np.mean(ab.astype(np.float64) ** 2))
    assert np.max(np.abs(ab.astype(np.float64) - a - b)) <= 1e-4 * rms


@pytest.mark.gpu
@pytest.mark.parametrize("nz,ny,nx,kw", [
    (1, 1, 40, {}),                        # one row: columns of length 1
    (1, 40, 1, {}),                        # one column: every pixel on the zero-weight edge
    (1, 1, 1, {}),
    (2, 16, 16, {}),                       # one strip covering the image
    (2, 7, 9, {"tilt_axis_deg": 30.0}),
    (3, 5, 3, {"tilt_axis_deg": 90.0}),
])
def test_ctf_degenerate_shapes(gpu, rel, nz, ny, nx, kw):
    print(f"case: {nz}x{ny}x{nx} {kw}", flush=True)
    P = np.random.default_rng(nz * 100 + ny + nx).normal(size=(nz, ny, nx)).astype(np.float32)
    tilt = np.linspace(-30.0, 30.0, nz).astype(np.float32)
    g = gpu.ctf_phaseflip(P, tilt, pixel_a=3.0, **kw)
    width = gpu.ctf_strip_width(gpu.ctf_params(pixel_a=3.0, **kw), nx)
    o = O.ctf_phaseflip(P, tilt, 3.0, strip_width=width, **kw)
    assert rel(g, o) <= TOL
    same = o == P
    assert np.array_equal(g[same], P[same])
