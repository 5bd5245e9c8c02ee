"""MRC I/O, local-alignment file parsing and the `tilt` command line.

The MRC cases restate crates/imod-mrc/src/tests.rs (roundtrip_header :13-36,
roundtrip_float_file :68-89, read_big_endian_float_file :113-149,
read_big_endian_short_file :151-181, roundtrip_byte_file :39-66) with files
written here; the big-endian files are built byte by byte as those tests do."""
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle as O
import tiltgpu as tg

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TILT = os.path.join(REPO, "imod-rs_amd", "_build", "tilt")


def _header_bytes(nx, ny, nz, mode, big=False, label=b""):
    """A 1024-byte MRC2014 header in the layout of crates/imod-mrc/src/header.rs."""
    e = ">" if big else "<"
    h = bytearray(1024)
    struct.pack_into(e + "10i", h, 0, nx, ny, nz, mode, 0, 0, 0, nx, ny, nz)
    struct.pack_into(e + "6f", h, 40, nx, ny, nz, 90, 90, 90)
    struct.pack_into(e + "3i", h, 64, 1, 2, 3)
    struct.pack_into(e + "i", h, 108, 20140)
    h[208:212] = b"MAP "
    h[212:216] = b"\x11\x11\x00\x00" if big else b"\x44\x44\x00\x00"
    if label:
        struct.pack_into(e + "i", h, 220, 1)
        h[224:224 + len(label)] = label
    return bytes(h)


def test_roundtrip_header_and_float_file(tmp_path):
    p = str(tmp_path / "a.mrc")
    vol = np.arange(2 * 3 * 4, dtype=np.float32).reshape(2, 3, 4) * 0.5 - 1.5
    tg.mrc_write(p, vol, label="Test label written by imod-mrc")
    assert os.path.getsize(p) == 1024 + vol.nbytes  # header is 1024 bytes
    info = tg.mrc_info(p)
    assert (info.nx, info.ny, info.nz, info.mode) == (4, 3, 2, 2)
    assert info.is_imod == 1 and info.swapped == 0
    assert info.label0.decode() == "Test label written by imod-mrc"
    assert info.amin == -1.5 and info.amax == vol.max()
    assert info.amean == pytest.approx(float(vol.mean()), abs=1e-6)
    assert np.array_equal(tg.mrc_read(p), vol)


def test_read_big_endian_float_file(tmp_path):
    p = tmp_path / "be.mrc"
    data = np.arange(6 * 5 * 2, dtype=np.float32).reshape(2, 5, 6) / 7.0
    p.write_bytes(_header_bytes(6, 5, 2, 2, big=True) + data.astype(">f4").tobytes())
    info = tg.mrc_info(str(p))
    assert info.swapped == 1 and (info.nx, info.ny, info.nz) == (6, 5, 2)
    assert np.max(np.abs(tg.mrc_read(str(p)) - data)) < 1e-6


def test_read_big_endian_short_and_byte_files(tmp_path):
    s = (np.arange(20, dtype=np.int16) - 7).reshape(1, 4, 5)
    p = tmp_path / "s.mrc"
    p.write_bytes(_header_bytes(5, 4, 1, 1, big=True) + s.astype(">i2").tobytes())
    assert np.array_equal(tg.mrc_read(str(p)), s.astype(np.float32))
    b = np.arange(12, dtype=np.uint8).reshape(2, 2, 3) * 20
    q = tmp_path / "b.mrc"
    q.write_bytes(_header_bytes(3, 2, 2, 0) + b.tobytes())
    assert np.array_equal(tg.mrc_read(str(q)), b.astype(np.float32))
    h = np.array([0.5, -2.0, 65504.0, 6e-8], np.float16).reshape(1, 1, 4)
    r = tmp_path / "h.mrc"
    r.write_bytes(_header_bytes(4, 1, 1, 12) + h.tobytes())
    assert np.array_equal(tg.mrc_read(str(r)), h.astype(np.float32))


def test_mrc_errors(tmp_path):
    p = tmp_path / "short.mrc"
    p.write_bytes(b"\0" * 100)
    with pytest.raises(tg.TiltGpuError):
        tg.mrc_info(str(p))
    with pytest.raises(tg.TiltGpuError):
        tg.mrc_info(str(tmp_path / "missing.mrc"))


def test_local_file_parse_matches_oracle(tmp_path):
    p = tmp_path / "local.txt"
    p.write_text("# view x y dx dy\n"
                 "0 10 5 1.5 -0.5\n0 30 5 2.0 0.25\n0 10 25 -1 1\n0 30 25 0 0\n"
                 "1 20.004 5 3 3\n1 20 5 4 4\n"  # merges within 0.01; later patch wins
                 "2 1 1\n"                       # short line: skipped
                 "x 7 7 0.5 0.5\n"               # bad view -> 0
                 "9 1 1 1 1\n")                  # view >= nz: dropped
    got = tg.parse_local_file(str(p), 4)
    ref = O.build_local_grids(O.parse_local_file(str(p)), 4).flat()
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    for a, b in zip(got[2:], ref[2:]):
        assert np.array_equal(a, b[: len(a)])


def test_cli_usage_errors():
    r = subprocess.run([TILT], capture_output=True, text=True)
    assert r.returncode == 2 and "required arguments" in r.stderr
    r = subprocess.run([TILT, "-i", "a", "-o", "b", "-t", "c", "--bogus"], capture_output=True,
                       text=True)
    assert r.returncode == 2 and "unexpected argument" in r.stderr


def test_cli_missing_files_exit_1(tmp_path):
    t = tmp_path / "a.tlt"
    r = subprocess.run([TILT, "-i", "a.mrc", "-o", "b.mrc", "-t", str(t)], capture_output=True,
                       text=True)
    assert r.returncode == 1 and "tilt file" in r.stderr
    t.write_text("0\n")
    r = subprocess.run([TILT, "-i", str(tmp_path / "nope.mrc"), "-o", "b.mrc", "-t", str(t)],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "Error" in r.stderr


@pytest.mark.gpu
def test_cli_end_to_end_matches_oracle(gpu, tmp_path, rel):
    """`tilt` with --xtilt-file, --local-file and a few options against the oracle
    on the same MRC input; header statistics included."""
    ang = np.linspace(-60, 60, 25).astype(np.float32)
    rng = np.random.default_rng(3)
    sx, sy = rng.normal(0, 1.5, 25).astype(np.float32), rng.normal(0, 1.5, 25).astype(np.float32)
    P = gpu.synth_series(96, 14, ang, vol_nz=40, n_ellipsoids=25, n_beads=4, seed=3,
                         xtilt=np.full(25, 0.5, np.float32), shift_x=sx, shift_y=sy,
                         axis=(2.0, 9.0), bead_radius=2.0)
    inp, out = str(tmp_path / "in.mrc"), str(tmp_path / "out.rec")
    gpu.mrc_write(inp, P)
    (tmp_path / "a.tlt").write_text("".join(f"{a:10.2f}\n" for a in ang))
    (tmp_path / "xt").write_text("0.5\n" * 25)
    (tmp_path / "loc").write_text("".join(f"{v} 48 7 {sx[v]} {sy[v]}\n" for v in range(25)))
    cmd = [TILT, "-i", inp, "-o", out, "-t", str(tmp_path / "a.tlt"), "-z", "40",
           "--xtilt-file", str(tmp_path / "xt"), "--local-file", str(tmp_path / "loc"),
           "--densweight", "2", "--outadd=1", "--slab-rows", "5", "--upload-views", "4"]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    vol = gpu.mrc_read(out)
    info = gpu.mrc_info(out)
    angs = O.read_tilt_file(str(tmp_path / "a.tlt"))
    grids = O.build_local_grids(O.parse_local_file(str(tmp_path / "loc")), 25)
    ref, (mn, mx, mean) = O.tilt_run(P, angs, O.TiltOptions(thickness=40, densweight=2.0,
                                                            outadd=1.0),
                                     xtilt=O.load_per_view_file(str(tmp_path / "xt")),
                                     local=grids)
    # file sections are Y rows, each [z][x] (main.rs:710-716)
    assert (info.nx, info.ny, info.nz) == (96, 40, 14)
    assert (info.mx, info.my, info.mz) == (96, 40, 14)
    assert info.label0.decode() == "tilt: 25 projections, thickness 40"
    assert rel(vol, ref) <= 1e-4
    scale = float(np.sqrt(np.mean(ref.astype(np.float64) ** 2)))
    assert abs(info.amin - mn) <= 1e-4 * scale and abs(info.amax - mx) <= 1e-4 * scale
    assert abs(info.amean - mean) <= 1e-4 * scale


@pytest.mark.gpu
@pytest.mark.parametrize("args,oopts,zf", [
    (["--sirt", "2", "--sirt-relax", "0.7", "--fill-mode", "edge"],
     dict(sirt=2, sirt_relax=0.7, fill_mode="edge"), True),
    (["--log", "--hamming", "--radial", "0.3,0.1", "--exact-filter-size", "50",
      "--cosine-interp", "cubic", "--max-cosine-tilt", "45", "-w", "80"],
     dict(log=True, hamming=True, radial=(0.3, 0.1), exact_filter_size=50,
          cosine_interp="cubic", max_cosine_tilt=45.0, width=80), False),
    (["--no-weight", "--no-cosine-stretch", "--fill-mode", "zero", "--fake-sirt", "1",
      "--outscale", "0.5"],
     dict(no_weight=True, no_cosine_stretch=True, fill_mode="zero", fake_sirt=1,
          outscale=0.5), False),
])
def test_cli_options_reach_the_library(gpu, tmp_path, rel, args, oopts, zf):
    """Every option of the `tilt` surface (main.rs:18-120) is parsed and passed
    on: each option set gives the oracle's reconstruction of the same file."""
    ang = np.linspace(-55, 55, 17).astype(np.float32)
    P = gpu.synth_series(64, 6, ang, vol_nz=24, n_ellipsoids=12, n_beads=3, seed=9,
                         axis=(2.0, 7.0), bead_radius=2.0)
    P = P - P.min() + 0.5  # positive, for --log
    inp, out = str(tmp_path / "in.mrc"), str(tmp_path / "out.rec")
    gpu.mrc_write(inp, P)
    (tmp_path / "a.tlt").write_text("".join(f"{a:.6f}\n" for a in ang))
    cmd = [TILT, "-i", inp, "-o", out, "-t", str(tmp_path / "a.tlt"), "-z", "24", *args]
    zfac = None
    if zf:
        zfac = np.linspace(0.9, 1.1, 17).astype(np.float32)
        (tmp_path / "zf").write_text("".join(f"{v:.6f}\n" for v in zfac))
        cmd += ["--zfactor-file", str(tmp_path / "zf")]
        zfac = O.load_per_view_file(str(tmp_path / "zf"))
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    vol = gpu.mrc_read(out)
    ref, _ = O.tilt_run(P, O.read_tilt_file(str(tmp_path / "a.tlt")),
                        O.TiltOptions(thickness=24, **oopts), zfactor=zfac)
    assert vol.shape == ref.shape
    assert rel(vol, ref) <= 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("storage,geometry", [("f16", "f64"), ("bf16", "f32")])
def test_cli_storage_and_geometry_options(gpu, tmp_path, storage, geometry):
    """The two additions, --storage and --geometry, reach the library: the file
    written by `tilt` equals the in-memory reconstruction with the same options
    bit for bit (chunked upload and slab read-back change nothing)."""
    ang = np.linspace(-50, 50, 13).astype(np.float32)
    P = gpu.synth_series(64, 5, ang, vol_nz=20, n_ellipsoids=10, n_beads=2, seed=4,
                         axis=(2.0, 7.0), bead_radius=2.0)
    inp, out = str(tmp_path / "in.mrc"), str(tmp_path / "out.rec")
    gpu.mrc_write(inp, P)
    (tmp_path / "a.tlt").write_text("".join(f"{a:.6f}\n" for a in ang))
    cmd = [TILT, "-i", inp, "-o", out, "-t", str(tmp_path / "a.tlt"), "-z", "20",
           "--storage", storage, "--geometry", geometry, "--upload-views", "5", "--slab-rows", "2"]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    angs = np.array(O.read_tilt_file(str(tmp_path / "a.tlt")), np.float32)
    ref, _ = gpu.reconstruct(P, angs, gpu.TiltOptions(thickness=20, storage=storage,
                                                      geometry=geometry))
    assert np.array_equal(gpu.mrc_read(out), ref)


@pytest.mark.gpu
def test_cli_section_transforms(gpu, tmp_path, rel):
    """`tilt --xf-file` applies the newstack -x transforms on the GPU before the
    reconstruction: the result equals the oracle's reconstruction of the
    oracle-transformed stack."""
    ang = np.linspace(-50, 50, 11).astype(np.float32)
    P = gpu.synth_series(80, 10, ang, vol_nz=30, n_ellipsoids=15, n_beads=3, seed=4,
                         axis=(2.0, 8.0), bead_radius=2.0)
    inp, out = str(tmp_path / "in.mrc"), str(tmp_path / "out.rec")
    gpu.mrc_write(inp, P)
    (tmp_path / "a.tlt").write_text("".join(f"{a:10.2f}\n" for a in ang))
    xf = [[0.9961947, -0.0871557, 0.0871557, 0.9961947, 2.25, -1.5],
          [1.0, 0.0, 0.0, 1.0, -3.0, 0.75]]
    (tmp_path / "a.xf").write_text("".join("%12.7f%12.7f%12.7f%12.7f%12.3f%12.3f\n" % tuple(x)
                                           for x in xf))
    cmd = [TILT, "-i", inp, "-o", out, "-t", str(tmp_path / "a.tlt"), "-z", "30",
           "--xf-file", str(tmp_path / "a.xf"), "--xf-fill", "0.5"]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    vol = gpu.mrc_read(out)
    aligned = O.apply_xf_stack(P, O.read_xf_file(str(tmp_path / "a.xf")), fill=0.5)
    ref, _ = O.tilt_run(aligned, O.read_tilt_file(str(tmp_path / "a.tlt")),
                        O.TiltOptions(thickness=30))
    assert rel(vol, ref) <= 1e-4
