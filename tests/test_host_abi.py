"""CPU tests of libtiltgpu's host side: the C ABI loads and exports every
symbol include/tiltgpu.h declares, the host-computed tables equal the oracle's
bit for bit, the text parsers follow the reference, the phantom generator is
deterministic, and errors come back as status codes (no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle as O
import tiltgpu as tg

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(REPO, "include", "tiltgpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tiltgpu_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = tg.lib()
    names = declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    # and the Python binding declares a signature for each of them
    bound = {n for n, _, _ in tg.SIGNATURES}
    assert set(names) <= bound, set(names) - bound


def test_version_and_device_count():
    assert b"gfx950" in tg.lib().tiltgpu_version()
    assert tg.device_count() >= 0


@pytest.mark.parametrize("xt,zf", [(None, None), (0.5, None), (None, 1.05), (-2.0, 0.9)])
def test_proj_params_bitwise_equal_oracle(xt, zf):
    ang = np.linspace(-61.3, 59.7, 47).astype(np.float32)
    xtilt = None if xt is None else np.full(47, xt, np.float32)
    zfac = None if zf is None else np.linspace(zf, 1.0, 40).astype(np.float32)  # fewer entries
    g = tg.proj_params(ang, xtilt, zfac)
    tab = O.proj_params(ang, xtilt, zfac, prec="f32")
    got = np.array([[v.cos_t, v.sin_t, v.cos_xt, v.sin_xt, v.zfactor] for v in g], np.float32).T
    assert np.array_equal(got, tab)


@pytest.mark.parametrize("n,efs,ham", [(2048, 0, False), (768, 0, False), (513, 0, True),
                                       (256, 100, False), (100, 40, True)])
def test_radial_weights_bitwise_equal_oracle(n, efs, ham):
    a = tg.radial_weights(n, 0.35, 0.05, efs, ham)
    b = O.radial_weights(n, 0.35, 0.05, exact_filter_size=efs, hamming=ham, prec="f32")
    assert np.array_equal(a, b)


def test_read_tilt_file_matches_oracle(tmp_path):
    golden = os.path.join(REPO, "tests", "golden", "tilt_roundtrip.tlt")
    assert np.array_equal(tg.read_tilt_file(golden), O.read_tilt_file(golden))
    p = tmp_path / "a.tlt"
    p.write_text("  -60.0 \n\nabc\n1e1\n0x10\n 3,5\n+2.5\n-inf\n1.\n.5\n-.5e1\n")
    a, b = tg.read_tilt_file(str(p)), O.read_tilt_file(str(p))
    assert len(a) == len(b) == 7
    assert np.array_equal(a, b)


def test_read_per_view_file(tmp_path):
    p = tmp_path / "xt"
    p.write_text("# comment\n0.5\n\n 0.25 \n")
    assert list(tg.read_per_view_file(str(p))) == [0.5, 0.25]
    p.write_text("0.5\nbad\n")
    with pytest.raises(tg.TiltGpuError) as e:
        tg.read_per_view_file(str(p))
    assert e.value.code == tg.ERR_IO
    with pytest.raises(tg.TiltGpuError):
        tg.read_tilt_file(str(tmp_path / "missing.tlt"))


def test_synth_series_deterministic():
    ang = np.linspace(-60, 60, 9).astype(np.float32)
    a = tg.synth_series(96, 12, ang, vol_nz=40, n_ellipsoids=20, n_beads=3, seed=5, nthreads=1)
    b = tg.synth_series(96, 12, ang, vol_nz=40, n_ellipsoids=20, n_beads=3, seed=5, nthreads=4)
    c = tg.synth_series(96, 12, ang, vol_nz=40, n_ellipsoids=20, n_beads=3, seed=6, nthreads=4)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.shape == (9, 12, 96) and np.all(np.isfinite(a))


def test_synth_series_bead_chord():
    """One bead of radius r and density d, no noise: the projection through its
    centre is the chord 2*r*d at every tilt."""
    ang = np.array([-45.0, 0.0, 30.0], np.float32)
    P = tg.synth_series(64, 64, ang, vol_nz=64, n_ellipsoids=0, n_beads=1, seed=1,
                        noise_rel=0.0, bead_radius=6.0, bead_density=2.0)
    for v in range(3):
        assert P[v].max() == pytest.approx(2 * 6.0 * 2.0, rel=2e-2)


def test_errors_are_status_codes():
    L = tg.lib()
    h = C.c_void_p()
    d = tg.Dims(0, 10, 10, 10, 10)
    rc = L.tiltgpu_create(C.byref(d), None, C.byref(h))
    assert rc == tg.ERR_ARG and b"invalid dimensions" in L.tiltgpu_last_error()
    assert L.tiltgpu_run(None) == tg.ERR_STATE
    if tg.device_count() == 0:
        d = tg.Dims(64, 8, 5, 64, 16)
        rc = L.tiltgpu_create(C.byref(d), None, C.byref(h))
        assert rc == tg.ERR_NODEV and not h.value


def test_gpu_session_new_returns_none_without_gpu():
    if tg.device_count() > 0:
        pytest.skip("a GPU is visible")
    P = np.zeros((3, 4, 16), np.float32)
    s = tg.GpuSession.new(P, (np.zeros(3, np.float32),), np.zeros(3, np.float32), 16, 4, 3, 16, 8)
    assert s is None  # GpuSession::new -> None, main.rs:746-749 falls back
