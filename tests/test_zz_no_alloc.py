"""The allocation-free backprojection entry point (tiltgpu_reconstruct_rows_noalloc)
against the seam session it varies (tiltgpu_bp_session_new +
tiltgpu_reconstruct_rows): bit for bit, on two calls in a row, the second with a
smaller input, and its size limit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(tg, nx, ny, nz, xtilt, seed):
    rng = np.random.default_rng(seed)
    P = rng.normal(size=(nz, ny, nx)).astype(np.float32)
    ang = np.linspace(-58, 58, nz).astype(np.float32)
    params = tg.proj_params(ang, None if xtilt is None else np.full(nz, xtilt, np.float32))
    means = P.reshape(nz, -1).mean(axis=1).astype(np.float32)
    return P, params, means


def _both(tg, P, params, means, out_nx, out_nz, y0, y1, fill):
    nz, ny, nx = P.shape
    print(f"noalloc case: in {nx}x{ny}x{nz} out {out_nx}x{out_nz} rows [{y0},{y1}) fill {fill}",
          flush=True)
    s = tg.GpuSession.new(P, params, means, nx, ny, nz, out_nx, out_nz, fill_mode=fill)
    assert s is not None
    ref = s.reconstruct_rows(y0, y1)
    s.close()
    d_p = tg.DeviceArray(P.shape)
    d_o = tg.DeviceArray((y1 - y0, out_nz, out_nx))
    try:
        d_p.upload(P)
        tg.reconstruct_rows_noalloc(d_p, params, means, nx, ny, nz, out_nx, out_nz, y0, y1, d_o,
                                    fill_mode=fill)
        got = d_o.download()
    finally:
        d_p.close()
        d_o.close()
    return got, ref


def test_noalloc_equals_seam_session_two_calls(gpu):
    # first call: two-row geometry (x-axis tilt), wider output, edge fill
    P, prm, m = _inputs(gpu, 96, 12, 25, 0.7, 1)
    got, ref = _both(gpu, P, prm, m, 100, 40, 2, 10, "edge")
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    # second call, smaller: one-row geometry, mean fill (the scratch keeps the
    # first call's tables; everything read must be rewritten)
    P, prm, m = _inputs(gpu, 64, 6, 15, None, 2)
    got, ref = _both(gpu, P, prm, m, 64, 24, 0, 6, "mean")
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_noalloc_rejects_geometry_beyond_scratch(gpu):
    P, prm, m = _inputs(gpu, 16, 2, 400, None, 3)
    d_p = gpu.DeviceArray(P.shape)
    d_o = gpu.DeviceArray((1, 8192, 16))
    try:
        with pytest.raises(gpu.TiltGpuError) as ei:
            gpu.reconstruct_rows_noalloc(d_p, prm, m, 16, 2, 400, 16, 8192, 0, 1, d_o)
        assert ei.value.code == gpu.ERR_ARG and "scratch" in str(ei.value)
    finally:
        d_p.close()
        d_o.close()
