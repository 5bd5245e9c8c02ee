"""The forward projection's sliding lane windows (imod-rs_amd/csrc/kernels_fp.hip,
SlideMap) rest on two claims about the reference's coordinate chain
px = ((dx*cos_t) + zc) + cx in FP32 (main.rs:858-860): for each z row the bins a
voxel row reaches lie inside the live windows w0 .. w0+ncol-1 (w0 = floor(lo/bw),
lo = max(floor(px of voxel 0), 0), bw = ceil((cos_t*(out_nx-1) + 4)/(ncol-1))),
and w0 moves one way as z advances.  Checked here over random geometries with
the same FP32 operations, CPU only."""
import numpy as np


def _bins_reached(px, in_nx):
    fl = np.floor(px)
    ok = (fl >= 0) & (fl + 1 < in_nx)
    f = fl[ok].astype(np.int64)
    return np.concatenate([f, f + 1])


def test_live_windows_cover_every_reached_bin_and_move_one_way():
    rng = np.random.default_rng(7)
    f32 = np.float32
    checked = 0
    for _ in range(400):
        in_nx = int(rng.integers(16, 2049))
        out_nx = int(rng.integers(8, 2 * in_nx + 1))
        nz = int(rng.integers(1, 400))
        B = int(rng.choice([4, 8, 16]))
        ncol = 64 * ((in_nx + 64 * B - 1) // (64 * B))
        t = rng.uniform(-87.0, 87.0)
        zf = f32(rng.choice([1.0, rng.uniform(0.3, 3.5)]))
        ct, st = f32(np.cos(np.radians(t))), f32(np.sin(np.radians(t)))
        cx, ocx = f32(in_nx / 2.0), f32(out_nx / 2.0)
        need = (ct * f32(out_nx - 1) + f32(4.0)) / f32(ncol - 1)
        bw = int(np.ceil(need))
        if bw >= B:
            continue  # the fixed map serves this view
        bw = max(bw, 1)
        checked += 1
        dxs = (np.arange(out_nx, dtype=f32) - ocx).astype(f32)
        a = (dxs * ct).astype(f32)
        prev = None
        for oz in range(nz):
            zc = f32(f32(f32(oz) - f32(nz / 2.0)) * st) * zf
            px = ((a + zc).astype(f32) + cx).astype(f32)
            assert np.all(np.diff(px) >= 0)  # non-decreasing in ox
            lo = max(int(np.floor(px[0])), 0)
            w0 = lo // bw
            bins = _bins_reached(px, in_nx)
            if bins.size:
                assert bins.min() >= w0 * bw
                assert bins.max() < (w0 + ncol) * bw
            if prev is not None:
                assert (w0 - prev) * np.sign(st * zf) >= 0 or w0 == prev
            prev = w0
    assert checked >= 100  # geometries that take the sliding map
