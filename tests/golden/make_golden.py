"""Writes the small golden fixtures under tests/golden/.

Sources (values only; nothing from the reference is run or copied):
* min_max_mean.json   -- the known answer of test_min_max_mean,
                         crates/imod-math/src/tests.rs:17-23.
* tilt_roundtrip.tlt  -- angles -60..60 step 3 written in write_tilt_file's
                         format "{:10.2}" (crates/imod-transforms/src/tilt.rs:27-33),
                         the input of tilt_file_roundtrip (tests.rs:118-131).
* generate_tilt_angles.json -- generate_tilt_angles(-60, 3, 41) expectations
                         (crates/imod-transforms/src/tests.rs:133-140).
* bp_bruteforce.npz   -- a tiny backprojection computed by brute force in FP64
                         with numpy.interp (textbook linear interpolation), for
                         the oracle's FP64 and FP32 forms to match.
* geometry_bruteforce.npz -- backprojection and forward projection of the
                         two-row geometry (x-axis tilt, z factors, local
                         alignment grids) under all three fill modes, computed
                         in FP64 as an explicit sparse sampling operator built
                         from tent (hat) weights, numpy.interp for the local
                         grids and for Edge fill.  The forward projection is
                         that operator's transpose.  Rules restated from
                         apps/tilt/src/main.rs (values only; no reference code
                         is run): coordinates :1015-1049, row selection with
                         the last-row fallback :1051-1066, fill :386-406, local
                         grid lerp with clamped ends :171-228, FP :866-889.

    python tests/golden/make_golden.py
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def bp_bruteforce(P, angles_deg, out_nx, out_nz, fill_value):
    """Unweighted, unstretched backprojection of one-row geometry in FP64:
    V[iy, oz, ox] = (1/nz) * sum_p interp(P[p, iy], proj_x), where
    proj_x = (ox - out_nx/2) cos t + (oz - out_nz/2) sin t + in_nx/2 and
    samples with floor(proj_x) outside [0, in_nx-2] take fill_value[p]."""
    nz, ny, nx = P.shape
    V = np.zeros((ny, out_nz, out_nx))
    ox = np.arange(out_nx)[None, :] - out_nx / 2.0
    oz = np.arange(out_nz)[:, None] - out_nz / 2.0
    for p, deg in enumerate(angles_deg):
        t = np.deg2rad(float(deg))
        px = ox * np.cos(t) + oz * np.sin(t) + nx / 2.0
        inside = (np.floor(px) >= 0) & (np.floor(px) <= nx - 2)
        for iy in range(ny):
            v = np.interp(px, np.arange(nx), P[p, iy].astype(np.float64))
            V[iy] += np.where(inside, v, fill_value[p])
    return V / nz


def local_shift_bruteforce(xs, ys, gdx, gdy, x, y):
    """Bilinear interpolation of a rectilinear patch grid, clamped at its ends
    (main.rs:171-228): separable numpy.interp along x for each grid row, then
    along y.  gdx, gdy: [len(ys)][len(xs)].  An empty grid gives no shift."""
    if len(xs) == 0 or len(ys) == 0:
        return np.zeros_like(x), np.zeros_like(x)
    xs = np.asarray(xs, np.float64)
    ys = np.asarray(ys, np.float64)
    rx = np.array([np.interp(x, xs, np.asarray(gdx[j], np.float64)) for j in range(len(ys))])
    ry = np.array([np.interp(x, xs, np.asarray(gdy[j], np.float64)) for j in range(len(ys))])
    # rx[j] is row j at every point; interpolate across rows point by point
    out_x = np.array([np.interp(y[i], ys, rx[:, i]) for i in range(x.size)])
    out_y = np.array([np.interp(y[i], ys, ry[:, i]) for i in range(x.size)])
    return out_x, out_y


def sampling_operator(nx, ny, nz, out_nx, out_nz, angles, xtilt, zf, grids, fill, means):
    """The WBP of one series as an explicit affine map in FP64:
    vec(V) = A @ vec(P) + b, V [ny][out_nz][out_nx], P [nz][ny][nx].

    Per voxel (iy, oz, ox) and view p:
      px = dx cos t cos xt + dz sin t zf + nx/2 + ldx(ox, iy),  dx = ox - out_nx/2,
      sy = iy + dx sin xt + ldy(ox, iy),                        dz = oz - out_nz/2.
    Two-row path (|dx sin xt| > 1e-6 or a local file present) when rows
    floor(sy) and floor(sy)+1 both exist: tent weights 1 - |sy - r| over them.
    Otherwise row floor(sy) alone with weight 1 when it exists (this is also the
    last-row fallback), nothing when it does not.  Within a row: tent weights
    1 - |px - j| over pixels floor(px), floor(px)+1 when both exist; otherwise the
    fill: Mean adds the view mean, Zero nothing, Edge the end pixel (the clamped
    numpy.interp of the row).  Everything is scaled by 1/nz.
    grids: None or per view (xs, ys, gdx [ny][nx], gdy)."""
    import scipy.sparse as sp
    nvox = ny * out_nz * out_nx
    iy, oz, ox = np.meshgrid(np.arange(ny), np.arange(out_nz), np.arange(out_nx), indexing="ij")
    iy, oz, ox = iy.ravel(), oz.ravel(), ox.ravel()
    vox = np.arange(nvox)
    rows, cols, vals = [], [], []
    b = np.zeros(nvox)
    for p in range(len(angles)):
        t = np.deg2rad(np.float64(angles[p]))
        xt = np.deg2rad(np.float64(xtilt[p]))
        dx = ox - out_nx / 2.0
        dz = oz - out_nz / 2.0
        px = dx * np.cos(t) * np.cos(xt) + dz * np.sin(t) * np.float64(zf[p]) + nx / 2.0
        yoff = dx * np.sin(xt)
        sy = iy + yoff
        if grids is not None:
            xs, ys, gdx, gdy = grids[p]
            lx, ly = local_shift_bruteforce(xs, ys, gdx, gdy, ox.astype(np.float64),
                                            iy.astype(np.float64))
            px = px + lx
            sy = sy + ly
        two = (np.abs(yoff) > 1e-6) | (grids is not None)
        r0 = np.floor(sy).astype(np.int64)
        both = two & (r0 >= 0) & (r0 + 1 < ny)
        one = ~both & (r0 >= 0) & (r0 < ny)
        row_terms = []  # (mask, row index, row weight)
        row_terms.append((both, r0, 1.0 - np.abs(sy - r0)))
        row_terms.append((both, r0 + 1, 1.0 - np.abs(sy - (r0 + 1))))
        row_terms.append((one, r0, np.ones(nvox)))
        j0 = np.floor(px).astype(np.int64)
        inside = (j0 >= 0) & (j0 + 1 < nx)
        for m, r, wr in row_terms:
            base = (p * ny + r) * nx
            for j in (j0, j0 + 1):
                w = wr * (1.0 - np.abs(px - j))
                k = m & inside
                rows.append(vox[k]); cols.append(base[k] + j[k]); vals.append(w[k] / nz)
            k = m & ~inside
            if fill == "mean":
                b[k] += wr[k] * float(means[p]) / nz
            elif fill == "edge":
                je = np.where(px < 0, 0, nx - 1)
                rows.append(vox[k]); cols.append(base[k] + je[k]); vals.append(wr[k] / nz)
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(nvox, nz * ny * nx))
    return A, b


def geometry_cases():
    """Three geometries of the two-row path, each reconstructed with all three
    fills; sizes small enough that every rule is hit (rows at both ends, samples
    left and right of the row, views with and without patches)."""
    rng = np.random.default_rng(4321)
    nz, ny, nx, out_nx, out_nz = 7, 6, 20, 24, 10
    ang = (np.linspace(-54, 54, nz) + rng.uniform(-1, 1, nz)).astype(np.float32)
    cases = {}
    # x-axis tilt 3 deg on every view but one (its y offset is 0: one-row path)
    xt = np.full(nz, 3.0, np.float32)
    xt[3] = 0.0
    cases["xtilt"] = dict(angles=ang, xtilt=xt, zf=np.ones(nz, np.float32), patches=None)
    # per-view x-axis tilt and z factor, one patch per view (a constant shift)
    xt = rng.uniform(-2, 2, nz).astype(np.float32)
    zf = rng.uniform(0.9, 1.1, nz).astype(np.float32)
    pts = [(v, 12.0, 3.0, float(rng.normal(0, 1.5)), float(rng.normal(0, 1.5)))
           for v in range(nz)]
    cases["xtilt_zf_local1"] = dict(angles=ang, xtilt=xt, zf=zf,
                                    patches=np.array(pts, np.float64))
    # a 3 x 3 grid per view that covers only part of the slice (clamped outside);
    # view 2 has no patches; no x-axis tilt
    pts = []
    for v in range(nz):
        if v == 2:
            continue
        for y in (1.0, 3.0, 4.5):
            for x in (3.0, 11.0, 18.0):
                pts.append((v, x, y, float(rng.normal(0, 1.0)), float(rng.normal(0, 0.8))))
    cases["grid3x3"] = dict(angles=ang, xtilt=np.zeros(nz, np.float32),
                            zf=np.ones(nz, np.float32), patches=np.array(pts, np.float64))
    P = rng.normal(size=(nz, ny, nx)).astype(np.float32) + 0.5
    Vin = rng.normal(size=(ny, out_nz, out_nx)).astype(np.float32)
    return dict(nz=nz, ny=ny, nx=nx, out_nx=out_nx, out_nz=out_nz, P=P, Vin=Vin), cases


def grids_from_patches(patches, nz):
    """Per view (xs, ys, gdx, gdy) from distinct patch positions (no merging
    needed: the fixtures' positions are at least 1 apart)."""
    out = []
    for v in range(nz):
        sel = [q for q in patches if int(q[0]) == v]
        xs = sorted({np.float32(q[1]) for q in sel})
        ys = sorted({np.float32(q[2]) for q in sel})
        gdx = np.zeros((len(ys), len(xs)))
        gdy = np.zeros((len(ys), len(xs)))
        for q in sel:
            gdx[ys.index(np.float32(q[2])), xs.index(np.float32(q[1]))] = np.float32(q[3])
            gdy[ys.index(np.float32(q[2])), xs.index(np.float32(q[1]))] = np.float32(q[4])
        out.append((xs, ys, gdx, gdy))
    return out


def write_geometry_fixture(path):
    dims, cases = geometry_cases()
    nz, ny, nx = dims["nz"], dims["ny"], dims["nx"]
    out_nx, out_nz = dims["out_nx"], dims["out_nz"]
    P, Vin = dims["P"], dims["Vin"]
    means = P.reshape(nz, -1).astype(np.float64).mean(axis=1)
    arrays = {k: np.asarray(v) for k, v in dims.items()}
    for name, c in cases.items():
        grids = None if c["patches"] is None else grids_from_patches(c["patches"], nz)
        for fill in ("mean", "zero", "edge"):
            A, b = sampling_operator(nx, ny, nz, out_nx, out_nz, c["angles"], c["xtilt"],
                                     c["zf"], grids, fill, means)
            V = (A @ P.astype(np.float64).ravel() + b).reshape(ny, out_nz, out_nx)
            arrays[f"{name}/bp_{fill}"] = V
            if fill == "zero":
                # forward_project is the transpose of the zero-fill sampling
                # without the 1/nz (main.rs:866-889)
                R = nz * (A.T @ Vin.astype(np.float64).ravel())
                arrays[f"{name}/fp"] = R.reshape(nz, ny, nx)
        for k in ("angles", "xtilt", "zf"):
            arrays[f"{name}/{k}"] = np.asarray(c[k], np.float32)
        arrays[f"{name}/patches"] = (np.zeros((0, 5)) if c["patches"] is None
                                     else np.asarray(c["patches"], np.float64))
    np.savez(path, **arrays)


def main():
    with open(os.path.join(HERE, "min_max_mean.json"), "w") as fh:
        json.dump({"input": [3.0, 1.0, 4.0, 1.0, 5.0, 9.0], "min": 1.0, "max": 9.0,
                   "mean": 23.0 / 6.0, "mean_tol": 1e-5}, fh, indent=1)
    with open(os.path.join(HERE, "tilt_roundtrip.tlt"), "w") as fh:
        for a in range(-60, 61, 3):
            fh.write(f"{float(a):10.2f}\n")
    with open(os.path.join(HERE, "generate_tilt_angles.json"), "w") as fh:
        json.dump({"start": -60.0, "increment": 3.0, "count": 41,
                   "checks": [[0, -60.0, 1e-6], [20, 0.0, 1e-4], [40, 60.0, 1e-4]]}, fh)
    rng = np.random.default_rng(1234)
    P = rng.normal(size=(5, 3, 12)).astype(np.float32)
    ang = np.array([-50.0, -20.0, 0.0, 25.0, 55.0], np.float32)
    means = P.reshape(5, -1).astype(np.float64).mean(axis=1).astype(np.float32)
    V = bp_bruteforce(P, ang, 10, 7, means)
    np.savez(os.path.join(HERE, "bp_bruteforce.npz"), P=P, angles=ang, means=means, V=V,
             out_nx=10, out_nz=7)
    write_geometry_fixture(os.path.join(HERE, "geometry_bruteforce.npz"))


if __name__ == "__main__":
    main()
