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


if __name__ == "__main__":
    main()
