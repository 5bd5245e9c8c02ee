"""GPU parity at the benchmark configurations' own sizes, against the FP32
oracle (the reference's CPU path restated in its operation order):

* the 2048^2 x 121 -> 2048^2 x 512 WBP with x-axis tilt 0.5 deg and per-view
  shifts as a local file (one patch per view), FP32 and FP16 storage, on output
  rows [992, 1056), with the oracle run on the whole stack so that every
  coordinate is the reference's;
* SIRT-20 of 1024^2 x 61 -> 1024^2 x 300, FP32 and BF16 storage, every iterate
  of output rows [504, 520) (single-row geometry: rows evolve independently, so
  the oracle's row subset is the whole run's).

Bar: max |gpu - oracle| <= 1e-4 * RMS for FP32 storage; the 16-bit storage
forms at the tolerances stated in their tests.
"""
import os
import sys

import numpy as np
import pytest

import oracle as O

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
import bench  # noqa: E402  (input generators of the benchmark configurations)

pytestmark = pytest.mark.gpu
TOL = 1e-4
C1_ROWS = (992, 1056)
C2_ROWS = (504, 520)
NTHREADS = 16


def _errs(x, ref):
    ref = np.asarray(ref, np.float64)
    d = np.asarray(x, np.float64) - ref
    rms = float(np.sqrt(np.mean(ref ** 2)))
    return float(np.max(np.abs(d))) / rms, float(np.sqrt(np.mean(d ** 2))) / rms


@pytest.fixture(scope="module")
def c1(gpu):
    print("c1: generating 2048x2048x121 inputs", flush=True)
    P, ang, xt, local = bench.c1_inputs(gpu, nthreads=NTHREADS)
    nv = P.shape[0]
    patches = [(v, float(local[2][v]), float(local[3][v]), float(local[4][v]),
                float(local[5][v])) for v in range(nv)]
    grids = O.build_local_grids(patches, nv)
    print(f"c1: FP32 oracle on rows {C1_ROWS} of the whole stack", flush=True)
    ref, _ = O.tilt_run(P, ang, O.TiltOptions(thickness=512, radial=(0.35, 0.05)), xtilt=xt,
                        local=grids, prec="f32", nthreads=NTHREADS, rows=C1_ROWS)
    return P, ang, xt, local, ref


def _c1_gpu_rows(tg, P, ang, xt, local, storage):
    opts = tg.TiltOptions(thickness=512, radial=(0.35, 0.05), storage=storage)
    with tg.Session(P.shape[2], P.shape[1], P.shape[0], opts) as s:
        s.set_angles(ang)
        s.set_xtilt(xt)
        s.set_local(*local)
        s.upload(P)
        s.run()
        return s.read_rows(*C1_ROWS, stats=False)[0]


def test_c1_wbp_f32_storage_vs_f32_oracle(gpu, c1):
    P, ang, xt, local, ref = c1
    print("c1: GPU WBP, FP32 storage", flush=True)
    got = _c1_gpu_rows(gpu, P, ang, xt, local, "f32")
    mx, rms = _errs(got, ref)
    print(f"c1 f32: max|d|/rms {mx:.3e}, rms(d)/rms {rms:.3e}", flush=True)
    assert got.shape == ref.shape == (64, 512, 2048)
    assert mx <= TOL


def test_c1_wbp_f16_storage_vs_f32_oracle(gpu, c1):
    """FP16 storage (per-view scale, FP32 accumulation) at its stated bar:
    max 5e-3 and RMS 5e-4 of the volume RMS."""
    P, ang, xt, local, ref = c1
    print("c1: GPU WBP, FP16 storage", flush=True)
    got = _c1_gpu_rows(gpu, P, ang, xt, local, "f16")
    mx, rms = _errs(got, ref)
    print(f"c1 f16: max|d|/rms {mx:.3e}, rms(d)/rms {rms:.3e}", flush=True)
    assert mx <= 5e-3 and rms <= 5e-4


@pytest.fixture(scope="module")
def c2(gpu):
    print("c2: generating 1024x1024x61 inputs", flush=True)
    P, ang = bench.c2_inputs(gpu, nthreads=NTHREADS)
    print(f"c2: FP32 oracle SIRT-20 trace of rows {C2_ROWS}", flush=True)
    ref, res = O.sirt_trace(P, ang, O.TiltOptions(thickness=300, sirt=20, sirt_relax=1.0),
                            C2_ROWS, prec="f32", nthreads=NTHREADS)
    return P, ang, ref


def _c2_gpu_trace(tg, P, ang, storage):
    opts = tg.TiltOptions(thickness=300, sirt=20, sirt_relax=1.0, storage=storage)
    nv, ny, nx = P.shape
    with tg.Session(nx, ny, nv, opts) as s:
        s.set_angles(ang)
        s.set_sirt_trace(*C2_ROWS)
        s.upload(P)
        s.run()
        return s.sirt_trace()


def test_c2_sirt_f32_storage_every_iterate_vs_f32_oracle(gpu, c2):
    P, ang, ref = c2
    print("c2: GPU SIRT-20, FP32 storage", flush=True)
    tr = _c2_gpu_trace(gpu, P, ang, "f32")
    assert tr.shape == ref.shape == (21, 16, 300, 1024)
    errs = [_errs(tr[k], ref[k]) for k in range(21)]
    print("c2 f32 per iterate max|d|/rms: " + " ".join(f"{e[0]:.2e}" for e in errs), flush=True)
    assert max(e[0] for e in errs) <= TOL


def test_c2_sirt_bf16_storage_every_iterate_vs_f32_oracle(gpu, c2):
    """BF16 storage of the filtered projections and residuals (8 mantissa bits,
    per-view scale): every iterate within RMS 1e-2 and max 1e-1 of the volume
    RMS of the FP32 oracle's."""
    P, ang, ref = c2
    print("c2: GPU SIRT-20, BF16 storage", flush=True)
    tr = _c2_gpu_trace(gpu, P, ang, "bf16")
    errs = [_errs(tr[k], ref[k]) for k in range(21)]
    print("c2 bf16 per iterate max|d|/rms: " + " ".join(f"{e[0]:.2e}" for e in errs)
          + "; rms(d)/rms: " + " ".join(f"{e[1]:.2e}" for e in errs), flush=True)
    assert max(e[1] for e in errs) <= 1e-2
    assert max(e[0] for e in errs) <= 1e-1
