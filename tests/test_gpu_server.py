"""Continuous-batching server on the GPU: every request's result is the same
(bit for bit) alone, in a mix, and as a plain single-series session; admission
respects the device-memory budget; slabs of a resident series equal the rows of
the full reconstruction; results agree with the oracle."""
import threading
import time

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def series(tg, nx, nviews, thick, seed, storage="f32", sirt=0):
    ang = np.linspace(-60, 60, nviews).astype(np.float32)
    P = tg.synth_series(nx, nx // 4, ang, vol_nz=thick, n_ellipsoids=20, n_beads=4, seed=seed,
                        axis=(2.0, 9.0), bead_radius=2.0)
    return P, ang, tg.TiltOptions(thickness=thick, sirt=sirt, storage=storage)


def mix(tg):
    return [series(tg, 128, 31, 48, 1), series(tg, 256, 41, 72, 2, sirt=3),
            series(tg, 96, 25, 40, 3, storage="f16"), series(tg, 192, 35, 64, 4),
            series(tg, 160, 29, 100, 5, sirt=2), series(tg, 128, 21, 130, 6)]


def test_server_alone_equals_mix_equals_session(gpu, rel):
    reqs = mix(gpu)
    print("case: 6 requests (WBP, SIRT-2/3, f16), alone / mix / session", flush=True)
    with gpu.Server(8 << 30) as srv:
        alone = []
        for P, ang, o in reqs:
            rid = srv.submit(P, ang, o)
            alone.append(srv.wait(rid)[0])
        ids = [srv.submit(P, ang, o) for P, ang, o in reqs]
        mixed = [srv.wait(i)[0] for i in ids]
        st = srv.stats()
    assert st["completed"] == 12 and st["failed"] == 0
    assert st["max_active"] >= 2  # the mix really ran concurrently
    for (P, ang, o), a, m in zip(reqs, alone, mixed):
        assert np.array_equal(a, m)
        s, _ = gpu.reconstruct(P, ang, o)
        assert np.array_equal(a, s)
        if o.storage == "f32":
            ref, _ = O.tilt_run(P, ang, O.TiltOptions(thickness=o.thickness, sirt=o.sirt))
            assert rel(a, ref) <= 1e-4


def test_server_budget_admission(gpu):
    P, ang, o = series(gpu, 128, 31, 48, 7)
    one = 2 * P.nbytes + P.shape[1] * 48 * 128 * 4 + (1 << 20)
    with gpu.Server(int(one * 1.5)) as srv:  # room for one request at a time
        ids = [srv.submit(P, ang, o) for _ in range(4)]
        outs = [srv.wait(i)[0] for i in ids]
        st = srv.stats()
        assert st["max_active"] == 1 and st["completed"] == 4
        for x in outs[1:]:
            assert np.array_equal(x, outs[0])
        big, bang, bo = series(gpu, 512, 61, 128, 8)
        with pytest.raises(gpu.TiltGpuError) as e:
            srv.submit(big, bang, bo)
        assert e.value.code == gpu.ERR_BUDGET


def test_server_concurrent_submitters(gpu):
    reqs = mix(gpu)
    ref = [gpu.reconstruct(P, a, o)[0] for P, a, o in reqs]
    got = [None] * len(reqs)
    with gpu.Server(8 << 30) as srv:
        def work(k):
            time.sleep(0.002 * k)
            rid = srv.submit(*reqs[k])
            got[k] = srv.wait(rid)[0]
        th = [threading.Thread(target=work, args=(k,)) for k in range(len(reqs))]
        for t in th:
            t.start()
        for t in th:
            t.join(timeout=120)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


def test_server_slabs_of_resident_series(gpu, rel):
    ang = np.linspace(-60, 60, 31).astype(np.float32)
    rng = np.random.default_rng(9)
    sx, sy = rng.normal(0, 1.5, 31).astype(np.float32), rng.normal(0, 1.5, 31).astype(np.float32)
    xt = np.full(31, 0.5, np.float32)
    P = gpu.synth_series(128, 40, ang, vol_nz=48, n_ellipsoids=25, n_beads=4, seed=9, xtilt=xt,
                         shift_x=sx, shift_y=sy, axis=(2.0, 9.0), bead_radius=2.0)
    grids = O.build_local_grids([(v, 64.0, 20.0, float(sx[v]), float(sy[v])) for v in range(31)], 31)
    opts = gpu.TiltOptions(thickness=48)
    full, _ = gpu.reconstruct(P, ang, opts, xtilt=xt, local=grids)
    with gpu.Server(4 << 30) as srv:
        sid = srv.add_series(P, ang, opts, xtilt=xt, local=grids)
        ids = [srv.submit_slab(sid, y0, y0 + 8) for y0 in (0, 8, 16, 24, 32)]
        ids.append(srv.submit_slab(sid, 13, 29))
        slabs = [srv.wait(i)[0] for i in ids]
    for k, y0 in enumerate((0, 8, 16, 24, 32)):
        assert np.array_equal(slabs[k], full[y0:y0 + 8])
    assert np.array_equal(slabs[5], full[13:29])
    ref, _ = O.tilt_run(P, ang, O.TiltOptions(thickness=48), xtilt=xt, local=grids)
    assert rel(np.concatenate(slabs[:5]), ref) <= 1e-4


def test_server_rejects_bad_requests(gpu):
    with gpu.Server(1 << 30) as srv:
        with pytest.raises(gpu.TiltGpuError):
            srv.submit_slab(3, 0, 4)  # no such series
        P, ang, o = series(gpu, 128, 31, 48, 10)
        with pytest.raises(gpu.TiltGpuError):
            srv.submit(P, ang[:5], o)  # stretch needs an angle per section
        ok = srv.submit(P, ang, o)
        assert srv.wait(ok)[0].shape == (32, 48, 128)


def test_server_requests_with_geometry(gpu, rel):
    """Served reconstructions carrying x-axis tilts, z factors and local grids
    (tiltgpu_server_submit_geometry): WBP and SIRT bit for bit the plain
    session's, in a mix with plain requests, and at the bar against the oracle."""
    ang = np.linspace(-60, 60, 31).astype(np.float32)
    rng = np.random.default_rng(21)
    sx, sy = rng.normal(0, 1.5, 31).astype(np.float32), rng.normal(0, 1.5, 31).astype(np.float32)
    xt = np.linspace(-1.0, 1.0, 31).astype(np.float32)
    zf = np.linspace(0.95, 1.05, 31).astype(np.float32)
    P = gpu.synth_series(128, 32, ang, vol_nz=48, n_ellipsoids=25, n_beads=4, seed=21, xtilt=xt,
                         shift_x=sx, shift_y=sy, axis=(2.0, 9.0), bead_radius=2.0)
    patches = []
    for v in range(31):
        for x in (20.0, 100.0):
            for y in (6.0, 26.0):
                patches.append((v, x, y, float(sx[v] + rng.normal(0, 0.3)),
                                float(sy[v] + rng.normal(0, 0.3))))
    grids = O.build_local_grids(patches, 31)
    wbp = gpu.TiltOptions(thickness=48)
    sirt = gpu.TiltOptions(thickness=48, sirt=2)
    plain = mix(gpu)[:3]
    print("case: served WBP and SIRT-2 with x-tilt, z factor, 2x2 local grids, in a mix", flush=True)
    with gpu.Server(8 << 30) as srv:
        ids = [srv.submit(P, ang, wbp, xtilt=xt, zfactor=zf, local=grids)]
        ids += [srv.submit(Q, a, o) for Q, a, o in plain]
        ids.append(srv.submit(P, ang, sirt, xtilt=xt, zfactor=zf, local=grids))
        got = [srv.wait(i)[0] for i in ids]
    ref_w, _ = gpu.reconstruct(P, ang, wbp, xtilt=xt, zfactor=zf, local=grids)
    assert np.array_equal(got[0], ref_w)
    for (Q, a, o), g in zip(plain, got[1:4]):
        assert np.array_equal(g, gpu.reconstruct(Q, a, o)[0])
    ref_s, _ = gpu.reconstruct(P, ang, sirt, xtilt=xt, zfactor=zf, local=grids)
    assert np.array_equal(got[4], ref_s)
    o_w, _ = O.tilt_run(P, ang, O.TiltOptions(thickness=48), xtilt=xt, zfactor=zf, local=grids)
    assert rel(got[0], o_w) <= 1e-4
    o_s, _ = O.tilt_run(P, ang, O.TiltOptions(thickness=48, sirt=2), xtilt=xt, zfactor=zf,
                        local=grids)
    assert rel(got[4], o_s) <= 1e-4


def test_server_batched_sizes_against_oracle(gpu, rel):
    """Requests of the batched configuration's own sizes through one server
    launch set: a 768-wide SIRT-10 (97 views, 211 thick), a 1024-wide WBP (121
    views, 300 thick), a 512-wide SIRT-10 with FP16 storage and a 256-wide WBP.
    FP32 results match the FP32 oracle on row bands at the 1e-4 bar (SIRT: rows
    evolve independently in the single-row geometry, so the oracle's band is the
    whole run's); every result equals a plain session bit for bit."""
    def make(w, nv, th, seed):
        ang = np.linspace(-60, 60, nv).astype(np.float32)
        P = gpu.synth_series(w, w, ang, vol_nz=th, n_ellipsoids=int(min(400, 200 * (w / 512) ** 2)),
                             n_beads=30, seed=seed, noise_rel=0.1)
        return P, ang
    reqs = [(make(768, 97, 211, 31), gpu.TiltOptions(thickness=211, sirt=10)),
            (make(1024, 121, 300, 32), gpu.TiltOptions(thickness=300)),
            (make(512, 61, 150, 33), gpu.TiltOptions(thickness=150, sirt=10, storage="f16")),
            (make(256, 41, 64, 34), gpu.TiltOptions(thickness=64))]
    print("case: served 768 SIRT-10, 1024 WBP, 512 SIRT-10 f16, 256 WBP", flush=True)
    with gpu.Server(32 << 30) as srv:
        ids = [srv.submit(P, ang, o) for (P, ang), o in reqs]
        got = [srv.wait(i)[0] for i in ids]
        st = srv.stats()
    assert st["completed"] == 4 and st["failed"] == 0
    for ((P, ang), o), g in zip(reqs, got):
        assert np.array_equal(g, gpu.reconstruct(P, ang, o)[0])
    (P, ang), _ = reqs[0]
    tr, _ = O.sirt_trace(P, ang, O.TiltOptions(thickness=211, sirt=10), (380, 388), prec="f32",
                         nthreads=16)
    assert rel(got[0][380:388], tr[10]) <= 1e-4
    (P, ang), _ = reqs[1]
    ref, _ = O.tilt_run(P, ang, O.TiltOptions(thickness=300), prec="f32", nthreads=16,
                        rows=(508, 516))
    assert rel(got[1][508:516], ref) <= 1e-4
    (P, ang), _ = reqs[3]
    ref, _ = O.tilt_run(P, ang, O.TiltOptions(thickness=64), prec="f32", nthreads=16)
    assert rel(got[3], ref) <= 1e-4
