// AddressSanitizer / UndefinedBehaviorSanitizer driver for the host-side code
// of libtiltgpu (text parsers, MRC I/O, the synthetic series, the launch-shape
// table) and for the oracle.  Built with -fsanitize=address,undefined by
// tests/test_sanitizers.py and run on small valid and malformed inputs; a clean
// run exits 0 with no sanitizer report.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tiltgpu.h"
#include "../../imod-rs_amd/csrc/tuning.h"

extern "C" {
typedef struct {
    int in_nx, in_ny, nz, out_nx, out_nz;
    float cutoff, falloff;
    int no_weight, exact_filter_size, hamming;
    int no_cosine_stretch, cosine_cubic;
    float max_cosine_tilt;
    int log_;
    int fill_mode;
    float densweight, outscale, outadd;
    int sirt;
    float sirt_relax;
    int fake_sirt;
    int nthreads;
} or_opts;
void or_tilt_run_f32(const or_opts*, float*, const float*, int, const float*, int, const float*,
                     int, const int*, const int*, const float*, const float*, const float*,
                     const float*, float*, float*);
void or_tilt_run_f64(const or_opts*, float*, const float*, int, const float*, int, const float*,
                     int, const int*, const int*, const float*, const float*, const float*,
                     const float*, float*, float*);
void or_apply_xf_stack_f32(const float*, int, int, int, const float*, int, float, float*, int);
void or_ctf_phaseflip_f32(const float*, int, int, int, const float*, int, const float*, float,
                          const double*, int, float*, int);
void or_ctf_phaseflip_f64(const float*, int, int, int, const float*, int, const float*, float,
                          const double*, int, float*, int);
}

// libtiltgpu's error recorder, which the host sources call; the device part of
// the library (where it lives) is not built here.
namespace tiltgpu_sess {
int fail(int code, const char*, ...) { return code; }
}  // namespace tiltgpu_sess

static int g_fail = 0;
#define CHECK(c)                                                       \
    do {                                                               \
        if (!(c)) {                                                    \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            g_fail = 1;                                                \
        }                                                              \
    } while (0)

static std::string put(const std::string& dir, const char* name, const char* text) {
    const std::string p = dir + "/" + name;
    FILE* f = std::fopen(p.c_str(), "wb");
    std::fputs(text, f);
    std::fclose(f);
    return p;
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string dir = argv[1];

    // ---- text parsers: count-only calls, then filled, then malformed input
    const std::string tlt = put(dir, "a.tlt", "-60\n  -30.5 \nnot a number\n\n0\n30\n60.0\n");
    int n = tiltgpu_read_tilt_file(tlt.c_str(), nullptr, 0);
    CHECK(n == 5);
    std::vector<float> ang(n);
    CHECK(tiltgpu_read_tilt_file(tlt.c_str(), ang.data(), n) == 5 && ang[1] == -30.5f);
    CHECK(tiltgpu_read_tilt_file(tlt.c_str(), ang.data(), 2) == 5);  // cap below the count
    CHECK(tiltgpu_read_tilt_file((dir + "/missing").c_str(), nullptr, 0) < 0);
    const std::string pv = put(dir, "x.txt", "# xtilt\n0.5\n\n-0.25\n");
    CHECK(tiltgpu_read_per_view_file(pv.c_str(), nullptr, 0) == 2);
    CHECK(tiltgpu_read_per_view_file(put(dir, "bad.txt", "1\nq\n").c_str(), nullptr, 0) < 0);
    const std::string xf = put(dir, "a.xf", "1 0 0 1 2 3\n0.99 -0.1 0.1 0.99 -1 0.5\nshort 1 2\n");
    CHECK(tiltgpu_read_xf_file(xf.c_str(), nullptr, 0) == 2);
    std::vector<float> xfv(12);
    CHECK(tiltgpu_read_xf_file(xf.c_str(), xfv.data(), 2) == 2 && xfv[5] == 3.0f);
    const std::string df = put(dir, "d.txt", "# v d1 d2 a\n0 -2500\n3 -2600 -2900 12\n7 -2800 -3100\n");
    n = tiltgpu_read_defocus_file(df.c_str(), nullptr, nullptr, 0);
    CHECK(n == 3);
    std::vector<int32_t> dv(n);
    std::vector<float> de(3 * n), interp(3 * 10);
    CHECK(tiltgpu_read_defocus_file(df.c_str(), dv.data(), de.data(), n) == 3);
    CHECK(tiltgpu_interpolate_defocus(dv.data(), de.data(), n, 10, interp.data()) == 10);
    CHECK(tiltgpu_interpolate_defocus(dv.data(), de.data(), n, 2, interp.data()) == 2);
    CHECK(tiltgpu_interpolate_defocus(nullptr, nullptr, 0, 10, interp.data()) == 0);
    CHECK(tiltgpu_read_defocus_file(put(dir, "db.txt", "0 -1\n1 x\n").c_str(), nullptr, nullptr, 0) == -102);
    const std::string lf = put(dir, "l.txt",
                               "# view x y dx dy\n0 10 10 1 2\n0 50 10 3 4\n0 10 40 5 6\n"
                               "0 50 40 7 8\n1 30 30 -1 -1\n5 1 1 1 1\nshort line\n");
    const int nv = 9;  // views of the series below; the local grids cover all of them
    int32_t sizes[3] = {0, 0, 0};
    CHECK(tiltgpu_parse_local_file(lf.c_str(), nv, sizes, nullptr, nullptr, nullptr, nullptr,
                                   nullptr, nullptr) == 0);
    std::vector<int32_t> lnx(nv), lny(nv);
    std::vector<float> lxs(sizes[0] + 1), lys(sizes[1] + 1), ldx(sizes[2] + 1), ldy(sizes[2] + 1);
    CHECK(tiltgpu_parse_local_file(lf.c_str(), nv, sizes, lnx.data(), lny.data(), lxs.data(),
                                   lys.data(), ldx.data(), ldy.data()) == 0);
    CHECK(lnx[0] == 2 && lny[0] == 2 && lnx[1] == 1 && lnx[2] == 0 && lnx[5] == 1);

    // ---- MRC: write, header, read back, truncated file
    const int nx = 24, ny = 10, nz = 3;
    std::vector<float> vol(static_cast<size_t>(nx) * ny * nz);
    for (size_t i = 0; i < vol.size(); ++i) vol[i] = std::sin(0.37f * static_cast<float>(i));
    const std::string mrc = dir + "/v.mrc";
    CHECK(tiltgpu_mrc_write_float(mrc.c_str(), vol.data(), nx, ny, nz, 48.f, 20.f, 3.f,
                                  "written under the sanitizers") == 0);
    tiltgpu_mrc_info info;
    CHECK(tiltgpu_mrc_info_read(mrc.c_str(), &info) == 0 && info.nx == nx && info.nz == nz);
    std::vector<float> back(vol.size());
    CHECK(tiltgpu_mrc_read_sections(mrc.c_str(), 0, nz, back.data(),
                                    static_cast<int64_t>(back.size())) == 0);
    CHECK(std::memcmp(back.data(), vol.data(), vol.size() * 4) == 0);
    CHECK(tiltgpu_mrc_read_sections(mrc.c_str(), 0, nz, back.data(), 10) < 0);  // cap too small
    {
        FILE* f = std::fopen(mrc.c_str(), "rb");
        std::vector<char> bytes(1024 + 100);
        const size_t got = std::fread(bytes.data(), 1, bytes.size(), f);
        std::fclose(f);
        FILE* g = std::fopen((dir + "/t.mrc").c_str(), "wb");
        std::fwrite(bytes.data(), 1, got, g);
        std::fclose(g);
        CHECK(tiltgpu_mrc_read_sections((dir + "/t.mrc").c_str(), 0, nz, back.data(),
                                        static_cast<int64_t>(back.size())) < 0);
    }

    // ---- synthetic series and the launch-shape table
    tiltgpu_phantom ph;
    std::memset(&ph, 0, sizeof ph);
    ph.vol_nx = 32;
    ph.vol_ny = 8;
    ph.vol_nz = 16;
    ph.n_ellipsoids = 5;
    ph.n_beads = 2;
    ph.axis_min = 2.f;
    ph.axis_max = 6.f;
    ph.bead_radius = 1.5f;
    ph.bead_density = 3.f;
    ph.noise_rel = 0.1f;
    ph.seed = 7;
    std::vector<float> angles(nv), proj(static_cast<size_t>(nv) * 8 * 32);
    for (int v = 0; v < nv; ++v) angles[v] = -60.f + 15.f * static_cast<float>(v);
    CHECK(tiltgpu_synth_series(&ph, 32, 8, nv, angles.data(), nullptr, nullptr, nullptr, 2,
                               proj.data()) == 0);
    CHECK(tiltgpu::tuned_tile_z(1, 1, 0, 2048, 512) >= 0 && tiltgpu::tuned_fp_bins(1024, nullptr) > 0);

    // ---- oracle: WBP with x-tilt and local grids, SIRT, fake-SIRT, both precisions
    or_opts o;
    std::memset(&o, 0, sizeof o);
    o.in_nx = 32;
    o.in_ny = 3;
    o.nz = nv;
    o.out_nx = 32;
    o.out_nz = 12;
    o.cutoff = 0.35f;
    o.falloff = 0.05f;
    o.max_cosine_tilt = 80.f;
    o.densweight = o.outscale = 1.f;
    o.sirt_relax = 1.f;
    o.nthreads = 2;
    std::vector<float> P(static_cast<size_t>(nv) * 3 * 32);
    for (int v = 0; v < nv; ++v) std::memcpy(&P[static_cast<size_t>(v) * 96], &proj[static_cast<size_t>(v) * 256], 96 * 4);
    std::vector<float> xt(nv, 0.5f), out(static_cast<size_t>(3) * 12 * 32), stats(3);
    for (int variant = 0; variant < 4; ++variant) {
        or_opts q = o;
        q.sirt = variant == 1 ? 2 : 0;
        q.fake_sirt = variant == 2 ? 1 : 0;
        q.fill_mode = variant;
        q.cosine_cubic = variant == 3;
        std::vector<float> Pc = P;
        or_tilt_run_f32(&q, Pc.data(), angles.data(), nv, xt.data(), nv, nullptr, 0,
                        variant == 3 ? lnx.data() : nullptr, variant == 3 ? lny.data() : nullptr,
                        variant == 3 ? lxs.data() : nullptr, variant == 3 ? lys.data() : nullptr,
                        variant == 3 ? ldx.data() : nullptr, variant == 3 ? ldy.data() : nullptr,
                        out.data(), stats.data());
        Pc = P;
        or_tilt_run_f64(&q, Pc.data(), angles.data(), nv, xt.data(), nv, nullptr, 0, nullptr,
                        nullptr, nullptr, nullptr, nullptr, nullptr, out.data(), stats.data());
        CHECK(std::isfinite(stats[0]) && std::isfinite(stats[2]));
    }

    // ---- oracle: section transforms and the CTF phase flip (astigmatic, rotated axis)
    std::vector<float> xo(vol.size());
    or_apply_xf_stack_f32(vol.data(), nx, ny, nz, xfv.data(), 2, 0.f, xo.data(), 2);
    const double phys[7] = {0.0197, 2.7e7, 0.07, 0.0, 0.3, 0.5, 3.0};
    const float vdef[9] = {-2500, -2800, 20, -2600, -2600, 0, -2700, -3000, 40};
    const float tilt3[3] = {-30, 0, 30};
    or_ctf_phaseflip_f32(vol.data(), nx, ny, nz, tilt3, 3, vdef, -3000.f, phys, 8, xo.data(), 2);
    or_ctf_phaseflip_f64(vol.data(), nx, ny, nz, tilt3, 2, nullptr, -3000.f, phys, 16, xo.data(), 1);
    // ---- tiltalign local corrections: a small case, an empty one, bad arguments
    {
        const int nb = 4, nvw = 3;
        std::vector<double> obs(static_cast<size_t>(nb) * nvw * 2), beads(nb * 3);
        for (int b = 0; b < nb; ++b) {
            beads[3 * b] = 10.0 * b - 15.0;
            beads[3 * b + 1] = 7.0 * b - 10.0;
            beads[3 * b + 2] = 3.0 * b;
            for (int v = 0; v < nvw; ++v) {
                obs[(b * nvw + v) * 2] = (b == 2 && v == 1) ? NAN : beads[3 * b] + 0.3 * v;
                obs[(b * nvw + v) * 2 + 1] = beads[3 * b + 1] - 0.2 * v;
            }
        }
        tiltgpu_align_view av[3] = {{0.1, -0.2, 0.01, 1.0, -0.5}, {0, 0, 0, 1, 0}, {0.2, 0.1, -0.02, 1.01, 0.5}};
        int32_t ov[64], opx[64], opy[64];
        double odx[64], ody[64];
        const int n = tiltgpu_local_corrections(obs.data(), beads.data(), av, nb, nvw, 40, 0.5f,
                                                50.0, 40.0, 0, 0.0, 0.0, 64, ov, opx, opy, odx, ody);
        CHECK(n > 0 && n <= 64);
        CHECK(tiltgpu_local_corrections(obs.data(), beads.data(), av, nb, nvw, 40, 0.5f, 50.0, 40.0,
                                        5, 0.0, 0.0, 2, ov, opx, opy, odx, ody) >= 0);
        CHECK(tiltgpu_local_corrections(nullptr, nullptr, av, 0, nvw, 40, 0.5f, 50.0, 40.0, 0, 0.0,
                                        0.0, 0, nullptr, nullptr, nullptr, nullptr, nullptr) == 0);
        CHECK(tiltgpu_local_corrections(nullptr, beads.data(), av, nb, nvw, 40, 0.5f, 50.0, 40.0, 0,
                                        0.0, 0.0, 4, ov, opx, opy, odx, ody) == TILTGPU_ERR_ARG);
        CHECK(tiltgpu_write_local_corrections((dir + "/local.txt").c_str(), n > 64 ? 64 : n, ov,
                                              opx, opy, odx, ody) == TILTGPU_OK);
        CHECK(tiltgpu_write_local_corrections((dir + "/no/such/dir/l.txt").c_str(), 0, ov, opx,
                                              opy, odx, ody) == TILTGPU_ERR_IO);
    }
    std::fprintf(stderr, "host_asan_driver: %s\n", g_fail ? "FAILED" : "ok");
    return g_fail;
}
