// Synthetic tilt series for benchmarks and tests: a seeded phantom of random
// ellipsoids plus gold beads, projected analytically (exact chord lengths) with
// the `tilt` geometry of imod-rs apps/tilt/src/main.rs:1015-1045:
//   u = ((x - ocx)*cos t)*cos xt + (z - cz)*sin t + cx + shift_x
//   v = y + (x - ocx)*sin xt + shift_y
// so a reconstruction with the same angles / x-tilt / local shifts lines up.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "../../include/tiltgpu.h"

namespace {

struct Rng {  // xoshiro256** seeded through splitmix64
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (auto& v : s) {
            z += 0x9e3779b97f4a7c15ull;
            uint64_t x = z;
            x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
            x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
            v = x ^ (x >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return static_cast<double>(next() >> 11) * (1.0 / 9007199254740992.0); }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    double normal() {
        double u1 = uniform();
        while (u1 <= 1e-300) u1 = uniform();
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

struct Body {
    double c[3];
    double M[3][3];  // quadratic form: (p-c)^T M (p-c) <= 1
    double rmax;
    double density;
};

void make_ellipsoid(Rng& r, const tiltgpu_phantom& ph, Body& b) {
    b.c[0] = r.uniform(0, ph.vol_nx);
    b.c[1] = r.uniform(0, ph.vol_ny);
    b.c[2] = r.uniform(0, ph.vol_nz);
    double a[3];
    for (double& v : a) v = r.uniform(ph.axis_min, ph.axis_max);
    // uniform random rotation from a unit quaternion
    double q[4];
    double n2 = 0;
    do {
        n2 = 0;
        for (double& v : q) {
            v = r.normal();
            n2 += v * v;
        }
    } while (n2 < 1e-12);
    const double inv = 1.0 / std::sqrt(n2);
    for (double& v : q) v *= inv;
    const double w = q[0], x = q[1], y = q[2], z = q[3];
    const double R[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)},
                            {2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)},
                            {2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)}};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0;
            for (int k = 0; k < 3; ++k) s += R[i][k] * R[j][k] / (a[k] * a[k]);
            b.M[i][j] = s;
        }
    b.rmax = std::max(a[0], std::max(a[1], a[2]));
    b.density = r.uniform(-1.0, 1.0);
}

void make_bead(Rng& r, const tiltgpu_phantom& ph, Body& b) {
    b.c[0] = r.uniform(0, ph.vol_nx);
    b.c[1] = r.uniform(0, ph.vol_ny);
    b.c[2] = r.uniform(0, ph.vol_nz);
    const double inv = 1.0 / (static_cast<double>(ph.bead_radius) * ph.bead_radius);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) b.M[i][j] = i == j ? inv : 0.0;
    b.rmax = ph.bead_radius;
    b.density = ph.bead_density;
}

void project_view(const std::vector<Body>& bodies, const tiltgpu_phantom& ph, int nx, int ny,
                  double theta_deg, double xt_deg, double sx, double sy, float* img) {
    std::fill(img, img + static_cast<size_t>(nx) * ny, 0.f);
    const double t = theta_deg * M_PI / 180.0, xt = xt_deg * M_PI / 180.0;
    const double ct = std::cos(t), st = std::sin(t), cxt = std::cos(xt), sxt = std::sin(xt);
    const double ocx = ph.vol_nx / 2.0, cz = ph.vol_nz / 2.0, cx = nx / 2.0;
    const double cu = ct * cxt;
    if (std::fabs(cu) < 1e-9) return;
    // ray direction: null vector of the projection map, normalised
    double d[3] = {st, -sxt * st, -ct * cxt};
    const double dn = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (double& v : d) v /= dn;
    const double eu = std::sqrt(cu * cu + st * st), ev = std::sqrt(sxt * sxt + 1.0);
    std::vector<double> row(static_cast<size_t>(nx) * ny, 0.0);
    for (const Body& b : bodies) {
        const double uc = ((b.c[0] - ocx) * ct) * cxt + (b.c[2] - cz) * st + cx + sx;
        const double vc = b.c[1] + (b.c[0] - ocx) * sxt + sy;
        const int u0 = std::max(0, static_cast<int>(std::floor(uc - b.rmax * eu)) - 1);
        const int u1 = std::min(nx - 1, static_cast<int>(std::ceil(uc + b.rmax * eu)) + 1);
        const int v0 = std::max(0, static_cast<int>(std::floor(vc - b.rmax * ev)) - 1);
        const int v1 = std::min(ny - 1, static_cast<int>(std::ceil(vc + b.rmax * ev)) + 1);
        if (u0 > u1 || v0 > v1) continue;
        double Md[3];
        for (int i = 0; i < 3; ++i) Md[i] = b.M[i][0] * d[0] + b.M[i][1] * d[1] + b.M[i][2] * d[2];
        const double A = d[0] * Md[0] + d[1] * Md[1] + d[2] * Md[2];
        for (int v = v0; v <= v1; ++v) {
            for (int u = u0; u <= u1; ++u) {
                // point on the ray with z = cz
                const double x = ocx + (u - cx - sx) / cu;
                const double y = v - sy - (x - ocx) * sxt;
                const double w[3] = {x - b.c[0], y - b.c[1], cz - b.c[2]};
                double Mw[3];
                for (int i = 0; i < 3; ++i)
                    Mw[i] = b.M[i][0] * w[0] + b.M[i][1] * w[1] + b.M[i][2] * w[2];
                const double B = d[0] * Mw[0] + d[1] * Mw[1] + d[2] * Mw[2];
                const double C = w[0] * Mw[0] + w[1] * Mw[1] + w[2] * Mw[2] - 1.0;
                const double disc = B * B - A * C;
                if (disc > 0) row[static_cast<size_t>(v) * nx + u] += b.density * 2.0 * std::sqrt(disc) / A;
            }
        }
    }
    for (size_t i = 0; i < row.size(); ++i) img[i] = static_cast<float>(row[i]);
}

template <typename F>
void parallel_views(int n, int nthreads, F fn) {
    std::atomic<int> next{0};
    auto work = [&]() {
        for (;;) {
            const int i = next.fetch_add(1);
            if (i >= n) break;
            fn(i);
        }
    };
    const int nt = std::max(1, std::min(nthreads, n));
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(work);
    work();
    for (auto& x : th) x.join();
}

}  // namespace

extern "C" int tiltgpu_synth_series(const tiltgpu_phantom* ph, int32_t nx, int32_t ny,
                                    int32_t n_views, const float* angles, const float* xtilt,
                                    const float* shift_x, const float* shift_y,
                                    int32_t nthreads, float* out) {
    if (!ph || !angles || !out || nx < 1 || ny < 1 || n_views < 1) return TILTGPU_ERR_ARG;
    Rng r(ph->seed);
    std::vector<Body> bodies(static_cast<size_t>(ph->n_ellipsoids) + ph->n_beads);
    for (int i = 0; i < ph->n_ellipsoids; ++i) make_ellipsoid(r, *ph, bodies[i]);
    for (int i = 0; i < ph->n_beads; ++i) make_bead(r, *ph, bodies[ph->n_ellipsoids + i]);
    const size_t per = static_cast<size_t>(nx) * ny;
    if (nthreads < 1) nthreads = static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
    parallel_views(n_views, nthreads, [&](int v) {
        project_view(bodies, *ph, nx, ny, angles[v], xtilt ? xtilt[v] : 0.0,
                     shift_x ? shift_x[v] : 0.0, shift_y ? shift_y[v] : 0.0, out + v * per);
    });
    if (ph->noise_rel > 0) {
        std::vector<double> s1(n_views), s2(n_views);
        parallel_views(n_views, nthreads, [&](int v) {
            double a = 0, b = 0;
            for (size_t i = 0; i < per; ++i) {
                const double x = out[v * per + i];
                a += x;
                b += x * x;
            }
            s1[v] = a;
            s2[v] = b;
        });
        double a = 0, b = 0;
        for (int v = 0; v < n_views; ++v) {
            a += s1[v];
            b += s2[v];
        }
        const double N = static_cast<double>(per) * n_views;
        const double mean = a / N;
        const double sd = std::sqrt(std::max(0.0, b / N - mean * mean));
        const double sigma = ph->noise_rel * sd;
        parallel_views(n_views, nthreads, [&](int v) {
            Rng nr(ph->seed * 0x100000001b3ull + 0x51ed27ull + static_cast<uint64_t>(v) * 0x9e3779b97f4a7c15ull);
            for (size_t i = 0; i < per; ++i)
                out[v * per + i] = static_cast<float>(out[v * per + i] + sigma * nr.normal());
        });
    }
    return TILTGPU_OK;
}
