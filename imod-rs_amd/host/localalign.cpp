// tiltalign's local-alignment output: per-patch mean residuals of the bead
// observations against the solved global alignment, and the file `tilt`
// reads with --local-file.  Host code: a few thousand residuals per series.
#include <cmath>
#include <cstdio>
#include <limits>

#include "../../include/tiltgpu.h"

namespace tiltgpu_sess {
int fail(int code, const char* fmt, ...);
}
using tiltgpu_sess::fail;

namespace {

// project (apps/tiltalign/src/main.rs:169-184): the bead's position in a view
// under the view's tilt, in-plane rotation, magnification and shift.
void project(const double* bead, const tiltgpu_align_view& vp, double* px, double* py) {
    const double cos_tilt = std::cos(vp.tilt), sin_tilt = std::sin(vp.tilt);
    const double cos_rot = std::cos(vp.rotation), sin_rot = std::sin(vp.rotation);
    const double xp = bead[0] * cos_tilt + bead[2] * sin_tilt;
    const double yp = bead[1];
    *px = vp.mag * (cos_rot * xp - sin_rot * yp) + vp.dx;
    *py = vp.mag * (sin_rot * xp + cos_rot * yp) + vp.dy;
}

// `(x).ceil() as usize`: Rust's saturating float-to-unsigned cast.
size_t ceil_as_usize(double x) {
    const double c = std::ceil(x);
    if (!(c > 0.0)) return 0;  // negative, zero and NaN saturate to 0
    if (c >= 1.8446744073709552e19) return std::numeric_limits<size_t>::max();
    return static_cast<size_t>(c);
}

}  // namespace

extern "C" {

int tiltgpu_local_corrections(const double* obs, const double* beads,
                              const tiltgpu_align_view* views, int32_t n_beads, int32_t n_views,
                              int32_t patch_size, float overlap, double cx, double cy,
                              int32_t skip_edge, double img_w, double img_h, int32_t cap,
                              int32_t* out_view, int32_t* out_px, int32_t* out_py,
                              double* out_dx, double* out_dy) {
    if (n_beads < 0 || n_views < 0 || patch_size < 0 || cap < 0 || skip_edge < 0 ||
        (n_beads > 0 && (!obs || !beads)) || (n_views > 0 && !views) ||
        (cap > 0 && (!out_view || !out_px || !out_py || !out_dx || !out_dy)))
        return fail(TILTGPU_ERR_ARG, "bad arguments");
    // bounding box of the beads in image coordinates (main.rs:880-893)
    double min_x = std::numeric_limits<double>::max(), max_x = std::numeric_limits<double>::lowest();
    double min_y = std::numeric_limits<double>::max(), max_y = std::numeric_limits<double>::lowest();
    for (int b = 0; b < n_beads; ++b) {
        const double bx = beads[3 * b] + cx, by = beads[3 * b + 1] + cy;
        if (bx < min_x) min_x = bx;
        if (bx > max_x) max_x = bx;
        if (by < min_y) min_y = by;
        if (by > max_y) max_y = by;
    }
    const double ps = static_cast<double>(patch_size);
    const double step = ps * (1.0 - static_cast<double>(overlap));
    if (step < 1.0) return 0;
    const size_t npx = ceil_as_usize((max_x - min_x) / step) + 1;
    const size_t npy = ceil_as_usize((max_y - min_y) / step) + 1;
    // the --skip-edge-patches filter (main.rs:1374-1402): the patch centre is
    // approximated from its grid indices as if the grid started at 0
    const double margin = static_cast<double>(skip_edge);
    const double w = img_w > 0.0 ? img_w : cx * 2.0;
    const double h = img_h > 0.0 ? img_h : cy * 2.0;
    int64_t n = 0;
    if (n_beads == 0) return 0;
    for (size_t py = 0; py < npy; ++py) {
        for (size_t px = 0; px < npx; ++px) {
            const double pcx = min_x + static_cast<double>(px) * step + ps / 2.0;
            const double pcy = min_y + static_cast<double>(py) * step + ps / 2.0;
            const double half = ps / 2.0;
            if (margin > 0.0) {
                const double ex = static_cast<double>(px) * step + ps / 2.0;
                const double ey = static_cast<double>(py) * step + ps / 2.0;
                if (!(ex >= margin && ex <= w - margin && ey >= margin && ey <= h - margin)) continue;
            }
            for (int vi = 0; vi < n_views; ++vi) {
                double sum_dx = 0.0, sum_dy = 0.0;
                size_t count = 0;
                for (int bi = 0; bi < n_beads; ++bi) {
                    const double* o = obs + (static_cast<int64_t>(bi) * n_views + vi) * 2;
                    if (std::isnan(o[0])) continue;  // no observation of this bead here
                    const double ix = o[0] + cx, iy = o[1] + cy;
                    if (std::fabs(ix - pcx) <= half && std::fabs(iy - pcy) <= half) {
                        double qx, qy;
                        project(beads + 3 * bi, views[vi], &qx, &qy);
                        sum_dx += o[0] - qx;
                        sum_dy += o[1] - qy;
                        ++count;
                    }
                }
                if (count >= 2) {
                    if (n < cap) {
                        out_view[n] = vi;
                        out_px[n] = static_cast<int32_t>(px);
                        out_py[n] = static_cast<int32_t>(py);
                        out_dx[n] = sum_dx / static_cast<double>(count);
                        out_dy[n] = sum_dy / static_cast<double>(count);
                    }
                    ++n;
                }
            }
        }
    }
    if (n > INT32_MAX) return fail(TILTGPU_ERR_ARG, "too many corrections");
    return static_cast<int>(n);
}

int tiltgpu_write_local_corrections(const char* path, int32_t n, const int32_t* view,
                                    const int32_t* px, const int32_t* py, const double* dx,
                                    const double* dy) {
    if (!path || n < 0 || (n > 0 && (!view || !px || !py || !dx || !dy)))
        return fail(TILTGPU_ERR_ARG, "bad arguments");
    FILE* f = std::fopen(path, "w");
    if (!f) return fail(TILTGPU_ERR_IO, "cannot create %s", path);
    bool ok = std::fprintf(f, "# Local alignment corrections: view patch_x patch_y dx dy\n") > 0;
    for (int32_t i = 0; i < n && ok; ++i)
        ok = std::fprintf(f, "%d %d %d %.4f %.4f\n", view[i], px[i], py[i], dx[i], dy[i]) > 0;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? TILTGPU_OK : fail(TILTGPU_ERR_IO, "write to %s failed", path);
}

}  // extern "C"
