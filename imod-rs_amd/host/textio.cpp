// Text inputs of `tilt`: .tlt angles and per-view files.
//   read_tilt_file      crates/imod-transforms/src/tilt.rs:6-23
//   load_per_view_file  apps/tilt/src/main.rs:310-326
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/tiltgpu.h"
#include "textio.h"

namespace tiltgpu_host {

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) ++a;
    while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) --b;
    return s.substr(a, b - a);
}

// str::parse::<f32>: an optional sign, decimal digits with an optional point and
// exponent, or inf / infinity / nan; nothing else (no hex, no spaces).
bool parse_f32(const std::string& t, float* out) {
    if (t.empty()) return false;
    size_t i = 0;
    if (t[i] == '+' || t[i] == '-') ++i;
    std::string rest = t.substr(i);
    std::string low;
    for (char c : rest) low += static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    if (low == "inf" || low == "infinity" || low == "nan") {
        float v = low == "nan" ? NAN : INFINITY;
        *out = (t[0] == '-') ? -v : v;
        return true;
    }
    bool digits = false, point = false;
    size_t j = 0;
    for (; j < rest.size(); ++j) {
        const char c = rest[j];
        if (std::isdigit(static_cast<unsigned char>(c))) {
            digits = true;
        } else if (c == '.' && !point) {
            point = true;
        } else {
            break;
        }
    }
    if (!digits) return false;
    if (j < rest.size()) {
        if (rest[j] != 'e' && rest[j] != 'E') return false;
        ++j;
        if (j < rest.size() && (rest[j] == '+' || rest[j] == '-')) ++j;
        bool ed = false;
        for (; j < rest.size(); ++j) {
            if (!std::isdigit(static_cast<unsigned char>(rest[j]))) return false;
            ed = true;
        }
        if (!ed) return false;
    }
    errno = 0;
    *out = std::strtof(t.c_str(), nullptr);
    return true;
}

bool parse_usize(const std::string& t, long long* out) {
    if (t.empty()) return false;
    size_t i = (t[0] == '+') ? 1 : 0;
    if (i == t.size()) return false;
    for (size_t j = i; j < t.size(); ++j)
        if (!std::isdigit(static_cast<unsigned char>(t[j]))) return false;
    *out = std::strtoll(t.c_str() + i, nullptr, 10);
    return true;
}

}  // namespace tiltgpu_host

using namespace tiltgpu_host;

extern "C" int tiltgpu_read_tilt_file(const char* path, float* out, int32_t cap) {
    std::ifstream f(path);
    if (!f) return TILTGPU_ERR_IO;
    std::string line;
    int n = 0;
    while (std::getline(f, line)) {
        const std::string t = trim(line);
        if (t.empty()) continue;
        float v;
        if (!parse_f32(t, &v)) continue;  // unparsable lines are skipped (tilt.rs:17-19)
        if (n < cap && out) out[n] = v;
        ++n;
    }
    return n;
}

extern "C" int tiltgpu_read_per_view_file(const char* path, float* out, int32_t cap) {
    std::ifstream f(path);
    if (!f) return TILTGPU_ERR_IO;
    std::string line;
    int n = 0;
    while (std::getline(f, line)) {
        const std::string t = trim(line);
        if (t.empty() || t[0] == '#') continue;
        float v;
        if (!parse_f32(t, &v)) return TILTGPU_ERR_IO;  // main.rs:320-322 exits on a parse error
        if (n < cap && out) out[n] = v;
        ++n;
    }
    return n;
}

// read_xf_file, crates/imod-transforms/src/xf.rs:147-174: blank lines skipped;
// of each line the whitespace-separated tokens that parse as f32 are kept, and
// a line with at least six of them is one transform (a11 a12 a21 a22 dx dy).
extern "C" int tiltgpu_read_xf_file(const char* path, float* out, int32_t cap) {
    std::ifstream f(path);
    if (!f) return TILTGPU_ERR_IO;
    std::string line;
    int n = 0;
    while (std::getline(f, line)) {
        const std::string t = trim(line);
        if (t.empty()) continue;
        float vals[6];
        int k = 0;
        size_t i = 0;
        while (i < t.size() && k < 6) {
            while (i < t.size() && std::isspace(static_cast<unsigned char>(t[i]))) ++i;
            size_t j = i;
            while (j < t.size() && !std::isspace(static_cast<unsigned char>(t[j]))) ++j;
            if (j > i) {
                float v;
                if (parse_f32(t.substr(i, j - i), &v)) vals[k++] = v;
            }
            i = j;
        }
        if (k < 6) continue;
        if (n < cap && out)
            for (int c = 0; c < 6; ++c) out[6 * n + c] = vals[c];
        ++n;
    }
    return n;
}

// ------------------------------------------------------------ defocus files
// read_defocus_file, apps/ctfphaseflip/src/main.rs:103-157: '#' and blank lines
// skipped; "view defocus1 [defocus2 [astig]]" (nm, nm, degrees; one value means
// defocus2 = defocus1, three mean astig = 0).  A malformed line makes the whole
// file an error (the reference exits), reported as -(line number) - 100.
extern "C" int tiltgpu_read_defocus_file(const char* path, int32_t* views, float* entries,
                                         int32_t cap) {
    std::ifstream f(path);
    if (!f) return TILTGPU_ERR_IO;
    std::string line;
    int n = 0, ln = 0;
    while (std::getline(f, line)) {
        ++ln;
        const std::string t = trim(line);
        if (t.empty() || t[0] == '#') continue;
        std::vector<std::string> fl;
        size_t i = 0;
        while (i < t.size()) {
            while (i < t.size() && std::isspace(static_cast<unsigned char>(t[i]))) ++i;
            size_t j = i;
            while (j < t.size() && !std::isspace(static_cast<unsigned char>(t[j]))) ++j;
            if (j > i) fl.push_back(t.substr(i, j - i));
            i = j;
        }
        long long view;
        float d1, d2, aa = 0.0f;
        if (fl.size() < 2 || !parse_usize(fl[0], &view) || !parse_f32(fl[1], &d1))
            return -100 - ln;
        d2 = d1;
        if (fl.size() >= 3 && !parse_f32(fl[2], &d2)) return -100 - ln;
        if (fl.size() >= 4 && !parse_f32(fl[3], &aa)) return -100 - ln;
        if (n < cap && views && entries) {
            views[n] = static_cast<int32_t>(std::min<long long>(view, INT32_MAX));
            entries[3 * n] = d1;
            entries[3 * n + 1] = d2;
            entries[3 * n + 2] = aa;
        }
        ++n;
    }
    return n;
}

// interpolate_defocus, apps/ctfphaseflip/src/main.rs:164-198: listed views
// (< n_views) take their entry (a later line for the same view wins); a gap
// takes the f32 linear interpolation between the nearest listed views around
// it, or the nearest one where it lies beyond the ends.  Returns the row count
// written to out [n_views][3] (0 when nothing is listed).
extern "C" int tiltgpu_interpolate_defocus(const int32_t* views, const float* entries, int32_t n,
                                           int32_t n_views, float* out) {
    if (n < 0 || n_views < 0 || (n > 0 && (!views || !entries)) || (n_views > 0 && !out))
        return TILTGPU_ERR_ARG;
    if (n == 0 || n_views == 0) return 0;
    std::vector<char> have(n_views, 0);
    for (int k = 0; k < 3 * n_views; ++k) out[k] = 0.0f;
    for (int e = 0; e < n; ++e) {
        const int v = views[e];
        if (v < 0 || v >= n_views) continue;
        for (int c = 0; c < 3; ++c) out[3 * v + c] = entries[3 * e + c];
        have[v] = 1;
    }
    for (int v = 0; v < n_views; ++v) {
        if (have[v]) continue;
        int b = -1, a = -1;
        for (int i = v - 1; i >= 0; --i)
            if (have[i]) {
                b = i;
                break;
            }
        for (int i = v + 1; i < n_views; ++i)
            if (have[i]) {
                a = i;
                break;
            }
        if (b >= 0 && a >= 0) {
            const float t = static_cast<float>(v - b) / static_cast<float>(a - b);
            for (int c = 0; c < 3; ++c) {
                const float lo = out[3 * b + c], hi = out[3 * a + c];
                const float d = hi - lo;
                const float td = t * d;
                out[3 * v + c] = lo + td;
            }
        } else if (b >= 0 || a >= 0) {
            const int src = b >= 0 ? b : a;
            for (int c = 0; c < 3; ++c) out[3 * v + c] = out[3 * src + c];
        }
    }
    return n_views;
}

// ------------------------------------------------------------ local alignment
// load_local_alignment, apps/tilt/src/main.rs:232-307: lines "view x y dx dy"
// ('#' and blank lines skipped, short lines skipped, each field unwrap_or(0)),
// grouped by view < nz into grids of sorted unique xs / ys (a position within
// 0.01 of the last kept one merges), dx/dy row-major [iy*nx + ix].
namespace tiltgpu_host {

struct LocalPatch {
    long long view;
    float x, y, dx, dy;
};

static std::vector<float> sorted_unique(std::vector<float> v) {
    std::sort(v.begin(), v.end());
    std::vector<float> out;
    for (float a : v)
        if (out.empty() || !(std::fabs(a - out.back()) < 0.01f)) out.push_back(a);
    return out;
}

bool load_local_grids(const std::string& path, int nz, LocalGrids* g, std::string* err) {
    std::ifstream f(path);
    if (!f) {
        *err = "cannot open " + path;
        return false;
    }
    std::vector<LocalPatch> patches;
    std::string line;
    while (std::getline(f, line)) {
        const std::string t = trim(line);
        if (t.empty() || t[0] == '#') continue;
        std::vector<std::string> parts;
        size_t i = 0;
        while (i < t.size()) {
            while (i < t.size() && std::isspace(static_cast<unsigned char>(t[i]))) ++i;
            size_t j = i;
            while (j < t.size() && !std::isspace(static_cast<unsigned char>(t[j]))) ++j;
            if (j > i) parts.push_back(t.substr(i, j - i));
            i = j;
        }
        if (parts.size() < 5) continue;
        LocalPatch p{};
        if (!parse_usize(parts[0], &p.view)) p.view = 0;
        float* dst[4] = {&p.x, &p.y, &p.dx, &p.dy};
        for (int k = 0; k < 4; ++k)
            if (!parse_f32(parts[k + 1], dst[k])) *dst[k] = 0.f;
        patches.push_back(p);
    }
    g->nx.assign(nz, 0);
    g->ny.assign(nz, 0);
    g->xs.clear();
    g->ys.clear();
    g->dx.clear();
    g->dy.clear();
    std::vector<std::vector<const LocalPatch*>> by_view(nz);
    for (const auto& p : patches)
        if (p.view >= 0 && p.view < nz) by_view[p.view].push_back(&p);
    for (int v = 0; v < nz; ++v) {
        if (by_view[v].empty()) continue;
        std::vector<float> xs, ys;
        for (auto* p : by_view[v]) {
            xs.push_back(p->x);
            ys.push_back(p->y);
        }
        xs = sorted_unique(xs);
        ys = sorted_unique(ys);
        const int nx = static_cast<int>(xs.size()), ny = static_cast<int>(ys.size());
        std::vector<float> gdx(static_cast<size_t>(nx) * ny, 0.f), gdy(gdx.size(), 0.f);
        for (auto* p : by_view[v]) {
            int ix = 0, iy = 0;
            for (int k = 0; k < nx; ++k)
                if (std::fabs(xs[k] - p->x) < 0.01f) {
                    ix = k;
                    break;
                }
            for (int k = 0; k < ny; ++k)
                if (std::fabs(ys[k] - p->y) < 0.01f) {
                    iy = k;
                    break;
                }
            gdx[static_cast<size_t>(iy) * nx + ix] = p->dx;
            gdy[static_cast<size_t>(iy) * nx + ix] = p->dy;
        }
        g->nx[v] = nx;
        g->ny[v] = ny;
        g->xs.insert(g->xs.end(), xs.begin(), xs.end());
        g->ys.insert(g->ys.end(), ys.begin(), ys.end());
        g->dx.insert(g->dx.end(), gdx.begin(), gdx.end());
        g->dy.insert(g->dy.end(), gdy.begin(), gdy.end());
    }
    return true;
}

}  // namespace tiltgpu_host

extern "C" int tiltgpu_parse_local_file(const char* path, int32_t nz, int32_t* sizes, int32_t* nx,
                                        int32_t* ny, float* xs, float* ys, float* dx, float* dy) {
    if (!path || nz < 0 || !sizes) return TILTGPU_ERR_ARG;
    tiltgpu_host::LocalGrids g;
    std::string err;
    if (!tiltgpu_host::load_local_grids(path, nz, &g, &err)) return TILTGPU_ERR_IO;
    sizes[0] = static_cast<int32_t>(g.xs.size());
    sizes[1] = static_cast<int32_t>(g.ys.size());
    sizes[2] = static_cast<int32_t>(g.dx.size());
    if (!nx) return TILTGPU_OK;
    if (!ny || (sizes[0] && !xs) || (sizes[1] && !ys) || (sizes[2] && (!dx || !dy)))
        return TILTGPU_ERR_ARG;
    std::copy(g.nx.begin(), g.nx.end(), nx);
    std::copy(g.ny.begin(), g.ny.end(), ny);
    std::copy(g.xs.begin(), g.xs.end(), xs);
    std::copy(g.ys.begin(), g.ys.end(), ys);
    std::copy(g.dx.begin(), g.dx.end(), dx);
    std::copy(g.dy.begin(), g.dy.end(), dy);
    return TILTGPU_OK;
}
