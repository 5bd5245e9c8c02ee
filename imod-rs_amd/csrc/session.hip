// libtiltgpu session: the C ABI of include/tiltgpu.h over the gfx950 kernels.
//
// A session holds one tilt series and its volume resident in HBM and runs the
// pipeline of imod-rs apps/tilt/src/main.rs:601-1090 on one HIP stream:
//   log + per-view means -> R-weighting + cosine stretch (+ FP16/BF16 storage)
//   -> WBP (output transform fused)  or  SIRT: FP -> (O - R) filter -> BP += relax
//   -> fake-SIRT low-pass -> output transform.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "session_internal.h"
#include "tuning.h"

namespace tiltgpu_sess {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

bool forced_alloc_failure() {
    static std::mutex m;
    static std::string seen;
    static long count = 0;
    const char* e = std::getenv("TILTGPU_FAIL_ALLOC_AT");
    std::lock_guard<std::mutex> lk(m);
    if (!e || !*e) {
        seen.clear();
        return false;
    }
    if (seen != e) {  // a new setting restarts the count
        seen = e;
        count = 0;
    }
    return ++count == std::atol(e);
}

}  // namespace tiltgpu_sess

using namespace tiltgpu_sess;

namespace tiltgpu_sess {

float f32_proj_cos(float deg) {
    const float rad = deg * static_cast<float>(M_PI) / 180.0f;
    return cosf(rad);
}

// 1/cos(theta) for the cosine stretch, 0 when the view is skipped
// (main.rs:449-459).
float stretch_factor(float deg, float max_tilt) {
    if (std::fabs(deg) > max_tilt) return 0.f;
    const float ct = f32_proj_cos(deg);
    if (std::fabs(ct) < 1e-6f) return 0.f;
    return 1.0f / ct;
}

int pick_pairs_per_wg(int n) { return std::max(1, 2048 / std::max(1, n)); }

// Upper bound on the LDS entries one backprojection tile needs for a view whose
// local shifts span ldx_span x ldy_span pixels.
void tile_bounds(const tiltgpu_view_params& g, int tile_z, int tile_x, double ldx_span,
                 double ldy_span, int* seg, int* rows) {
    const double sx = std::fabs(static_cast<double>(g.cos_t) * g.cos_xt) * (tile_x - 1) +
                      std::fabs(static_cast<double>(g.sin_t) * g.zfactor) * (tile_z - 1) +
                      ldx_span + 4.0;
    const double sy = std::fabs(static_cast<double>(g.sin_xt)) * (tile_x - 1) + ldy_span + 3.0;
    *seg = static_cast<int>(std::min(sx, 1e9));
    *rows = static_cast<int>(std::min(sy, 1e9));
}

// Extent of a view's local-alignment shifts (0 without a grid).
static void local_spans(const tiltgpu_session* s, int view, double* ldx_span, double* ldy_span) {
    *ldx_span = *ldy_span = 0;
    if (!s || !s->has_local || s->lnx[view] <= 0 || s->lny[view] <= 0) return;
    const int ng = s->lnx[view] * s->lny[view];
    const float* gx = s->ldx.data() + s->lgoff[view];
    const float* gy = s->ldy.data() + s->lgoff[view];
    float xmn = gx[0], xmx = gx[0], ymn = gy[0], ymx = gy[0];
    for (int i = 1; i < ng; ++i) {
        xmn = std::min(xmn, gx[i]);
        xmx = std::max(xmx, gx[i]);
        ymn = std::min(ymn, gy[i]);
        ymx = std::max(ymx, gy[i]);
    }
    *ldx_span = static_cast<double>(xmx) - xmn;
    *ldy_span = static_cast<double>(ymx) - ymn;
}

bool needs_two_rows(const tiltgpu_view_params* vp, int nv, bool has_local) {
    // (sin_xt == 0 implies cos_xt == 1 exactly; the second test keeps the
    // single-row kernels, which rely on it, safe for any table)
    bool two = has_local;
    for (int v = 0; v < nv; ++v)
        if (vp[v].sin_xt != 0.0f || vp[v].cos_xt != 1.0f) two = true;
    return two;
}

int choose_bp_tile(const tiltgpu_session* local_src, const tiltgpu_view_params* vp, int nv,
                   bool two_row, bool has_local, bool half, int in_nx, int out_nz, int forced,
                   BpTileChoice* out, int forced_vps, int forced_xpl, bool allow_shapes) {
    // Candidates in order: the measured shape (tile depth, views per stage, x per
    // lane), then the depths by padded depth x (1 + per-view overhead / voxels per
    // lane) at shape (1, 1); the first whose per-view LDS bound fits is taken.
    const int max_lds_bytes = 64 * 1024;
    struct Cand {
        int tz, vps, xpl;
    };
    Cand order[5];
    int norder = 0;
    int depths[4] = {256, 128, 64, 32};
    auto cost = [&](int tz) {
        const double pad = static_cast<double>((out_nz + tz - 1) / tz * tz);
        return pad * (1.0 + 32.0 / tz);
    };
    for (int i = 1; i < 4; ++i)  // stable insertion sort by cost (no temporary buffer)
        for (int j = i; j > 0 && cost(depths[j]) < cost(depths[j - 1]); --j)
            std::swap(depths[j], depths[j - 1]);
    out->source = kTuneHeuristic;
    if (!allow_shapes) forced_vps = forced_xpl = 1;
    int t_vps = 1, t_xpl = 1;
    const int tuned = allow_shapes
                          ? tuned_bp_shape(two_row ? 1 : 0, has_local ? 1 : 0, half ? 1 : 0, in_nx,
                                           out_nz, &t_vps, &t_xpl)
                          : tuned_tile_z(two_row ? 1 : 0, has_local ? 1 : 0, half ? 1 : 0, in_nx,
                                         out_nz);
    const int fv = forced_vps > 0 ? forced_vps : 0, fx = forced_xpl > 0 ? forced_xpl : 0;
    if (forced > 0) {
        order[norder++] = {forced, fv ? fv : 1, fx ? fx : 1};
        out->source = kTuneForced;
    } else {
        if (tuned > 0) {
            order[norder++] = {tuned, fv ? fv : t_vps, fx ? fx : t_xpl};
            out->source = kTuneTable;
        }
        for (int i = 0; i < 4; ++i) order[norder++] = {depths[i], fv ? fv : 1, fx ? fx : 1};
        if (fv || fx) out->source = kTuneForced;
    }
    for (int i = 0; i < norder; ++i) {
        const Cand c = order[i];
        if (!bp_shape_supported(c.tz, c.vps, c.xpl)) continue;
        int cap = 1;
        for (int v = 0; v < nv; ++v) {
            int seg, rows;
            double lx, ly;
            local_spans(has_local ? local_src : nullptr, v, &lx, &ly);
            tile_bounds(vp[v], c.tz, kBpTileX * c.xpl, lx, ly, &seg, &rows);
            const int need = two_row ? seg * rows : seg;
            cap = std::max(cap, need);
        }
        cap = (cap + 63) / 64 * 64;
        if (bp_lds_bytes(cap, two_row, c.vps) <= static_cast<size_t>(max_lds_bytes)) {
            out->tile_z = c.tz;
            out->lds_cap = cap;
            out->vps = c.vps;
            out->xpl = c.xpl;
            return TILTGPU_OK;
        }
    }
    if (forced > 0 || fv || fx)
        return fail(TILTGPU_ERR_ARG,
                    "launch shape tile depth %d, %d views per stage, %d x per lane is not "
                    "built or does not fit the LDS for this geometry",
                    forced > 0 ? forced : 0, fv ? fv : 1, fx ? fx : 1);
    return fail(TILTGPU_ERR_ARG,
                "geometry needs more LDS per backprojection tile than supported "
                "(local-alignment shifts vary too much within 64 pixels)");
}

int setup_geometry(tiltgpu_session* s, const std::vector<tiltgpu_view_params>& vp) {
    const int nv = static_cast<int>(vp.size());
    s->nviews = nv;
    s->two_row = needs_two_rows(vp.data(), nv, s->has_local);
    const int forced = s->tune_tile_z > 0 ? s->tune_tile_z : s->tile_z_override;
    BpTileChoice tc{};
    int rc0 = choose_bp_tile(s, vp.data(), nv, s->two_row, s->has_local,
                             s->o.storage != TILTGPU_STORE_F32, s->d.in_nx, s->d.out_nz, forced,
                             &tc, s->tune_vps, s->tune_xpl, !s->fixed_shape);
    if (rc0) return rc0;
    s->tile_z = tc.tile_z;
    s->bp_vps = tc.vps;
    s->bp_xpl = tc.xpl;
    s->lds_cap = tc.lds_cap;
    s->tile_source = tc.source;
    const int out_nz = s->d.out_nz;
    s->nz_pad = (out_nz + s->tile_z - 1) / s->tile_z * s->tile_z;
    std::vector<float4> g4(std::max(nv, 1));
    std::vector<BpView> bv(std::max(nv, 1));
    std::vector<float> zc(static_cast<size_t>(std::max(nv, 1)) * s->nz_pad);
    const float center_z = static_cast<float>(out_nz) / 2.0f;
    for (int v = 0; v < nv; ++v) {
        g4[v] = make_float4(vp[v].cos_t, vp[v].sin_t, vp[v].cos_xt, vp[v].sin_xt);
        BpView& b = bv[v];
        std::memset(&b, 0, sizeof b);
        b.cos_t = vp[v].cos_t;
        b.sin_t = vp[v].sin_t;
        b.cos_xt = vp[v].cos_xt;
        b.sin_xt = vp[v].sin_xt;
        b.shift_mode = 0;
        if (s->has_local && s->lnx[v] > 0) {
            if (s->lnx[v] == 1 && s->lny[v] == 1) {
                // ViewPatches::interpolate of a one-patch grid returns that patch
                b.shift_mode = 1;
                b.sdx = s->ldx[s->lgoff[v]];
                b.sdy = s->ldy[s->lgoff[v]];
            } else {
                b.shift_mode = 2;
            }
        }
        for (int oz = 0; oz < s->nz_pad; ++oz) {
            const float dz = static_cast<float>(oz) - center_z;
            // main.rs:1015-1016: z_contrib = dz * sin_t * zfactor
            zc[static_cast<size_t>(v) * s->nz_pad + oz] = dz * vp[v].sin_t * vp[v].zfactor;
        }
    }
    dev_free(s, s->vgeom);
    dev_free(s, s->views);
    dev_free(s, s->zc);
    int rc;
    if ((rc = dev_alloc(s, &s->vgeom, g4.size())) || (rc = dev_alloc(s, &s->views, bv.size())) ||
        (rc = dev_alloc(s, &s->zc, zc.size())))
        return rc;
    HIP_TRY(hipMemcpy(s->vgeom, g4.data(), g4.size() * sizeof(float4), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->views, bv.data(), bv.size() * sizeof(BpView), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->zc, zc.data(), zc.size() * sizeof(float), hipMemcpyHostToDevice));
    s->geometry_ready = true;
    return TILTGPU_OK;
}

int upload_local(tiltgpu_session* s) {
    if (!s->has_local) return TILTGPU_OK;
    int rc;
    auto up_i = [&](int** dp, const std::vector<int>& h) -> int {
        dev_free(s, *dp);
        int r = dev_alloc(s, dp, h.size());
        if (r) return r;
        if (!h.empty()) HIP_TRY(hipMemcpy(*dp, h.data(), h.size() * 4, hipMemcpyHostToDevice));
        return TILTGPU_OK;
    };
    auto up_f = [&](float** dp, const std::vector<float>& h) -> int {
        dev_free(s, *dp);
        int r = dev_alloc(s, dp, h.size());
        if (r) return r;
        if (!h.empty()) HIP_TRY(hipMemcpy(*dp, h.data(), h.size() * 4, hipMemcpyHostToDevice));
        return TILTGPU_OK;
    };
    if ((rc = up_i(&s->d_lnx, s->lnx)) || (rc = up_i(&s->d_lny, s->lny)) ||
        (rc = up_i(&s->d_lxoff, s->lxoff)) || (rc = up_i(&s->d_lyoff, s->lyoff)) ||
        (rc = up_i(&s->d_lgoff, s->lgoff)) || (rc = up_f(&s->d_lxs, s->lxs)) ||
        (rc = up_f(&s->d_lys, s->lys)) || (rc = up_f(&s->d_ldx, s->ldx)) ||
        (rc = up_f(&s->d_ldy, s->ldy)))
        return rc;
    return TILTGPU_OK;
}

// The view tables and stretch factors in FP64 on the device, rounded once to
// FP32 (TILTGPU_GEOMETRY_F64).  Leaves the angles and z factors on the device
// for the z-term table (dev_angles, dev_zfac).
static int geometry_f64_tables(tiltgpu_session* s, int nv, std::vector<tiltgpu_view_params>& vp,
                               std::vector<float>& st, float** dev_angles, float** dev_zfac) {
    const int nz = s->d.nz;
    const int na = static_cast<int>(s->angles.size());
    const int nxt = s->has_xtilt ? static_cast<int>(s->xtilt.size()) : 0;
    const int nzf = s->has_zfac ? static_cast<int>(s->zfac.size()) : 0;
    float *d_xt = nullptr, *d_view = nullptr, *d_st = nullptr;
    int rc;
    if ((rc = dev_alloc(s, dev_angles, std::max(na, 1))) ||
        (rc = dev_alloc(s, &d_xt, std::max(nxt, 1))) ||
        (rc = dev_alloc(s, dev_zfac, std::max(nzf, 1))) ||
        (rc = dev_alloc(s, &d_view, 5 * std::max(nv, 1))) ||
        (rc = dev_alloc(s, &d_st, std::max(nz, 1))))
        return rc;
    HIP_TRY(hipMemcpy(*dev_angles, s->angles.data(), na * 4, hipMemcpyHostToDevice));
    if (nxt) HIP_TRY(hipMemcpy(d_xt, s->xtilt.data(), nxt * 4, hipMemcpyHostToDevice));
    if (nzf) HIP_TRY(hipMemcpy(*dev_zfac, s->zfac.data(), nzf * 4, hipMemcpyHostToDevice));
    HIP_TRY(launch_geometry_f64(*dev_angles, na, nxt ? d_xt : nullptr, nxt, nzf ? *dev_zfac : nullptr,
                                nzf, nv, nz, s->o.max_cosine_tilt, s->o.no_cosine_stretch ? 0 : 1,
                                d_view, d_st, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    static_assert(sizeof(tiltgpu_view_params) == 5 * sizeof(float), "view params layout");
    HIP_TRY(hipMemcpy(vp.data(), d_view, nv * sizeof(tiltgpu_view_params), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(st.data(), d_st, nz * 4, hipMemcpyDeviceToHost));
    dev_free(s, d_xt);
    dev_free(s, d_view);
    dev_free(s, d_st);
    return TILTGPU_OK;
}

// Geometry of the full pipeline from the angle / xtilt / zfactor inputs.
int prepare_pipeline_geometry(tiltgpu_session* s) {
    if (!s->has_angles) return fail(TILTGPU_ERR_STATE, "tilt angles not set");
    const int nz = s->d.nz;
    const int nv = std::min(nz, static_cast<int>(s->angles.size()));
    if (!s->o.no_cosine_stretch && static_cast<int>(s->angles.size()) < nz)
        return fail(TILTGPU_ERR_ARG,
                    "cosine stretch needs one tilt angle per section (%zu angles, %d "
                    "sections); the reference panics here (main.rs:450)",
                    s->angles.size(), nz);
    std::vector<tiltgpu_view_params> vp(std::max(nv, 1));
    // stretch factors for every section (main.rs:449 indexes tilt_angles[p], p < nz)
    std::vector<float> st(std::max(nz, 1), 0.f);
    const bool f64 = s->o.geometry == TILTGPU_GEOMETRY_F64;
    float *d_ang = nullptr, *d_zf = nullptr;
    int rc;
    if (f64) {
        if ((rc = geometry_f64_tables(s, nv, vp, st, &d_ang, &d_zf))) return rc;
    } else {
        rc = tiltgpu_proj_params(s->angles.data(), nv, s->has_xtilt ? s->xtilt.data() : nullptr,
                                 static_cast<int32_t>(s->xtilt.size()),
                                 s->has_zfac ? s->zfac.data() : nullptr,
                                 static_cast<int32_t>(s->zfac.size()), vp.data());
        if (rc) return rc;
        if (!s->o.no_cosine_stretch)
            for (int p = 0; p < nz; ++p) st[p] = stretch_factor(s->angles[p], s->o.max_cosine_tilt);
    }
    vp.resize(nv);
    s->vp_host = vp;
    rc = setup_geometry(s, vp);
    if (f64) {
        // the z terms in FP64 as well (setup_geometry wrote the FP32 ones)
        hipError_t e = hipSuccess;
        if (rc == TILTGPU_OK) {
            e = launch_zc_f64(d_ang, s->has_zfac ? d_zf : nullptr,
                              s->has_zfac ? static_cast<int>(s->zfac.size()) : 0, nv, s->nz_pad,
                              s->d.out_nz, s->zc, s->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        }
        dev_free(s, d_ang);
        dev_free(s, d_zf);
        if (e != hipSuccess) {
            s->geometry_ready = false;
            return fail(TILTGPU_ERR_HIP, "FP64 z-term table: %s", hipGetErrorString(e));
        }
    }
    if (rc) return rc;
    if ((rc = upload_local(s))) return rc;
    dev_free(s, s->stretch);
    if ((rc = dev_alloc(s, &s->stretch, st.size()))) return rc;
    HIP_TRY(hipMemcpy(s->stretch, st.data(), st.size() * 4, hipMemcpyHostToDevice));
    return TILTGPU_OK;
}

int prepare_filter(tiltgpu_session* s) {
    const int n = s->d.in_nx;
    if (filter_plan(n, s->radix, &s->nradix) != 0)
        return fail(TILTGPU_ERR_ARG, "projection width %d is not supported by the LDS FFT", n);
    size_t lds_limit = kFilterMaxLds;
    int optin = 0;
    if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, s->device) ==
            hipSuccess &&
        optin > 0)
        lds_limit = std::min(lds_limit, static_cast<size_t>(optin));
    if (filter_lds_bytes(n, pick_pairs_per_wg(n)) > lds_limit)
        return fail(TILTGPU_ERR_ARG,
                    "projection width %d too large for the LDS FFT (its rows need %zu bytes of "
                    "LDS per workgroup, the device allows %zu)",
                    n, filter_lds_bytes(n, pick_pairs_per_wg(n)), lds_limit);
    std::vector<float> w(n);
    int rc = tiltgpu_radial_weights(n, s->o.exact_filter_size, s->o.radial_cutoff,
                                    s->o.radial_falloff, s->o.hamming, w.data());
    if (rc) return rc;
    std::vector<float2> tw(n);
    for (int t = 0; t < n; ++t) {
        const double a = -2.0 * M_PI * static_cast<double>(t) / static_cast<double>(n);
        tw[t] = make_float2(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a)));
    }
    dev_free(s, s->weights);
    dev_free(s, s->tw);
    if ((rc = dev_alloc(s, &s->weights, n)) || (rc = dev_alloc(s, &s->tw, n))) return rc;
    HIP_TRY(hipMemcpy(s->weights, w.data(), n * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->tw, tw.data(), n * sizeof(float2), hipMemcpyHostToDevice));
    dev_free(s, s->tw8);
    s->tw8 = nullptr;
    if (const int m = pow2_twiddles(n, nullptr)) {
        std::vector<float2> t8(m);
        pow2_twiddles(n, t8.data());
        if ((rc = dev_alloc(s, &s->tw8, m))) return rc;
        HIP_TRY(hipMemcpy(s->tw8, t8.data(), m * sizeof(float2), hipMemcpyHostToDevice));
    }
    return TILTGPU_OK;
}

int check_kernel_error(tiltgpu_session* s) {
    int e = 0;
    HIP_TRY(hipMemcpy(&e, s->err, sizeof(int), hipMemcpyDeviceToHost));
    if (e) {
        int z = 0;
        HIP_TRY(hipMemcpy(s->err, &z, sizeof(int), hipMemcpyHostToDevice));
        return fail(TILTGPU_ERR_KERNEL,
                    "a backprojection tile needed more LDS entries than its bound (%d)",
                    s->lds_cap);
    }
    return TILTGPU_OK;
}

BpParams make_bp(tiltgpu_session* s, const void* proj, const float* scale, int y0, int y1,
                 int out_mode, float* out) {
    BpParams p{};
    p.proj = proj;
    p.view_scale = scale;
    p.views = s->views;
    p.zc_tab = s->zc;
    p.means = s->means;
    p.out = out;
    p.in_nx = s->d.in_nx;
    p.in_ny = s->d.in_ny;
    p.nviews = s->nviews;
    p.out_nx = s->d.out_nx;
    p.out_nz = s->d.out_nz;
    p.nz_pad = s->nz_pad;
    p.y0 = y0;
    p.y1 = y1;
    const int tile_x = kBpTileX * s->bp_xpl;
    p.ntx = (s->d.out_nx + tile_x - 1) / tile_x;
    p.ntz = s->nz_pad / s->tile_z;
    p.cx = static_cast<float>(s->d.in_nx) / 2.0f;
    p.ocx = static_cast<float>(s->d.out_nx) / 2.0f;
    p.inv_n = 1.0f / static_cast<float>(s->d.nz);
    p.fill_mode = s->o.fill_mode;
    p.out_mode = out_mode;
    p.dw = s->o.densweight;
    p.os = s->o.outscale;
    p.oa = s->o.outadd;
    p.relax = s->o.sirt_relax;
    p.has_local = s->has_local ? 1 : 0;
    p.lnx = s->d_lnx;
    p.lny = s->d_lny;
    p.lxoff = s->d_lxoff;
    p.lyoff = s->d_lyoff;
    p.lgoff = s->d_lgoff;
    p.lxs = s->d_lxs;
    p.lys = s->d_lys;
    p.ldx = s->d_ldx;
    p.ldy = s->d_ldy;
    p.lds_cap = s->lds_cap;
    p.err = s->err;
    return p;
}

int run_bp(tiltgpu_session* s, const void* proj, const float* scale, int y0, int y1,
           int out_mode, float* out) {
    BpParams p = make_bp(s, proj, scale, y0, y1, out_mode, out);
    // the staged segment of every (view, tile) of this launch shape, computed
    // once per launch instead of by the four waves of each workgroup per view
    const size_t nseg = static_cast<size_t>(p.nviews) * p.ntz * p.ntx;
    if (nseg > s->seg_cap) {
        dev_free(s, s->seg_tab);
        s->seg_tab = nullptr;
        s->seg_cap = 0;
        int rc = dev_alloc(s, &s->seg_tab, nseg);
        if (rc) return rc;
        s->seg_cap = nseg;
    }
    Scope sc(s, kClsBp);
    if (nseg > 0) {
        HIP_TRY(launch_bp_segments(p, s->tile_z, s->bp_xpl, s->seg_tab, s->stream));
        p.seg_tab = s->seg_tab;
    }
    HIP_TRY(launch_backproject(p, s->seam ? TILTGPU_STORE_F32 : s->o.storage, s->two_row ? 1 : 0,
                               s->tile_z, s->bp_vps, s->bp_xpl, s->stream));
    s->n_bp++;
    return TILTGPU_OK;
}

// Filter (and optionally stretch) rows of `src` (minus `sub`) into s->filt and,
// for FP16/BF16 storage, convert into s->store.
FilterParams make_filter(tiltgpu_session* s, const float* src, const float* sub, bool stretch) {
    const int n = s->d.in_nx;
    FilterParams f{};
    f.src = src;
    f.src_sub = sub;
    f.dst = s->filt;
    f.weights = s->o.no_weight ? nullptr : s->weights;
    f.tw = s->tw;
    f.stretch = (stretch && !s->o.no_cosine_stretch) ? s->stretch : nullptr;
    f.n = n;
    f.nrows = s->d.nz * s->d.in_ny;
    f.rows_per_view = s->d.in_ny;
    f.pairs_per_wg = pick_pairs_per_wg(n);
    f.cubic = s->o.cosine_cubic;
    f.nradix = s->nradix;
    f.n_div = make_fastdiv(static_cast<uint32_t>(n));
    f.tw8 = s->tw8;
    f.log2n = 0;
    while ((1 << f.log2n) < n) ++f.log2n;
    int ns = 1;
    for (int i = 0; i < s->nradix; ++i) {
        f.radix[i] = s->radix[i];
        f.m_div[i] = make_fastdiv(static_cast<uint32_t>(n / s->radix[i]));
        f.ns_div[i] = make_fastdiv(static_cast<uint32_t>(ns));
        f.r_div[i] = make_fastdiv(static_cast<uint32_t>(s->radix[i]));
        ns *= s->radix[i];
    }
    return f;
}

// Per-workgroup partial buffer for the residual sum of squares of a filter launch.
static int ensure_partial(tiltgpu_session* s, int nblk) {
    if (s->partial_n < nblk) {
        dev_free(s, s->partial);
        int rc = dev_alloc(s, &s->partial, nblk);
        if (rc) return rc;
        s->partial_n = nblk;
    }
    return TILTGPU_OK;
}

int run_filter(tiltgpu_session* s, const float* src, const float* sub, bool stretch,
               double* sq_out, double* row_sum) {
    FilterParams f = make_filter(s, src, sub, stretch);
    f.row_sum = row_sum;
    const int nblk = filter_blocks(f);
    if (sq_out) {
        int rc = ensure_partial(s, nblk);
        if (rc) return rc;
        f.sq_partial = s->partial;
    }
    {
        Scope sc(s, kClsFilter);
        HIP_TRY(launch_filter(f, s->stream));
        if (s->o.storage != TILTGPU_STORE_F32)
            HIP_TRY(launch_convert_storage(s->filt, s->store, s->scale, s->absmax,
                                           static_cast<int64_t>(f.n) * s->d.in_ny, s->d.nz,
                                           s->o.storage, s->stream));
        s->n_filter++;
    }
    if (sq_out) {
        Scope sc(s, kClsOther);
        HIP_TRY(launch_sum_partials(s->partial, nblk, sq_out, 1, s->stream));
    }
    return TILTGPU_OK;
}

FpParams make_fp(tiltgpu_session* s) {
    FpParams p{};
    p.vol = s->vol;
    p.sino = s->sino;
    p.vgeom = s->vgeom;
    p.zc_tab = s->zc;
    p.in_nx = s->d.in_nx;
    p.in_ny = s->d.in_ny;
    p.nviews = s->nviews;
    p.out_nx = s->d.out_nx;
    p.out_nz = s->d.out_nz;
    p.nz_pad = s->nz_pad;
    p.cx = static_cast<float>(s->d.in_nx) / 2.0f;
    p.ocx = static_cast<float>(s->d.out_nx) / 2.0f;
    int rows = 1, slide = 1;
    const int bins = tuned_fp_bins(s->d.in_nx, nullptr, &rows, &slide);
    p.bins_per_lane = s->tune_fp_bins > 0 ? s->tune_fp_bins : bins;
    // volume rows per workgroup: the table's choice for the width, or
    // TILTGPU_FP_ROWS (1 or 2); two rows only for rows up to kFpRows2Nx wide
    if (const char* e = std::getenv("TILTGPU_FP_ROWS")) {
        const int r = std::atoi(e);
        if (r == 1 || r == 2) rows = r;
    }
    p.rows_per_wg = rows == 2 && s->d.out_nx <= kFpRows2Nx ? 2 : 1;
    // lane window map (same bits either way): the table's choice for the width,
    // or TILTGPU_FP_SLIDE (0 fixed, 1 sliding)
    p.slide_windows = slide;
    if (const char* e = std::getenv("TILTGPU_FP_SLIDE")) p.slide_windows = std::atoi(e) != 0;
    return p;
}

// --log and per-view means (main.rs:607-623)
const float* stack_src(const tiltgpu_session* s) { return s->d_xf ? s->aligned : s->raw; }

// The transformed stack from the uploaded one, once per upload or transform set.
int ensure_aligned(tiltgpu_session* s) {
    if (!s->d_xf || s->aligned_ok) return TILTGPU_OK;
    const int64_t nproj = static_cast<int64_t>(s->d.nz) * s->d.in_ny * s->d.in_nx;
    int rc;
    if (!s->aligned && (rc = dev_alloc(s, &s->aligned, nproj))) return rc;
    Scope sc(s, kClsPrep);
    HIP_TRY(launch_apply_xf(s->raw, s->aligned, s->d.in_nx, s->d.in_ny, s->d.nz, s->d_xf,
                            s->xf_fill, s->stream));
    s->aligned_ok = true;
    return TILTGPU_OK;
}

int run_means(tiltgpu_session* s) {
    const int nz = s->d.nz;
    const int64_t nxy = static_cast<int64_t>(s->d.in_nx) * s->d.in_ny;
    const int nblk = static_cast<int>(std::min<int64_t>(64, (nxy + 4095) / 4096));
    int rc = ensure_partial(s, nblk * nz);
    if (rc) return rc;
    Scope sc(s, kClsPrep);
    HIP_TRY(launch_log_means(stack_src(s), s->logd, nxy, nz, s->o.log, s->partial, nblk, s->means,
                             s->stream));
    return TILTGPU_OK;
}

int run_forward(tiltgpu_session* s) {
    Scope sc(s, kClsFp);
    const int64_t total = static_cast<int64_t>(s->d.nz) * s->d.in_ny * s->d.in_nx;
    if (s->two_row) {
        if (!s->fp_acc) {
            int rc = dev_alloc(s, &s->fp_acc, static_cast<size_t>(total));
            if (rc) return rc;
            if ((rc = dev_alloc(s, &s->fp_fixed, 2))) return rc;
        }
        FpScatterParams p{};
        p.nz_sections = s->d.nz;
        p.acc = s->fp_acc;
        p.scale_exp = reinterpret_cast<const int*>(s->fp_fixed + 1);
        p.vol = s->vol;
        p.sino = s->sino;
        p.vgeom = s->vgeom;
        p.zc_tab = s->zc;
        p.in_nx = s->d.in_nx;
        p.in_ny = s->d.in_ny;
        p.nviews = s->nviews;
        p.out_nx = s->d.out_nx;
        p.out_nz = s->d.out_nz;
        p.nz_pad = s->nz_pad;
        p.cx = static_cast<float>(s->d.in_nx) / 2.0f;
        p.ocx = static_cast<float>(s->d.out_nx) / 2.0f;
        p.has_local = s->has_local ? 1 : 0;
        p.lnx = s->d_lnx;
        p.lny = s->d_lny;
        p.lxoff = s->d_lxoff;
        p.lyoff = s->d_lyoff;
        p.lgoff = s->d_lgoff;
        p.lxs = s->d_lxs;
        p.lys = s->d_lys;
        p.ldx = s->d_ldx;
        p.ldy = s->d_ldy;
        HIP_TRY(launch_forward_scatter(p, s->fp_fixed, reinterpret_cast<int*>(s->fp_fixed + 1),
                                       s->stream));
    } else {
        // views >= nviews are never projected (main.rs:835 iterates proj_params)
        if (s->nviews < s->d.nz) {
            const int64_t per = static_cast<int64_t>(s->d.in_ny) * s->d.in_nx;
            HIP_TRY(launch_fill_f32(s->sino + s->nviews * per, (s->d.nz - s->nviews) * per, 0.f,
                                    s->stream));
        }
        HIP_TRY(launch_forward_gather(make_fp(s), s->stream));
    }
    s->n_fp++;
    return TILTGPU_OK;
}

int validate_dims(const tiltgpu_dims* d) {
    if (!d) return fail(TILTGPU_ERR_ARG, "dims is NULL");
    if (d->in_nx < 2 || d->in_ny < 1 || d->nz < 1 || d->out_nx < 1 || d->out_nz < 1)
        return fail(TILTGPU_ERR_ARG, "invalid dimensions in=%dx%dx%d out_nx=%d out_nz=%d",
                    d->in_nx, d->in_ny, d->nz, d->out_nx, d->out_nz);
    return TILTGPU_OK;
}

int init_device(tiltgpu_session* s, hipStream_t external) {
    if (const char* tz = std::getenv("TILTGPU_TILE_Z")) {
        const int v = std::atoi(tz);
        if (v == 32 || v == 64 || v == 128 || v == 256) s->tile_z_override = v;
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        return fail(TILTGPU_ERR_NODEV, "no GPU visible to HIP");
    HIP_TRY(hipGetDevice(&s->device));
    if (external) {
        s->stream = external;
        s->own_stream = false;
    } else {
        HIP_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    }
    int rc;
    if ((rc = dev_alloc(s, &s->err, 1))) return rc;
    HIP_TRY(hipMemset(s->err, 0, sizeof(int)));
    return TILTGPU_OK;
}

}  // namespace tiltgpu_sess

namespace tiltgpu_sess {

int session_create(const tiltgpu_dims* dims, const tiltgpu_options* opts, hipStream_t stream,
                   tiltgpu_session** out, BufferPool* pool) {
    if (!out) return fail(TILTGPU_ERR_ARG, "out is NULL");
    *out = nullptr;
    int rc = validate_dims(dims);
    if (rc) return rc;
    auto* s = new tiltgpu_session();
    s->pool = pool;
    s->timing = pool == nullptr;
    s->d = *dims;
    if (opts)
        s->o = *opts;
    else
        tiltgpu_default_options(&s->o);
    if (s->o.storage < 0 || s->o.storage > 2) {
        delete s;
        return fail(TILTGPU_ERR_ARG, "unknown storage %d", s->o.storage);
    }
    if (s->o.geometry != TILTGPU_GEOMETRY_F32 && s->o.geometry != TILTGPU_GEOMETRY_F64) {
        delete s;
        return fail(TILTGPU_ERR_ARG, "unknown geometry precision %d", s->o.geometry);
    }
    if (s->o.fill_mode < 0 || s->o.fill_mode > 2) s->o.fill_mode = TILTGPU_FILL_MEAN;
    if ((rc = init_device(s, stream)) || (rc = prepare_filter(s))) {
        tiltgpu_destroy(s);
        return rc;
    }
    const int64_t nproj = static_cast<int64_t>(s->d.nz) * s->d.in_ny * s->d.in_nx;
    const int64_t nvol = static_cast<int64_t>(s->d.in_ny) * s->d.out_nz * s->d.out_nx;
    if ((rc = dev_alloc(s, &s->raw, nproj)) || (rc = dev_alloc(s, &s->filt, nproj)) ||
        (rc = dev_alloc(s, &s->vol, nvol)) || (rc = dev_alloc(s, &s->means, s->d.nz)) ||
        (rc = dev_alloc(s, &s->scale, s->d.nz)) || (rc = dev_alloc(s, &s->absmax, s->d.nz))) {
        tiltgpu_destroy(s);
        return rc;
    }
    if (s->o.log && (rc = dev_alloc(s, &s->logd, nproj))) {
        tiltgpu_destroy(s);
        return rc;
    }
    if (s->o.storage != TILTGPU_STORE_F32) {
        unsigned short* st = nullptr;
        if ((rc = dev_alloc(s, &st, nproj))) {
            tiltgpu_destroy(s);
            return rc;
        }
        s->store = st;
    }
    if (s->o.sirt > 0) {
        if ((rc = dev_alloc(s, &s->sino, nproj)) || (rc = dev_alloc(s, &s->resid, s->o.sirt))) {
            tiltgpu_destroy(s);
            return rc;
        }
    }
    if (s->o.fake_sirt > 0 && (rc = dev_alloc(s, &s->vol_tmp, nvol))) {
        tiltgpu_destroy(s);
        return rc;
    }
    *out = s;
    return TILTGPU_OK;
}

}  // namespace tiltgpu_sess

// ======================================================================= ABI

extern "C" {

const char* tiltgpu_last_error(void) { return g_last_error.c_str(); }

const char* tiltgpu_version(void) { return "tiltgpu 0.1 (gfx950)"; }

int tiltgpu_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

void tiltgpu_default_options(tiltgpu_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->radial_cutoff = 0.35f;
    o->radial_falloff = 0.05f;
    o->max_cosine_tilt = 80.0f;
    o->fill_mode = TILTGPU_FILL_MEAN;
    o->densweight = 1.0f;
    o->outscale = 1.0f;
    o->outadd = 0.0f;
    o->sirt_relax = 1.0f;
    o->storage = TILTGPU_STORE_F32;
}

int tiltgpu_geometry_tables(tiltgpu_session* s, tiltgpu_view_params* views, float* stretch,
                            float* zc) {
    if (!s) return fail(TILTGPU_ERR_ARG, "null session");
    int rc;
    if (!s->geometry_ready && (rc = prepare_pipeline_geometry(s))) return rc;
    const int nv = s->nviews, out_nz = s->d.out_nz;
    if (views && nv) std::memcpy(views, s->vp_host.data(), nv * sizeof(tiltgpu_view_params));
    if (stretch) HIP_TRY(hipMemcpy(stretch, s->stretch, s->d.nz * sizeof(float), hipMemcpyDeviceToHost));
    if (zc && nv && out_nz)
        HIP_TRY(hipMemcpy2D(zc, out_nz * sizeof(float), s->zc, s->nz_pad * sizeof(float),
                            out_nz * sizeof(float), nv, hipMemcpyDeviceToHost));
    return TILTGPU_OK;
}

int tiltgpu_proj_params(const float* angles, int32_t n, const float* xtilt, int32_t n_xtilt,
                        const float* zfactor, int32_t n_zfactor, tiltgpu_view_params* out) {
    if (n < 0 || (n > 0 && (!angles || !out))) return fail(TILTGPU_ERR_ARG, "bad arguments");
    const float pi = static_cast<float>(M_PI);
    for (int i = 0; i < n; ++i) {
        // main.rs:688-697, all in f32
        const float rad = angles[i] * pi / 180.0f;
        const float xt_deg = (xtilt && i < n_xtilt) ? xtilt[i] : 0.0f;
        const float xt_rad = xt_deg * pi / 180.0f;
        const float zf = (zfactor && i < n_zfactor) ? zfactor[i] : 1.0f;
        out[i].cos_t = cosf(rad);
        out[i].sin_t = sinf(rad);
        out[i].cos_xt = cosf(xt_rad);
        out[i].sin_xt = sinf(xt_rad);
        out[i].zfactor = zf;
    }
    return TILTGPU_OK;
}

int tiltgpu_radial_weights(int32_t n, int32_t exact_filter_size, float cutoff, float falloff,
                           int32_t hamming, float* out) {
    if (n < 1 || !out) return fail(TILTGPU_ERR_ARG, "bad arguments");
    const int nxc = exact_filter_size > 0 ? exact_filter_size : n / 2 + 1;
    const float pi = static_cast<float>(M_PI);
    for (int i = 0; i < n; ++i) {
        // main.rs:351-369
        const int bin = (i <= n / 2) ? i : n - i;
        const float w = static_cast<float>(bin) / static_cast<float>(nxc);
        float weight;
        if (hamming) {
            float h;
            if (cutoff > 1e-12f && w <= cutoff)
                h = 0.54f + 0.46f * cosf(pi * w / cutoff);
            else if (w > cutoff)
                h = 0.0f;
            else
                h = 1.0f;
            weight = w * h;
        } else if (w <= cutoff) {
            weight = w;
        } else {
            const float dd = w - cutoff;
            weight = w * expf(-dd * dd / (2.0f * falloff * falloff));
        }
        out[i] = weight;
    }
    return TILTGPU_OK;
}

int tiltgpu_create(const tiltgpu_dims* dims, const tiltgpu_options* opts,
                   tiltgpu_session** out) {
    return session_create(dims, opts, nullptr, out);
}

int tiltgpu_set_angles(tiltgpu_session* s, const float* a, int32_t n) {
    if (!s || n < 0 || (n > 0 && !a)) return fail(TILTGPU_ERR_ARG, "bad arguments");
    s->angles.assign(a, a + n);
    s->has_angles = true;
    s->geometry_ready = false;
    s->filtered = false;
    return TILTGPU_OK;
}

int tiltgpu_set_xtilt(tiltgpu_session* s, const float* a, int32_t n) {
    if (!s || n < 0 || (n > 0 && !a)) return fail(TILTGPU_ERR_ARG, "bad arguments");
    s->xtilt.assign(a, a + n);
    s->has_xtilt = true;
    s->geometry_ready = false;
    s->filtered = false;
    return TILTGPU_OK;
}

int tiltgpu_set_zfactor(tiltgpu_session* s, const float* a, int32_t n) {
    if (!s || n < 0 || (n > 0 && !a)) return fail(TILTGPU_ERR_ARG, "bad arguments");
    s->zfac.assign(a, a + n);
    s->has_zfac = true;
    s->geometry_ready = false;
    s->filtered = false;
    return TILTGPU_OK;
}

int tiltgpu_set_local_alignment(tiltgpu_session* s, const int32_t* nx, const int32_t* ny,
                                const float* xs, const float* ys, const float* dx,
                                const float* dy) {
    if (!s || !nx || !ny) return fail(TILTGPU_ERR_ARG, "bad arguments");
    const int nz = s->d.nz;
    s->lnx.assign(nx, nx + nz);
    s->lny.assign(ny, ny + nz);
    s->lxoff.assign(nz + 1, 0);
    s->lyoff.assign(nz + 1, 0);
    s->lgoff.assign(nz + 1, 0);
    for (int v = 0; v < nz; ++v) {
        if (nx[v] < 0 || ny[v] < 0 || (nx[v] == 0) != (ny[v] == 0))
            return fail(TILTGPU_ERR_ARG, "bad local grid size for view %d", v);
        s->lxoff[v + 1] = s->lxoff[v] + nx[v];
        s->lyoff[v + 1] = s->lyoff[v] + ny[v];
        s->lgoff[v + 1] = s->lgoff[v] + nx[v] * ny[v];
    }
    const int tx = s->lxoff[nz], ty = s->lyoff[nz], tg = s->lgoff[nz];
    if ((tx && !xs) || (ty && !ys) || (tg && (!dx || !dy)))
        return fail(TILTGPU_ERR_ARG, "local grid arrays are NULL");
    s->lxs.assign(xs, xs + tx);
    s->lys.assign(ys, ys + ty);
    s->ldx.assign(dx, dx + tg);
    s->ldy.assign(dy, dy + tg);
    s->has_local = true;
    s->geometry_ready = false;
    s->filtered = false;
    return TILTGPU_OK;
}

int tiltgpu_upload_projections(tiltgpu_session* s, const float* host, int64_t count) {
    if (!s || !host) return fail(TILTGPU_ERR_ARG, "bad arguments");
    const int64_t nproj = static_cast<int64_t>(s->d.nz) * s->d.in_ny * s->d.in_nx;
    if (count != nproj)
        return fail(TILTGPU_ERR_ARG, "expected %lld projection values, got %lld",
                    static_cast<long long>(nproj), static_cast<long long>(count));
    HIP_TRY(hipMemcpyAsync(s->raw, host, nproj * sizeof(float), hipMemcpyHostToDevice,
                           s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->uploaded = true;
    s->aligned_ok = false;
    s->view_up.assign(s->d.nz, 1);
    s->views_up = s->d.nz;
    s->filtered = false;
    return TILTGPU_OK;
}

int tiltgpu_upload_views(tiltgpu_session* s, int32_t v0, int32_t nv, const float* host) {
    if (!s || !host || s->seam) return fail(TILTGPU_ERR_ARG, "bad arguments");
    if (v0 < 0 || nv < 1 || v0 + nv > s->d.nz)
        return fail(TILTGPU_ERR_ARG, "views [%d, %d) outside the %d sections", v0, v0 + nv,
                    s->d.nz);
    const int64_t plane = static_cast<int64_t>(s->d.in_ny) * s->d.in_nx;
    HIP_TRY(hipMemcpyAsync(s->raw + v0 * plane, host, nv * plane * sizeof(float),
                           hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (s->view_up.size() != static_cast<size_t>(s->d.nz)) {
        s->view_up.assign(s->d.nz, 0);
        s->views_up = 0;
    }
    for (int v = v0; v < v0 + nv; ++v)
        if (!s->view_up[v]) {
            s->view_up[v] = 1;
            ++s->views_up;
        }
    s->uploaded = s->views_up == s->d.nz;
    s->aligned_ok = false;
    s->filtered = false;
    return TILTGPU_OK;
}

int tiltgpu_set_section_transforms(tiltgpu_session* s, const float* xf, int32_t n, float fill) {
    if (!s || s->seam || n < 0 || (n > 0 && !xf)) return fail(TILTGPU_ERR_ARG, "bad arguments");
    s->aligned_ok = false;
    s->filtered = false;
    if (n == 0) {
        dev_free(s, s->d_xf);
        dev_free(s, s->aligned);
        s->d_xf = s->aligned = nullptr;
        return TILTGPU_OK;
    }
    // section z takes transform min(z, n-1) (apps/newstack/src/main.rs:1281-1283);
    // LinearTransform::inverse in f32 (crates/imod-transforms/src/xf.rs:95-109)
    std::vector<float> inv(6 * static_cast<size_t>(s->d.nz));
    for (int z = 0; z < s->d.nz; ++z) {
        const float* t = xf + 6 * std::min(z, n - 1);
        const float a11 = t[0], a12 = t[1], a21 = t[2], a22 = t[3], dx = t[4], dy = t[5];
        const float det = a11 * a22 - a12 * a21;
        const float i11 = a22 / det, i12 = -a12 / det, i21 = -a21 / det, i22 = a11 / det;
        float* o = inv.data() + 6 * z;
        o[0] = i11;
        o[1] = i12;
        o[2] = i21;
        o[3] = i22;
        o[4] = -(i11 * dx + i12 * dy);
        o[5] = -(i21 * dx + i22 * dy);
    }
    int rc;
    if (!s->d_xf && (rc = dev_alloc(s, &s->d_xf, inv.size()))) return rc;
    HIP_TRY(hipMemcpy(s->d_xf, inv.data(), inv.size() * sizeof(float), hipMemcpyHostToDevice));
    s->xf_fill = fill;
    return TILTGPU_OK;
}

int tiltgpu_read_projections(tiltgpu_session* s, float* out) {
    if (!s || !out || s->seam) return fail(TILTGPU_ERR_ARG, "bad arguments");
    if (!s->uploaded) return fail(TILTGPU_ERR_STATE, "projections not uploaded");
    int rc;
    if ((rc = ensure_aligned(s))) return rc;
    const int64_t nproj = static_cast<int64_t>(s->d.nz) * s->d.in_ny * s->d.in_nx;
    HIP_TRY(hipMemcpyAsync(out, stack_src(s), nproj * sizeof(float), hipMemcpyDeviceToHost,
                           s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    collect_timing(s);
    return TILTGPU_OK;
}

void* tiltgpu_host_alloc(int64_t bytes) {
    if (bytes <= 0) {
        fail(TILTGPU_ERR_ARG, "bad size");
        return nullptr;
    }
    void* p = nullptr;
    const hipError_t e = hipHostMalloc(&p, static_cast<size_t>(bytes), hipHostMallocDefault);
    if (e != hipSuccess) {
        fail(TILTGPU_ERR_HIP, "hipHostMalloc(%lld bytes) failed: %s", static_cast<long long>(bytes),
             hipGetErrorString(e));
        return nullptr;
    }
    return p;
}

void tiltgpu_host_free(void* p) {
    if (p) (void)hipHostFree(p);
}

void* tiltgpu_device_alloc(int64_t bytes) {
    if (bytes <= 0) {
        fail(TILTGPU_ERR_ARG, "bad size");
        return nullptr;
    }
    void* p = nullptr;
    const hipError_t e = hipMalloc(&p, static_cast<size_t>(bytes));
    if (e != hipSuccess) {
        fail(TILTGPU_ERR_HIP, "hipMalloc(%lld bytes) failed: %s", static_cast<long long>(bytes),
             hipGetErrorString(e));
        return nullptr;
    }
    return p;
}

void tiltgpu_device_free(void* p) {
    if (p) (void)hipFree(p);
}

int tiltgpu_copy_to_device(void* dst, const void* src, int64_t bytes) {
    if (!dst || !src || bytes < 0) return fail(TILTGPU_ERR_ARG, "bad arguments");
    HIP_TRY(hipMemcpy(dst, src, static_cast<size_t>(bytes), hipMemcpyHostToDevice));
    return TILTGPU_OK;
}

int tiltgpu_copy_to_host(void* dst, const void* src, int64_t bytes) {
    if (!dst || !src || bytes < 0) return fail(TILTGPU_ERR_ARG, "bad arguments");
    HIP_TRY(hipMemcpy(dst, src, static_cast<size_t>(bytes), hipMemcpyDeviceToHost));
    return TILTGPU_OK;
}

// Per-view means (main.rs:615-623), then the R-weighting + cosine stretch of
// O.  Without --log the power-of-two filter path (rows 256-2048 wide) sums each
// row as it loads it, in f64, and the means follow from those row sums: the
// stack is read once for both.  Otherwise (and with --log, whose kernel writes
// the logged stack) the log/means kernel reads it first.
static int run_means_and_filter(tiltgpu_session* s, const float* O) {
    int rc;
    const bool fused_means = !s->o.log && s->tw8 && s->d.in_nx >= 256;
    if (fused_means && !s->row_sum &&
        (rc = dev_alloc(s, &s->row_sum, static_cast<size_t>(s->d.nz) * s->d.in_ny)))
        return rc;
    if (!fused_means && (rc = run_means(s))) return rc;
    if ((rc = run_filter(s, O, nullptr, true, nullptr, fused_means ? s->row_sum : nullptr)))
        return rc;
    if (fused_means) {
        Scope sc(s, kClsPrep);
        HIP_TRY(launch_means_from_rows(s->row_sum, s->d.in_ny,
                                       static_cast<int64_t>(s->d.in_nx) * s->d.in_ny, s->d.nz,
                                       s->means, s->stream));
    }
    return TILTGPU_OK;
}

int tiltgpu_run_rows(tiltgpu_session* s, int32_t y0, int32_t y1) {
    if (!s || s->seam) return fail(TILTGPU_ERR_STATE, "not a pipeline session");
    if (!s->uploaded) return fail(TILTGPU_ERR_STATE, "projections not uploaded");
    if (s->o.sirt > 0 || s->o.fake_sirt > 0)
        return fail(TILTGPU_ERR_ARG, "row ranges are reconstructed by WBP only (sirt and "
                                     "fake_sirt couple the rows)");
    if (y0 < 0 || y1 > s->d.in_ny || y0 >= y1) return fail(TILTGPU_ERR_ARG, "bad row range");
    int rc;
    if (!s->geometry_ready && (rc = prepare_pipeline_geometry(s))) return rc;
    if (!s->filtered) {
        if ((rc = ensure_aligned(s))) return rc;
        const float* O = s->o.log ? s->logd : stack_src(s);
        if ((rc = run_means_and_filter(s, O))) return rc;
        s->filtered = true;
    }
    const void* proj = s->o.storage == TILTGPU_STORE_F32 ? static_cast<const void*>(s->filt)
                                                         : s->store;
    const int64_t plane = static_cast<int64_t>(s->d.out_nz) * s->d.out_nx;
    if ((rc = run_bp(s, proj, s->scale, y0, y1, 1, s->vol + y0 * plane))) return rc;
    HIP_TRY(hipStreamSynchronize(s->stream));
    collect_timing(s);
    s->ran = true;
    return check_kernel_error(s);
}

int tiltgpu_run(tiltgpu_session* s) {
    if (!s || s->seam) return fail(TILTGPU_ERR_STATE, "not a pipeline session");
    if (!s->uploaded) return fail(TILTGPU_ERR_STATE, "projections not uploaded");
    int rc;
    if (!s->geometry_ready && (rc = prepare_pipeline_geometry(s))) return rc;
    if ((rc = ensure_aligned(s))) return rc;
    const float* O = s->o.log ? s->logd : stack_src(s);
    // log + means, R-weighting + cosine stretch (+ storage conversion)
    if ((rc = run_means_and_filter(s, O))) return rc;
    s->filtered = s->o.sirt == 0;  // SIRT filters its residuals into the same buffers
    const void* proj = s->o.storage == TILTGPU_STORE_F32 ? static_cast<const void*>(s->filt)
                                                         : s->store;
    const int out_ny = s->d.in_ny;
    const int trows = s->trace_y1 - s->trace_y0;
    const int64_t plane = static_cast<int64_t>(s->d.out_nz) * s->d.out_nx;
    if (trows > 0) {
        dev_free(s, s->trace);
        s->trace = nullptr;
        if ((rc = dev_alloc(s, &s->trace, (s->o.sirt + 1) * trows * plane))) return rc;
    }
    auto keep_trace = [&](int k) -> int {
        if (trows <= 0) return TILTGPU_OK;
        HIP_TRY(hipMemcpyAsync(s->trace + k * trows * plane, s->vol + s->trace_y0 * plane,
                               trows * plane * sizeof(float), hipMemcpyDeviceToDevice,
                               s->stream));
        return TILTGPU_OK;
    };
    const bool fused_transform = s->o.sirt == 0 && s->o.fake_sirt == 0 && trows <= 0;
    if ((rc = run_bp(s, proj, s->scale, 0, out_ny, fused_transform ? 1 : 0, s->vol))) return rc;
    if ((rc = keep_trace(0))) return rc;
    // SIRT (main.rs:919-953)
    s->sirt_done = 0;
    for (int it = 0; it < s->o.sirt; ++it) {
        if ((rc = run_forward(s))) return rc;
        if ((rc = run_filter(s, O, s->sino, false, s->resid + it))) return rc;
        if ((rc = run_bp(s, proj, s->scale, 0, out_ny, 2, s->vol))) return rc;
        if ((rc = keep_trace(it + 1))) return rc;
        s->sirt_done++;
    }
    // fake-SIRT low-pass (main.rs:955-962, 976-985)
    const int onz = s->d.out_nz, onx = s->d.out_nx;
    for (int k = 0; k < s->o.fake_sirt; ++k) {
        Scope sc(s, kClsOther);
        HIP_TRY(launch_lowpass_x(s->vol, s->vol_tmp, out_ny, onz, onx, s->stream));
        HIP_TRY(launch_lowpass_z(s->vol_tmp, s->vol, out_ny, onz, onx, s->stream));
        HIP_TRY(launch_lowpass_y(s->vol, s->vol_tmp, out_ny, onz, onx, s->stream));
        std::swap(s->vol, s->vol_tmp);
    }
    if (!fused_transform) {
        Scope sc(s, kClsOther);
        HIP_TRY(launch_transform(s->vol, static_cast<int64_t>(out_ny) * onz * onx,
                                 s->o.densweight, s->o.outscale, s->o.outadd, s->stream));
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    collect_timing(s);
    if ((rc = check_kernel_error(s))) return rc;
    s->ran = true;
    return TILTGPU_OK;
}

int tiltgpu_read_rows(tiltgpu_session* s, int32_t y0, int32_t y1, float* out, float* stats) {
    if (!s || !out) return fail(TILTGPU_ERR_ARG, "bad arguments");
    if (!s->ran) return fail(TILTGPU_ERR_STATE, "nothing reconstructed yet");
    if (y0 < 0 || y1 > s->d.in_ny || y0 > y1) return fail(TILTGPU_ERR_ARG, "bad row range");
    const int64_t per = static_cast<int64_t>(s->d.out_nz) * s->d.out_nx;
    const int rows = y1 - y0;
    if (rows == 0) return TILTGPU_OK;
    if (stats) {
        if (s->stats_rows < rows) {
            dev_free(s, s->stats);
            int rc = dev_alloc(s, &s->stats, 3 * static_cast<size_t>(rows));
            if (rc) return rc;
            s->stats_rows = rows;
        }
        Scope sc(s, kClsOther);
        HIP_TRY(launch_slice_stats(s->vol + y0 * per, rows, per, s->stats, s->stream));
    }
    HIP_TRY(hipMemcpyAsync(out, s->vol + y0 * per, rows * per * sizeof(float),
                           hipMemcpyDeviceToHost, s->stream));
    if (stats)
        HIP_TRY(hipMemcpyAsync(stats, s->stats, 3 * rows * sizeof(float), hipMemcpyDeviceToHost,
                               s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    collect_timing(s);
    return TILTGPU_OK;
}

int tiltgpu_sirt_residuals(tiltgpu_session* s, double* out, int32_t n) {
    if (!s || !out || n < 0) return fail(TILTGPU_ERR_ARG, "bad arguments");
    const int k = std::min<int>(n, s->sirt_done);
    if (k > 0) HIP_TRY(hipMemcpy(out, s->resid, k * sizeof(double), hipMemcpyDeviceToHost));
    return k;
}

int tiltgpu_set_tuning(tiltgpu_session* s, int32_t bp_tile_z, int32_t fp_bins_per_lane) {
    if (!s || s->seam) return fail(TILTGPU_ERR_STATE, "not a pipeline session");
    if (bp_tile_z != 0 && bp_tile_z != 32 && bp_tile_z != 64 && bp_tile_z != 128 &&
        bp_tile_z != 256)
        return fail(TILTGPU_ERR_ARG, "tile depth %d is not 0, 32, 64, 128 or 256", bp_tile_z);
    if (fp_bins_per_lane != 0 && fp_bins_per_lane != 4 && fp_bins_per_lane != 8 &&
        fp_bins_per_lane != 16)
        return fail(TILTGPU_ERR_ARG, "bins per lane %d is not 0, 4, 8 or 16", fp_bins_per_lane);
    s->tune_tile_z = bp_tile_z;
    s->tune_fp_bins = fp_bins_per_lane;
    s->geometry_ready = false;  // the tile depth sets the padded depth of the view tables
    return TILTGPU_OK;
}

int tiltgpu_get_tuning(tiltgpu_session* s, int32_t* out4) {
    if (!s || s->seam || !out4) return fail(TILTGPU_ERR_ARG, "bad arguments");
    int rc;
    if (!s->geometry_ready && (rc = prepare_pipeline_geometry(s))) return rc;
    int fp_src = kTuneForced;
    const int bins =
        s->tune_fp_bins > 0 ? s->tune_fp_bins : tuned_fp_bins(s->d.in_nx, &fp_src);
    out4[0] = s->tile_z;
    out4[1] = s->tile_source;
    out4[2] = bins;
    out4[3] = fp_src;
    return TILTGPU_OK;
}

int tiltgpu_set_bp_shape(tiltgpu_session* s, int32_t views_per_stage, int32_t x_per_lane) {
    if (!s || s->seam) return fail(TILTGPU_ERR_STATE, "not a pipeline session");
    if (views_per_stage < 0 || views_per_stage > 2 || x_per_lane < 0 || x_per_lane > 2)
        return fail(TILTGPU_ERR_ARG, "views per stage %d / x per lane %d not in 0..2",
                    views_per_stage, x_per_lane);
    s->tune_vps = views_per_stage;
    s->tune_xpl = x_per_lane;
    s->geometry_ready = false;
    return TILTGPU_OK;
}

int tiltgpu_get_bp_shape(tiltgpu_session* s, int32_t* out3) {
    if (!s || s->seam || !out3) return fail(TILTGPU_ERR_ARG, "bad arguments");
    int rc;
    if (!s->geometry_ready && (rc = prepare_pipeline_geometry(s))) return rc;
    out3[0] = s->tile_z;
    out3[1] = s->bp_vps;
    out3[2] = s->bp_xpl;
    return TILTGPU_OK;
}

int tiltgpu_set_sirt_trace(tiltgpu_session* s, int32_t y0, int32_t y1) {
    if (!s || s->seam) return fail(TILTGPU_ERR_STATE, "not a pipeline session");
    if (y0 < 0 || y1 < y0 || y1 > s->d.in_ny) return fail(TILTGPU_ERR_ARG, "bad row range");
    s->trace_y0 = y0;
    s->trace_y1 = y1;
    return TILTGPU_OK;
}

int tiltgpu_read_sirt_trace(tiltgpu_session* s, int32_t k, float* out) {
    if (!s || !out) return fail(TILTGPU_ERR_ARG, "bad arguments");
    const int trows = s->trace_y1 - s->trace_y0;
    if (trows <= 0 || !s->trace) return fail(TILTGPU_ERR_STATE, "no SIRT trace recorded");
    if (k < 0 || k > s->sirt_done) return fail(TILTGPU_ERR_ARG, "trace %d not recorded", k);
    const int64_t n = static_cast<int64_t>(trows) * s->d.out_nz * s->d.out_nx;
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(out, s->trace + k * n, n * sizeof(float), hipMemcpyDeviceToHost));
    return TILTGPU_OK;
}

int tiltgpu_reproject(tiltgpu_session* s, const float* vol, float* out) {
    if (!s || !vol || !out) return fail(TILTGPU_ERR_ARG, "bad arguments");
    if (s->seam) return fail(TILTGPU_ERR_STATE, "not a pipeline session");
    int rc;
    if (!s->geometry_ready && (rc = prepare_pipeline_geometry(s))) return rc;
    const int64_t nproj = static_cast<int64_t>(s->d.nz) * s->d.in_ny * s->d.in_nx;
    const int64_t nvol = static_cast<int64_t>(s->d.in_ny) * s->d.out_nz * s->d.out_nx;
    if (!s->sino && (rc = dev_alloc(s, &s->sino, nproj))) return rc;
    HIP_TRY(hipMemcpyAsync(s->vol, vol, nvol * 4, hipMemcpyHostToDevice, s->stream));
    if ((rc = run_forward(s))) return rc;
    HIP_TRY(hipMemcpyAsync(out, s->sino, nproj * 4, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    collect_timing(s);
    return check_kernel_error(s);
}

int tiltgpu_read_filtered(tiltgpu_session* s, float* out) {
    if (!s || !out) return fail(TILTGPU_ERR_ARG, "bad arguments");
    const int64_t nproj = static_cast<int64_t>(s->d.nz) * s->d.in_ny * s->d.in_nx;
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (s->o.storage == TILTGPU_STORE_F32 || s->seam) {
        HIP_TRY(hipMemcpy(out, s->filt, nproj * 4, hipMemcpyDeviceToHost));
        return TILTGPU_OK;
    }
    std::vector<unsigned short> h(nproj);
    std::vector<float> sc(s->d.nz);
    HIP_TRY(hipMemcpy(h.data(), s->store, nproj * 2, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(sc.data(), s->scale, s->d.nz * 4, hipMemcpyDeviceToHost));
    const int64_t nxy = static_cast<int64_t>(s->d.in_ny) * s->d.in_nx;
    for (int64_t i = 0; i < nproj; ++i) {
        float v;
        if (s->o.storage == TILTGPU_STORE_BF16) {
            const uint32_t u = static_cast<uint32_t>(h[i]) << 16;
            std::memcpy(&v, &u, 4);
        } else {
            _Float16 hv;
            std::memcpy(&hv, &h[i], 2);
            v = static_cast<float>(hv);
        }
        out[i] = v * sc[i / nxy];
    }
    return TILTGPU_OK;
}

int tiltgpu_read_means(tiltgpu_session* s, float* out) {
    if (!s || !out) return fail(TILTGPU_ERR_ARG, "bad arguments");
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(out, s->means, s->d.nz * 4, hipMemcpyDeviceToHost));
    return TILTGPU_OK;
}

int tiltgpu_synchronize(tiltgpu_session* s) {
    if (!s) return fail(TILTGPU_ERR_ARG, "bad arguments");
    HIP_TRY(hipStreamSynchronize(s->stream));
    collect_timing(s);
    return TILTGPU_OK;
}

int tiltgpu_get_timing(tiltgpu_session* s, tiltgpu_timing* t) {
    if (!s || !t) return fail(TILTGPU_ERR_ARG, "bad arguments");
    HIP_TRY(hipStreamSynchronize(s->stream));
    collect_timing(s);
    t->prep_ms = s->ms[kClsPrep];
    t->filter_ms = s->ms[kClsFilter];
    t->bp_ms = s->ms[kClsBp];
    t->fp_ms = s->ms[kClsFp];
    t->other_ms = s->ms[kClsOther];
    t->bp_launches = s->n_bp;
    t->filter_launches = s->n_filter;
    t->fp_launches = s->n_fp;
    return TILTGPU_OK;
}

int tiltgpu_reset_timing(tiltgpu_session* s) {
    if (!s) return fail(TILTGPU_ERR_ARG, "bad arguments");
    HIP_TRY(hipStreamSynchronize(s->stream));
    collect_timing(s);
    for (double& m : s->ms) m = 0;
    s->n_bp = s->n_filter = s->n_fp = 0;
    return TILTGPU_OK;
}

int64_t tiltgpu_device_bytes(const tiltgpu_session* s) { return s ? s->bytes : 0; }

void tiltgpu_destroy(tiltgpu_session* s) {
    if (!s) return;
    // a pooled (server) session is destroyed only after its work has completed
    if (s->stream && !s->pool) (void)hipStreamSynchronize(s->stream);
    for (auto& r : s->pending) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (hipEvent_t e : s->event_pool) (void)hipEventDestroy(e);
    for (size_t i = 0; i < s->allocs.size(); ++i) {
        if (s->pool)
            s->pool->put(s->allocs[i], s->alloc_bytes[i]);
        else
            (void)hipFree(s->allocs[i]);
    }
    if (s->stream && s->own_stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

// ----------------------------------------------------- GpuSession-shaped seam

int tiltgpu_bp_session_new(const float* projections, const tiltgpu_view_params* params,
                           const float* proj_means, int32_t n_views, int32_t in_nx,
                           int32_t in_ny, int32_t nz, int32_t out_nx, int32_t out_nz,
                           int32_t fill_mode, tiltgpu_session** out) {
    if (!out) return fail(TILTGPU_ERR_ARG, "out is NULL");
    *out = nullptr;
    tiltgpu_dims d{in_nx, in_ny, nz, out_nx, out_nz};
    int rc = validate_dims(&d);
    if (rc) return rc;
    if (!projections || !proj_means || n_views < 0 || n_views > nz || (n_views && !params))
        return fail(TILTGPU_ERR_ARG, "bad arguments");
    auto* s = new tiltgpu_session();
    s->seam = true;
    s->d = d;
    tiltgpu_default_options(&s->o);
    s->o.fill_mode = (fill_mode >= 0 && fill_mode <= 2) ? fill_mode : TILTGPU_FILL_MEAN;
    const int64_t nproj = static_cast<int64_t>(nz) * in_ny * in_nx;
    std::vector<tiltgpu_view_params> vp(params, params + n_views);
    if ((rc = init_device(s, nullptr)) || (rc = dev_alloc(s, &s->filt, nproj)) ||
        (rc = dev_alloc(s, &s->means, nz)) || (rc = setup_geometry(s, vp))) {
        tiltgpu_destroy(s);
        return rc;
    }
    hipError_t e = hipMemcpy(s->filt, projections, nproj * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->means, proj_means, nz * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        tiltgpu_destroy(s);
        return fail(TILTGPU_ERR_HIP, "upload failed: %s", hipGetErrorString(e));
    }
    s->uploaded = true;
    *out = s;
    return TILTGPU_OK;
}

int tiltgpu_reconstruct_rows(tiltgpu_session* s, int32_t y0, int32_t y1, float* out) {
    if (!s || !out) return fail(TILTGPU_ERR_ARG, "bad arguments");
    if (!s->seam) return fail(TILTGPU_ERR_STATE, "not a backprojection seam session");
    if (y0 < 0 || y1 > s->d.in_ny || y0 >= y1) return fail(TILTGPU_ERR_ARG, "bad row range");
    const int rows = y1 - y0;
    const int64_t per = static_cast<int64_t>(s->d.out_nz) * s->d.out_nx;
    if (!s->vol || s->stats_rows < rows) {
        // the seam keeps a slab buffer sized by the largest request so far
        dev_free(s, s->vol);
        int rc = dev_alloc(s, &s->vol, rows * per);
        if (rc) return rc;
        s->stats_rows = rows;
    }
    int rc = run_bp(s, s->filt, nullptr, y0, y1, 0, s->vol);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(out, s->vol, rows * per * 4, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    collect_timing(s);
    return check_kernel_error(s);
}

int tiltgpu_reconstruct_row(tiltgpu_session* s, int32_t iy, float* out) {
    return tiltgpu_reconstruct_rows(s, iy, iy + 1, out);
}

// One timed launch (after a warm-up) of an issue-rate mix over 8 waves per SIMD;
// ms and the wave instructions issued per SIMD.
static int run_issue_mix(int mix, int32_t iters, double* ms_out, double* instr_per_simd) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        return fail(TILTGPU_ERR_NODEV, "no GPU visible to HIP");
    int dev = 0, cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
        cus = prop.multiProcessorCount;
    const int blocks = cus * 8;  // 8 workgroups of 4 waves per CU: 8 waves per SIMD
    float* out = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = TILTGPU_OK;
    if (hipMalloc(&out, static_cast<size_t>(blocks) * 256 * 4) != hipSuccess ||
        hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        rc = fail(TILTGPU_ERR_HIP, "issue calibration: allocation failed");
    } else {
        float ms = 0.f;
        if (launch_issue_rate(mix, out, blocks, iters, st) != hipSuccess ||  // warm-up
            hipEventRecord(e0, st) != hipSuccess ||
            launch_issue_rate(mix, out, blocks, iters, st) != hipSuccess ||
            hipEventRecord(e1, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            rc = fail(TILTGPU_ERR_HIP, "issue calibration: launch failed");
        *ms_out = ms;
        // waves per SIMD (8) x lane-loop iterations x instructions per iteration
        *instr_per_simd = 8.0 * iters * issue_rate_instructions(mix);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (out) (void)hipFree(out);
    return rc;
}

int tiltgpu_fma_calibration(int32_t iters, double* tflops) {
    if (iters < 1 || !tflops) return fail(TILTGPU_ERR_ARG, "bad arguments");
    double ms = 0, n = 0;
    int rc = run_issue_mix(0, iters, &ms, &n);
    if (rc) return rc;
    int dev = 0, cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
        cus = prop.multiProcessorCount;
    // n wave FMAs per SIMD, 4 SIMDs per CU, 64 lanes, 2 FLOP each
    *tflops = n * 4.0 * cus * 64.0 * 2.0 / (ms * 1e-3) / 1e12;
    return TILTGPU_OK;
}

int tiltgpu_issue_calibration(int32_t mix, int32_t iters, double* ns_per_instr) {
    if (mix < 0 || mix > 4 || iters < 1 || !ns_per_instr)
        return fail(TILTGPU_ERR_ARG, "bad arguments");
    double ms = 0, n = 0;
    int rc = run_issue_mix(mix, iters, &ms, &n);
    if (rc) return rc;
    *ns_per_instr = ms * 1e6 / n;
    return TILTGPU_OK;
}

int tiltgpu_copy_calibration(int64_t bytes, int32_t iters, double* ms_per_copy) {
    if (bytes < 16 || iters < 1 || !ms_per_copy) return fail(TILTGPU_ERR_ARG, "bad arguments");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        return fail(TILTGPU_ERR_NODEV, "no GPU visible to HIP");
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bytes = bytes / 16 * 16;
    int rc = TILTGPU_OK;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess ||
        hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        rc = fail(TILTGPU_ERR_HIP, "copy calibration: allocation failed");
    } else {
        float ms = 0.f;
        if (launch_fill_f32(static_cast<float*>(a), bytes / 4, 1.0f, st) != hipSuccess ||
            hipEventRecord(e0, st) != hipSuccess)
            rc = fail(TILTGPU_ERR_HIP, "copy calibration: launch failed");
        for (int i = 0; rc == TILTGPU_OK && i < iters; ++i)
            if (launch_copy(a, b, bytes, st) != hipSuccess)
                rc = fail(TILTGPU_ERR_HIP, "copy calibration: launch failed");
        if (rc == TILTGPU_OK && (hipEventRecord(e1, st) != hipSuccess ||
                                 hipStreamSynchronize(st) != hipSuccess ||
                                 hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = fail(TILTGPU_ERR_HIP, "copy calibration: timing failed");
        *ms_per_copy = static_cast<double>(ms) / iters;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

}  // extern "C"
