// Internal: the session object and the helpers shared by the single-series
// ABI (session.hip) and the batching server (server.hip).  Not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/tiltgpu.h"
#include "tiltgpu_internal.h"

namespace tiltgpu_sess {

// Records the calling thread's last error message; returns code.
int fail(int code, const char* fmt, ...);

// Failure injection for tests of the error paths: with TILTGPU_FAIL_ALLOC_AT=n
// in the environment, the n-th device allocation counted since the variable
// last changed is refused (TILTGPU_ERR_BUDGET), as an exhausted device would.
bool forced_alloc_failure();

#define HIP_TRY(expr)                                                                 \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return ::tiltgpu_sess::fail(TILTGPU_ERR_HIP, "%s failed: %s (%s:%d)", #expr, \
                                        hipGetErrorString(e_), __FILE__, __LINE__);   \
    } while (0)

struct TimedRegion {
    hipEvent_t a, b;
    int cls;
};

}  // namespace tiltgpu_sess

using tiltgpu_sess::TimedRegion;
using namespace tiltgpu;

enum { kClsPrep = 0, kClsFilter, kClsBp, kClsFp, kClsOther, kNumCls };

namespace tiltgpu_sess {
// Device-buffer provider a session may borrow (the server's buffer cache):
// get() returns a buffer of at least `bytes`, put() takes one back.  Neither
// synchronises the device, unlike hipFree.
struct BufferPool {
    virtual void* get(size_t bytes) = 0;
    virtual void put(void* p, size_t bytes) = 0;
    virtual ~BufferPool() = default;
};
}  // namespace tiltgpu_sess

struct tiltgpu_session {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = true;  // false when a server lends its stream
    tiltgpu_sess::BufferPool* pool = nullptr;  // server sessions: buffers come from its cache
    bool timing = true;      // per-stage event timing (off for server sessions)
    tiltgpu_dims d{};
    tiltgpu_options o{};
    bool seam = false;  // created by tiltgpu_bp_session_new

    std::vector<float> angles, xtilt, zfac;
    bool has_xtilt = false, has_zfac = false, has_angles = false;
    int nviews = 0;
    bool geometry_ready = false;
    std::vector<tiltgpu_view_params> vp_host;  // the view tables of the prepared geometry

    // local alignment (host copy)
    bool has_local = false;
    std::vector<int> lnx, lny, lxoff, lyoff, lgoff;
    std::vector<float> lxs, lys, ldx, ldy;

    // device buffers
    float* raw = nullptr;      // uploaded projections (never modified)
    float* logd = nullptr;     // --log output (when enabled)
    float* filt = nullptr;     // filtered + stretched, f32
    void* store = nullptr;     // f16 / bf16 copy of filt
    float* scale = nullptr;    // [nz] dequantisation scale
    unsigned* absmax = nullptr;
    float* vol = nullptr;      // [out_ny][out_nz][out_nx]
    float* vol_tmp = nullptr;  // fake-SIRT ping-pong
    float* sino = nullptr;     // SIRT reprojection [nz][in_ny][in_nx]
    unsigned long long* fp_acc = nullptr;  // two-row reprojection fixed-point sums
    unsigned* fp_fixed = nullptr;          // [0]: volume absmax bits, [1]: fixed-point exponent
    float4* vgeom = nullptr;
    BpView* views = nullptr;
    float* zc = nullptr;
    float* means = nullptr;
    double* partial = nullptr;
    int partial_n = 0;
    double* row_sum = nullptr;  // per-row f64 sums of the stack (means fused into the filter)
    int2* seg_tab = nullptr;    // staged segments of the backprojection tiles (launch_bp_segments)
    size_t seg_cap = 0;
    float* stretch = nullptr;
    float* weights = nullptr;
    float2* tw = nullptr;
    float2* tw8 = nullptr;
    // newstack -x section transforms: inverses [nz][6] on the device; the
    // transformed stack `aligned` is made from raw when first needed
    float* d_xf = nullptr;
    float xf_fill = 0.f;
    float* aligned = nullptr;
    bool aligned_ok = false;     // per-pass twiddles of the power-of-two filter plan, or null
    int* err = nullptr;
    double* resid = nullptr;
    float* trace = nullptr;  // [sirt+1][trace_y1-trace_y0][out_nz][out_nx]
    int trace_y0 = 0, trace_y1 = 0;
    float* stats = nullptr;
    int stats_rows = 0;
    int* d_lnx = nullptr;
    int* d_lny = nullptr;
    int* d_lxoff = nullptr;
    int* d_lyoff = nullptr;
    int* d_lgoff = nullptr;
    float* d_lxs = nullptr;
    float* d_lys = nullptr;
    float* d_ldx = nullptr;
    float* d_ldy = nullptr;

    int nz_pad = 0;
    int tile_z = 64;
    int tile_z_override = 0;  // TILTGPU_TILE_Z (tuning): force the backprojection tile depth
    int tune_tile_z = 0;      // tiltgpu_set_tuning: forced tile depth (0: table / heuristic)
    int tune_vps = 0;         // tiltgpu_set_bp_shape: forced views per stage (0: table)
    int tune_xpl = 0;         // tiltgpu_set_bp_shape: forced x per lane (0: table)
    int bp_vps = 1, bp_xpl = 1;  // backprojection launch shape in use
    bool fixed_shape = false;    // shape (1, 1) only (the server's batched kernel)
    int tune_fp_bins = 0;     // tiltgpu_set_tuning: forced FP bins per lane (0: table / rule)
    int tile_source = 0;      // kTuneHeuristic / kTuneTable / kTuneForced
    int lds_cap = 0;
    bool two_row = false;
    int radix[kMaxRadix];
    int nradix = 0;
    bool uploaded = false;
    std::vector<uint8_t> view_up;  // views present after tiltgpu_upload_views
    int views_up = 0;
    bool filtered = false;  // filt/store hold this upload's R-weighted, stretched views
    bool ran = false;
    int sirt_done = 0;
    int64_t bytes = 0;
    std::vector<void*> allocs;
    std::vector<size_t> alloc_bytes;  // parallel to allocs

    std::vector<TimedRegion> pending;
    std::vector<hipEvent_t> event_pool;
    double ms[kNumCls] = {0, 0, 0, 0, 0};
    int64_t n_bp = 0, n_filter = 0, n_fp = 0;
};


namespace tiltgpu_sess {

template <typename T>
inline int dev_alloc(tiltgpu_session* s, T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    void* q = nullptr;
    const size_t bytes = count * sizeof(T);
    if (forced_alloc_failure())
        return fail(TILTGPU_ERR_BUDGET, "device buffer of %zu bytes refused (TILTGPU_FAIL_ALLOC_AT)",
                    bytes);
    if (s->pool) {
        q = s->pool->get(bytes);
        if (!q) return fail(TILTGPU_ERR_BUDGET, "device buffer of %zu bytes unavailable", bytes);
    } else {
        hipError_t e = hipMalloc(&q, bytes);
        if (e != hipSuccess)
            return fail(e == hipErrorOutOfMemory ? TILTGPU_ERR_BUDGET : TILTGPU_ERR_HIP,
                        "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    }
    s->allocs.push_back(q);
    s->alloc_bytes.push_back(bytes);
    s->bytes += static_cast<int64_t>(bytes);
    *p = static_cast<T*>(q);
    return TILTGPU_OK;
}

template <typename T>
inline void dev_free(tiltgpu_session* s, T*& p) {
    if (!p) return;
    auto it = std::find(s->allocs.begin(), s->allocs.end(), static_cast<void*>(p));
    size_t bytes = 0;
    if (it != s->allocs.end()) {
        const size_t k = static_cast<size_t>(it - s->allocs.begin());
        bytes = s->alloc_bytes[k];
        s->allocs.erase(it);
        s->alloc_bytes.erase(s->alloc_bytes.begin() + static_cast<std::ptrdiff_t>(k));
        s->bytes -= static_cast<int64_t>(bytes);
    }
    if (s->pool)
        s->pool->put(p, bytes);  // stream order keeps earlier uses ahead of any reuse
    else
        (void)hipFree(p);
    p = nullptr;
}

inline hipEvent_t take_event(tiltgpu_session* s) {
    if (!s->event_pool.empty()) {
        hipEvent_t e = s->event_pool.back();
        s->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

struct Scope {
    tiltgpu_session* s;
    TimedRegion r;
    Scope(tiltgpu_session* s_, int cls) : s(s_) {
        if (!s->timing) return;
        r.a = take_event(s);
        r.b = take_event(s);
        r.cls = cls;
        (void)hipEventRecord(r.a, s->stream);
    }
    ~Scope() {
        if (!s->timing) return;
        (void)hipEventRecord(r.b, s->stream);
        s->pending.push_back(r);
    }
};

inline void collect_timing(tiltgpu_session* s) {
    for (auto& r : s->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) s->ms[r.cls] += ms;
        s->event_pool.push_back(r.a);
        s->event_pool.push_back(r.b);
    }
    s->pending.clear();
}


int setup_geometry(tiltgpu_session* s, const std::vector<tiltgpu_view_params>& vp);
// Backprojection tile depth and LDS capacity of a geometry (allocates nothing):
// the tuned table's depth for the kernel variant, else the padded-depth rule,
// falling back to shallower tiles until the per-view LDS bound fits 64 KiB;
// forced > 0 pins the depth.  local_src supplies the local-shift extents when
// has_local.
struct BpTileChoice {
    int tile_z, lds_cap, source;
    int vps = 1, xpl = 1;  // views per LDS stage, x positions per lane
};
// forced_vps / forced_xpl: 0 = measured table; allow_shapes false keeps (1, 1).
int choose_bp_tile(const tiltgpu_session* local_src, const tiltgpu_view_params* vp, int nv,
                   bool two_row, bool has_local, bool half, int in_nx, int out_nz, int forced,
                   BpTileChoice* out, int forced_vps = 0, int forced_xpl = 0,
                   bool allow_shapes = true);
// Whether a view table needs the two-row backprojection (x-axis tilt or local shifts).
bool needs_two_rows(const tiltgpu_view_params* vp, int nv, bool has_local);
int upload_local(tiltgpu_session* s);
int prepare_pipeline_geometry(tiltgpu_session* s);
int prepare_filter(tiltgpu_session* s);
int check_kernel_error(tiltgpu_session* s);
BpParams make_bp(tiltgpu_session* s, const void* proj, const float* scale, int y0, int y1,
                 int out_mode, float* out);
FilterParams make_filter(tiltgpu_session* s, const float* src, const float* sub, bool stretch);
FpParams make_fp(tiltgpu_session* s);
int run_bp(tiltgpu_session* s, const void* proj, const float* scale, int y0, int y1,
           int out_mode, float* out);
int run_filter(tiltgpu_session* s, const float* src, const float* sub, bool stretch,
               double* sq_out, double* row_sum = nullptr);
int run_forward(tiltgpu_session* s);
int run_means(tiltgpu_session* s);
const float* stack_src(const tiltgpu_session* s);  // the transformed stack, or raw
int ensure_aligned(tiltgpu_session* s);
int validate_dims(const tiltgpu_dims* d);
// Allocates the session's buffers on `stream` (nullptr: create its own stream).
// pool: optional buffer provider (a server's cache); its sessions also skip
// per-stage event timing and never synchronise on destruction.
int session_create(const tiltgpu_dims* dims, const tiltgpu_options* opts, hipStream_t stream,
                   tiltgpu_session** out, BufferPool* pool = nullptr);
int pick_pairs_per_wg(int n);

}  // namespace tiltgpu_sess
