// Allocation-free backprojection: the computation of tiltgpu_bp_session_new +
// tiltgpu_reconstruct_rows on caller-owned device buffers, with the per-view
// tables it needs kept in a fixed static device array instead of buffers the
// library allocates.
#include <hip/hip_runtime.h>

#include <mutex>

#include "session_internal.h"

// The scratch area: one per GPU, held until the process ends.  Layout of one
// call (byte offsets, each region 256-byte aligned):
//   err    int               kernel precondition flag (cleared every call)
//   params [nv] view params   copied from the caller's host table
//   means  [nz] f32           copied from the caller's host table
//   views  [nv] BpView
//   zc     [nv][nz_pad] f32   (dz * sin_t) * zfactor
static __device__ __align__(256) unsigned char scratch[TILTGPU_NOALLOC_SCRATCH_BYTES];

namespace tiltgpu_sess {
namespace {

constexpr size_t kScratchBytes = static_cast<size_t>(TILTGPU_NOALLOC_SCRATCH_BYTES);

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

// The device address of this GPU's scratch array (looked up once per device,
// with that device current).  The array's name is a host-side stand-in and is
// never passed to a kernel or a copy.
int scratch_address(int dev, unsigned char** out) {
    if (dev < 0 || dev >= kMaxDevices) return fail(TILTGPU_ERR_ARG, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(TILTGPU_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s",
                        hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_ptr[dev];
    return TILTGPU_OK;
}

// Builds the per-view BpView table and the z terms from the copied view
// parameters, exactly as setup_geometry does on the host (FP32, no fused
// multiply-add: the library is built with -ffp-contract=off), and clears the
// error flag.  One thread per (view, oz) entry.
__global__ void noalloc_tables_kernel(const tiltgpu_view_params* __restrict__ vp, int nv,
                                      int nz_pad, int out_nz, BpView* __restrict__ views,
                                      float* __restrict__ zc, int* __restrict__ err) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t n = static_cast<int64_t>(nv) * nz_pad;
    if (i == 0) *err = 0;
    if (i < nv) {
        BpView b;
        b.cos_t = vp[i].cos_t;
        b.sin_t = vp[i].sin_t;
        b.cos_xt = vp[i].cos_xt;
        b.sin_xt = vp[i].sin_xt;
        b.sdx = 0.f;
        b.sdy = 0.f;
        b.shift_mode = 0;
        b.pad = 0;
        views[i] = b;
    }
    if (i < n) {
        const int v = static_cast<int>(i / nz_pad);
        const int oz = static_cast<int>(i - static_cast<int64_t>(v) * nz_pad);
        const float center_z = static_cast<float>(out_nz) / 2.0f;
        const float dz = static_cast<float>(oz) - center_z;
        zc[i] = dz * vp[v].sin_t * vp[v].zfactor;
    }
}

}  // namespace
}  // namespace tiltgpu_sess

using namespace tiltgpu_sess;

extern "C" int tiltgpu_reconstruct_rows_noalloc(
    const float* d_projections, const tiltgpu_view_params* params, const float* proj_means,
    int32_t n_views, int32_t in_nx, int32_t in_ny, int32_t nz, int32_t out_nx, int32_t out_nz,
    int32_t fill_mode, int32_t y0, int32_t y1, float* d_out, void* stream) {
    tiltgpu_dims d{in_nx, in_ny, nz, out_nx, out_nz};
    int rc = validate_dims(&d);
    if (rc) return rc;
    if (!d_projections || !proj_means || !d_out || n_views < 0 || n_views > nz ||
        (n_views && !params))
        return fail(TILTGPU_ERR_ARG, "bad arguments");
    if (y0 < 0 || y1 > in_ny || y0 >= y1) return fail(TILTGPU_ERR_ARG, "bad row range");
    const int fill = (fill_mode >= 0 && fill_mode <= 2) ? fill_mode : TILTGPU_FILL_MEAN;
    const int nv = n_views;
    const bool two_row = needs_two_rows(params, nv, false);
    BpTileChoice tc{};
    if ((rc = choose_bp_tile(nullptr, params, nv, two_row, false, false, in_nx, out_nz, 0, &tc, 1, 1,
                             false)))
        return rc;
    const int nz_pad = (out_nz + tc.tile_z - 1) / tc.tile_z * tc.tile_z;
    const size_t off_params = 256;
    const size_t off_means = off_params + align256(sizeof(tiltgpu_view_params) * (nv ? nv : 1));
    const size_t off_views = off_means + align256(sizeof(float) * static_cast<size_t>(nz));
    const size_t off_zc = off_views + align256(sizeof(BpView) * (nv ? nv : 1));
    const size_t need = off_zc + align256(sizeof(float) * static_cast<size_t>(nv ? nv : 1) * nz_pad);
    if (need > kScratchBytes)
        return fail(TILTGPU_ERR_ARG,
                    "allocation-free backprojection: %d views x %d padded slices need %zu bytes "
                    "of scratch, more than its %zu (use tiltgpu_bp_session_new)",
                    nv, nz_pad, need, kScratchBytes);
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    unsigned char* base = nullptr;
    if ((rc = scratch_address(dev, &base))) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    auto* err = reinterpret_cast<int*>(base);
    auto* vp = reinterpret_cast<tiltgpu_view_params*>(base + off_params);
    auto* means = reinterpret_cast<float*>(base + off_means);
    auto* views = reinterpret_cast<BpView*>(base + off_views);
    auto* zc = reinterpret_cast<float*>(base + off_zc);
    if (nv)
        HIP_TRY(hipMemcpyAsync(vp, params, sizeof(tiltgpu_view_params) * nv,
                               hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(means, proj_means, sizeof(float) * nz, hipMemcpyHostToDevice, st));
    {
        const int64_t n = std::max<int64_t>(1, static_cast<int64_t>(nv) * nz_pad);
        const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
        hipLaunchKernelGGL(noalloc_tables_kernel, dim3(blocks), dim3(256), 0, st, vp, nv, nz_pad,
                           out_nz, views, zc, err);
        HIP_TRY(hipGetLastError());
    }
    BpParams p{};
    p.proj = d_projections;
    p.view_scale = nullptr;
    p.views = views;
    p.zc_tab = zc;
    p.means = means;
    p.out = d_out;
    p.in_nx = in_nx;
    p.in_ny = in_ny;
    p.nviews = nv;
    p.out_nx = out_nx;
    p.out_nz = out_nz;
    p.nz_pad = nz_pad;
    p.y0 = y0;
    p.y1 = y1;
    p.ntx = (out_nx + kBpTileX - 1) / kBpTileX;
    p.ntz = nz_pad / tc.tile_z;
    p.cx = static_cast<float>(in_nx) / 2.0f;
    p.ocx = static_cast<float>(out_nx) / 2.0f;
    p.inv_n = 1.0f / static_cast<float>(nz);
    p.fill_mode = fill;
    p.out_mode = 0;
    p.dw = 1.f;
    p.os = 1.f;
    p.oa = 0.f;
    p.relax = 1.f;
    p.has_local = 0;
    p.lds_cap = tc.lds_cap;
    p.err = err;
    HIP_TRY(launch_backproject(p, TILTGPU_STORE_F32, two_row ? 1 : 0, tc.tile_z, 1, 1, st));
    int e = 0;
    HIP_TRY(hipMemcpyAsync(&e, err, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (e)
        return fail(TILTGPU_ERR_KERNEL,
                    "a backprojection tile needed more LDS entries than its bound (%d)",
                    tc.lds_cap);
    return TILTGPU_OK;
}
