// Internal declarations shared by the libtiltgpu translation units: kernel
// parameter blocks and the host-side launch functions.  Not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tiltgpu {

constexpr int kWave = 64;        // CDNA wavefront width
constexpr int kBpThreads = 256;  // 4 waves per backprojection workgroup
constexpr int kBpTileX = 64;     // one wave spans 64 output x positions
constexpr int kMaxRadix = 20;

// floor(x) as an int in one instruction (|x| well inside the int range).
__device__ __forceinline__ int cvt_floor(float x) {
    int i;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(i) : "v"(x));
    return i;
}

// Workgroup id -> logical work item.  Workgroups are dispatched round-robin over
// the 8 XCDs; giving each XCD a contiguous run of logical items keeps items that
// read the same rows (projection rows, volume slices) on one XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblk) {
    const int X = 8;
    const int q = nblk / X, r = nblk % X;
    const int xcd = bid % X, j = bid / X;
    return (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + j;
}

// Per-view constants of the backprojection, read through the scalar cache.
// shift_mode: 0 = no local shift, 1 = constant shift (sdx, sdy) (a one-patch
// grid: ViewPatches::interpolate returns its single value), 2 = patch grid
// interpolated per (ox, iy).
struct alignas(32) BpView {
    float cos_t, sin_t, cos_xt, sin_xt;
    float sdx, sdy;
    int shift_mode;
    int pad;
};

// Backprojection of one series (one launch covers Y rows [y0, y1)).
//
// Layout in HBM:
//   proj    [nz][in_ny][in_nx]  storage type (f32, f16 or bf16)
//   out     [y1-y0][out_nz][out_nx] f32 (the MRC section order)
//   zc_tab  [nviews][nz_pad] f32: (dz * sin_t) * zfactor, dz = oz - out_nz/2
//   views   [nviews] BpView
struct BpParams {
    const void* proj;
    const float* view_scale;  // [nz] dequantisation scale (f16/bf16), else null
    const BpView* views;
    const float* zc_tab;
    const float* means;       // [nz] per-view means (Mean fill)
    float* out;
    int in_nx, in_ny, nviews, out_nx, out_nz, nz_pad;
    int y0, y1;
    int ntx, ntz;             // tiles along x (64 * xpl wide) and z (tile_z deep)
    float cx, ocx, inv_n;
    int fill_mode;
    int out_mode;             // 0: acc/nz, 1: output transform, 2: out += relax*acc/nz
    float dw, os, oa, relax;
    // local alignment grids (has_local != 0)
    int has_local;
    const int* lnx;
    const int* lny;
    const int* lxoff;
    const int* lyoff;
    const int* lgoff;
    const float* lxs;
    const float* lys;
    const float* ldx;
    const float* ldy;
    int lds_cap;              // capacity of one LDS buffer in entries
    int* err;                 // set to 1 when a tile needs more than lds_cap entries
    // optional: staged segment [lo, hi] of every (view, tile z, tile x) of this
    // launch's shape, for views without a patch grid ((p * ntz + tz) * ntx + tx);
    // null: each workgroup computes its segments itself
    const int2* seg_tab;
};

// Persistent batched backprojection: the tiles of nreq series (descs[r] covers
// tiles [tile_start[r], tile_start[r+1])) claimed from 8 per-XCD chunk counters.
struct BpBatch {
    const BpParams* descs;   // [nreq] device copies
    const int* tile_start;   // [nreq + 1]
    int nreq, total, chunk;  // chunk = ceil(total / 8)
    int* counters;           // [8], zero before the launch
};

// R-weighting filter + cosine stretch of rows of length n.
// Division by a runtime constant d >= 1 as a multiply-high and shifts
// (Granlund-Montgomery, exact for every 32-bit dividend).
struct FastDiv {
    uint32_t mul, shift, d;
    __host__ __device__ uint32_t div(uint32_t x) const {
        if (d == 1) return x;
        const uint32_t t = static_cast<uint32_t>((static_cast<uint64_t>(x) * mul) >> 32);
        return (t + ((x - t) >> 1)) >> (shift - 1);
    }
};
inline FastDiv make_fastdiv(uint32_t d) {
    FastDiv f{0, 0, d};
    if (d > 1) {
        uint32_t l = 0;
        while ((1ull << l) < d) ++l;
        f.shift = l;
        f.mul = static_cast<uint32_t>((((1ull << l) - d) << 32) / d + 1);
    }
    return f;
}

struct FilterParams {
    const float* src;       // [nrows][n]
    const float* src_sub;   // optional: rows are src - src_sub (SIRT residual)
    float* dst;             // [nrows][n]
    const float* weights;   // [n] radial weights W(k)
    const float2* tw;       // [n] exp(-2*pi*i*t/n)
    const float* stretch;   // [nviews] 1/cos(theta) or 0 (no stretch); null = none
    double* sq_partial;     // optional: per-workgroup sum of squares of the input rows
    double* row_sum;        // optional: f64 sum of each input row (power-of-two path, n >= 256)
    int n, nrows, rows_per_view;
    int pairs_per_wg;
    int cubic;
    int nradix;
    int radix[kMaxRadix];
    FastDiv n_div;                  // by n
    FastDiv m_div[kMaxRadix];       // by n / radix[s] (butterflies per row of pass s)
    FastDiv ns_div[kMaxRadix];      // by radix[0] * ... * radix[s-1] (pass s stride)
    FastDiv r_div[kMaxRadix];       // by radix[s] (outputs per butterfly, generic passes)
    // Power-of-two plan (64 <= n <= 2048): log2 n and the per-pass twiddle table
    // (pow2_twiddles); tw8 == null selects the mixed-radix passes above.
    int log2n;
    const float2* tw8;
};

// Per-pass twiddles of the power-of-two plan (radix-8 passes, then one
// radix-4 or radix-2 pass): for pass s of radix R and span NS, entries
// [off_s + k*R + r] = exp(-2*pi*i*r*k/(NS*R)), k < NS, r < R; returns the entry
// count (0 when n is not a power of two in [min_n, max_n]).  out may be null.
int pow2_twiddles(int n, float2* out, int max_n = 2048, int min_n = 64);

// Ordered-walk forward projection (no x-axis tilt, no local alignment).
// Layout: vol [in_ny][out_nz][out_nx] f32, sino [nz][in_ny][in_nx] f32.
struct FpParams {
    const float* vol;     // [ny][out_nz][out_nx]
    float* sino;          // [nz][ny][in_nx]
    const float4* vgeom;
    const float* zc_tab;
    int in_nx, in_ny, nviews, out_nx, out_nz, nz_pad;
    float cx, ocx;
    int bins_per_lane;    // 4, 8 or 16 (tuned_fp_bins)
    int rows_per_wg;      // volume rows walked per workgroup: 1, or 2 (rows up to kFpRows2Nx wide)
    int slide_windows;    // 1: each view's lane windows narrow to its cos and follow the
                          // z row's bins (staged walks); 0: fixed windows of bins_per_lane
};
// Widest volume row the two-rows-per-workgroup walk stages (wider rows walk one row).
constexpr int kFpRows2Nx = 2048;

// Scatter-form forward projection (x-axis tilt and/or local alignment).
struct FpScatterParams {
    const float* vol;
    float* sino;
    const float4* vgeom;
    const float* zc_tab;
    int in_nx, in_ny, nviews, out_nx, out_nz, nz_pad;
    int nz_sections;  // sections of the sinogram (>= nviews)
    float cx, ocx;
    int has_local;
    const int* lnx;
    const int* lny;
    const int* lxoff;
    const int* lyoff;
    const int* lgoff;
    const float* lxs;
    const float* lys;
    const float* ldx;
    const float* ldy;
    // fixed-point accumulation: acc [nz][in_ny][in_nx] 64-bit sums of the terms
    // rounded to multiples of 2^-(*scale_exp); scale_exp is set on the device
    unsigned long long* acc;
    const int* scale_exp;
};

// Several series in one filter / forward-projection launch: workgroups
// [block_start[r], block_start[r+1]) run series r.
struct FilterBatch {
    const FilterParams* descs;
    const int* block_start;
    int nreq;
};
struct FpBatch {
    const FpParams* descs;
    const int* block_start;
    int nreq;
};

// ---- launchers (kernels_*.hip); all asynchronous on `st` ----
// Launch shape: tile depth tile_z, vps views staged per barrier (1, 2) and xpl
// x positions per lane (tile x 64 * xpl); bp_shape_supported lists the built
// combinations.  The batched kernel runs shape (1, 1).
// The staged segments of launch_backproject's tiles (BpParams::seg_tab) for
// tile depth tile_z and xpl x positions per lane: nviews * ntz * ntx entries.
hipError_t launch_bp_segments(const BpParams& p, int tile_z, int xpl, int2* seg, hipStream_t st);
hipError_t launch_backproject(const BpParams& p, int storage, int two_row, int tile_z, int vps,
                              int xpl, hipStream_t st);
bool bp_shape_supported(int tile_z, int vps, int xpl);
size_t bp_lds_bytes(int lds_cap, int two_row, int vps = 1);
hipError_t launch_backproject_batch(const BpBatch& b, int storage, int two_row, int local,
                                    int tile_z, int grid, size_t lds, hipStream_t st);

hipError_t launch_filter(const FilterParams& p, hipStream_t st);
hipError_t launch_filter_batch(const FilterBatch& b, int total_blocks, size_t lds, hipStream_t st);
int filter_blocks(const FilterParams& p);
size_t filter_lds_bytes(int n, int pairs_per_wg);
// Largest dynamic LDS a filter workgroup may take: the 160 KiB of a gfx950 CU
// less 256 bytes for the kernels' own static LDS (rows up to 10224 wide on the
// mixed-radix path).
constexpr size_t kFilterMaxLds = 160 * 1024 - 256;
// Lets `kernel` launch with `lds` bytes of dynamic LDS (more than the default
// 64 KiB, up to kFilterMaxLds); remembers what it granted (thread-safe).
hipError_t lds_opt_in(const void* kernel, size_t lds);
int filter_plan(int n, int* radix, int* nradix);  // 0 ok, -1 unsupported n

hipError_t launch_log_means(const float* proj, float* logd, int64_t nxy, int nz, int do_log,
                            double* partial, int nblk, float* means, hipStream_t st);
// Per-view means from per-row f64 sums (row_sum [nz][rows]): each view's rows
// summed in a fixed order, mean = sum / (rows * n).
hipError_t launch_means_from_rows(const double* row_sum, int rows, int64_t nxy, int nz,
                                  float* means, hipStream_t st);
hipError_t launch_convert_storage(const float* src, void* dst, float* scale,
                                  unsigned* absmax_bits, int64_t nxy, int nz, int storage,
                                  hipStream_t st);
hipError_t launch_forward_gather(const FpParams& p, hipStream_t st);
// rows_per_wg: the walk form of every series of the batch (FpParams::rows_per_wg)
hipError_t launch_forward_gather_batch(const FpBatch& b, int total_blocks, size_t lds,
                                       int rows_per_wg, hipStream_t st);
size_t fp_lds_bytes(const FpParams& p);  // dynamic LDS of one series' workgroups
int fp_gather_blocks(const FpParams& p);
// Two-row forward projection, deterministic: the terms of main.rs:866-889 as
// 64-bit fixed point summed with integer atomics (order-independent), then
// converted to FP32.  absmax_scratch: one unsigned; scale_exp: one int.
hipError_t launch_forward_scatter(const FpScatterParams& p, unsigned* absmax_scratch,
                                  int* scale_exp, hipStream_t st);
hipError_t launch_sum_partials(const double* partial, int n, double* out, int sqrt_result,
                               hipStream_t st);
hipError_t launch_transform(float* v, int64_t n, float dw, float os, float oa,
                            hipStream_t st);
hipError_t launch_lowpass_x(const float* src, float* dst, int ny, int nz, int nx,
                            hipStream_t st);
hipError_t launch_lowpass_z(const float* src, float* dst, int ny, int nz, int nx,
                            hipStream_t st);
hipError_t launch_lowpass_y(const float* src, float* dst, int ny, int nz, int nx,
                            hipStream_t st);
hipError_t launch_slice_stats(const float* v, int rows, int64_t per_row, float* stats,
                              hipStream_t st);
hipError_t launch_fill_f32(float* p, int64_t n, float value, hipStream_t st);
hipError_t launch_copy(const void* src, void* dst, int64_t bytes, hipStream_t st);
// VALU issue-rate calibration kernels (kernels_misc.hip): mix 0..4, and the
// wave instructions one lane-loop iteration of a mix issues.
hipError_t launch_issue_rate(int mix, float* out, int blocks, int iters, hipStream_t st);
int issue_rate_instructions(int mix);
hipError_t launch_geometry_f64(const float* angles, int n_angles, const float* xtilt,
                               int n_xtilt, const float* zfac, int n_zfac, int nviews, int nsec,
                               float max_tilt, int stretch, float* view_out, float* stretch_out,
                               hipStream_t st);
hipError_t launch_apply_xf(const float* in, float* out, int nx, int ny, int nz, const float* inv,
                           float fill, hipStream_t st);
hipError_t launch_zc_f64(const float* angles, const float* zfac, int n_zfac, int nviews,
                         int nz_pad, int out_nz, float* zc, hipStream_t st);

}  // namespace tiltgpu
