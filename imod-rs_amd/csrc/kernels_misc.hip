// Memory-bound helper kernels for gfx950: --log and per-view means, FP16/BF16
// storage conversion, forward projection (SIRT), output transform, fake-SIRT
// low-pass, per-section statistics.
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <algorithm>
#include <cfloat>

#include "tiltgpu_internal.h"

namespace tiltgpu {

template <typename T>
__device__ __forceinline__ T block_sum(T v, T* red) {
    const int tid = threadIdx.x;
    red[tid] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    const T r = red[0];
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------- log + means
// --log (main.rs:607-613): v = log10(max(v, 0.001)) into `logd`; then the f64 sum
// of each view (min_max_mean, crates/imod-math/src/stats.rs:844-861) as
// per-block partials [nz][nblk], summed in a fixed order by means_finalize.
__global__ void __launch_bounds__(256) log_sum_kernel(const float* P, float* logd,
                                                      int64_t nxy, int do_log,
                                                      double* partial) {
    __shared__ double red[256];
    const int view = blockIdx.y;
    const int64_t chunk = (nxy + gridDim.x - 1) / gridDim.x;
    const int64_t i0 = static_cast<int64_t>(blockIdx.x) * chunk;
    const int64_t i1 = min(nxy, i0 + chunk);
    const float* base = P + static_cast<int64_t>(view) * nxy;
    float* lbase = logd + static_cast<int64_t>(view) * nxy;
    double s = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        float v = base[i];
        if (do_log) {
            v = log10f(fmaxf(v, 0.001f));
            lbase[i] = v;
        }
        s += static_cast<double>(v);
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) partial[static_cast<int64_t>(view) * gridDim.x + blockIdx.x] = s;
}

__global__ void means_finalize(const double* partial, int nblk, int64_t nxy, float* means) {
    const int view = blockIdx.x;
    if (threadIdx.x != 0) return;
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += partial[static_cast<int64_t>(view) * nblk + b];
    means[view] = nxy > 0 ? static_cast<float>(s / static_cast<double>(nxy)) : 0.f;
}

hipError_t launch_log_means(const float* proj, float* logd, int64_t nxy, int nz, int do_log,
                            double* partial, int nblk, float* means, hipStream_t st) {
    if (nz <= 0) return hipSuccess;
    hipLaunchKernelGGL(log_sum_kernel, dim3(nblk, nz), dim3(256), 0, st, proj, logd, nxy,
                       do_log, partial);
    hipLaunchKernelGGL(means_finalize, dim3(nz), dim3(64), 0, st, partial, nblk, nxy, means);
    return hipGetLastError();
}

// -------------------------------------------------------- storage conversion
// Per-view absolute maximum, scale = absmax / 2^15, then FP16 (RNE) or BF16
// (RNE) values v / scale.  Dequantisation is h * scale.
__global__ void __launch_bounds__(256) absmax_kernel(const float* src, int64_t nxy,
                                                     unsigned* absmax_bits) {
    __shared__ float red[256];
    const int view = blockIdx.y;
    const float* base = src + static_cast<int64_t>(view) * nxy;
    float m = 0.f;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < nxy;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        m = fmaxf(m, fabsf(base[i]));
    const int tid = threadIdx.x;
    red[tid] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] = fmaxf(red[tid], red[tid + s]);
        __syncthreads();
    }
    if (tid == 0) atomicMax(absmax_bits + view, __float_as_uint(red[0]));
}

__device__ __forceinline__ unsigned short f32_to_bf16_rne(float f) {
    const unsigned u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<unsigned short>((u >> 16) | 0x40u);
    const unsigned r = u + 0x7fffu + ((u >> 16) & 1u);
    return static_cast<unsigned short>(r >> 16);
}

__global__ void __launch_bounds__(256) convert_kernel(const float* src, unsigned short* dst,
                                                      const unsigned* absmax_bits,
                                                      float* scale_out, int64_t nxy,
                                                      int storage) {
    const int view = blockIdx.y;
    const float amax = __uint_as_float(absmax_bits[view]);
    const float scale = amax > 0.f ? amax / 32768.0f : 1.0f;
    if (blockIdx.x == 0 && threadIdx.x == 0) scale_out[view] = scale;
    const float inv = 1.0f / scale;
    const float* s = src + static_cast<int64_t>(view) * nxy;
    unsigned short* d = dst + static_cast<int64_t>(view) * nxy;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < nxy;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const float v = s[i] * inv;
        d[i] = storage == 1 ? __half_as_ushort(__float2half_rn(v)) : f32_to_bf16_rne(v);
    }
}

__global__ void zero_u32(unsigned* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

hipError_t launch_convert_storage(const float* src, void* dst, float* scale,
                                  unsigned* absmax_bits, int64_t nxy, int nz, int storage,
                                  hipStream_t st) {
    if (nz <= 0) return hipSuccess;
    hipLaunchKernelGGL(zero_u32, dim3((nz + 255) / 256), dim3(256), 0, st, absmax_bits, nz);
    const int gx = static_cast<int>(std::min<int64_t>(64, (nxy + 255) / 256));
    hipLaunchKernelGGL(absmax_kernel, dim3(gx, nz), dim3(256), 0, st, src, nxy, absmax_bits);
    hipLaunchKernelGGL(convert_kernel, dim3(gx, nz), dim3(256), 0, st, src,
                       static_cast<unsigned short*>(dst), absmax_bits, scale, nxy, storage);
    return hipGetLastError();
}

// ------------------------------------------------- forward projection (SIRT)
// forward_project, main.rs:831-896, in gather form for the single-row geometry
// (no x-axis tilt, no local alignment).  One lane owns one sinogram bin
// (view, iy, x) and visits the voxels whose floor(proj_x) is x-1 or x, in the
// reference's per-bin order (oz ascending, then ox ascending), adding
// val*frac / val*(1-frac) with the reference's rounding: the result equals
// the serial scatter bit for bit.
__device__ __forceinline__ void fp_gather_bin(const FpParams& P, int x, int iy, int p) {
    if (x >= P.in_nx) return;
    const float4 g = P.vgeom[p];
    const float c = g.x * g.z;
    const float* zc = P.zc_tab + static_cast<int64_t>(p) * P.nz_pad;
    const float* vrow = P.vol + static_cast<int64_t>(iy) * P.out_nz * P.out_nx;
    const float fx = static_cast<float>(x);
    const float fn = static_cast<float>(P.in_nx);
    float acc = 0.f;
    for (int oz = 0; oz < P.out_nz; ++oz) {
        const float zcv = zc[oz];
        int lo = 0, hi = P.out_nx - 1;
        if (fabsf(c) > 1e-4f) {
            const float t0 = P.ocx + ((fx - 1.0f) - P.cx - zcv) / c;
            const float t1 = P.ocx + ((fx + 1.0f) - P.cx - zcv) / c;
            const float tl = fminf(t0, t1), th = fmaxf(t0, t1);
            const float fnx = static_cast<float>(P.out_nx);
            lo = static_cast<int>(fminf(fmaxf(floorf(tl) - 1.0f, 0.0f), fnx));
            hi = static_cast<int>(fmaxf(fminf(ceilf(th) + 1.0f, fnx - 1.0f), -1.0f));
        }
        const float* vz = vrow + static_cast<int64_t>(oz) * P.out_nx;
        for (int ox = lo; ox <= hi; ++ox) {
            const float dx = static_cast<float>(ox) - P.ocx;
            const float px = ((dx * g.x) * g.z + zcv) + P.cx;
            const float fl = floorf(px);
            if (fl >= 0.f && fl + 1.f < fn) {
                const float frac = px - fl;
                if (fl == fx)
                    acc += vz[ox] * (1.0f - frac);
                else if (fl == fx - 1.0f)
                    acc += vz[ox] * frac;
            }
        }
    }
    P.sino[(static_cast<int64_t>(p) * P.in_ny + iy) * P.in_nx + x] = acc;
}

__global__ void __launch_bounds__(256) fp_gather_kernel(FpParams P) {
    fp_gather_bin(P, blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y, blockIdx.z);
}

// Several series in one launch: workgroup b of series r covers 256 bins of one
// (view, row); per-bin work and order are those of a launch of its own.
__global__ void __launch_bounds__(256) fp_gather_batch_kernel(FpBatch B) {
    int a = 0, c = B.nreq - 1;
    const int b = blockIdx.x;
    while (a < c) {
        const int m = (a + c + 1) / 2;
        if (B.block_start[m] <= b)
            a = m;
        else
            c = m - 1;
    }
    const FpParams& P = B.descs[a];
    const int local = b - B.block_start[a];
    const int nbx = (P.in_nx + 255) / 256;
    const int bx = local % nbx;
    const int rest = local / nbx;
    fp_gather_bin(P, bx * 256 + static_cast<int>(threadIdx.x), rest % P.in_ny, rest / P.in_ny);
}

int fp_gather_blocks(const FpParams& p) {
    return ((p.in_nx + 255) / 256) * p.in_ny * p.nviews;
}

hipError_t launch_forward_gather_batch(const FpBatch& b, int total_blocks, hipStream_t st) {
    if (total_blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(fp_gather_batch_kernel, dim3(total_blocks), dim3(256), 0, st, b);
    return hipGetLastError();
}

hipError_t launch_forward_gather(const FpParams& p, hipStream_t st) {
    dim3 grid((p.in_nx + 255) / 256, p.in_ny, p.nviews);
    if (p.nviews <= 0 || p.in_ny <= 0) return hipSuccess;
    hipLaunchKernelGGL(fp_gather_kernel, grid, dim3(256), 0, st, p);
    return hipGetLastError();
}

// Two-row geometry (x-axis tilt / local alignment): scatter with float atomics.
// Same coordinates and weights as main.rs:867-889; the summation order per bin
// is not the reference's (results agree to rounding, not bit for bit).
__device__ __forceinline__ void fb_find_bracket(const float* pos, int len, float val, int& idx,
                                                float& frac) {
    idx = 0;
    frac = 0.f;
    if (len == 1) return;
    int lo = 0, hi = len - 1;
    if (val <= pos[lo]) return;
    if (val >= pos[hi]) {
        idx = hi;
        return;
    }
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (pos[mid] <= val)
            lo = mid;
        else
            hi = mid;
    }
    const float span = pos[hi] - pos[lo];
    idx = lo;
    frac = span > 1e-12f ? (val - pos[lo]) / span : 0.f;
}

__device__ bool fp_local_shift(const FpScatterParams& P, int view, float x, float y,
                               float& odx, float& ody) {
    const int nx = P.lnx[view], ny = P.lny[view];
    odx = 0.f;
    ody = 0.f;
    if (nx == 0) return false;
    if (ny == 0) return true;
    const float* xs = P.lxs + P.lxoff[view];
    const float* ys = P.lys + P.lyoff[view];
    const float* gx = P.ldx + P.lgoff[view];
    const float* gy = P.ldy + P.lgoff[view];
    int ix0, iy0;
    float fx, fy;
    fb_find_bracket(xs, nx, x, ix0, fx);
    fb_find_bracket(ys, ny, y, iy0, fy);
    const int ix1 = min(ix0 + 1, nx - 1);
    const int iy1 = min(iy0 + 1, ny - 1);
    const float d00x = gx[iy0 * nx + ix0], d00y = gy[iy0 * nx + ix0];
    const float d10x = gx[iy0 * nx + ix1], d10y = gy[iy0 * nx + ix1];
    const float d01x = gx[iy1 * nx + ix0], d01y = gy[iy1 * nx + ix0];
    const float d11x = gx[iy1 * nx + ix1], d11y = gy[iy1 * nx + ix1];
    const float ax = d00x + fx * (d10x - d00x), bx = d01x + fx * (d11x - d01x);
    const float ay = d00y + fx * (d10y - d00y), by = d01y + fx * (d11y - d01y);
    odx = ax + fy * (bx - ax);
    ody = ay + fy * (by - ay);
    return true;
}

__global__ void __launch_bounds__(256) fp_scatter_kernel(FpScatterParams P) {
    const int ox = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y;
    const int p = blockIdx.z;
    if (ox >= P.out_nx) return;
    const float4 g = P.vgeom[p];
    const float dx = static_cast<float>(ox) - P.ocx;
    const float a = (dx * g.x) * g.z;
    const float yo = dx * g.w;
    float ldx = 0.f, ldy = 0.f;
    bool shifted = false;
    if (P.has_local)
        shifted = fp_local_shift(P, p, static_cast<float>(ox), static_cast<float>(iy), ldx, ldy);
    float sy = static_cast<float>(iy) + yo;
    if (shifted) sy = sy + ldy;
    const float fl_y = floorf(sy);
    const float fy = sy - fl_y;
    const float fny = static_cast<float>(P.in_ny);
    const bool two = P.has_local || fabsf(yo) > 1e-6f;
    int mode = 0;  // 0 none, 1 single row, 2 two rows
    if (fl_y >= 0.f && fl_y + 1.f < fny && two)
        mode = 2;
    else if (fl_y >= 0.f && fl_y < fny)
        mode = 1;
    if (mode == 0) return;
    const int sy0 = static_cast<int>(fl_y);
    const float* zc = P.zc_tab + static_cast<int64_t>(p) * P.nz_pad;
    const float fn = static_cast<float>(P.in_nx);
    const float* vcol = P.vol + static_cast<int64_t>(iy) * P.out_nz * P.out_nx + ox;
    float* r0 = P.sino + (static_cast<int64_t>(p) * P.in_ny + sy0) * P.in_nx;
    float* r1 = r0 + P.in_nx;
    const float w0 = 1.0f - fy, w1 = fy;
    for (int oz = 0; oz < P.out_nz; ++oz) {
        float px = (a + zc[oz]) + P.cx;
        if (shifted) px = px + ldx;
        const float fl = floorf(px);
        if (!(fl >= 0.f && fl + 1.f < fn)) continue;
        const int px0 = static_cast<int>(fl);
        const float frac = px - fl;
        const float val = vcol[static_cast<int64_t>(oz) * P.out_nx];
        if (mode == 2) {
            atomicAdd(r0 + px0, val * (1.0f - frac) * w0);
            atomicAdd(r0 + px0 + 1, val * frac * w0);
            atomicAdd(r1 + px0, val * (1.0f - frac) * w1);
            atomicAdd(r1 + px0 + 1, val * frac * w1);
        } else {
            atomicAdd(r0 + px0, val * (1.0f - frac));
            atomicAdd(r0 + px0 + 1, val * frac);
        }
    }
}

hipError_t launch_forward_scatter(const FpScatterParams& p, hipStream_t st) {
    if (p.nviews <= 0 || p.in_ny <= 0) return hipSuccess;
    dim3 grid((p.out_nx + 255) / 256, p.in_ny, p.nviews);
    hipLaunchKernelGGL(fp_scatter_kernel, grid, dim3(256), 0, st, p);
    return hipGetLastError();
}

// ----------------------------------------------------------- small helpers
__global__ void sum_partials_kernel(const double* partial, int n, double* out, int sq) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += partial[i];
    *out = sq ? sqrt(s) : s;
}

hipError_t launch_sum_partials(const double* partial, int n, double* out, int sqrt_result,
                               hipStream_t st) {
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(64), 0, st, partial, n, out,
                       sqrt_result);
    return hipGetLastError();
}

// Output transform (main.rs:965-969): v = (v * densweight) * outscale + outadd.
__global__ void transform_kernel(float* v, int64_t n, float dw, float os, float oa) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        v[i] = (v[i] * dw) * os + oa;
}

hipError_t launch_transform(float* v, int64_t n, float dw, float os, float oa,
                            hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int g = static_cast<int>(std::min<int64_t>(4096, (n + 255) / 256));
    hipLaunchKernelGGL(transform_kernel, dim3(g), dim3(256), 0, st, v, n, dw, os, oa);
    return hipGetLastError();
}

__global__ void fill_kernel(float* p, int64_t n, float value) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        p[i] = value;
}

hipError_t launch_fill_f32(float* p, int64_t n, float value, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int g = static_cast<int>(std::min<int64_t>(4096, (n + 255) / 256));
    hipLaunchKernelGGL(fill_kernel, dim3(g), dim3(256), 0, st, p, n, value);
    return hipGetLastError();
}

// lowpass_3d_volume, main.rs:493-538: 3-tap box along X, then Z, then Y, with
// clamped edges, each pass reading the previous pass's complete output.
__global__ void lowpass_x_kernel(const float* s, float* d, int ny, int nz, int nx) {
    const int64_t total = static_cast<int64_t>(ny) * nz * nx;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int ox = static_cast<int>(i % nx);
        const int64_t b = i - ox;
        const int x0 = ox > 0 ? ox - 1 : 0, x2 = ox + 1 < nx ? ox + 1 : nx - 1;
        d[i] = (s[b + x0] + s[b + ox] + s[b + x2]) / 3.0f;
    }
}
__global__ void lowpass_z_kernel(const float* s, float* d, int ny, int nz, int nx) {
    const int64_t total = static_cast<int64_t>(ny) * nz * nx;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int ox = static_cast<int>(i % nx);
        const int oz = static_cast<int>((i / nx) % nz);
        const int64_t b = i - ox - static_cast<int64_t>(oz) * nx;
        const int z0 = oz > 0 ? oz - 1 : 0, z2 = oz + 1 < nz ? oz + 1 : nz - 1;
        d[i] = (s[b + static_cast<int64_t>(z0) * nx + ox] + s[b + static_cast<int64_t>(oz) * nx + ox] +
                s[b + static_cast<int64_t>(z2) * nx + ox]) / 3.0f;
    }
}
__global__ void lowpass_y_kernel(const float* s, float* d, int ny, int nz, int nx) {
    const int64_t ns = static_cast<int64_t>(nz) * nx;
    const int64_t total = static_cast<int64_t>(ny) * ns;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int iy = static_cast<int>(i / ns);
        const int64_t r = i - static_cast<int64_t>(iy) * ns;
        const int y0 = iy > 0 ? iy - 1 : 0, y2 = iy + 1 < ny ? iy + 1 : ny - 1;
        d[i] = (s[y0 * ns + r] + s[static_cast<int64_t>(iy) * ns + r] + s[y2 * ns + r]) / 3.0f;
    }
}

static int grid_for(int64_t n) { return static_cast<int>(std::min<int64_t>(8192, (n + 255) / 256)); }

hipError_t launch_lowpass_x(const float* src, float* dst, int ny, int nz, int nx,
                            hipStream_t st) {
    const int64_t n = static_cast<int64_t>(ny) * nz * nx;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(lowpass_x_kernel, dim3(grid_for(n)), dim3(256), 0, st, src, dst, ny, nz, nx);
    return hipGetLastError();
}
hipError_t launch_lowpass_z(const float* src, float* dst, int ny, int nz, int nx,
                            hipStream_t st) {
    const int64_t n = static_cast<int64_t>(ny) * nz * nx;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(lowpass_z_kernel, dim3(grid_for(n)), dim3(256), 0, st, src, dst, ny, nz, nx);
    return hipGetLastError();
}
hipError_t launch_lowpass_y(const float* src, float* dst, int ny, int nz, int nx,
                            hipStream_t st) {
    const int64_t n = static_cast<int64_t>(ny) * nz * nx;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(lowpass_y_kernel, dim3(grid_for(n)), dim3(256), 0, st, src, dst, ny, nz, nx);
    return hipGetLastError();
}

// Per-section statistics (min_max_mean of each output Y row, main.rs:1080):
// stats[3*row] = min, max, mean (f64 sum -> f32).
__global__ void __launch_bounds__(256) slice_stats_kernel(const float* v, int64_t per_row,
                                                          float* stats) {
    __shared__ double reds[256];
    __shared__ float redmin[256], redmax[256];
    const int row = blockIdx.x;
    const float* b = v + static_cast<int64_t>(row) * per_row;
    float mn = FLT_MAX, mx = -FLT_MAX;
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < per_row; i += blockDim.x) {
        const float x = b[i];
        mn = fminf(mn, x);
        mx = fmaxf(mx, x);
        s += static_cast<double>(x);
    }
    const int tid = threadIdx.x;
    reds[tid] = s;
    redmin[tid] = mn;
    redmax[tid] = mx;
    __syncthreads();
    for (int k = blockDim.x / 2; k > 0; k >>= 1) {
        if (tid < k) {
            reds[tid] += reds[tid + k];
            redmin[tid] = fminf(redmin[tid], redmin[tid + k]);
            redmax[tid] = fmaxf(redmax[tid], redmax[tid + k]);
        }
        __syncthreads();
    }
    if (tid == 0) {
        stats[3 * row] = per_row ? redmin[0] : 0.f;
        stats[3 * row + 1] = per_row ? redmax[0] : 0.f;
        stats[3 * row + 2] = per_row ? static_cast<float>(reds[0] / static_cast<double>(per_row)) : 0.f;
    }
}

hipError_t launch_slice_stats(const float* v, int rows, int64_t per_row, float* stats,
                              hipStream_t st) {
    if (rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(slice_stats_kernel, dim3(rows), dim3(256), 0, st, v, per_row, stats);
    return hipGetLastError();
}

}  // namespace tiltgpu

namespace tiltgpu {

// Plain streaming copy (16 B per lane per iteration): the known-size reference
// against which HBM byte counters are calibrated.
__global__ void __launch_bounds__(256) copy_kernel(const float4* __restrict__ src,
                                                   float4* __restrict__ dst, int64_t n4) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n4;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

hipError_t launch_copy(const void* src, void* dst, int64_t bytes, hipStream_t st) {
    const int64_t n4 = bytes / 16;
    if (n4 <= 0) return hipSuccess;
    hipLaunchKernelGGL(copy_kernel, dim3(8192), dim3(256), 0, st,
                       static_cast<const float4*>(src), static_cast<float4*>(dst), n4);
    return hipGetLastError();
}

}  // namespace tiltgpu
