// Backprojection kernels for gfx950: the weighted-backprojection hot loop of
// imod-rs apps/tilt/src/main.rs:1011-1073 (== backproject_all, :759-826).
//
// Work decomposition.  A workgroup (4 waves) owns one output Y row iy and a
// 64 (x) by 4*ZPT (z) tile of that row's XZ slice.  Lane l of every wave owns
// ox = tile_x*64 + l; wave w owns ZPT consecutive oz values, so the per-(oz,view)
// term (dz*sin_t)*zfactor is wave-uniform and comes from a table through
// scalar loads.  FP32 accumulators live in registers for all views.
//
// LDS staging.  Per view, the workgroup stages the projection x-segment the
// tile can touch as (v0, v1-v0) pairs -- or, for the two-row (x-axis tilt /
// local alignment) path, (v0_r, d_r, v0_r+1, d_r+1) quads -- with one padded
// entry on each side that holds (fill, 0).  Clamping floor(proj_x) into
// [-1, in_nx-1] then reproduces sample_row_with_fill (main.rs:386-406)
// without a per-sample select: v = v0 + frac*d is the lerp inside the row and
// exactly the fill value outside it.
//
// Exactness.  The file is compiled with -ffp-contract=off, so the coordinate
// chain proj_x = ((dx*cos_t)*cos_xt + (dz*sin_t)*zf) + cx [+ ldx] and
// sample_y = (iy + dx*sin_xt) [+ ldy] round exactly as the reference's, and
// every in-range / fill / row-selection decision is the reference's.  Only the
// interpolation arithmetic (explicit fmaf) and the accumulation differ by
// rounding.
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <climits>

#include "tiltgpu_internal.h"

namespace tiltgpu {

// Workgroup id -> logical tile.  Workgroups are dispatched round-robin over the
// 8 XCDs; giving each XCD a contiguous run of logical tiles keeps the tiles of
// one Y row (which read the same projection rows) on one XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblk) {
    const int X = 8;
    const int q = nblk / X, r = nblk % X;
    const int xcd = bid % X, j = bid / X;
    return (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + j;
}

template <int ST>
__device__ __forceinline__ float load_proj(const void* base, long i, float scale) {
    if constexpr (ST == 0) {
        return static_cast<const float*>(base)[i];
    } else if constexpr (ST == 1) {
        const unsigned short u = static_cast<const unsigned short*>(base)[i];
        return __half2float(__ushort_as_half(u)) * scale;
    } else {
        const unsigned short u = static_cast<const unsigned short*>(base)[i];
        return __uint_as_float(static_cast<unsigned>(u) << 16) * scale;
    }
}

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

// ViewPatches::find_bracket, main.rs:200-228.
__device__ __forceinline__ void find_bracket(const float* pos, int len, float val, int& idx,
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

// ViewPatches::interpolate, main.rs:172-198 (lerp a + t*(b-a), unfused).
// Returns false when the view has no patches (no shift is applied then).
template <typename PP>
__device__ __forceinline__ bool local_shift(const PP& P, int view, float x, float y,
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
    find_bracket(xs, nx, x, ix0, fx);
    find_bracket(ys, ny, y, iy0, fy);
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

// One staged entry of row `row` (pointer to its first element) at x:
// (v[x], v[x+1]-v[x]) inside, (fill, 0) on the padded sides.
template <int ST>
__device__ __forceinline__ float2 stage_entry(const void* proj, long row_off, int x, int n,
                                              float scale, int fill_mode, float mean) {
    if (x >= 0 && x <= n - 2) {
        const float u0 = load_proj<ST>(proj, row_off + x, scale);
        const float u1 = load_proj<ST>(proj, row_off + x + 1, scale);
        return make_float2(u0, u1 - u0);
    }
    float f;
    if (fill_mode == 0)
        f = mean;
    else if (fill_mode == 1)
        f = 0.f;
    else
        f = load_proj<ST>(proj, row_off + (x < 0 ? 0 : n - 1), scale);
    return make_float2(f, 0.f);
}

template <int ZPT, bool TWO, bool LOCAL, int ST>
__global__ void __launch_bounds__(kBpThreads) bp_kernel(BpParams P) {
    extern __shared__ float4 lds4[];
    constexpr int TZ = 4 * ZPT;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int logical = xcd_remap(blockIdx.x, gridDim.x);
    const int tx = logical % P.ntx;
    const int rest = logical / P.ntx;
    const int tz = rest % P.ntz;
    const int iy = P.y0 + rest / P.ntz;
    const int ox = tx * kBpTileX + lane;
    const int zt0 = tz * TZ;
    const int zw0 = zt0 + wave * ZPT;
    const int n = P.in_nx;
    const float nm1 = static_cast<float>(n - 1);
    const float dx = static_cast<float>(ox) - P.ocx;
    const long nxy = static_cast<long>(n) * P.in_ny;

    float acc[ZPT];
#pragma unroll
    for (int k = 0; k < ZPT; ++k) acc[k] = 0.f;
    int staged = 0;  // views staged so far (workgroup-uniform): selects the LDS buffer

    for (int p = 0; p < P.nviews; ++p) {
        const float4 g = P.vgeom[p];
        const float a = (dx * g.x) * g.z;
        float ldx = 0.f, ldy = 0.f;
        bool shifted = false;
        if constexpr (LOCAL) shifted = local_shift(P, p, static_cast<float>(ox),
                                                   static_cast<float>(iy), ldx, ldy);
        const float* zc = P.zc_tab + static_cast<long>(p) * P.nz_pad;
        const float z_first = zc[zt0], z_last = zc[zt0 + TZ - 1];
        float pa = (a + z_first) + P.cx;
        float pb = (a + z_last) + P.cx;
        if constexpr (LOCAL) {
            if (shifted) {
                pa = pa + ldx;
                pb = pb + ldx;
            }
        }
        float lo = fminf(pa, pb), hi = fmaxf(pa, pb);
        if constexpr (LOCAL) {
            lo = wave_min(lo);
            hi = wave_max(hi);
        } else {
            // proj_x is monotonic in ox for a fixed oz: the extremes are at lanes 0, 63.
            lo = fminf(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(lo), 0)),
                       __int_as_float(__builtin_amdgcn_readlane(__float_as_int(lo), 63)));
            hi = fmaxf(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(hi), 0)),
                       __int_as_float(__builtin_amdgcn_readlane(__float_as_int(hi), 63)));
        }
        const int seg_lo = static_cast<int>(fminf(fmaxf(floorf(lo), -1.f), nm1));
        const int seg_hi = static_cast<int>(fminf(fmaxf(floorf(hi), -1.f), nm1));
        const int L = seg_hi - seg_lo + 1;

        int my_row = -1;  // projection row pair start for this lane, -1 = no contribution
        float fy_eff = 0.f;
        int r_lo = 0, NR = 1;
        if constexpr (TWO) {
            const float yo = dx * g.w;
            float sy = static_cast<float>(iy) + yo;
            if constexpr (LOCAL) {
                if (shifted) sy = sy + ldy;
            }
            const float fl = floorf(sy);
            const float fy = sy - fl;
            const bool two = LOCAL || fabsf(yo) > 1e-6f;
            const float fny = static_cast<float>(P.in_ny);
            if (two && fl >= 0.f && fl + 1.f < fny) {
                my_row = static_cast<int>(fl);
                fy_eff = fy;
            } else if (fl >= 0.f && fl < fny) {
                my_row = static_cast<int>(fl);
                fy_eff = 0.f;
            }
            const int rmin = wave_min_i(my_row >= 0 ? my_row : INT_MAX);
            const int rmax = wave_max_i(my_row);
            if (rmax < 0) continue;  // no lane of this tile samples this view
            r_lo = rmin;
            NR = rmax - rmin + 1;
        }
        const int total = NR * L;
        if (total > P.lds_cap) {
            if (tid == 0) atomicOr(P.err, 1);
            return;
        }
        // buffer offset in entries (float4 entries for two rows, float2 for one)
        const int buf = (staged & 1) * P.lds_cap;
        ++staged;
        const float scale = (ST == 0 || P.view_scale == nullptr) ? 1.f : P.view_scale[p];
        const float mean = P.means[p];
        // -- stage the view into LDS buffer (staged & 1).  That buffer was last read
        //    by the compute of the view staged two steps ago, and every thread has
        //    passed the barrier of the view staged one step ago, so no thread still
        //    reads it.
        if constexpr (TWO) {
            for (int e = tid; e < total; e += kBpThreads) {
                const int ri = e / L;
                const int x = seg_lo + (e - ri * L);
                const int r = r_lo + ri;
                const long off0 = static_cast<long>(p) * nxy + static_cast<long>(r) * n;
                const float2 e0 = stage_entry<ST>(P.proj, off0, x, n, scale, P.fill_mode, mean);
                float2 e1 = make_float2(0.f, 0.f);
                if (r + 1 < P.in_ny)
                    e1 = stage_entry<ST>(P.proj, off0 + n, x, n, scale, P.fill_mode, mean);
                lds4[buf + e] = make_float4(e0.x, e0.y, e1.x, e1.y);
            }
        } else {
            float2* lds2 = reinterpret_cast<float2*>(lds4);
            const long off0 = static_cast<long>(p) * nxy + static_cast<long>(iy) * n;
            for (int e = tid; e < total; e += kBpThreads)
                lds2[buf + e] =
                    stage_entry<ST>(P.proj, off0, seg_lo + e, n, scale, P.fill_mode, mean);
        }
        __syncthreads();

        // -- accumulate this view into the ZPT voxels of the lane
        if constexpr (TWO) {
            if (my_row >= 0) {
                const float4* E = lds4 + buf + (my_row - r_lo) * L - seg_lo;
#pragma unroll
                for (int k = 0; k < ZPT; ++k) {
                    float px = (a + zc[zw0 + k]) + P.cx;
                    if constexpr (LOCAL) {
                        if (shifted) px = px + ldx;
                    }
                    const float fl = floorf(px);
                    const int i = static_cast<int>(fminf(fmaxf(fl, -1.f), nm1));
                    const float f = px - fl;
                    const float4 q = E[i];
                    const float s0 = fmaf(f, q.y, q.x);
                    const float s1 = fmaf(f, q.w, q.z);
                    acc[k] += fmaf(fy_eff, s1 - s0, s0);
                }
            }
        } else {
            const float2* E = reinterpret_cast<const float2*>(lds4) + buf - seg_lo;
#pragma unroll
            for (int k = 0; k < ZPT; ++k) {
                const float px = (a + zc[zw0 + k]) + P.cx;
                const float fl = floorf(px);
                const int i = static_cast<int>(fminf(fmaxf(fl, -1.f), nm1));
                const float f = px - fl;
                const float2 q = E[i];
                acc[k] += fmaf(f, q.y, q.x);
            }
        }
        // No second barrier: the next staged view goes to the other buffer.
    }

    if (ox >= P.out_nx) return;
    const long slice = static_cast<long>(P.out_nz) * P.out_nx;
    float* dst = P.out + static_cast<long>(iy - P.y0) * slice + ox;
#pragma unroll
    for (int k = 0; k < ZPT; ++k) {
        const int oz = zw0 + k;
        if (oz < P.out_nz) {
            const float v = acc[k] * P.inv_n;
            float* o = dst + static_cast<long>(oz) * P.out_nx;
            if (P.out_mode == 0)
                *o = v;
            else if (P.out_mode == 1)
                *o = (v * P.dw) * P.os + P.oa;
            else
                *o = *o + P.relax * v;
        }
    }
}

size_t bp_lds_bytes(int lds_cap, int two_row) {
    return static_cast<size_t>(2) * lds_cap * (two_row ? 16 : 8);
}

template <int ZPT, bool TWO, bool LOCAL, int ST>
static hipError_t launch_bp_t(const BpParams& p, hipStream_t st) {
    const long nblk = static_cast<long>(p.ntx) * p.ntz * (p.y1 - p.y0);
    if (nblk <= 0) return hipSuccess;
    const size_t lds = bp_lds_bytes(p.lds_cap, TWO);
    hipLaunchKernelGGL((bp_kernel<ZPT, TWO, LOCAL, ST>), dim3(static_cast<unsigned>(nblk)),
                       dim3(kBpThreads), lds, st, p);
    return hipGetLastError();
}

template <int ZPT>
static hipError_t launch_bp_z(const BpParams& p, int storage, int two_row, hipStream_t st) {
    const bool local = p.has_local != 0;
    if (!two_row && local) return hipErrorInvalidValue;
#define TILTGPU_BP_CASE(TW, LC, S)                         \
    if (two_row == TW && local == LC && storage == S)      \
        return launch_bp_t<ZPT, TW, LC, S>(p, st);
    TILTGPU_BP_CASE(false, false, 0)
    TILTGPU_BP_CASE(false, false, 1)
    TILTGPU_BP_CASE(false, false, 2)
    TILTGPU_BP_CASE(true, false, 0)
    TILTGPU_BP_CASE(true, false, 1)
    TILTGPU_BP_CASE(true, false, 2)
    TILTGPU_BP_CASE(true, true, 0)
    TILTGPU_BP_CASE(true, true, 1)
    TILTGPU_BP_CASE(true, true, 2)
#undef TILTGPU_BP_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_backproject(const BpParams& p, int storage, int two_row, int tile_z,
                              hipStream_t st) {
    switch (tile_z) {
        case 32: return launch_bp_z<8>(p, storage, two_row, st);
        case 64: return launch_bp_z<16>(p, storage, two_row, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace tiltgpu
