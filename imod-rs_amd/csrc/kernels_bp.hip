// Backprojection kernels for gfx950: the weighted-backprojection hot loop of
// imod-rs apps/tilt/src/main.rs:1011-1073 (== backproject_all, :759-826).
//
// Work decomposition.  A workgroup (4 waves) owns one output Y row iy and a
// 64 (x) by 4*ZPT (z) tile of that row's XZ slice.  Lane l of every wave owns
// ox = tile_x*64 + l; wave w owns ZPT consecutive oz values, so the per-(oz,view)
// term (dz*sin_t)*zfactor is wave-uniform and arrives through scalar loads.
// FP32 accumulators live in registers for all views.
//
// LDS staging.  Per view the workgroup stages one entry for every x in
// [floor(min proj_x), floor(max proj_x)] of the tile:
//   one-row geometry:  (v0, d0)            d0 = v[x+1]-v[x]
//   two-row geometry:  (v0, d0, dv, dd)    rows r and r+1, dv = v0'-v0, dd = d0'-d0
// for 0 <= x <= in_nx-2, and (fill, 0[, fill'-fill, 0]) for x outside that range
// (left fill for x < 0, right fill for x >= in_nx-1).  Indexing the entries with
// floor(proj_x) then reproduces sample_row_with_fill (main.rs:386-406) with no
// clamp and no per-sample select: v = v0 + frac*d is the lerp inside the row and
// exactly the fill value outside it.  The next view's entries are loaded into
// registers before the current view is accumulated, so the global-load latency
// hides behind the interpolation work; one barrier per view.
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

#define TILTGPU_CONSTANT __attribute__((address_space(4)))

namespace tiltgpu {

template <typename T>
__device__ __forceinline__ const TILTGPU_CONSTANT T* cptr(const T* p) {
    return (const TILTGPU_CONSTANT T*)(p);
}


// Raw projection sample as stored (f32 bits or 16-bit code); decoded after the
// compute of the previous view so the loads stay in flight meanwhile.
template <int ST>
__device__ __forceinline__ unsigned load_raw(const void* base, long i) {
    if constexpr (ST == 0)
        return __float_as_uint(static_cast<const float*>(base)[i]);
    else
        return static_cast<unsigned>(static_cast<const unsigned short*>(base)[i]);
}

template <int ST>
__device__ __forceinline__ float decode(unsigned u, float scale) {
    if constexpr (ST == 0)
        return __uint_as_float(u);
    else if constexpr (ST == 1)
        return __half2float(__ushort_as_half(static_cast<unsigned short>(u))) * scale;
    else
        return __uint_as_float(u << 16) * scale;
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
__device__ __forceinline__ float lane_f(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
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

// ViewPatches::interpolate, main.rs:172-198 (lerp a + t*(b-a), unfused), for a
// view with a patch grid.
__device__ __forceinline__ void grid_shift(const BpParams& P, int view, float x, float y,
                                           float& odx, float& ody) {
    const int nx = P.lnx[view], ny = P.lny[view];
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
}

constexpr int kSlots = 2;  // staged entries prefetched per thread (one view per stage)

// Staging geometry of one view for the workgroup's tile: wave-uniform and equal
// in all four waves (they hold the same x positions), so it lives in SGPRs.
struct Stg {
    int seg_lo, L, r_lo, NR, total, shift;
    float inv_L;  // 1/L (row of a linear entry index, two-row geometry)
};

// Per-view state of one lane: for each of its XPL x positions (ox + 64*j)
// a, ldx, fy and row, plus the view's staging geometry.
template <int XPL>
struct Vs : Stg {
    float a[XPL];    // (dx*cos_t)*cos_xt
    float ldx[XPL];  // x shift (0 when none)
    float fy[XPL];   // weight of row r+1 (0: one row)
    int row[XPL];    // row pair start r, or -1 (no contribution)
};

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// x offset of a lane within the 64-wide tile of the two-row kernel.  gfx950
// serves a ds_read_b128 16 lanes per cycle from 64 four-byte banks, in the lane
// groups {0-3, 12-15, 20-23, 24-27}, {4-11, 16-19, 28-31} and both + 32 (a
// 4-lane block b joins the group of the parity of popcount(b); measured with
// SQ_LDS_BANK_CONFLICT on single-wave reads).  Consecutive x per lane puts
// entries 16 apart into one group whenever the stride |cos| lies in (1/2, 1):
// 4 conflict cycles per read.  Giving each group 16 consecutive x positions
// makes its entries, at stride <= 1, at most 16 consecutive float4s: distinct
// banks.  Lanes 0 and 63 keep x offsets 0 and 63 (the extremes view_setup
// reads); a voxel's arithmetic does not depend on its lane.  The one-row
// kernel's ds_read_b64 is served in the consecutive halves 0-31 and 32-63,
// conflict-free at any stride <= 1, so it keeps lane = x.
__device__ __forceinline__ int b128_lane_x(int lane) {
    const int b = (lane >> 2) & 7;
    return (lane & 32) | ((__builtin_popcount(b) & 1) << 4) | ((b >> 1) << 2) | (lane & 3);
}

// floor(proj_x) at the extremes of z of one x position: the bounds of its
// staged entries (floor is monotonic).  With local shifts the shift is always
// added: ldx = +0 without one, and x + 0 == x for every x but -0 (same floor
// and fraction).
template <bool LOCAL>
__device__ __forceinline__ void seg_bounds(float a, float z_first, float z_last, float cx,
                                           float lx, int& flo, int& fhi) {
    float pa = (a + z_first) + cx;
    float pb = (a + z_last) + cx;
    if (LOCAL) {
        pa = pa + lx;
        pb = pb + lx;
    }
    flo = min(flo, cvt_floor(fminf(pa, pb)));
    fhi = max(fhi, cvt_floor(fmaxf(pa, pb)));
}

template <bool TWO, bool LOCAL, int XPL>
__device__ __forceinline__ Vs<XPL> view_setup(const BpParams& P, int p, const float (&dx)[XPL],
                                              const int (&ox)[XPL], int iy, int zt0, int TZ,
                                              int tile) {
    Vs<XPL> s;
    const TILTGPU_CONSTANT BpView* V = cptr(P.views) + p;
    const float cos_t = V->cos_t, cos_xt = V->cos_xt, sin_xt = V->sin_xt;
    const int mode = LOCAL ? V->shift_mode : 0;
    s.shift = mode;
    // the segment from the launch's table (views without a patch grid)
    const bool tab = P.seg_tab != nullptr && mode != 2;  // uniform
    float ldy[XPL];
    int flo = INT_MAX, fhi = INT_MIN;
#pragma unroll
    for (int j = 0; j < XPL; ++j) {
        s.a[j] = (dx[j] * cos_t) * cos_xt;
        float lx = 0.f, ly = 0.f;
        if (LOCAL) {
            if (mode == 1) {
                lx = V->sdx;
                ly = V->sdy;
            } else if (mode == 2) {
                grid_shift(P, p, static_cast<float>(ox[j]), static_cast<float>(iy), lx, ly);
            }
        }
        s.ldx[j] = lx;
        ldy[j] = ly;
        if (!tab) {
            const TILTGPU_CONSTANT float* zc = cptr(P.zc_tab) + static_cast<long>(p) * P.nz_pad;
            seg_bounds<LOCAL>(s.a[j], zc[zt0], zc[zt0 + TZ - 1], P.cx, lx, flo, fhi);
        }
    }
    int seg_lo, seg_hi;
    if (tab) {
        const TILTGPU_CONSTANT int* sg = reinterpret_cast<const TILTGPU_CONSTANT int*>(
            cptr(P.seg_tab) + static_cast<long>(p) * (P.ntz * P.ntx) + tile);
        seg_lo = sg[0];
        seg_hi = sg[1];
    } else if (mode == 2) {
        seg_lo = uni(wave_min_i(flo));
        seg_hi = uni(wave_max_i(fhi));
    } else {
        // proj_x is monotonic in ox (a constant shift keeps it so): the extremes
        // are at x offset 0 (lane 0, j = 0) and 64*XPL-1 (lane 63, j = XPL-1)
        seg_lo = min(__builtin_amdgcn_readlane(flo, 0), __builtin_amdgcn_readlane(flo, 63));
        seg_hi = max(__builtin_amdgcn_readlane(fhi, 0), __builtin_amdgcn_readlane(fhi, 63));
    }
    s.seg_lo = seg_lo;
    s.L = seg_hi - seg_lo + 1;
    s.r_lo = iy;
#pragma unroll
    for (int j = 0; j < XPL; ++j) {
        s.row[j] = iy;
        s.fy[j] = 0.f;
    }
    int NR = 1;
    if constexpr (TWO) {
        const float fny = static_cast<float>(P.in_ny);
        int ilo = INT_MAX, ihi = INT_MIN;  // floor(sy) over the lane's x positions
        int rlo = INT_MAX, rhi = -1;       // rows the lane uses
#pragma unroll
        for (int j = 0; j < XPL; ++j) {
            const float yo = dx[j] * sin_xt;
            float sy = static_cast<float>(iy) + yo;
            if (LOCAL) sy = sy + ldy[j];
            const float fl = floorf(sy);
            const int ifl = static_cast<int>(fl);
            const bool two = LOCAL || fabsf(yo) > 1e-6f;
            s.row[j] = -1;
            if (two && fl >= 0.f && fl + 1.f < fny) {
                s.row[j] = ifl;
                s.fy[j] = sy - fl;
            } else if (fl >= 0.f && fl < fny) {
                s.row[j] = ifl;
            }
            ilo = min(ilo, ifl);
            ihi = max(ihi, ifl);
            if (s.row[j] >= 0) {
                rlo = min(rlo, s.row[j]);
                rhi = max(rhi, s.row[j]);
            }
        }
        int rmin, rmax;
        if (mode == 2) {
            rmin = uni(wave_min_i(rlo));
            rmax = uni(wave_max_i(rhi));
        } else {
            // floor(sy) is monotonic in ox; clamp the extremes into the valid rows
            const int f0 = min(__builtin_amdgcn_readlane(ilo, 0), __builtin_amdgcn_readlane(ilo, 63));
            const int f1 = max(__builtin_amdgcn_readlane(ihi, 0), __builtin_amdgcn_readlane(ihi, 63));
            const int lo_r = max(f0, 0), hi_r = min(f1, P.in_ny - 1);
            rmin = lo_r <= hi_r ? lo_r : INT_MAX;
            rmax = lo_r <= hi_r ? hi_r : -1;
        }
        if (rmax < 0) {
            NR = 0;
            s.r_lo = 0;
        } else {
            s.r_lo = rmin;
            NR = rmax - rmin + 1;
        }
    }
    s.NR = NR;
    s.total = NR * s.L;
    s.inv_L = __builtin_amdgcn_rcpf(static_cast<float>(max(s.L, 1)));
    return s;
}

// Staged entry e (0 <= e < total) of a view is row ri = e / L of the staged rows
// and x = seg_lo + e % L; it lands at LDS index e.  The row comes from a float
// product: with e < 2^16 the error of (e + 0.5) * rcp(L) is below 1e-3 / L,
// far inside the 0.5 / L margin to the nearest integer.
template <bool TWO>
__device__ __forceinline__ void entry_pos(const Stg& s, int e, int& ri, int& x) {
    if constexpr (TWO) {
        ri = static_cast<int>(fmaf(static_cast<float>(e), s.inv_L, 0.5f * s.inv_L));
        x = s.seg_lo + (e - __mul24(ri, s.L));
    } else {
        ri = 0;
        x = s.seg_lo + e;
    }
}

// Raw samples of entry e: v[i0], v[i1] of row r (and of row r+1) with
// i0 = clamp(x, 0, n-1) and i1 = clamp(x+1, 0, n-1).  Outside [0, n-2] i0 == i1
// is the edge sample, so the difference of the pair is 0 there.  Row bases are
// wave-uniform; per lane only a 32-bit element offset is formed.  A row r+1
// past the last row (weight 0) is read as row r.
// min(max(x, 0), hi) in one v_med3_i32 (hi wave-uniform).
__device__ __forceinline__ int clamp_i(int x, int hi) {
    int r;
    asm("v_med3_i32 %0, %1, 0, %2" : "=v"(r) : "v"(x), "s"(hi));
    return r;
}

// Buffer resource over the samples of view p (range-checked to the view: an
// access outside it would read 0, never another allocation).  Word 3 is the
// gfx9-family raw-buffer format (32-bit data, no swizzle).
template <int ST>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t view_rsrc(const BpParams& P, int p) {
    constexpr long esz = ST == 0 ? 4 : 2;
    const long nxy = static_cast<long>(P.in_nx) * P.in_ny;
    const char* base = static_cast<const char*>(P.proj) + static_cast<long>(p) * nxy * esz;
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<char*>(base), 0,
                                             static_cast<int>(nxy * esz), 0x00020000);
}

template <int ST>
__device__ __forceinline__ unsigned load_at(__amdgpu_buffer_rsrc_t r, unsigned voff, int soff) {
    if constexpr (ST == 0)
        return __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
    else
        return static_cast<unsigned>(__builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0));
}

// Raw samples of entry e: v[i0], v[i1] of row r (and of row r+1) with
// i0 = clamp(x, 0, n-1) and i1 = clamp(x+1, 0, n-1).  Outside [0, n-2] i0 == i1
// is the edge sample, so the difference of the pair is 0 there.  The row
// offset is wave-uniform (scalar); per lane only a 32-bit byte offset is
// formed.  A row r+1 past the last row (weight 0) is read as row r.
template <bool TWO, int ST>
__device__ __forceinline__ void load_entry(const BpParams& P, const Stg& s, int p, int e,
                                           unsigned (&r)[4], int& x) {
    constexpr unsigned esz = ST == 0 ? 4u : 2u;
    const int n = P.in_nx;
    int ri;
    entry_pos<TWO>(s, e, ri, x);
    const int i0 = clamp_i(x, n - 1);
    const int i1 = clamp_i(x + 1, n - 1);
    const __amdgpu_buffer_rsrc_t rs = view_rsrc<ST>(P, p);
    const int srow = s.r_lo * n * static_cast<int>(esz);
    const int off = __mul24(ri, n);
    r[0] = load_at<ST>(rs, static_cast<unsigned>(off + i0) * esz, srow);
    r[1] = load_at<ST>(rs, static_cast<unsigned>(off + i1) * esz, srow);
    r[2] = r[3] = 0u;
    if constexpr (TWO) {
        if (s.r_lo + s.NR < P.in_ny) {  // every staged row has a row below it
            const int srow2 = srow + n * static_cast<int>(esz);
            r[2] = load_at<ST>(rs, static_cast<unsigned>(off + i0) * esz, srow2);
            r[3] = load_at<ST>(rs, static_cast<unsigned>(off + i1) * esz, srow2);
        } else {
            const int off2 = __mul24(min(ri + 1, P.in_ny - 1 - s.r_lo), n);
            r[2] = load_at<ST>(rs, static_cast<unsigned>(off2 + i0) * esz, srow);
            r[3] = load_at<ST>(rs, static_cast<unsigned>(off2 + i1) * esz, srow);
        }
    }
}

// Entry (v0, d0[, dv, dd]) from the raw samples: inside [0, n-2] the lerp
// coefficients of sample_row_with_fill, outside the fill value with slope 0
// (Edge fill: the edge sample itself, which the clamped loads delivered).
template <bool TWO, int ST>
__device__ __forceinline__ float4 make_entry(const BpParams& P, int x, const unsigned (&r)[4],
                                             float scale, float cfill) {
    const bool keep = P.fill_mode == 2 || static_cast<unsigned>(x) <= static_cast<unsigned>(P.in_nx - 2);
    const float a0 = decode<ST>(r[0], scale), a1 = decode<ST>(r[1], scale);
    const float v0 = keep ? a0 : cfill;
    const float d0 = a1 - a0;
    if constexpr (!TWO) {
        return make_float4(v0, d0, 0.f, 0.f);
    } else {
        const float b0 = decode<ST>(r[2], scale), b1 = decode<ST>(r[3], scale);
        const float v1 = keep ? b0 : cfill;
        const float d1 = b1 - b0;
        return make_float4(v0, d0, v1 - v0, d1 - d0);
    }
}

template <bool TWO, int ST>
__device__ __forceinline__ void put_entry(void* lds, int idx, float4 q) {
    if constexpr (TWO)
        static_cast<float4*>(lds)[idx] = q;
    else
        static_cast<float2*>(lds)[idx] = make_float2(q.x, q.y);
}

// Slot k of a thread is entry e = tid + 256*k; entries past the KS slots (more
// than 256*KS per view) are loaded and stored directly in stage_store.
template <bool TWO, int ST, int KS>
__device__ __forceinline__ void stage_load(const BpParams& P, const Stg& s, int p,
                                           unsigned (&raw)[KS][4], int (&xs)[KS]) {
#pragma unroll
    for (int k = 0; k < KS; ++k) {
        const int e = threadIdx.x + k * kBpThreads;
        if (e < s.total) load_entry<TWO, ST>(P, s, p, e, raw[k], xs[k]);
    }
}

template <bool TWO, int ST, int KS>
__device__ __forceinline__ void stage_store(const BpParams& P, const Stg& s, int p,
                                            const unsigned (&raw)[KS][4],
                                            const int (&xs)[KS], void* lds, int buf) {
    const float scale = (ST == 0 || P.view_scale == nullptr) ? 1.f : cptr(P.view_scale)[p];
    const float cfill = P.fill_mode == 0 ? cptr(P.means)[p] : 0.f;
#pragma unroll
    for (int k = 0; k < KS; ++k) {
        const int e = threadIdx.x + k * kBpThreads;
        if (e < s.total) put_entry<TWO, ST>(lds, buf + e, make_entry<TWO, ST>(P, xs[k], raw[k], scale, cfill));
    }
    for (int e = threadIdx.x + KS * kBpThreads; e < s.total; e += kBpThreads) {
        unsigned one[4];
        int x;
        load_entry<TWO, ST>(P, s, p, e, one, x);
        put_entry<TWO, ST>(lds, buf + e, make_entry<TWO, ST>(P, x, one, scale, cfill));
    }
}

// The kernel: one tile of one series per workgroup (BATCH = false, the
// workgroup id picks the tile), or a persistent loop over the tiles of several
// series (BATCH = true, continuous batching).  A tile is Y row iy, 64*XPL x
// positions and 4*ZPT z positions; its result depends only on (series, tile
// position of each voxel), so the same voxel computed alone, inside a batched
// launch or under another launch shape is bit-identical.
//
// Launch shape: XPL x positions per lane (tile x 64 or 128) and VPS views
// staged per barrier (1 or 2; with 2, the LDS holds two views per buffer and
// each thread prefetches one slot of entries per view).
//
// Batched tile claiming: the tiles are split into 8 contiguous chunks, one per
// XCD; a workgroup drains its own XCD's chunk (L2 locality: consecutive tiles
// share projection rows), then steals from the others.  Every workgroup visits
// every chunk once and exits when all are drained.
// Register budget: 4 waves per SIMD (<= 128 VGPRs) up to 32 accumulators per
// lane; 64 keep 3 (<= 168).
template <int ZPT, bool TWO, bool LOCAL, int ST, bool BATCH, int VPS, int XPL>
__global__ void __launch_bounds__(kBpThreads)
    __attribute__((amdgpu_waves_per_eu(ZPT * XPL > 32 ? 3 : 4, 8))) bp_kernel(BpParams P1, BpBatch B) {
    extern __shared__ float4 lds4[];
    __shared__ int s_tile;
    constexpr int TZ = 4 * ZPT;
    constexpr int KS = VPS == 1 ? kSlots : 1;
    int chunk_i = 0, lo = 0, hi = 0;
    if constexpr (BATCH) {
        lo = (blockIdx.x % 8) * B.chunk;
        hi = min(B.total, lo + B.chunk);
    }
    for (;;) {
        BpParams P;
        int logical;
        if constexpr (!BATCH) {
            P = P1;
            logical = xcd_remap(blockIdx.x, gridDim.x);
        } else {
            if (threadIdx.x == 0) s_tile = lo + atomicAdd(B.counters + (lo / max(B.chunk, 1)), 1);
            __syncthreads();
            const int t = s_tile;
            __syncthreads();
            if (t >= hi) {
                if (++chunk_i == 8) break;
                lo = ((blockIdx.x + chunk_i) % 8) * B.chunk;
                hi = min(B.total, lo + B.chunk);
                continue;
            }
            // series of tile t: last r with tile_start[r] <= t
            int ra = 0, rb = B.nreq - 1;
            while (ra < rb) {
                const int m = (ra + rb + 1) / 2;
                if (cptr(B.tile_start)[m] <= t)
                    ra = m;
                else
                    rb = m - 1;
            }
            // the descriptor is workgroup-uniform: copy it through the scalar cache
            const TILTGPU_CONSTANT unsigned* src =
                reinterpret_cast<const TILTGPU_CONSTANT unsigned*>(cptr(B.descs) + ra);
            unsigned* dst = reinterpret_cast<unsigned*>(&P);
#pragma unroll
            for (int i = 0; i < static_cast<int>(sizeof(BpParams) / 4); ++i) dst[i] = src[i];
            logical = t - cptr(B.tile_start)[ra];
        }
        const int tid = threadIdx.x;
        const int lane = tid & (kWave - 1);
        const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
        const int tx = logical % P.ntx;
        const int rest = logical / P.ntx;
        const int tz = rest % P.ntz;
        const int iy = P.y0 + rest / P.ntz;
        const int lx = TWO ? b128_lane_x(lane) : lane;
        int ox[XPL];
        float dx[XPL];
#pragma unroll
        for (int j = 0; j < XPL; ++j) {
            ox[j] = (tx * XPL + j) * kBpTileX + lx;
            dx[j] = static_cast<float>(ox[j]) - P.ocx;
        }
        const int zt0 = tz * TZ;
        const int zw0 = zt0 + wave * ZPT;
        const int seg_tile = tz * P.ntx + tx;

        float acc[XPL][ZPT];
#pragma unroll
        for (int j = 0; j < XPL; ++j)
#pragma unroll
            for (int k = 0; k < ZPT; ++k) acc[j][k] = 0.f;

        if (P.nviews > 0) {
            unsigned raw[VPS][KS][4];
            int xs[VPS][KS];
            Vs<XPL> cur[VPS];
#pragma unroll
            for (int v = 0; v < VPS; ++v) {
                if (v < P.nviews) {
                    cur[v] = view_setup<TWO, LOCAL, XPL>(P, v, dx, ox, iy, zt0, TZ, seg_tile);
                    stage_load<TWO, ST, KS>(P, cur[v], v, raw[v], xs[v]);
                }
            }
            const int ngroups = (P.nviews + VPS - 1) / VPS;
            for (int g = 0; g < ngroups; ++g) {
                // buffer set (g & 1) was last read by the compute of group g-2; every
                // thread has passed the barrier of group g-1 since, so it is free.
                // Staging (the stores of this group, the next group's setup and loads)
                // runs at raised wave priority, so the four waves reach the barrier
                // sooner while other workgroups' waves accumulate; the accumulate
                // runs at normal priority.  2048^2 x 121 -> 2048^2 x 512 launch: -1.2 to -2%.
                __builtin_amdgcn_s_setprio(2);
                const int set = (g & 1) * VPS * P.lds_cap;
#pragma unroll
                for (int v = 0; v < VPS; ++v) {
                    const int p = g * VPS + v;
                    if (p >= P.nviews) break;
                    if (cur[v].total > P.lds_cap) {
                        // host bound violated: report it (the host fails the call) and
                        // skip the view rather than overrun LDS
                        if (tid == 0) atomicOr(P.err, 1);
                        cur[v].total = 0;
                        cur[v].NR = 0;
                        cur[v].L = 0;
#pragma unroll
                        for (int j = 0; j < XPL; ++j) cur[v].row[j] = -1;
                    }
                    stage_store<TWO, ST, KS>(P, cur[v], p, raw[v], xs[v], lds4, set + v * P.lds_cap);
                }
                // the next group's loads go out before the barrier: they touch only
                // registers and global memory, and their latency then overlaps the
                // wait for the other waves' stores
                Vs<XPL> nxt[VPS];
#pragma unroll
                for (int v = 0; v < VPS; ++v) {
                    nxt[v] = cur[v];
                    const int p = (g + 1) * VPS + v;
                    if (p < P.nviews) {
                        nxt[v] = view_setup<TWO, LOCAL, XPL>(P, p, dx, ox, iy, zt0, TZ, seg_tile);
                        stage_load<TWO, ST, KS>(P, nxt[v], p, raw[v], xs[v]);
                    }
                }
                __syncthreads();
                __builtin_amdgcn_s_setprio(0);
                // a wave whose z run lies wholly past the volume (the padded
                // last tile) only stages: its sums would never be stored
#pragma unroll
                for (int v = 0; v < VPS; ++v) {
                    const int p = g * VPS + v;
                    if (p >= P.nviews || zw0 >= P.out_nz) break;
                    const Vs<XPL>& c = cur[v];
                    const int buf = set + v * P.lds_cap;
                    const TILTGPU_CONSTANT float* zc =
                        cptr(P.zc_tab) + static_cast<long>(p) * P.nz_pad + zw0;
                    if constexpr (TWO) {
#pragma unroll
                        for (int j = 0; j < XPL; ++j) {
                            if (c.row[j] >= 0) {
                                const float4* E = lds4 + buf + __mul24(c.row[j] - c.r_lo, c.L) - c.seg_lo;
                                const float a = c.a[j], fy = c.fy[j], ldx = c.ldx[j];
                                // the fences keep the scheduler from hoisting every voxel's
                                // 16-byte LDS read at once (registers, not issue, limit it).
                                // With local shifts every view takes the shifted chain: a view
                                // without one has ldx = +0, and px + 0 == px for every px but
                                // -0 (which becomes +0: same floor, same fraction).
#pragma unroll
                                for (int k = 0; k < ZPT; ++k) {
                                    float px = (a + zc[k]) + P.cx;
                                    if constexpr (LOCAL) px = px + ldx;
                                    const float4 q = E[cvt_floor(px)];
                                    const float f = __builtin_amdgcn_fractf(px);
                                    acc[j][k] += fmaf(fy, fmaf(f, q.w, q.z), fmaf(f, q.y, q.x));
                                    if ((k & 15) == 15) __builtin_amdgcn_sched_barrier(0);
                                }
                            }
                        }
                    } else if (c.total > 0) {
                        const float2* E = reinterpret_cast<const float2*>(lds4) + buf - c.seg_lo;
#pragma unroll
                        for (int j = 0; j < XPL; ++j) {
                            const float a = c.a[j];
#pragma unroll
                            for (int k = 0; k < ZPT; ++k) {
                                const float px = (a + zc[k]) + P.cx;
                                const float2 q = E[cvt_floor(px)];
                                const float f = __builtin_amdgcn_fractf(px);
                                acc[j][k] += fmaf(f, q.y, q.x);
                            }
                        }
                    }
                }
#pragma unroll
                for (int v = 0; v < VPS; ++v) cur[v] = nxt[v];
            }
        }

        // no early return: in the persistent batched kernel the workgroup goes on
        // to further tiles and its barriers
        const long slice = static_cast<long>(P.out_nz) * P.out_nx;
#pragma unroll
        for (int j = 0; j < XPL; ++j) {
            float* dst = P.out + static_cast<long>(iy - P.y0) * slice + ox[j];
#pragma unroll
            for (int k = 0; k < ZPT; ++k) {
                const int oz = zw0 + k;
                if (oz < P.out_nz && ox[j] < P.out_nx) {
                    const float v = acc[j][k] * P.inv_n;
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
        if constexpr (!BATCH) break;
    }
}

size_t bp_lds_bytes(int lds_cap, int two_row, int vps) {
    return static_cast<size_t>(2) * vps * lds_cap * (two_row ? 16 : 8);
}

bool bp_shape_supported(int tile_z, int vps, int xpl) {
    if (vps == 1 && xpl == 1) return tile_z == 32 || tile_z == 64 || tile_z == 128 || tile_z == 256;
    if (vps == 2 && xpl == 1) return tile_z == 32 || tile_z == 64 || tile_z == 128;
    if (vps == 1 && xpl == 2) return tile_z == 32 || tile_z == 64;
    return false;
}

template <int ZPT, bool TWO, bool LOCAL, int ST, int VPS, int XPL>
static hipError_t launch_bp_t(const BpParams& p, hipStream_t st) {
    const long nblk = static_cast<long>(p.ntx) * p.ntz * (p.y1 - p.y0);
    if (nblk <= 0) return hipSuccess;
    const size_t lds = bp_lds_bytes(p.lds_cap, TWO, VPS);
    hipLaunchKernelGGL((bp_kernel<ZPT, TWO, LOCAL, ST, false, VPS, XPL>),
                       dim3(static_cast<unsigned>(nblk)), dim3(kBpThreads), lds, st, p, BpBatch{});
    return hipGetLastError();
}

template <int ZPT, int VPS, int XPL>
static hipError_t launch_bp_z(const BpParams& p, int storage, int two_row, hipStream_t st) {
    const bool local = p.has_local != 0;
    if (!two_row && local) return hipErrorInvalidValue;
#define TILTGPU_BP_CASE(TW, LC, S)                         \
    if (two_row == TW && local == LC && storage == S)      \
        return launch_bp_t<ZPT, TW, LC, S, VPS, XPL>(p, st);
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

template <int ZPT, bool TWO, bool LOCAL, int ST>
static hipError_t launch_bp_batch_t(const BpBatch& b, int grid, size_t lds, hipStream_t st) {
    hipLaunchKernelGGL((bp_kernel<ZPT, TWO, LOCAL, ST, true, 1, 1>), dim3(grid), dim3(kBpThreads),
                       lds, st, BpParams{}, b);
    return hipGetLastError();
}

template <int ZPT>
static hipError_t launch_bp_batch_z(const BpBatch& b, int storage, int two_row, int local,
                                    int grid, size_t lds, hipStream_t st) {
    if (!two_row && local) return hipErrorInvalidValue;
#define TILTGPU_BPB_CASE(TW, LC, S)                          \
    if (two_row == TW && local == LC && storage == S)        \
        return launch_bp_batch_t<ZPT, TW, LC, S>(b, grid, lds, st);
    TILTGPU_BPB_CASE(0, 0, 0)
    TILTGPU_BPB_CASE(0, 0, 1)
    TILTGPU_BPB_CASE(0, 0, 2)
    TILTGPU_BPB_CASE(1, 0, 0)
    TILTGPU_BPB_CASE(1, 0, 1)
    TILTGPU_BPB_CASE(1, 0, 2)
    TILTGPU_BPB_CASE(1, 1, 0)
    TILTGPU_BPB_CASE(1, 1, 1)
    TILTGPU_BPB_CASE(1, 1, 2)
#undef TILTGPU_BPB_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_backproject_batch(const BpBatch& b, int storage, int two_row, int local,
                                    int tile_z, int grid, size_t lds, hipStream_t st) {
    if (b.total <= 0) return hipSuccess;
    switch (tile_z) {
        case 32: return launch_bp_batch_z<8>(b, storage, two_row, local, grid, lds, st);
        case 64: return launch_bp_batch_z<16>(b, storage, two_row, local, grid, lds, st);
        case 128: return launch_bp_batch_z<32>(b, storage, two_row, local, grid, lds, st);
        case 256: return launch_bp_batch_z<64>(b, storage, two_row, local, grid, lds, st);
        default: return hipErrorInvalidValue;
    }
}

// The staged segment of every (view, tile z, tile x) for views without a patch
// grid: the floors of proj_x at the x positions of lanes 0 and 63 (each of
// their xpl positions) and the tile's first and last z, exactly as view_setup
// evaluates them (proj_x is monotonic in x and in z there, so these bound every
// voxel of the tile).  One thread per entry; patch-grid views keep computing
// theirs in the workgroup.
__global__ void __launch_bounds__(256) bp_segment_kernel(BpParams P, int TZ, int xpl, int local,
                                                         int2* seg) {
    const long n = static_cast<long>(P.nviews) * P.ntz * P.ntx;
    const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int tx = static_cast<int>(i % P.ntx);
    const long rest = i / P.ntx;
    const int tz = static_cast<int>(rest % P.ntz);
    const int p = static_cast<int>(rest / P.ntz);
    const BpView V = P.views[p];
    const int mode = local ? V.shift_mode : 0;
    int flo = INT_MAX, fhi = INT_MIN;
    if (mode != 2) {
        const float* zc = P.zc_tab + static_cast<long>(p) * P.nz_pad;
        const float z_first = zc[tz * TZ], z_last = zc[tz * TZ + TZ - 1];
        const float lx = mode == 1 ? V.sdx : 0.f;
        for (int j = 0; j < xpl; ++j) {
            for (int e = 0; e < 2; ++e) {
                const int ox = (tx * xpl + j) * kBpTileX + (e ? kBpTileX - 1 : 0);
                const float dx = static_cast<float>(ox) - P.ocx;
                const float a = (dx * V.cos_t) * V.cos_xt;
                if (local)
                    seg_bounds<true>(a, z_first, z_last, P.cx, lx, flo, fhi);
                else
                    seg_bounds<false>(a, z_first, z_last, P.cx, lx, flo, fhi);
            }
        }
    }
    seg[i] = make_int2(flo, fhi);
}

hipError_t launch_bp_segments(const BpParams& p, int tile_z, int xpl, int2* seg, hipStream_t st) {
    const long n = static_cast<long>(p.nviews) * p.ntz * p.ntx;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(bp_segment_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256),
                       0, st, p, tile_z, xpl, p.has_local, seg);
    return hipGetLastError();
}

hipError_t launch_backproject(const BpParams& p, int storage, int two_row, int tile_z, int vps,
                              int xpl, hipStream_t st) {
    if (!bp_shape_supported(tile_z, vps, xpl)) return hipErrorInvalidValue;
    if (vps == 2) {
        switch (tile_z) {
            case 32: return launch_bp_z<8, 2, 1>(p, storage, two_row, st);
            case 64: return launch_bp_z<16, 2, 1>(p, storage, two_row, st);
            default: return launch_bp_z<32, 2, 1>(p, storage, two_row, st);
        }
    }
    if (xpl == 2) {
        if (tile_z == 32) return launch_bp_z<8, 1, 2>(p, storage, two_row, st);
        return launch_bp_z<16, 1, 2>(p, storage, two_row, st);
    }
    switch (tile_z) {
        case 32: return launch_bp_z<8, 1, 1>(p, storage, two_row, st);
        case 64: return launch_bp_z<16, 1, 1>(p, storage, two_row, st);
        case 128: return launch_bp_z<32, 1, 1>(p, storage, two_row, st);
        default: return launch_bp_z<64, 1, 1>(p, storage, two_row, st);
    }
}

}  // namespace tiltgpu
