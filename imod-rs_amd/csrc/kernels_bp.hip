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

constexpr int kSlots = 2;  // staged entries prefetched per thread

// Per-view state of one lane and of its wave (the wave-level values are equal in
// all four waves: they hold the same 64 ox positions).
struct Vs {
    float a;      // (dx*cos_t)*cos_xt
    float ldx;    // x shift (0 when none)
    float fy;     // weight of row r+1 (0: one row)
    int row;      // row pair start r, or -1 (no contribution)
    int seg_lo, L, r_lo, NR, total, shift;
};

template <bool TWO, bool LOCAL>
__device__ __forceinline__ Vs view_setup(const BpParams& P, int p, float dx, int ox, int iy,
                                         int zt0, int TZ) {
    Vs s;
    const TILTGPU_CONSTANT BpView* V = cptr(P.views) + p;
    const float cos_t = V->cos_t, cos_xt = V->cos_xt, sin_xt = V->sin_xt;
    const int mode = LOCAL ? V->shift_mode : 0;
    s.shift = mode;
    s.a = (dx * cos_t) * cos_xt;
    float ldx = 0.f, ldy = 0.f;
    if (LOCAL) {
        if (mode == 1) {
            ldx = V->sdx;
            ldy = V->sdy;
        } else if (mode == 2) {
            grid_shift(P, p, static_cast<float>(ox), static_cast<float>(iy), ldx, ldy);
        }
    }
    s.ldx = ldx;
    const TILTGPU_CONSTANT float* zc = cptr(P.zc_tab) + static_cast<long>(p) * P.nz_pad;
    const float z_first = zc[zt0], z_last = zc[zt0 + TZ - 1];
    float pa = (s.a + z_first) + P.cx;
    float pb = (s.a + z_last) + P.cx;
    if (mode) {
        pa = pa + ldx;
        pb = pb + ldx;
    }
    float lo = fminf(pa, pb), hi = fmaxf(pa, pb);
    if (mode == 2) {
        lo = wave_min(lo);
        hi = wave_max(hi);
    } else {
        // proj_x is monotonic in ox (a constant shift keeps it so): extremes at lanes 0, 63
        lo = fminf(lane_f(lo, 0), lane_f(lo, 63));
        hi = fmaxf(lane_f(hi, 0), lane_f(hi, 63));
    }
    s.seg_lo = cvt_floor(lo);
    s.L = cvt_floor(hi) - s.seg_lo + 1;
    s.row = iy;
    s.fy = 0.f;
    s.r_lo = iy;
    int NR = 1;
    if constexpr (TWO) {
        const float yo = dx * sin_xt;
        float sy = static_cast<float>(iy) + yo;
        if (mode) sy = sy + ldy;
        const float fl = floorf(sy);
        const float fny = static_cast<float>(P.in_ny);
        const bool two = LOCAL || fabsf(yo) > 1e-6f;
        s.row = -1;
        if (two && fl >= 0.f && fl + 1.f < fny) {
            s.row = static_cast<int>(fl);
            s.fy = sy - fl;
        } else if (fl >= 0.f && fl < fny) {
            s.row = static_cast<int>(fl);
        }
        int rmin, rmax;
        if (mode == 2) {
            rmin = wave_min_i(s.row >= 0 ? s.row : INT_MAX);
            rmax = wave_max_i(s.row);
        } else {
            // floor(sy) is monotonic in ox; clamp the extremes into the valid rows
            const float f0 = lane_f(fl, 0), f1 = lane_f(fl, 63);
            const float lo_r = fmaxf(fminf(f0, f1), 0.f);
            const float hi_r = fminf(fmaxf(f0, f1), fny - 1.f);
            rmin = lo_r <= hi_r ? static_cast<int>(lo_r) : INT_MAX;
            rmax = lo_r <= hi_r ? static_cast<int>(hi_r) : -1;
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
    return s;
}

// Global loads for staged entry e of view p: v[x], v[x+1] of row r (and r+1);
// a padded entry loads the edge value instead (used by the Edge fill).
template <bool TWO, int ST>
__device__ __forceinline__ void load_entry(const BpParams& P, const Vs& s, int p, int iy,
                                           int ri, int xi, unsigned (&r)[4]) {
    const int n = P.in_nx;
    const long nxy = static_cast<long>(n) * P.in_ny;
    const int x = s.seg_lo + xi;
    const int row = TWO ? s.r_lo + ri : iy;
    const long off = static_cast<long>(p) * nxy + static_cast<long>(row) * n;
    const int i0 = x < 0 ? 0 : (x > n - 2 ? n - 1 : x);
    const int i1 = (x >= 0 && x <= n - 2) ? x + 1 : i0;
    r[0] = load_raw<ST>(P.proj, off + i0);
    r[1] = load_raw<ST>(P.proj, off + i1);
    r[2] = r[3] = 0u;
    if constexpr (TWO) {
        if (row + 1 < P.in_ny) {  // a row past the end is only used with weight 0
            r[2] = load_raw<ST>(P.proj, off + n + i0);
            r[3] = load_raw<ST>(P.proj, off + n + i1);
        }
    }
}

template <bool TWO, int ST>
__device__ __forceinline__ float4 make_entry(const BpParams& P, const Vs& s, int ri, int xi,
                                             const unsigned (&r)[4], float scale, float mean) {
    const int x = s.seg_lo + xi;
    const bool inside = x >= 0 && x <= P.in_nx - 2;
    const float a0 = decode<ST>(r[0], scale), a1 = decode<ST>(r[1], scale);
    float v0, d0;
    if (inside) {
        v0 = a0;
        d0 = a1 - a0;
    } else {
        v0 = P.fill_mode == 0 ? mean : (P.fill_mode == 1 ? 0.f : a0);
        d0 = 0.f;
    }
    if constexpr (!TWO) {
        return make_float4(v0, d0, 0.f, 0.f);
    } else {
        float v1 = 0.f, d1 = 0.f;
        if (s.r_lo + ri + 1 < P.in_ny) {
            const float b0 = decode<ST>(r[2], scale), b1 = decode<ST>(r[3], scale);
            if (inside) {
                v1 = b0;
                d1 = b1 - b0;
            } else {
                v1 = P.fill_mode == 0 ? mean : (P.fill_mode == 1 ? 0.f : b0);
            }
        }
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

// Staging work split: slot k of a thread is entry x = seg_lo + tid of row pair k
// (two-row geometry) or entry x = seg_lo + tid + 256*k (one-row geometry).
// Entries beyond the slots (segments wider than 256, more than kSlots row
// pairs) are loaded and stored directly in stage_store.
template <bool TWO>
__device__ __forceinline__ bool slot_entry(const Vs& s, int k, int& ri, int& xi) {
    if constexpr (TWO) {
        ri = k;
        xi = threadIdx.x;
        return ri < s.NR && xi < s.L;
    } else {
        ri = 0;
        xi = threadIdx.x + k * kBpThreads;
        return xi < s.L;
    }
}

template <bool TWO, int ST>
__device__ __forceinline__ void stage_load(const BpParams& P, const Vs& s, int p, int iy,
                                           unsigned (&raw)[kSlots][4]) {
#pragma unroll
    for (int k = 0; k < kSlots; ++k) {
        int ri, xi;
        if (slot_entry<TWO>(s, k, ri, xi)) load_entry<TWO, ST>(P, s, p, iy, ri, xi, raw[k]);
    }
}

template <bool TWO, int ST>
__device__ __forceinline__ void stage_store(const BpParams& P, const Vs& s, int p, int iy,
                                            const unsigned (&raw)[kSlots][4], void* lds,
                                            int buf) {
    const float scale = (ST == 0 || P.view_scale == nullptr) ? 1.f : cptr(P.view_scale)[p];
    const float mean = cptr(P.means)[p];
#pragma unroll
    for (int k = 0; k < kSlots; ++k) {
        int ri, xi;
        if (slot_entry<TWO>(s, k, ri, xi))
            put_entry<TWO, ST>(lds, buf + ri * s.L + xi,
                               make_entry<TWO, ST>(P, s, ri, xi, raw[k], scale, mean));
    }
    // entries beyond the prefetch slots: load and store directly
    for (int ri = 0; ri < s.NR; ++ri) {
        int xi0 = threadIdx.x;
        if (TWO ? ri < kSlots : true) xi0 += TWO ? kBpThreads : kSlots * kBpThreads;
        for (int xi = xi0; xi < s.L; xi += kBpThreads) {
            unsigned one[4];
            load_entry<TWO, ST>(P, s, p, iy, ri, xi, one);
            put_entry<TWO, ST>(lds, buf + ri * s.L + xi,
                               make_entry<TWO, ST>(P, s, ri, xi, one, scale, mean));
        }
    }
}

// The kernel: one tile of one series per workgroup (BATCH = false, the
// workgroup id picks the tile), or a persistent loop over the tiles of several
// series (BATCH = true, continuous batching).  A tile is Y row iy, 64 x positions
// and 4*ZPT z positions; its result depends only on (series, tile), so the same
// tile computed alone or inside a batched launch is bit-identical.
//
// Batched tile claiming: the tiles are split into 8 contiguous chunks, one per
// XCD; a workgroup drains its own XCD's chunk (L2 locality: consecutive tiles
// share projection rows), then steals from the others.  Every workgroup visits
// every chunk once and exits when all are drained.
// Register budget: 4 waves per SIMD (<= 128 VGPRs) up to 32 z per lane; the
// 64-z tile keeps 3 (<= 168).
template <int ZPT, bool TWO, bool LOCAL, int ST, bool BATCH>
__global__ void __launch_bounds__(kBpThreads)
    __attribute__((amdgpu_waves_per_eu(ZPT > 32 ? 3 : 4, 8))) bp_kernel(BpParams P1, BpBatch B) {
    extern __shared__ float4 lds4[];
    __shared__ int s_tile;
    constexpr int TZ = 4 * ZPT;
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
        const int ox = tx * kBpTileX + lane;
        const int zt0 = tz * TZ;
        const int zw0 = zt0 + wave * ZPT;
        const float dx = static_cast<float>(ox) - P.ocx;

        float acc[ZPT];
    #pragma unroll
        for (int k = 0; k < ZPT; ++k) acc[k] = 0.f;

        if (P.nviews > 0) {
            unsigned raw[kSlots][4];
            Vs cur = view_setup<TWO, LOCAL>(P, 0, dx, ox, iy, zt0, TZ);
            stage_load<TWO, ST>(P, cur, 0, iy, raw);
            for (int p = 0; p < P.nviews; ++p) {
                if (cur.total > P.lds_cap) {
                    // host bound violated: report it (the host fails the call) and skip
                    // the view rather than overrun LDS
                    if (tid == 0) atomicOr(P.err, 1);
                    cur.total = 0;
                    cur.NR = 0;
                    cur.L = 0;
                    cur.row = -1;
                }
                // buffer (p & 1) was last read by the compute of view p-2; every thread
                // has passed the barrier of view p-1 since, so it is free.
                const int buf = (p & 1) * P.lds_cap;
                stage_store<TWO, ST>(P, cur, p, iy, raw, lds4, buf);
                __syncthreads();
                Vs nxt = cur;
                if (p + 1 < P.nviews) {
                    nxt = view_setup<TWO, LOCAL>(P, p + 1, dx, ox, iy, zt0, TZ);
                    stage_load<TWO, ST>(P, nxt, p + 1, iy, raw);
                }
                const TILTGPU_CONSTANT float* zc =
                    cptr(P.zc_tab) + static_cast<long>(p) * P.nz_pad + zw0;
                if constexpr (TWO) {
                    if (cur.row >= 0) {
                        const float4* E = lds4 + buf + (cur.row - cur.r_lo) * cur.L - cur.seg_lo;
                        const float a = cur.a, fy = cur.fy, ldx = cur.ldx;
                        // the fences keep the scheduler from hoisting every voxel's
                        // 16-byte LDS read at once (registers, not issue, limit it)
                        if (LOCAL && cur.shift) {
    #pragma unroll
                            for (int k = 0; k < ZPT; ++k) {
                                const float px = ((a + zc[k]) + P.cx) + ldx;
                                const float4 q = E[cvt_floor(px)];
                                const float f = __builtin_amdgcn_fractf(px);
                                acc[k] += fmaf(fy, fmaf(f, q.w, q.z), fmaf(f, q.y, q.x));
                                if ((k & 15) == 15) __builtin_amdgcn_sched_barrier(0);
                            }
                        } else {
    #pragma unroll
                            for (int k = 0; k < ZPT; ++k) {
                                const float px = (a + zc[k]) + P.cx;
                                const float4 q = E[cvt_floor(px)];
                                const float f = __builtin_amdgcn_fractf(px);
                                acc[k] += fmaf(fy, fmaf(f, q.w, q.z), fmaf(f, q.y, q.x));
                                if ((k & 15) == 15) __builtin_amdgcn_sched_barrier(0);
                            }
                        }
                    }
                } else if (cur.total > 0) {
                    const float2* E = reinterpret_cast<const float2*>(lds4) + buf - cur.seg_lo;
                    const float a = cur.a;
    #pragma unroll
                    for (int k = 0; k < ZPT; ++k) {
                        const float px = (a + zc[k]) + P.cx;
                        const float2 q = E[cvt_floor(px)];
                        const float f = __builtin_amdgcn_fractf(px);
                        acc[k] += fmaf(f, q.y, q.x);
                    }
                }
                cur = nxt;
            }
        }

        // no early return: in the persistent batched kernel the workgroup goes on
        // to further tiles and its barriers
        const long slice = static_cast<long>(P.out_nz) * P.out_nx;
        float* dst = P.out + static_cast<long>(iy - P.y0) * slice + ox;
    #pragma unroll
        for (int k = 0; k < ZPT; ++k) {
            const int oz = zw0 + k;
            if (oz < P.out_nz && ox < P.out_nx) {
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
        if constexpr (!BATCH) break;
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
    hipLaunchKernelGGL((bp_kernel<ZPT, TWO, LOCAL, ST, false>), dim3(static_cast<unsigned>(nblk)),
                       dim3(kBpThreads), lds, st, p, BpBatch{});
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

template <int ZPT, bool TWO, bool LOCAL, int ST>
static hipError_t launch_bp_batch_t(const BpBatch& b, int grid, size_t lds, hipStream_t st) {
    hipLaunchKernelGGL((bp_kernel<ZPT, TWO, LOCAL, ST, true>), dim3(grid), dim3(kBpThreads), lds,
                       st, BpParams{}, b);
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

hipError_t launch_backproject(const BpParams& p, int storage, int two_row, int tile_z,
                              hipStream_t st) {
    switch (tile_z) {
        case 32: return launch_bp_z<8>(p, storage, two_row, st);
        case 64: return launch_bp_z<16>(p, storage, two_row, st);
        case 128: return launch_bp_z<32>(p, storage, two_row, st);
        case 256: return launch_bp_z<64>(p, storage, two_row, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace tiltgpu
