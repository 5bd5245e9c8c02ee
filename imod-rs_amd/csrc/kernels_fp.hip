// Forward projection of the SIRT iteration (main.rs:830-896) for gfx950.
//
// Single-row geometry (no x-axis tilt, no local alignment): an ordered voxel
// walk.  Each lane owns B consecutive sinogram bins of one (view, row) and keeps
// their sums in lane-private LDS slots.  For every z row, in ascending oz, it
// walks the voxels whose projection can land on its bins in ascending ox and
// adds v*(1-frac) to bin floor(px) and v*frac to bin floor(px)+1, exactly as the
// serial scatter of main.rs:877-887 does.  A bin therefore receives its terms
// in the reference's order (oz, then ox) and with the reference's roundings:
// the sinogram equals the serial scatter bit for bit.  Every voxel is evaluated
// once per lane window instead of once per bin it touches, and the loop has no
// per-bin range search.  For volume rows staged in LDS, each lane's voxel
// window of a z row is exact (trimmed by evaluating px at its ends), so a voxel
// needs no clamp and no validity test, and the sums of the two bins the lane is
// on ride in registers: a voxel costs one LDS store and one LDS load instead of
// a read-modify-write of two sums (the walk is then bound by VALU issue).
//
// Two-row geometry (x-axis tilt / local alignment): a scatter with the
// coordinates and weights of main.rs:840-889.  Each term is rounded to a
// multiple of 2^-k (k from the volume's largest magnitude, so that no bin sum
// can reach 2^62) and summed with 64-bit integer atomics, which are
// order-independent: the sinogram is the same bits on every run and launch
// shape.  The fixed-point sums are exact up to those term roundings (2^-31 of
// the volume's largest magnitude or finer), so they differ from the reference's
// sequential FP32 sums by the rounding of the latter.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>

#include "tiltgpu_internal.h"

namespace tiltgpu {

namespace {

constexpr int kFpWG = 512;          // threads per workgroup: 8 waves = 8 views
constexpr int kFpViewsPerWG = kFpWG / 64;
constexpr float kWalkMinC = 0.05f;  // views with cos_t*cos_xt below this use fp_bin_scan

// One bin by direct scan (main.rs:855-887 restricted to the voxels that can land
// on bin x, visited in the reference's order).  Used for views whose projected
// voxel spacing cos_t*cos_xt is tiny or negative, where the walk's window bound
// does not apply.
__device__ void fp_bin_scan(const FpParams& P, int x, int iy, int p, float* out) {
    const float4 g = P.vgeom[p];
    const float c = g.x * g.z;
    const float* zc = P.zc_tab + static_cast<int64_t>(p) * P.nz_pad;
    const float* vrow = P.vol + static_cast<int64_t>(iy) * P.out_nz * P.out_nx;
    const float fx = static_cast<float>(x);
    const float fn = static_cast<float>(P.in_nx);
    const float fnx = static_cast<float>(P.out_nx);
    float acc = 0.f;
    for (int oz = 0; oz < P.out_nz; ++oz) {
        const float zcv = zc[oz];
        int lo = 0, hi = P.out_nx - 1;
        if (fabsf(c) > 1e-4f) {
            const float t0 = P.ocx + ((fx - 1.0f) - P.cx - zcv) / c;
            const float t1 = P.ocx + ((fx + 1.0f) - P.cx - zcv) / c;
            const float tl = fminf(t0, t1), th = fmaxf(t0, t1);
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
    *out = acc;
}

// Per-lane constants of one walk window: bins x0 .. x0+B-1 of (view p, row iy).
// The lane's LDS slot s is at byte s * 4 * kFpWG + 4 * threadIdx.x.  In the
// clamped walk slots 2..B+1 hold the bins, and slots 0, 1, B+2 and B+3 absorb
// the terms of voxels that land outside the window; with exact windows slots
// 1..B hold the bins, slots 0 and B+1 take the terms of bins x0-1 and x0+B, and
// slot B+2 is the register walk's look-ahead past the window.
struct WalkLane {
    float ct, cxt, inv_c, cx, ocx, fnx, fin;
    float k_lo, k_hi;  // px = x0-1 and px = x0+B bound the voxels that reach the window
    float b_lo, b_hi;  // exact windows: max(x0-1, 0) <= px < min(x0+B, in_nx-1)
    float mbase;
    unsigned nvalid;   // in_nx - 1
    int lo, hi;        // x0 - 2, x0 + B: clamp of floor(px)
    unsigned qoff;     // byte offset of the slot of floor(px) = 0 (mod 2^32)
    int last;
    float margin;      // exact windows: bound of the inverted window ends (walk_view_bounds)
    bool regwalk;      // exact windows: register-carried sums apply (walk_view_bounds)
};

// first: the slot of bin x0 (2 with the clamped walk's two dummies below the
// window, 1 with an exact window's one).
__device__ __forceinline__ WalkLane walk_lane(const FpParams& P, float4 g, int x0, int B,
                                              int first) {
    WalkLane L;
    L.ct = g.x;
    L.cxt = g.z;
    L.inv_c = 1.0f / (g.x * g.z);
    L.cx = P.cx;
    L.ocx = P.ocx;
    L.fnx = static_cast<float>(P.out_nx);
    L.fin = static_cast<float>(P.in_nx);
    L.k_lo = static_cast<float>(x0 - 1) - P.cx;
    L.k_hi = static_cast<float>(x0 + B) - P.cx;
    L.b_lo = static_cast<float>(max(x0 - 1, 0));
    L.b_hi = static_cast<float>(min(x0 + B, P.in_nx - 1));
    // Bound on |computed - exact| of px and of the inverted window ends, in px
    // units: a few roundings of magnitudes below in_nx + out_nx + |zc|.
    L.mbase = 3e-7f * (4.f * L.fnx + 5.f * static_cast<float>(P.in_nx) + 64.f);
    L.nvalid = static_cast<unsigned>(P.in_nx - 1);
    L.lo = x0 - 2;
    L.hi = x0 + B;
    L.qoff = 4u * threadIdx.x + static_cast<unsigned>(first - x0) * (4u * kFpWG);
    L.last = P.out_nx - 1;
    return L;
}

// Voxel window of z row zc for the lane: [e0, e1) holds every voxel whose
// floor(px) lies in [x0-1, x0+B-1] (plus at most a few that land in the dummy
// slots).
__device__ __forceinline__ void walk_window(const WalkLane& L, float zcv, int& e0, int& e1) {
    const float margin = (L.mbase + 1.8e-6f * fabsf(zcv)) * L.inv_c + 1e-3f;
    const float tlo = (L.k_lo - zcv) * L.inv_c + L.ocx;
    const float thi = (L.k_hi - zcv) * L.inv_c + L.ocx;
    e0 = static_cast<int>(fminf(fmaxf(ceilf(tlo - margin), 0.f), L.fnx));
    e1 = static_cast<int>(fminf(fmaxf(floorf(thi + margin) + 1.f, 0.f), L.fnx));
}

// px of the voxel at dx = ox - ocx in the reference's operation order
// (main.rs:858-860; single-row geometry, so (dx*cos_t)*cos_xt is dx*cos_t).
__device__ __forceinline__ float walk_px(const WalkLane& L, float dx, float zcv) {
    return (dx * L.ct + zcv) + L.cx;
}

// The exact voxel window of z row zc: [e0, e1) holds exactly the voxels with
// b_lo <= px < b_hi, i.e. max(x0-1, 0) <= floor(px) <= min(x0+B-1, in_nx-2):
// the valid voxels (0 <= px0 < in_nx-1) that add a term to a bin of the lane's
// window.  px is non-decreasing in ox (each rounding step is monotonic and
// cos_t > 0).  The inverted bounds t(b_lo), t(b_hi) carry the walk_window
// margin m, so ceil(t(b_lo) - m) is the first voxel of the window or the one
// before it, and floor(t(b_hi) + m) + 1 its end or one past it, whenever
// 2m < 1: one evaluation of px, exactly as the walk computes it, settles each
// end.  (m stays far below 1/2 for every staged shape; the loops cover the
// rest.)
__device__ __forceinline__ void walk_window_exact(const WalkLane& L, float zcv, int& e0,
                                                  int& e1) {
    const float margin = L.margin;
    const float tlo = ((L.b_lo - L.cx) - zcv) * L.inv_c + L.ocx;
    const float thi = ((L.b_hi - L.cx) - zcv) * L.inv_c + L.ocx;
    e0 = static_cast<int>(fminf(fmaxf(ceilf(tlo - margin), 0.f), L.fnx));
    e1 = static_cast<int>(fminf(fmaxf(floorf(thi + margin) + 1.f, 0.f), L.fnx));
    if (margin < 0.5f) {  // uniform over the wave (one view, one z row)
        e0 += walk_px(L, static_cast<float>(e0) - L.ocx, zcv) < L.b_lo ? 1 : 0;
        e1 -= walk_px(L, static_cast<float>(e1 - 1) - L.ocx, zcv) >= L.b_hi ? 1 : 0;
    } else {
        while (e0 < e1 && walk_px(L, static_cast<float>(e0) - L.ocx, zcv) < L.b_lo) ++e0;
        while (e1 > e0 && walk_px(L, static_cast<float>(e1 - 1) - L.ocx, zcv) >= L.b_hi) --e1;
    }
}

// One voxel of an exact window: floor(px) lies in [x0-1, x0+B-1], so the pair
// of slots it adds to is in [0, B+1] with no clamp, and the voxel is valid.
template <int B>
__device__ __forceinline__ void walk_voxel_exact(const WalkLane& L, char* lds, float dx,
                                                 float zcv, float v) {
    const float px = walk_px(L, dx, zcv);
    const int fi = cvt_floor(px);
    const float frac = __builtin_amdgcn_fractf(px);
    float* q = reinterpret_cast<float*>(lds + (static_cast<unsigned>(fi) * (4u * kFpWG) + L.qoff));
    q[0] += v * (1.0f - frac);
    q[kFpWG] += v * frac;
}

// Register-carried sums for exact windows.  While floor(px) advances by at most
// one per voxel, the lane keeps the running sums of bins fi and fi+1 of its last
// voxel in registers (r0, r1) and the stored sum of bin fi+2 loaded one voxel
// ahead (la).  A voxel that stays on bin fi adds to r0 and r1; one that moves
// to fi+1 makes r1 and la the new pair.  Each voxel stores r0 -- the final sum
// of bin fi once the walk moves on, an intermediate one that a later store
// replaces otherwise -- so a voxel costs one LDS store and one LDS load instead
// of reading and writing two sums.  The terms and their order are those of
// walk_voxel_exact: each bin's sum is S + t1 + t2 + ... in voxel order.
// (Slot B+2, the look-ahead of the window's last bin, is loaded but never used.)
struct RegWalk {
    float r0, r1, la;
    float* q;  // slot of bin fi
    int fi;
};

// qb: the lane's slot of bin 0 (lds + qoff; LDS addresses wrap at 2^32)
__device__ __forceinline__ void reg_begin(const WalkLane& L, char* qb, float dx, float zcv,
                                          float v, RegWalk& W) {
    const float px = walk_px(L, dx, zcv);
    W.fi = cvt_floor(px);
    const float frac = __builtin_amdgcn_fractf(px);
    W.q = reinterpret_cast<float*>(qb + W.fi * (4 * kFpWG));
    W.r0 = W.q[0] + v * (1.0f - frac);
    W.r1 = W.q[kFpWG] + v * frac;
    W.la = W.q[2 * kFpWG];
}

__device__ __forceinline__ void reg_step(const WalkLane& L, char* qb, float dx, float zcv,
                                         float v, RegWalk& W) {
    const float px = walk_px(L, dx, zcv);
    const int fi = cvt_floor(px);
    const float frac = __builtin_amdgcn_fractf(px);
    const float a = v * (1.0f - frac);
    const float b = v * frac;
    W.q[0] = W.r0;
    float* q = reinterpret_cast<float*>(qb + fi * (4 * kFpWG));
    const float la = q[2 * kFpWG];
    const bool adv = fi != W.fi;  // fi == W.fi + 1
    const float r0 = adv ? W.r1 : W.r0;
    const float r1 = adv ? W.la : W.r1;
    W.r0 = r0 + a;
    W.r1 = r1 + b;
    W.la = la;
    W.q = q;
    W.fi = fi;
}

// floor(px) advances by at most one per voxel when cos_t < 1 - M*2^-21, M
// bounding |px| and its partial results (in_nx + out_nx + |zc|): each of px's
// three roundings is within half an ulp of M, so consecutive px differ by at
// most cos_t + 3 ulp(M) < 1.  Uniform over the wave.
__device__ __forceinline__ bool reg_walk_ok(const WalkLane& L, float zcv) {
    return L.ct < 1.0f - (L.fnx + L.fin + fabsf(zcv)) * 4.76837158203125e-7f;
}

// The exact walk's per-view bounds, taken at the largest |zc| of the view's z
// rows (zc = (dz*sin_t)*zfactor is monotonic in oz, so that is |zc| at the
// first or the last row): the walk_window margin, and reg_walk_ok, which then
// holds for every z row of the view.
__device__ __forceinline__ void walk_view_bounds(WalkLane& L, float zc_first, float zc_last) {
    const float mz = fmaxf(fabsf(zc_first), fabsf(zc_last));
    L.margin = (L.mbase + 1.8e-6f * mz) * L.inv_c + 1e-3f;
    L.regwalk = reg_walk_ok(L, mz);
}

// One voxel, main.rs:860-887 with px0 = floor(px): v*(1-frac) to bin px0 and
// v*frac to bin px0+1.  A voxel outside 0 <= px0 < in_nx-1 adds nothing in the
// reference; here it adds +0.0 (its value is replaced by 0), which leaves a sum
// that started at +0.0 bitwise unchanged.  frac = px - floor(px) is exact, so
// v_fract gives the reference's value wherever px >= 0.
template <int B>
__device__ __forceinline__ void walk_voxel(const WalkLane& L, char* lds, float dx, float zcv,
                                           float v) {
    // The walk serves the single-row geometry only: every view has sin_xt == 0,
    // hence xtilt == 0 and cos_xt == 1.0f exactly, and (dx*cos_t)*cos_xt is
    // dx*cos_t bit for bit.
    const float px = (dx * L.ct + zcv) + L.cx;
    const int fi = cvt_floor(px);
    const float frac = __builtin_amdgcn_fractf(px);
    const float w = static_cast<unsigned>(fi) < L.nvalid ? v : 0.0f;
    int m;  // clamp(fi, x0-2, x0+B) in one instruction
    asm("v_med3_i32 %0, %1, %2, %3" : "=v"(m) : "v"(fi), "v"(L.lo), "v"(L.hi));
    float* q = reinterpret_cast<float*>(lds + (static_cast<unsigned>(m) * (4u * kFpWG) + L.qoff));
    q[0] += w * (1.0f - frac);
    q[kFpWG] += w * frac;
}

// Walk of one z row.  STAGED: vz is the row in LDS; otherwise vz is the row in
// HBM and the next voxel is fetched one step ahead.
template <int B, bool STAGED>
__device__ __forceinline__ void walk_row(const WalkLane& L, char* lds,
                                         const float* __restrict__ vz, float zcv) {
    int e0, e1;
    if constexpr (STAGED)
        walk_window_exact(L, zcv, e0, e1);
    else
        walk_window(L, zcv, e0, e1);
    // dx = ox - ocx is exact in f32 and advances exactly by 1 per voxel
    float dx = static_cast<float>(e0) - L.ocx;
    const float dx_end = static_cast<float>(e1) - L.ocx;
    const float* vp = vz + e0;
    if constexpr (STAGED) {
        // two voxels per step (one paired LDS read), the odd one after; dx + 1
        // is exact, so every voxel sees the same dx as in a one-by-one walk
        const float* vend = vz + e1;
        const float* vlast = vend - 1;  // a pair starts before the last voxel
        if (L.regwalk) {
            if (vp < vend) {
                // An odd count starts with its first voxel; an even one with a
                // voxel of value 0 at the first voxel's px, whose +0.0 terms
                // leave the sums' bits unchanged.  The rest goes in pairs.
                char* const qb = lds + L.qoff;
                const bool odd = ((e1 - e0) & 1) != 0;
                RegWalk W;
                reg_begin(L, qb, dx, zcv, odd ? vp[0] : 0.0f, W);
                if (odd) {
                    ++vp;
                    dx += 1.0f;
                }
                for (; vp < vend; vp += 2) {
                    const float v0 = vp[0], v1 = vp[1];
                    reg_step(L, qb, dx, zcv, v0, W);
                    reg_step(L, qb, dx + 1.0f, zcv, v1, W);
                    dx += 2.0f;
                }
                W.q[0] = W.r0;
                W.q[kFpWG] = W.r1;
            }
        } else {
            for (; vp < vlast; vp += 2) {
                const float v0 = vp[0], v1 = vp[1];
                walk_voxel_exact<B>(L, lds, dx, zcv, v0);
                walk_voxel_exact<B>(L, lds, dx + 1.0f, zcv, v1);
                dx += 2.0f;
            }
            if (vp < vend) walk_voxel_exact<B>(L, lds, dx, zcv, vp[0]);
        }
        (void)dx_end;
    } else {
        const float* vlast = vz + L.last;
        float vn = *(vp < vlast ? vp : vlast);
        for (; dx < dx_end; dx += 1.0f) {
            const float v = vn;
            ++vp;
            vn = *(vp < vlast ? vp : vlast);
            walk_voxel<B>(L, lds, dx, zcv, v);
        }
    }
}

template <int B>
__device__ __forceinline__ void walk_store(const FpParams& P, const float* a, int x0, int iy,
                                           int p, int first) {
    float* row = P.sino + (static_cast<int64_t>(p) * P.in_ny + iy) * P.in_nx;
#pragma unroll
    for (int j = 0; j < B; ++j)
        if (x0 + j < P.in_nx) row[x0 + j] = a[(j + first) * kFpWG];
}

// ------------------------------------------------ sliding lane windows
//
// With fixed windows a lane owns bins [x0, x0+B) of its (view, row) for the
// whole walk, and at tilt t only about a fraction cos t of a view's lanes find
// bins of a z row in reach: the others idle while the wave walks.  A sliding
// map gives the view's ncol lane columns (64 per workgroup in x) windows of
//   bw = ceil((cos t * (out_nx-1) + 4) / (ncol-1)) <= B
// bins, and for each z row makes windows w0 .. w0+ncol-1 live, with
// w0 = floor(lo / bw) and lo = max(floor(px of voxel 0), 0).  px is
// non-decreasing in ox, so every bin the row reaches lies in
// [lo, lo + cos t*(out_nx-1) + 2]: inside the live windows, which reach
// lo + (ncol-1)*bw.  Window w is always walked by column w mod ncol, whose
// lane-private slots keep its sums, so each bin still receives its terms from
// one lane in the reference's (oz, ox) order: the same bits as fixed windows.
// px of voxel 0 is monotonic in z (zc is monotonic in oz and every rounding
// step is), so w0 moves one way: a window that leaves the live range is final;
// its lane writes it out, clears the slots and takes the window ncol further
// on.  Windows of a column never live get no terms and are written as zeros.
struct SlideMap {
    int bw;          // bins per window of this view
    int ncol;        // lane columns of the view
    int col;         // this lane's column
    int w;           // the lane's current window
    int wmin, wmax;  // first and last window the lane has held
    float inv_bw;
    bool slide;      // uniform over the wave (one view)
};

template <int B>
__device__ __forceinline__ SlideMap slide_map(const FpParams& P, float4 g, int nbx, int bx,
                                              int lane) {
    SlideMap M;
    M.ncol = 64 * nbx;
    M.col = bx * 64 + lane;
    M.bw = B;
    M.slide = false;
    if (P.slide_windows) {
        const float need = (g.x * g.z * static_cast<float>(P.out_nx - 1) + 4.0f) /
                           static_cast<float>(M.ncol - 1);
        const int b = static_cast<int>(ceilf(need));
        if (b < B) {
            M.bw = max(b, 1);
            M.slide = true;
        }
    }
    M.inv_bw = 1.0f / static_cast<float>(M.bw);
    M.w = M.col;
    M.wmin = M.wmax = M.col;
    return M;
}

// First live window of z row zc (uniform over the wave).
__device__ __forceinline__ int slide_w0(const WalkLane& L, const SlideMap& M, float zcv) {
    const int lo = max(cvt_floor(walk_px(L, -L.ocx, zcv)), 0);
    int w0 = static_cast<int>(static_cast<float>(lo) * M.inv_bw);
    if (w0 * M.bw > lo) --w0;
    if ((w0 + 1) * M.bw <= lo) ++w0;
    return w0;
}

// Exact-window bounds and slot offset of the lane's window [x0, x0+bw).
__device__ __forceinline__ void set_window(WalkLane& L, const FpParams& P, int x0, int bw) {
    L.b_lo = static_cast<float>(max(x0 - 1, 0));
    L.b_hi = static_cast<float>(min(x0 + bw, P.in_nx - 1));
    L.qoff = 4u * threadIdx.x + static_cast<unsigned>(1 - x0) * (4u * kFpWG);
}

// Write the bins of window w (slots 1..bw of each row's slot set, R1 floats
// apart) to the sinogram rows (row 0 at sino + rowoff, row 1 after it when
// two), and clear the slots if CLEAR.
template <int NR, bool CLEAR>
__device__ __forceinline__ void window_flush(const FpParams& P, float* a, int R1, int x0, int bw,
                                             int64_t rowoff, bool two) {
    float* row0 = P.sino + rowoff;
    float* row1 = two ? row0 + P.in_nx : nullptr;
    for (int j = 0; j < bw; ++j) {
        float* s = a + (j + 1) * kFpWG;
        if (x0 + j < P.in_nx) {
            row0[x0 + j] = s[0];
            if (NR == 2 && row1) row1[x0 + j] = s[R1];
        }
        if constexpr (CLEAR) {
            s[0] = 0.f;
            if (NR == 2) s[R1] = 0.f;
        }
    }
}

// Before z row zc: move the lane to its live window, writing out the windows
// that left the live range.
template <int NR>
__device__ __forceinline__ void slide_to(const FpParams& P, WalkLane& L, SlideMap& M, float* a,
                                         int R1, float zcv, int64_t rowoff, bool two) {
    const int w0 = slide_w0(L, M, zcv);
    if (M.w < w0 || M.w >= w0 + M.ncol) {
        do {
            window_flush<NR, true>(P, a, R1, M.w * M.bw, M.bw, rowoff, two);
            M.w += M.w < w0 ? M.ncol : -M.ncol;
        } while (M.w < w0 || M.w >= w0 + M.ncol);
        set_window(L, P, M.w * M.bw, M.bw);
        M.wmin = min(M.wmin, M.w);
        M.wmax = max(M.wmax, M.w);
    }
}

// First window of the lane at z row zc.
__device__ __forceinline__ void slide_start(const FpParams& P, WalkLane& L, SlideMap& M,
                                            float zcv) {
    const int w0 = slide_w0(L, M, zcv);
    int r = (M.col - w0) % M.ncol;
    if (r < 0) r += M.ncol;
    M.w = w0 + r;
    M.wmin = M.wmax = M.w;
    set_window(L, P, M.w * M.bw, M.bw);
}

// After the walk: the current window, and zeros for the column's windows that
// were never live.
template <int NR>
__device__ __forceinline__ void slide_finish(const FpParams& P, const SlideMap& M, float* a,
                                             int R1, int64_t rowoff, bool two) {
    window_flush<NR, false>(P, a, R1, M.w * M.bw, M.bw, rowoff, two);
    if (!M.slide) return;
    float* row0 = P.sino + rowoff;
    float* row1 = two ? row0 + P.in_nx : nullptr;
    for (int w = M.col; w * M.bw < P.in_nx; w += M.ncol) {
        if (w >= M.wmin && w <= M.wmax) continue;
        const int x0 = w * M.bw;
        for (int j = 0; j < M.bw && x0 + j < P.in_nx; ++j) {
            row0[x0 + j] = 0.f;
            if (NR == 2 && row1) row1[x0 + j] = 0.f;
        }
    }
}

// Unstaged form: each wave walks its own (view, row) reading the volume from
// HBM, or scans bin by bin for views with cos_t*cos_xt < kWalkMinC.
template <int B>
__device__ __forceinline__ void fp_walk_global(const FpParams& P, float* lds, int x0, int iy,
                                               int p) {
    float* a = lds + threadIdx.x;
    const float4 g = P.vgeom[p];
    if (!(g.x * g.z >= kWalkMinC)) {
        float* row = P.sino + (static_cast<int64_t>(p) * P.in_ny + iy) * P.in_nx;
        for (int j = 0; j < B; ++j)
            if (x0 + j < P.in_nx) fp_bin_scan(P, x0 + j, iy, p, row + x0 + j);
        return;
    }
    const WalkLane L = walk_lane(P, g, x0, B, 2);
#pragma unroll
    for (int s = 2; s < B + 2; ++s) a[s * kFpWG] = 0.f;
    const float* zc = P.zc_tab + static_cast<int64_t>(p) * P.nz_pad;
    const float* vrow = P.vol + static_cast<int64_t>(iy) * P.out_nz * P.out_nx;
    float zcv = zc[0];
    for (int oz = 0; oz < P.out_nz; ++oz) {
        const float zc_next = zc[min(oz + 1, P.out_nz - 1)];
        walk_row<B, false>(L, reinterpret_cast<char*>(lds),
                           vrow + static_cast<int64_t>(oz) * P.out_nx, zcv);
        zcv = zc_next;
    }
    walk_store<B>(P, a, x0, iy, p, 2);
}

// Rows of at most kStageNx voxels are staged in LDS: the 8 waves of a workgroup
// walk 8 views of the same volume row, which the workgroup reads from HBM once,
// coalesced, one z row ahead (registers) and double-buffered (LDS).
constexpr int kStageRegs = 4096 / kFpWG;
constexpr int kStageNx = kStageRegs * kFpWG;
__host__ __device__ constexpr int fp_stage_stride(int nx) { return nx + 4; }

template <int B>
__device__ __forceinline__ void fp_walk_block(const FpParams& P, float* lds, int blk) {
    // blk = ((iy * nvg) + vg) * nbx + bx: the view groups of one row are
    // adjacent, so the blocks that read one volume slice run together.
    const int nbx = (P.in_nx + 64 * B - 1) / (64 * B);
    const int nvg = (P.nviews + kFpViewsPerWG - 1) / kFpViewsPerWG;
    const int bx = blk % nbx;
    const int rest = blk / nbx;
    const int vg = rest % nvg;
    const int iy = rest / nvg;
    const int wave = static_cast<int>(threadIdx.x) >> 6;
    const int lane = static_cast<int>(threadIdx.x) & 63;
    // wave-uniform: view tables are read through the scalar cache
    const int p = __builtin_amdgcn_readfirstlane(vg * kFpViewsPerWG + wave);
    const int x0 = (bx * 64 + lane) * B;
    float* a = lds + threadIdx.x;
    // workgroup-uniform choice of the staged form
    bool staged = P.out_nx <= kStageNx;
    for (int w = 0; w < kFpViewsPerWG; ++w) {
        const int pw = vg * kFpViewsPerWG + w;
        if (pw < P.nviews) {
            const float4 gw = P.vgeom[pw];
            if (!(gw.x * gw.z >= kWalkMinC)) staged = false;
        }
    }
    if (!staged) {
        if (p >= P.nviews || x0 >= P.in_nx) return;
        fp_walk_global<B>(P, lds, x0, iy, p);
        return;
    }
    // Every thread takes part in staging and barriers; only live lanes walk.
    const bool wave_live = p < P.nviews;
    const int pv = wave_live ? p : 0;
    const float4 g = P.vgeom[pv];
    // exact windows: slot 0 and slot B+1 take the terms of bins x0-1 and x0+B
    WalkLane L = walk_lane(P, g, x0, B, 1);
    SlideMap M = slide_map<B>(P, g, nbx, bx, lane);
    const int64_t rowoff = (static_cast<int64_t>(pv) * P.in_ny + iy) * P.in_nx;
#pragma unroll
    for (int s = 1; s < B + 1; ++s) a[s * kFpWG] = 0.f;
    const float* zc = P.zc_tab + static_cast<int64_t>(pv) * P.nz_pad;
    walk_view_bounds(L, zc[0], zc[P.out_nz - 1]);
    if (M.slide) slide_start(P, L, M, zc[0]);
    const int nx = P.out_nx;
    const int nz = P.out_nz;
    const int t = static_cast<int>(threadIdx.x);
    float* stg = lds + (B + 3) * kFpWG;
    const int ns = fp_stage_stride(nx);
    const float* vrow = P.vol + static_cast<int64_t>(iy) * nz * nx;
    float r[kStageRegs];
    auto fetch = [&](int oz) {
        const float* src = vrow + static_cast<int64_t>(oz) * nx;
#pragma unroll
        for (int k = 0; k < kStageRegs; ++k)
            if (t + k * kFpWG < nx) r[k] = src[t + k * kFpWG];
    };
    auto put = [&](int buf) {
        float* dst = stg + buf * ns;
#pragma unroll
        for (int k = 0; k < kStageRegs; ++k)
            if (t + k * kFpWG < nx) dst[t + k * kFpWG] = r[k];
    };
    fetch(0);
    put(0);
    if (nz > 1) fetch(1);
    __syncthreads();
    float zcv = zc[0];
    for (int oz = 0; oz < nz; ++oz) {
        const float zc_next = zc[min(oz + 1, nz - 1)];
        if (oz + 1 < nz) put((oz + 1) & 1);
        if (oz + 2 < nz) fetch(oz + 2);
        if (wave_live) {
            if (M.slide && oz > 0) slide_to<1>(P, L, M, a, 0, zcv, rowoff, false);
            if (M.w * M.bw < P.in_nx)
                walk_row<B, true>(L, reinterpret_cast<char*>(lds), stg + (oz & 1) * ns, zcv);
        }
        zcv = zc_next;
        __syncthreads();
    }
    if (wave_live) slide_finish<1>(P, M, a, 0, rowoff, false);
}

// ------------------------------------------------ two volume rows per workgroup
//
// In the single-row geometry proj_x does not depend on the volume row, so a
// lane's window of a z row, and each voxel's px, floor, fraction and bin
// advance, serve two rows at once; only the products and the register-carried
// sums are per row.  Row 1's slots sit R1 = (B+3)*256 floats after row 0's (an
// immediate offset of the same address), and the two rows' staged z rows are
// interleaved as float2, so one paired read feeds two voxels of both rows.
// Each row's bins receive exactly the terms, in exactly the order, of the
// one-row walk: the sinogram is the same bits.

struct RegWalk2 {
    float r0[2], r1[2], la[2];
    float* q;
    int fi;
};

template <int B>
__device__ __forceinline__ void reg_begin2(const WalkLane& L, char* qb, float dx, float zcv,
                                           float2 v, RegWalk2& W) {
    constexpr int R1 = (B + 3) * kFpWG;
    const float px = walk_px(L, dx, zcv);
    W.fi = cvt_floor(px);
    const float frac = __builtin_amdgcn_fractf(px);
    const float omf = 1.0f - frac;
    W.q = reinterpret_cast<float*>(qb + W.fi * (4 * kFpWG));
    W.r0[0] = W.q[0] + v.x * omf;
    W.r1[0] = W.q[kFpWG] + v.x * frac;
    W.la[0] = W.q[2 * kFpWG];
    W.r0[1] = W.q[R1] + v.y * omf;
    W.r1[1] = W.q[R1 + kFpWG] + v.y * frac;
    W.la[1] = W.q[R1 + 2 * kFpWG];
}

template <int B>
__device__ __forceinline__ void reg_step2(const WalkLane& L, char* qb, float dx, float zcv,
                                          float2 v, RegWalk2& W) {
    constexpr int R1 = (B + 3) * kFpWG;
    const float px = walk_px(L, dx, zcv);
    const int fi = cvt_floor(px);
    const float frac = __builtin_amdgcn_fractf(px);
    const float omf = 1.0f - frac;
    const float a0 = v.x * omf, b0 = v.x * frac;
    const float a1 = v.y * omf, b1 = v.y * frac;
    W.q[0] = W.r0[0];
    W.q[R1] = W.r0[1];
    float* q = reinterpret_cast<float*>(qb + fi * (4 * kFpWG));
    const float la0 = q[2 * kFpWG];
    const float la1 = q[R1 + 2 * kFpWG];
    const bool adv = fi != W.fi;  // fi == W.fi + 1
    const float s00 = adv ? W.r1[0] : W.r0[0];
    const float s01 = adv ? W.la[0] : W.r1[0];
    const float s10 = adv ? W.r1[1] : W.r0[1];
    const float s11 = adv ? W.la[1] : W.r1[1];
    W.r0[0] = s00 + a0;
    W.r1[0] = s01 + b0;
    W.la[0] = la0;
    W.r0[1] = s10 + a1;
    W.r1[1] = s11 + b1;
    W.la[1] = la1;
    W.q = q;
    W.fi = fi;
}

template <int B>
__device__ __forceinline__ void walk_voxel_exact2(const WalkLane& L, char* lds, float dx,
                                                  float zcv, float2 v) {
    constexpr int R1 = (B + 3) * kFpWG;
    const float px = walk_px(L, dx, zcv);
    const int fi = cvt_floor(px);
    const float frac = __builtin_amdgcn_fractf(px);
    float* q = reinterpret_cast<float*>(lds + (static_cast<unsigned>(fi) * (4u * kFpWG) + L.qoff));
    q[0] += v.x * (1.0f - frac);
    q[kFpWG] += v.x * frac;
    q[R1] += v.y * (1.0f - frac);
    q[R1 + kFpWG] += v.y * frac;
}

// walk_row's staged form for two rows (vz: the z row of both, interleaved).
template <int B>
__device__ __forceinline__ void walk_row2(const WalkLane& L, char* lds,
                                          const float2* __restrict__ vz, float zcv) {
    constexpr int R1 = (B + 3) * kFpWG;
    int e0, e1;
    walk_window_exact(L, zcv, e0, e1);
    float dx = static_cast<float>(e0) - L.ocx;
    const float2* vp = vz + e0;
    const float2* vend = vz + e1;
    const float2* vlast = vend - 1;
    if (L.regwalk) {
        if (vp < vend) {
            // odd count: first voxel alone; even: a value-0 voxel first (walk_row)
            char* const qb = lds + L.qoff;
            const bool odd = ((e1 - e0) & 1) != 0;
            RegWalk2 W;
            reg_begin2<B>(L, qb, dx, zcv, odd ? vp[0] : make_float2(0.0f, 0.0f), W);
            if (odd) {
                ++vp;
                dx += 1.0f;
            }
            for (; vp < vend; vp += 2) {
                const float2 v0 = vp[0], v1 = vp[1];
                reg_step2<B>(L, qb, dx, zcv, v0, W);
                reg_step2<B>(L, qb, dx + 1.0f, zcv, v1, W);
                dx += 2.0f;
            }
            W.q[0] = W.r0[0];
            W.q[kFpWG] = W.r1[0];
            W.q[R1] = W.r0[1];
            W.q[R1 + kFpWG] = W.r1[1];
        }
    } else {
        for (; vp < vlast; vp += 2) {
            const float2 v0 = vp[0], v1 = vp[1];
            walk_voxel_exact2<B>(L, lds, dx, zcv, v0);
            walk_voxel_exact2<B>(L, lds, dx + 1.0f, zcv, v1);
            dx += 2.0f;
        }
        if (vp < vend) walk_voxel_exact2<B>(L, lds, dx, zcv, vp[0]);
    }
}

constexpr int kStageRegs2 = kFpRows2Nx / kFpWG;

template <int B>
__device__ __forceinline__ void fp_walk_block2(const FpParams& P, float* lds, int blk) {
    constexpr int R1 = (B + 3) * kFpWG;
    const int nbx = (P.in_nx + 64 * B - 1) / (64 * B);
    const int nvg = (P.nviews + kFpViewsPerWG - 1) / kFpViewsPerWG;
    const int bx = blk % nbx;
    const int rest = blk / nbx;
    const int vg = rest % nvg;
    const int iy = (rest / nvg) * 2;
    const bool two = iy + 1 < P.in_ny;  // uniform: the last row of an odd count walks alone
    const int wave = static_cast<int>(threadIdx.x) >> 6;
    const int lane = static_cast<int>(threadIdx.x) & 63;
    const int p = __builtin_amdgcn_readfirstlane(vg * kFpViewsPerWG + wave);
    const int x0 = (bx * 64 + lane) * B;
    bool staged = P.out_nx <= kFpRows2Nx;
    for (int w = 0; w < kFpViewsPerWG; ++w) {
        const int pw = vg * kFpViewsPerWG + w;
        if (pw < P.nviews) {
            const float4 gw = P.vgeom[pw];
            if (!(gw.x * gw.z >= kWalkMinC)) staged = false;
        }
    }
    if (!staged) {
        if (p >= P.nviews || x0 >= P.in_nx) return;
        fp_walk_global<B>(P, lds, x0, iy, p);
        if (two) fp_walk_global<B>(P, lds, x0, iy + 1, p);
        return;
    }
    const bool wave_live = p < P.nviews;
    const int pv = wave_live ? p : 0;
    const float4 g = P.vgeom[pv];
    WalkLane L = walk_lane(P, g, x0, B, 1);
    SlideMap M = slide_map<B>(P, g, nbx, bx, lane);
    const int64_t rowoff = (static_cast<int64_t>(pv) * P.in_ny + iy) * P.in_nx;
    float* a = lds + threadIdx.x;
#pragma unroll
    for (int s = 1; s < B + 1; ++s) {
        a[s * kFpWG] = 0.f;
        a[R1 + s * kFpWG] = 0.f;
    }
    const float* zc = P.zc_tab + static_cast<int64_t>(pv) * P.nz_pad;
    walk_view_bounds(L, zc[0], zc[P.out_nz - 1]);
    if (M.slide) slide_start(P, L, M, zc[0]);
    const int nx = P.out_nx;
    const int nz = P.out_nz;
    const int t = static_cast<int>(threadIdx.x);
    float2* stg = reinterpret_cast<float2*>(lds + 2 * R1);
    const float* vrow0 = P.vol + static_cast<int64_t>(iy) * nz * nx;
    const float* vrow1 = two ? vrow0 + static_cast<int64_t>(nz) * nx : vrow0;
    float r0[kStageRegs2], r1[kStageRegs2];
    auto fetch = [&](int oz) {
        const float* s0 = vrow0 + static_cast<int64_t>(oz) * nx;
        const float* s1 = vrow1 + static_cast<int64_t>(oz) * nx;
#pragma unroll
        for (int k = 0; k < kStageRegs2; ++k)
            if (t + k * kFpWG < nx) {
                r0[k] = s0[t + k * kFpWG];
                r1[k] = two ? s1[t + k * kFpWG] : 0.f;
            }
    };
    auto put = [&]() {
#pragma unroll
        for (int k = 0; k < kStageRegs2; ++k)
            if (t + k * kFpWG < nx) stg[t + k * kFpWG] = make_float2(r0[k], r1[k]);
    };
    // One staging buffer (a second barrier per z row): with 4 views per workgroup
    // the two slot sets and a double buffer took 39 KB at B = 8 (four workgroups
    // per CU) against 31 KB (five), which outweighed the barrier
    // (1024^2 x 61 -> z300 launch: -5%); with 8 views, 52 against 60 KB at B = 8
    // keeps three workgroups per CU instead of two.
    fetch(0);
    put();
    if (nz > 1) fetch(1);
    __syncthreads();
    float zcv = zc[0];
    for (int oz = 0; oz < nz; ++oz) {
        const float zc_next = zc[min(oz + 1, nz - 1)];
        if (wave_live) {
            if (M.slide && oz > 0) slide_to<2>(P, L, M, a, R1, zcv, rowoff, two);
            if (M.w * M.bw < P.in_nx) walk_row2<B>(L, reinterpret_cast<char*>(lds), stg, zcv);
        }
        zcv = zc_next;
        __syncthreads();
        if (oz + 1 < nz) {
            put();
            if (oz + 2 < nz) fetch(oz + 2);
            __syncthreads();
        }
    }
    if (wave_live) slide_finish<2>(P, M, a, R1, rowoff, two);
}

template <int RPW>
__device__ __forceinline__ void fp_block(const FpParams& P, float* lds, int blk) {
    if constexpr (RPW == 2) {
        switch (P.bins_per_lane) {
            case 4: fp_walk_block2<4>(P, lds, blk); break;
            case 8: fp_walk_block2<8>(P, lds, blk); break;
            default: fp_walk_block2<16>(P, lds, blk); break;
        }
        return;
    }
    switch (P.bins_per_lane) {
        case 4: fp_walk_block<4>(P, lds, blk); break;
        case 8: fp_walk_block<8>(P, lds, blk); break;
        default: fp_walk_block<16>(P, lds, blk); break;
    }
}

// RPW: volume rows per workgroup (FpParams::rows_per_wg), a kernel each so that
// the one-row walk keeps its own register allocation.
template <int RPW>
__global__ void __launch_bounds__(kFpWG) __attribute__((amdgpu_waves_per_eu(RPW == 2 ? 5 : 7))) fp_walk_kernel(FpParams P, int nblk) {
    extern __shared__ float fp_lds[];
    fp_block<RPW>(P, fp_lds, xcd_remap(blockIdx.x, nblk));
}

__host__ __device__ inline int fp_blocks(const FpParams& p) {
    const int B = p.bins_per_lane;
    const int nbx = (p.in_nx + 64 * B - 1) / (64 * B);
    const int nvg = (p.nviews + kFpViewsPerWG - 1) / kFpViewsPerWG;
    const int rows = p.rows_per_wg == 2 ? (p.in_ny + 1) / 2 : p.in_ny;
    return nbx * nvg * rows;
}

// Several series in one launch; a workgroup's work and order are those of a
// launch of its own.  Series r owns workgroups [block_start[r], block_start[r+1]),
// a range that starts at a multiple of 8 and is padded to one, so that its
// workgroups spread evenly over the 8 XCDs (a contiguous run of each series per
// XCD); padding workgroups exit at once.
template <int RPW>
__global__ void __launch_bounds__(kFpWG) __attribute__((amdgpu_waves_per_eu(RPW == 2 ? 5 : 7))) fp_walk_batch_kernel(FpBatch Bt, int nblk) {
    extern __shared__ float fp_lds[];
    const int b = blockIdx.x;
    if (b >= nblk) return;
    int lo = 0, hi = Bt.nreq - 1;
    while (lo < hi) {
        const int m = (lo + hi + 1) / 2;
        if (Bt.block_start[m] <= b)
            lo = m;
        else
            hi = m - 1;
    }
    const FpParams& P = Bt.descs[lo];
    const int local = b - Bt.block_start[lo];
    const int per = (Bt.block_start[lo + 1] - Bt.block_start[lo]) / 8;
    const int logical = (local % 8) * per + local / 8;
    if (logical >= fp_blocks(P)) return;
    fp_block<RPW>(P, fp_lds, logical);
}

// ---------------------------------------------------------------- two-row
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

// ViewPatches::interpolate (main.rs:166-224) on the device tables.
__device__ bool fp_local_shift(const FpScatterParams& P, int view, float x, float y, float& odx,
                               float& ody) {
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

// term * 2^k rounded to nearest as a 64-bit two's-complement integer, added with
// an integer atomic (the sum is independent of the order of the additions).
__device__ __forceinline__ void fixed_add(unsigned long long* a, float term, double scale) {
    const long long q = __double2ll_rn(static_cast<double>(term) * scale);
    atomicAdd(a, static_cast<unsigned long long>(q));
}

// k such that count_bound * max|V| * 2^k < 2^62 (0 for an all-zero volume).
__global__ void fp_fixed_exponent_kernel(const unsigned* absmax_bits, double count_bound,
                                         int* k_out) {
    const double bound = static_cast<double>(__uint_as_float(*absmax_bits)) * count_bound;
    int k = 0;
    if (bound > 0.0) {
        int e;
        (void)frexp(bound, &e);  // bound < 2^e
        k = 62 - e;
    }
    *k_out = max(-1000, min(1000, k));
}

__global__ void __launch_bounds__(256) fp_fixed_finalize_kernel(const unsigned long long* acc,
                                                               float* sino, int64_t n,
                                                               const int* k) {
    const double inv = ldexp(1.0, -*k);
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        sino[i] = static_cast<float>(static_cast<double>(static_cast<long long>(acc[i])) * inv);
}

__global__ void __launch_bounds__(256) fp_zero_u64_kernel(unsigned long long* a, int64_t n) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        a[i] = 0ull;
}

__global__ void __launch_bounds__(256) fp_absmax_kernel(const float* v, int64_t n,
                                                       unsigned* absmax_bits) {
    __shared__ float red[256];
    float m = 0.f;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        m = fmaxf(m, fabsf(v[i]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(absmax_bits, __float_as_uint(red[0]));
}

__global__ void fp_zero_u32_kernel(unsigned* p) { *p = 0u; }

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
    unsigned long long* a0 = P.acc + (static_cast<int64_t>(p) * P.in_ny + sy0) * P.in_nx;
    unsigned long long* a1 = a0 + P.in_nx;
    const double scale = ldexp(1.0, *P.scale_exp);
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
            fixed_add(a0 + px0, val * (1.0f - frac) * w0, scale);
            fixed_add(a0 + px0 + 1, val * frac * w0, scale);
            fixed_add(a1 + px0, val * (1.0f - frac) * w1, scale);
            fixed_add(a1 + px0 + 1, val * frac * w1, scale);
        } else {
            fixed_add(a0 + px0, val * (1.0f - frac), scale);
            fixed_add(a0 + px0 + 1, val * frac, scale);
        }
    }
}

}  // namespace

size_t fp_lds_bytes(const FpParams& p) {
    // unstaged workgroups: B + 4 slots per lane; staged: B + 3 slots and the
    // two staged volume rows (double buffer); two volume rows per workgroup:
    // two slot sets and one buffer of the rows' z rows, interleaved
    size_t b = static_cast<size_t>(p.bins_per_lane + 4) * kFpWG * sizeof(float);
    if (p.rows_per_wg == 2) {
        if (p.out_nx <= kFpRows2Nx)
            b = std::max(b, 2 * static_cast<size_t>(p.bins_per_lane + 3) * kFpWG * sizeof(float) +
                                static_cast<size_t>(fp_stage_stride(p.out_nx)) * sizeof(float2));
        return b;
    }
    if (p.out_nx <= kStageNx)
        b = std::max(b, static_cast<size_t>(p.bins_per_lane + 3) * kFpWG * sizeof(float) +
                            2 * static_cast<size_t>(fp_stage_stride(p.out_nx)) * sizeof(float));
    return b;
}

int fp_gather_blocks(const FpParams& p) { return fp_blocks(p); }

hipError_t launch_forward_gather(const FpParams& p, hipStream_t st) {
    if (p.nviews <= 0 || p.in_ny <= 0) return hipSuccess;
    if (p.bins_per_lane != 4 && p.bins_per_lane != 8 && p.bins_per_lane != 16)
        return hipErrorInvalidValue;
    if (p.rows_per_wg != 1 && p.rows_per_wg != 2) return hipErrorInvalidValue;
    const int nblk = fp_gather_blocks(p);
    const size_t lds = fp_lds_bytes(p);
    if (p.rows_per_wg == 2) {
        const hipError_t e = lds_opt_in(reinterpret_cast<const void*>(fp_walk_kernel<2>), lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(fp_walk_kernel<2>, dim3(nblk), dim3(kFpWG), lds, st, p, nblk);
    } else {
        const hipError_t e = lds_opt_in(reinterpret_cast<const void*>(fp_walk_kernel<1>), lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(fp_walk_kernel<1>, dim3(nblk), dim3(kFpWG), lds, st, p, nblk);
    }
    return hipGetLastError();
}

hipError_t launch_forward_gather_batch(const FpBatch& b, int total_blocks, size_t lds,
                                       int rows_per_wg, hipStream_t st) {
    if (total_blocks <= 0) return hipSuccess;
    if (rows_per_wg == 2) {
        const hipError_t e =
            lds_opt_in(reinterpret_cast<const void*>(fp_walk_batch_kernel<2>), lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(fp_walk_batch_kernel<2>, dim3(total_blocks), dim3(kFpWG), lds, st, b,
                           total_blocks);
    } else {
        const hipError_t e =
            lds_opt_in(reinterpret_cast<const void*>(fp_walk_batch_kernel<1>), lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(fp_walk_batch_kernel<1>, dim3(total_blocks), dim3(kFpWG), lds, st, b,
                           total_blocks);
    }
    return hipGetLastError();
}

hipError_t launch_forward_scatter(const FpScatterParams& p, unsigned* absmax_scratch,
                                  int* scale_exp, hipStream_t st) {
    // every section of the sinogram is written (views >= nviews stay zero)
    const int64_t nsino = static_cast<int64_t>(p.nz_sections) * p.in_ny * p.in_nx;
    if (nsino <= 0) return hipSuccess;
    const int64_t nvol = static_cast<int64_t>(p.in_ny) * p.out_nz * p.out_nx;
    const unsigned gs = static_cast<unsigned>(std::min<int64_t>(4096, (nsino + 255) / 256));
    hipLaunchKernelGGL(fp_zero_u64_kernel, dim3(gs), dim3(256), 0, st, p.acc, nsino);
    hipLaunchKernelGGL(fp_zero_u32_kernel, dim3(1), dim3(1), 0, st, absmax_scratch);
    const unsigned gv = static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(2048, (nvol + 255) / 256)));
    hipLaunchKernelGGL(fp_absmax_kernel, dim3(gv), dim3(256), 0, st, p.vol, nvol, absmax_scratch);
    // a bin receives at most one term per voxel of the volume
    const double count_bound = static_cast<double>(std::max<int64_t>(nvol, 1));
    hipLaunchKernelGGL(fp_fixed_exponent_kernel, dim3(1), dim3(1), 0, st, absmax_scratch,
                       count_bound, scale_exp);
    if (p.nviews > 0 && p.in_ny > 0) {
        dim3 grid((p.out_nx + 255) / 256, p.in_ny, p.nviews);
        hipLaunchKernelGGL(fp_scatter_kernel, grid, dim3(256), 0, st, p);
    }
    hipLaunchKernelGGL(fp_fixed_finalize_kernel, dim3(gs), dim3(256), 0, st, p.acc, p.sino, nsino,
                       static_cast<const int*>(scale_exp));
    return hipGetLastError();
}

}  // namespace tiltgpu
