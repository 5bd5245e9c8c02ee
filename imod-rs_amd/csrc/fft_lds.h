// Complex DFTs of rows held in LDS for gfx950: self-sorting Stockham passes of
// radix 4, 2, 3, 5 (unrolled butterflies) and a generic pass for any other
// prime, driven by the mixed-radix plan of a FilterParams (n, radix[], the
// FastDiv divisors and the twiddle table tw[t] = exp(-2*pi*i*t/n)).  Shared by
// the R-weighting filter and the CTF phase flip.
#pragma once

#include <hip/hip_runtime.h>

#include "tiltgpu_internal.h"

namespace tiltgpu {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
// multiply by -i
__device__ __forceinline__ float2 cmul_mi(float2 a) { return make_float2(a.y, -a.x); }

template <int R>
__device__ __forceinline__ void dft_small(float2* x);

template <>
__device__ __forceinline__ void dft_small<2>(float2* x) {
    const float2 a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = csub(a, b);
}

template <>
__device__ __forceinline__ void dft_small<4>(float2* x) {
    const float2 t0 = cadd(x[0], x[2]), t1 = csub(x[0], x[2]);
    const float2 t2 = cadd(x[1], x[3]), t3 = cmul_mi(csub(x[1], x[3]));
    x[0] = cadd(t0, t2);
    x[2] = csub(t0, t2);
    x[1] = cadd(t1, t3);
    x[3] = csub(t1, t3);
}

template <>
__device__ __forceinline__ void dft_small<3>(float2* x) {
    const float h = 0.86602540378443864676f;  // sin(2*pi/3)
    const float2 s = cadd(x[1], x[2]), d = csub(x[1], x[2]);
    const float2 m = make_float2(x[0].x - 0.5f * s.x, x[0].y - 0.5f * s.y);
    x[0] = cadd(x[0], s);
    x[1] = make_float2(m.x + h * d.y, m.y - h * d.x);
    x[2] = make_float2(m.x - h * d.y, m.y + h * d.x);
}

template <>
__device__ __forceinline__ void dft_small<5>(float2* x) {
    const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
    const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
    const float2 a1 = cadd(x[1], x[4]), b1 = csub(x[1], x[4]);
    const float2 a2 = cadd(x[2], x[3]), b2 = csub(x[2], x[3]);
    const float2 x0 = x[0];
    const float2 t1 = make_float2(x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y);
    const float2 t2 = make_float2(x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y);
    const float2 u1 = make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y);
    const float2 u2 = make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y);
    x[0] = make_float2(x0.x + a1.x + a2.x, x0.y + a1.y + a2.y);
    x[1] = make_float2(t1.x + u1.y, t1.y - u1.x);
    x[4] = make_float2(t1.x - u1.y, t1.y + u1.x);
    x[2] = make_float2(t2.x + u2.y, t2.y - u2.x);
    x[3] = make_float2(t2.x - u2.y, t2.y + u2.x);
}

// One Stockham pass of radix R over `pairs` complex rows of length n:
// out[(j/Ns)*Ns*R + j%Ns + q*Ns] = sum_r in[j + r*n/R] * w^(r*(j%Ns)) * W_R^(rq).
template <int R>
__device__ __forceinline__ void stockham_pass(const float2* __restrict__ in,
                                              float2* __restrict__ out, int n, int pairs,
                                              int Ns, const float2* __restrict__ tw,
                                              const FastDiv& mdiv, const FastDiv& nsdiv) {
    const int m = n / R;
    const int tstep = n / (Ns * R);
    for (int idx = threadIdx.x; idx < pairs * m; idx += blockDim.x) {
        const int q = static_cast<int>(mdiv.div(static_cast<uint32_t>(idx)));
        const int j = idx - q * m;
        const int jq = static_cast<int>(nsdiv.div(static_cast<uint32_t>(j)));
        const int k = j - jq * Ns;
        const float2* src = in + q * n;
        float2 x[R];
#pragma unroll
        for (int r = 0; r < R; ++r) x[r] = src[j + r * m];
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) x[r] = cmul(x[r], tw[r * k * tstep]);
        }
        dft_small<R>(x);
        float2* dst = out + q * n + jq * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) dst[r * Ns] = x[r];
    }
}

// Odd prime radix R (7, 11, 13) unrolled: one butterfly per work item, the
// R inputs in registers, and the R-point DFT from the symmetric sums
// a_r = x_r + x_{R-r}, b_r = x_r - x_{R-r} (r = 1 .. (R-1)/2):
//   y_q = x_0 + sum_r a_r cos(2 pi rq/R) - i b_r sin(2 pi rq/R),
// with cos / sin of 2 pi e/R read once per pass from the twiddle table
// (tw[e * n/R] = exp(-2 pi i e/R)).
template <int R>
__device__ __forceinline__ void stockham_pass_odd(const float2* __restrict__ in,
                                                  float2* __restrict__ out, int n, int pairs,
                                                  int Ns, const float2* __restrict__ tw,
                                                  const FastDiv& mdiv, const FastDiv& nsdiv) {
    constexpr int H = (R - 1) / 2;
    const int m = n / R;
    const int tstep = n / (Ns * R);
    float cs[H + 1], sn[H + 1];
#pragma unroll
    for (int e = 1; e <= H; ++e) {
        const float2 w = tw[e * m];
        cs[e] = w.x;
        sn[e] = -w.y;
    }
    for (int idx = threadIdx.x; idx < pairs * m; idx += blockDim.x) {
        const int q = static_cast<int>(mdiv.div(static_cast<uint32_t>(idx)));
        const int j = idx - q * m;
        const int jq = static_cast<int>(nsdiv.div(static_cast<uint32_t>(j)));
        const int k = j - jq * Ns;
        const float2* src = in + q * n;
        float2 x[R];
#pragma unroll
        for (int r = 0; r < R; ++r) x[r] = src[j + r * m];
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) x[r] = cmul(x[r], tw[r * k * tstep]);
        }
        float2 a[H + 1], b[H + 1];
        float2 y0 = x[0];
#pragma unroll
        for (int r = 1; r <= H; ++r) {
            a[r] = cadd(x[r], x[R - r]);
            b[r] = csub(x[r], x[R - r]);
            y0 = cadd(y0, a[r]);
        }
        float2* dst = out + q * n + jq * Ns * R + k;
        dst[0] = y0;
#pragma unroll
        for (int o = 1; o < R; ++o) {
            float yx = x[0].x, yy = x[0].y;
#pragma unroll
            for (int r = 1; r <= H; ++r) {
                const int e = (r * o) % R;
                const float c = e <= H ? cs[e] : cs[R - e];
                const float si = e <= H ? sn[e] : -sn[R - e];
                yx = fmaf(a[r].x, c, fmaf(b[r].y, si, yx));
                yy = fmaf(a[r].y, c, fmaf(-b[r].x, si, yy));
            }
            dst[o * Ns] = make_float2(yx, yy);
        }
    }
}

// Generic radix: any prime R the unrolled passes do not cover.  Each output of
// each butterfly is one work item (an O(R) sum with W_R^e = tw[e * n/R]), so a
// large prime spreads over the threads.
static __device__ __noinline__ void stockham_pass_generic(const float2* __restrict__ in,
                                                   float2* __restrict__ out, int n, int pairs,
                                                   int Ns, int R, const float2* __restrict__ tw,
                                                   FastDiv mdiv, FastDiv nsdiv, FastDiv rdiv) {
    const int m = n / R;
    const int tstep = n / (Ns * R);
    const int pstep = n / R;
    const int items = pairs * m * R;
    for (int w = threadIdx.x; w < items; w += blockDim.x) {
        const int idx = static_cast<int>(rdiv.div(static_cast<uint32_t>(w)));
        const int qq = w - idx * R;
        const int q = static_cast<int>(mdiv.div(static_cast<uint32_t>(idx)));
        const int j = idx - q * m;
        const int jq = static_cast<int>(nsdiv.div(static_cast<uint32_t>(j)));
        const int k = j - jq * Ns;
        const float2* src = in + q * n;
        float2 acc = make_float2(0.f, 0.f);
        int e = 0;  // (r * qq) mod R
        for (int r = 0; r < R; ++r) {
            float2 v = src[j + r * m];
            if (Ns > 1 && r > 0) v = cmul(v, tw[r * k * tstep]);
            acc = cadd(acc, cmul(v, tw[e * pstep]));
            e += qq;
            if (e >= R) e -= R;
        }
        out[q * n + jq * Ns * R + k + qq * Ns] = acc;
    }
}

// Forward DFT of `pairs` rows held in buf[0]; returns the buffer that holds the
// result (buf[0] or buf[1]).
static __device__ float2* fft_forward(float2* b0, float2* b1, const FilterParams& P, int pairs) {
    float2* in = b0;
    float2* out = b1;
    int Ns = 1;
    for (int s = 0; s < P.nradix; ++s) {
        const int R = P.radix[s];
        const FastDiv& md = P.m_div[s];
        const FastDiv& nd = P.ns_div[s];
        if (R == 4)
            stockham_pass<4>(in, out, P.n, pairs, Ns, P.tw, md, nd);
        else if (R == 2)
            stockham_pass<2>(in, out, P.n, pairs, Ns, P.tw, md, nd);
        else if (R == 3)
            stockham_pass<3>(in, out, P.n, pairs, Ns, P.tw, md, nd);
        else if (R == 5)
            stockham_pass<5>(in, out, P.n, pairs, Ns, P.tw, md, nd);
        else if (R == 7)
            stockham_pass_odd<7>(in, out, P.n, pairs, Ns, P.tw, md, nd);
        else if (R == 11)
            stockham_pass_odd<11>(in, out, P.n, pairs, Ns, P.tw, md, nd);
        else if (R == 13)
            stockham_pass_odd<13>(in, out, P.n, pairs, Ns, P.tw, md, nd);
        else
            stockham_pass_generic(in, out, P.n, pairs, Ns, R, P.tw, md, nd, P.r_div[s]);
        __syncthreads();
        Ns *= R;
        float2* t = in;
        in = out;
        out = t;
    }
    return in;
}

// The same with the twiddle table given separately (e.g. staged in LDS).
static __device__ __forceinline__ float2* fft_forward_tw(float2* b0, float2* b1,
                                                         const FilterParams& P, int pairs,
                                                         const float2* tw) {
    float2* in = b0;
    float2* out = b1;
    int Ns = 1;
    for (int s = 0; s < P.nradix; ++s) {
        const int R = P.radix[s];
        const FastDiv& md = P.m_div[s];
        const FastDiv& nd = P.ns_div[s];
        if (R == 4)
            stockham_pass<4>(in, out, P.n, pairs, Ns, tw, md, nd);
        else if (R == 2)
            stockham_pass<2>(in, out, P.n, pairs, Ns, tw, md, nd);
        else if (R == 3)
            stockham_pass<3>(in, out, P.n, pairs, Ns, tw, md, nd);
        else if (R == 5)
            stockham_pass<5>(in, out, P.n, pairs, Ns, tw, md, nd);
        else if (R == 7)
            stockham_pass_odd<7>(in, out, P.n, pairs, Ns, tw, md, nd);
        else if (R == 11)
            stockham_pass_odd<11>(in, out, P.n, pairs, Ns, tw, md, nd);
        else if (R == 13)
            stockham_pass_odd<13>(in, out, P.n, pairs, Ns, tw, md, nd);
        else
            stockham_pass_generic(in, out, P.n, pairs, Ns, R, tw, md, nd, P.r_div[s]);
        __syncthreads();
        Ns *= R;
        float2* t = in;
        in = out;
        out = t;
    }
    return in;
}

// ---------------------------------------------------------------------------
// n = 2^L, 64 <= n <= 2048: every thread holds 8 complex values of one row
// (n/8 threads per row).  Radix-8 Stockham passes with the butterfly in
// registers, then one radix-4 or radix-2 pass; sizes are compile-time, so
// indices are shifts and masks.  Each thread's twiddles of every pass are
// loaded once into registers from the per-pass table of pow2_twiddles.

// DFT8 by decimation in frequency: b_m = x_m + x_{m+4}, c_m = (x_m - x_{m+4}) W8^m,
// X[2m] = DFT4(b)[m], X[2m+1] = DFT4(c)[m].
__device__ __forceinline__ void dft8(float2 (&x)[8]) {
    const float h = 0.70710678118654752440f;  // 1/sqrt(2)
    float2 b[4], c[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        b[m] = cadd(x[m], x[m + 4]);
        c[m] = csub(x[m], x[m + 4]);
    }
    // W8 = (1 - i)/sqrt2, W8^2 = -i, W8^3 = -(1 + i)/sqrt2
    c[1] = make_float2((c[1].x + c[1].y) * h, (c[1].y - c[1].x) * h);
    c[2] = cmul_mi(c[2]);
    c[3] = make_float2((c[3].y - c[3].x) * h, -(c[3].x + c[3].y) * h);
    dft_small<4>(b);
    dft_small<4>(c);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        x[2 * m] = b[m];
        x[2 * m + 1] = c[m];
    }
}

template <int R>
__device__ __forceinline__ void dft_r(float2* x) {
    if constexpr (R == 8)
        dft8(*reinterpret_cast<float2(*)[8]>(x));
    else
        dft_small<R>(x);
}

// Fused operation on a pass's outputs before they are stored.
enum { kP2None = 0, kP2Weight = 1, kP2Scale = 2 };

// LDS slot of element i: one pad slot every 8 elements, which spreads the
// strided stores of the first passes over the banks.
__device__ __forceinline__ int p2_slot(int i) { return i + (i >> 3); }
__host__ __device__ constexpr int p2_row_slots(int n) { return n + n / 8; }

// Radix, span and per-thread butterflies of pass S, and the thread's twiddle
// count before pass S (compile time).
template <int N>
struct P2 {
    static constexpr int T = N / 8;
    static constexpr int radix(int s) {
        int ns = 1;
        for (int i = 0; i < s; ++i) ns *= (N / ns >= 8 ? 8 : N / ns);
        return N / ns >= 8 ? 8 : N / ns;
    }
    static constexpr int span(int s) {
        int ns = 1;
        for (int i = 0; i < s; ++i) ns *= radix(i);
        return ns;
    }
    static constexpr int passes() {
        int s = 0;
        while (span(s) < N) ++s;
        return s;
    }
    static constexpr int per(int s) { return (N / radix(s)) / T; }
    static constexpr int tw_regs(int s) { return span(s) > 1 ? per(s) * (radix(s) - 1) : 0; }
    static constexpr int tw_before(int s) {
        int c = 0;
        for (int i = 0; i < s; ++i) c += tw_regs(i);
        return c;
    }
    static constexpr int tw_total() { return tw_before(passes()); }
    static constexpr int tab_off(int s) {
        int c = 0;
        for (int i = 0; i < s; ++i) c += span(i) * radix(i);
        return c;
    }
};

// The thread's twiddles of every pass, loaded once (the same for the forward
// and the inverse transform).
template <int N, int S>
__device__ __forceinline__ void p2_load_tw(float2* twr, int j, const float2* __restrict__ tw8) {
    if constexpr (S < P2<N>::passes()) {
        constexpr int R = P2<N>::radix(S), NS = P2<N>::span(S), PER = P2<N>::per(S);
        if constexpr (NS > 1) {
            float2* w = twr + P2<N>::tw_before(S);
#pragma unroll
            for (int u = 0; u < PER; ++u) {
                const int k = (j + u * P2<N>::T) & (NS - 1);
#pragma unroll
                for (int r = 1; r < R; ++r)
                    w[u * (R - 1) + r - 1] = tw8[P2<N>::tab_off(S) + k * R + r];
            }
        }
        p2_load_tw<N, S + 1>(twr, j, tw8);
    }
}

// One in-place Stockham pass of radix R and span NS over a pair's padded row:
// thread j does the butterflies b = j + u*T (T = N/8, u < N/(R*T)); inputs
// b + r*N/R, outputs (b - k)*R + k + r*NS with k = b mod NS.  All reads finish
// (barrier) before any write.
// TWG: twr is the per-pass twiddle table itself (tw8), and the pass reads the
// thread's twiddles from it at its start instead of holding every pass's
// twiddles in registers across the transforms.
template <int N, int S, int FUSE, bool TWG = false>
__device__ __forceinline__ void p2_pass(float2* row, int j, const float2* twr,
                                        const float* __restrict__ weights, float inv_n) {
    constexpr int R = P2<N>::radix(S), NS = P2<N>::span(S), PER = P2<N>::per(S);
    constexpr int T = P2<N>::T, BF = N / R;
    float2 wl[PER][R - 1 > 0 ? R - 1 : 1];
    if constexpr (TWG && NS > 1) {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int k = (j + u * T) & (NS - 1);
#pragma unroll
            for (int r = 1; r < R; ++r) wl[u][r - 1] = twr[P2<N>::tab_off(S) + k * R + r];
        }
    }
    float2 x[PER][R];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int b = j + u * T;
#pragma unroll
        for (int r = 0; r < R; ++r) x[u][r] = row[p2_slot(b + r * BF)];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int b = j + u * T;
        const int k = b & (NS - 1);
        if constexpr (NS > 1) {
            const float2* w = TWG ? wl[u] : twr + P2<N>::tw_before(S) + u * (R - 1);
#pragma unroll
            for (int r = 1; r < R; ++r) x[u][r] = cmul(x[u][r], w[r - 1]);
        }
        dft_r<R>(x[u]);
        const int base = (b - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            float2 v = x[u][r];
            if constexpr (FUSE == kP2Weight) {
                // Z' = conj(W * X), the input of the inverse transform
                const float wt = weights[base + r * NS];
                v = make_float2(v.x * wt, -(v.y * wt));
            } else if constexpr (FUSE == kP2Scale) {
                // a = Re(Y) / n, b = -Im(Y) / n  (main.rs:376-379: re * (1/n))
                v = make_float2(v.x * inv_n, -v.y * inv_n);
            }
            row[p2_slot(base + r * NS)] = v;
        }
    }
    __syncthreads();
}

template <int N, int S, int FUSE, bool TWG = false>
__device__ __forceinline__ void p2_fft(float2* row, int j, const float2* twr,
                                       const float* __restrict__ weights, float inv_n) {
    if constexpr (S < P2<N>::passes()) {
        constexpr bool LAST = S + 1 == P2<N>::passes();
        p2_pass<N, S, LAST ? FUSE : kP2None, TWG>(row, j, twr, weights, inv_n);
        p2_fft<N, S + 1, FUSE, TWG>(row, j, twr, weights, inv_n);
    }
}

}  // namespace tiltgpu
