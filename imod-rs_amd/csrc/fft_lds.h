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

}  // namespace tiltgpu
