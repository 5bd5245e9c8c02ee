// R-weighting filter + cosine stretch for gfx950: apply_rweighting
// (imod-rs apps/tilt/src/main.rs:333-381) fused with apply_cosine_stretch
// (main.rs:446-488), one HBM read and one HBM write per projection row.
//
// Two real rows a, b are packed as one complex row z = a + i*b.  The radial
// weight W(k) is real and even (W(k) = W(n-k)), so W*DFT(z) is the DFT of
// (filtered a) + i*(filtered b) and one complex transform pair filters both
// rows exactly.  The inverse transform uses IDFT(Z) = conj(DFT(conj(Z))), so
// only the forward self-sorting Stockham kernel exists (radix 4, 2, 3, 5 and a
// generic radix for any other prime), run entirely in LDS with twiddles from a
// table computed in double on the host.
//
// The stretch then reads the filtered rows from LDS at
// src_x = (ox - c) * (1/cos theta) + c with the reference's rounding
// (-ffp-contract=off) and writes the row coalesced.
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
__device__ __noinline__ void stockham_pass_generic(const float2* __restrict__ in,
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
__device__ float2* fft_forward(float2* b0, float2* b1, const FilterParams& P, int pairs) {
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

// sample_row_cubic / cubic_interp_1d (main.rs:408-441) over an LDS row held in
// component `c` of float2 entries.
__device__ __forceinline__ float lds_at(const float2* row, int i, int c) {
    return c ? row[i].y : row[i].x;
}

__device__ float stretch_sample(const float2* row, int c, int n, float s, int cubic,
                                int ox) {
    const float center = static_cast<float>(n) / 2.0f;
    const float dxs = static_cast<float>(ox) - center;
    const float src_x = dxs * s + center;
    const float fl = floorf(src_x);
    const float fn = static_cast<float>(n);
    if (cubic) {
        const float frac = src_x - fl;
        if (fl >= 1.f && fl + 2.f < fn) {
            const int i = static_cast<int>(fl);
            const float p0 = lds_at(row, i - 1, c), p1 = lds_at(row, i, c);
            const float p2 = lds_at(row, i + 1, c), p3 = lds_at(row, i + 2, c);
            const float t = frac, t2 = t * t, t3 = t2 * t;
            const float a = -0.5f * p0 + 1.5f * p1 - 1.5f * p2 + 0.5f * p3;
            const float b = p0 - 2.5f * p1 + 2.0f * p2 - 0.5f * p3;
            const float cc = -0.5f * p0 + 0.5f * p2;
            return a * t3 + b * t2 + cc * t + p1;
        } else if (fl >= 0.f && fl + 1.f < fn) {
            const int i = static_cast<int>(fl);
            return lds_at(row, i, c) * (1.0f - frac) + lds_at(row, i + 1, c) * frac;
        } else if (fl >= 0.f && fl < fn) {
            return lds_at(row, static_cast<int>(fl), c);
        }
        return 0.f;
    }
    if (fl >= 0.f && fl + 1.f < fn) {
        const int i = static_cast<int>(fl);
        const float frac = src_x - fl;
        return lds_at(row, i, c) * (1.0f - frac) + lds_at(row, i + 1, c) * frac;
    } else if (fl >= 0.f && fl < fn) {
        return lds_at(row, static_cast<int>(fl), c);
    }
    return 0.f;
}

// Workgroup `blk` of one filter launch: row pairs [blk*ppw, (blk+1)*ppw).
__device__ __forceinline__ void filter_body(const FilterParams& P, int blk) {
    extern __shared__ float2 sm[];
    const int n = P.n;
    const int ppw = P.pairs_per_wg;
    float2* b0 = sm;
    float2* b1 = sm + ppw * n;
    const int pair0 = blk * ppw;
    const int total_pairs = (P.nrows + 1) / 2;
    const int pairs = min(ppw, total_pairs - pair0);
    const int tid = threadIdx.x;

    // ---- load two rows per pair (row a -> re, row b -> im), zero past the end
    double sq = 0.0;
    for (int idx = tid; idx < pairs * n; idx += blockDim.x) {
        const int q = static_cast<int>(P.n_div.div(static_cast<uint32_t>(idx)));
        const int i = idx - q * n;
        const int ra = 2 * (pair0 + q), rb = ra + 1;
        float va = P.src[static_cast<long>(ra) * n + i];
        if (P.src_sub) va = va - P.src_sub[static_cast<long>(ra) * n + i];
        float vb = 0.f;
        if (rb < P.nrows) {
            vb = P.src[static_cast<long>(rb) * n + i];
            if (P.src_sub) vb = vb - P.src_sub[static_cast<long>(rb) * n + i];
        }
        sq += static_cast<double>(va) * va + static_cast<double>(vb) * vb;
        b0[idx] = make_float2(va, vb);
    }
    if (P.sq_partial) {
        __shared__ double red[256];
        red[tid] = sq;
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
            if (tid < s) red[tid] += red[tid + s];
            __syncthreads();
        }
        if (tid == 0) P.sq_partial[blk] = red[0];
    }
    __syncthreads();

    const float2* filt = b0;  // filtered rows (a in .x, b in .y), before stretch
    float2* other = b1;
    if (P.weights) {
        float2* X = fft_forward(b0, b1, P, pairs);
        // Z' = conj(W * X)
        for (int idx = tid; idx < pairs * n; idx += blockDim.x) {
            const int i = idx - n * static_cast<int>(P.n_div.div(static_cast<uint32_t>(idx)));
            const float w = P.weights[i];
            const float2 v = X[idx];
            X[idx] = make_float2(v.x * w, -(v.y * w));
        }
        __syncthreads();
        float2* other_buf = (X == b0) ? b1 : b0;
        float2* Y = fft_forward(X, other_buf, P, pairs);
        // a = Re(Y) / n, b = -Im(Y) / n  (main.rs:376-379: re * (1/n))
        const float inv_n = 1.0f / static_cast<float>(n);
        for (int idx = tid; idx < pairs * n; idx += blockDim.x) {
            const float2 v = Y[idx];
            Y[idx] = make_float2(v.x * inv_n, -v.y * inv_n);
        }
        __syncthreads();
        filt = Y;
        other = (Y == b0) ? b1 : b0;
    }
    (void)other;

    // ---- cosine stretch (per row: its own view) and coalesced write
    for (int idx = tid; idx < pairs * n; idx += blockDim.x) {
        const int q = static_cast<int>(P.n_div.div(static_cast<uint32_t>(idx)));
        const int ox = idx - q * n;
        const float2* row = filt + q * n;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int r = 2 * (pair0 + q) + c;
            if (r >= P.nrows) break;
            const float s = P.stretch ? P.stretch[r / P.rows_per_view] : 0.f;
            float v;
            if (s != 0.f)
                v = stretch_sample(row, c, n, s, P.cubic, ox);
            else
                v = c ? row[ox].y : row[ox].x;
            P.dst[static_cast<long>(r) * n + ox] = v;
        }
    }
}

__global__ void __launch_bounds__(256) filter_kernel(FilterParams P) { filter_body(P, blockIdx.x); }

// Several series in one launch: workgroup b belongs to the series r with
// block_start[r] <= b < block_start[r+1]; each series keeps its own length,
// plan and options, so its rows are filtered exactly as in a launch of its own.
__global__ void __launch_bounds__(256) filter_batch_kernel(FilterBatch B) {
    int a = 0, c = B.nreq - 1;
    const int b = blockIdx.x;
    while (a < c) {
        const int m = (a + c + 1) / 2;
        if (B.block_start[m] <= b)
            a = m;
        else
            c = m - 1;
    }
    filter_body(B.descs[a], b - B.block_start[a]);
}

hipError_t launch_filter_batch(const FilterBatch& b, int total_blocks, size_t lds, hipStream_t st) {
    if (total_blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(filter_batch_kernel, dim3(total_blocks), dim3(256), lds, st, b);
    return hipGetLastError();
}

int filter_blocks(const FilterParams& p) {
    const int total_pairs = (p.nrows + 1) / 2;
    return (total_pairs + p.pairs_per_wg - 1) / p.pairs_per_wg;
}

size_t filter_lds_bytes(int n, int pairs_per_wg) {
    return static_cast<size_t>(2) * pairs_per_wg * n * sizeof(float2);
}

int filter_plan(int n, int* radix, int* nradix) {
    int m = n, k = 0;
    if (n < 2) return -1;
    while (m % 4 == 0 && k < kMaxRadix) { radix[k++] = 4; m /= 4; }
    while (m % 2 == 0 && k < kMaxRadix) { radix[k++] = 2; m /= 2; }
    for (int p = 3; p * p <= m && k < kMaxRadix; p += 2) {
        while (m % p == 0 && k < kMaxRadix) { radix[k++] = p; m /= p; }
    }
    if (m > 1 && k < kMaxRadix) { radix[k++] = m; m = 1; }  // the remaining prime factor
    if (m != 1) return -1;
    *nradix = k;
    return 0;
}

hipError_t launch_filter(const FilterParams& p, hipStream_t st) {
    const int total_pairs = (p.nrows + 1) / 2;
    const int nblk = (total_pairs + p.pairs_per_wg - 1) / p.pairs_per_wg;
    if (nblk <= 0) return hipSuccess;
    const size_t lds = filter_lds_bytes(p.n, p.pairs_per_wg);
    hipLaunchKernelGGL(filter_kernel, dim3(nblk), dim3(256), lds, st, p);
    return hipGetLastError();
}

}  // namespace tiltgpu
