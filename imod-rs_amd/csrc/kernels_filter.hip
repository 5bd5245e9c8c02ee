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

#include <mutex>
#include <vector>

#include <cmath>

#include "fft_lds.h"
#include "tiltgpu_internal.h"

namespace tiltgpu {

constexpr int kFilterThreads = 256;

__shared__ double s_wave_sum[kFilterThreads / 64];

// Sum over the 256 threads of a workgroup in a fixed order (waves by
// butterfly, then the four wave sums); all threads call it.
__device__ __forceinline__ double wg_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) s_wave_sum[threadIdx.x >> 6] = v;
    __syncthreads();
    const double t = (s_wave_sum[0] + s_wave_sum[1]) + (s_wave_sum[2] + s_wave_sum[3]);
    __syncthreads();
    return t;
}

__shared__ double2 s_wave_sum2[kFilterThreads / 64];

// Two such sums at once (one pair of barriers).
__device__ __forceinline__ double2 wg_sum2(double a, double b) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
    }
    if ((threadIdx.x & 63) == 0) s_wave_sum2[threadIdx.x >> 6] = make_double2(a, b);
    __syncthreads();
    const double2 w0 = s_wave_sum2[0], w1 = s_wave_sum2[1], w2 = s_wave_sum2[2], w3 = s_wave_sum2[3];
    __syncthreads();
    return make_double2((w0.x + w1.x) + (w2.x + w3.x), (w0.y + w1.y) + (w2.y + w3.y));
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

// ---------------------------------------------------------------- power of two
//
// n = 2^L, 64 <= n <= 2048: the register radix-8 passes of fft_lds.h over row
// pairs (n/8 threads per pair, 2048/n pairs per workgroup).  The weight
// multiply is fused into the last forward pass and the 1/n scale into the last
// inverse pass.

__device__ __forceinline__ float p2_at(const float2* row, int i, int c) {
    const float2 v = row[p2_slot(i)];
    return c ? v.y : v.x;
}

// stretch_sample over a padded LDS row (same arithmetic, main.rs:408-488).
__device__ float p2_stretch_sample(const float2* row, int c, int n, float s, int cubic, int ox) {
    const float center = static_cast<float>(n) / 2.0f;
    const float dxs = static_cast<float>(ox) - center;
    const float src_x = dxs * s + center;
    const float fl = floorf(src_x);
    const float fn = static_cast<float>(n);
    if (cubic) {
        const float frac = src_x - fl;
        if (fl >= 1.f && fl + 2.f < fn) {
            const int i = static_cast<int>(fl);
            const float p0 = p2_at(row, i - 1, c), p1 = p2_at(row, i, c);
            const float p2 = p2_at(row, i + 1, c), p3 = p2_at(row, i + 2, c);
            const float t = frac, t2 = t * t, t3 = t2 * t;
            const float a = -0.5f * p0 + 1.5f * p1 - 1.5f * p2 + 0.5f * p3;
            const float b = p0 - 2.5f * p1 + 2.0f * p2 - 0.5f * p3;
            const float cc = -0.5f * p0 + 0.5f * p2;
            return a * t3 + b * t2 + cc * t + p1;
        } else if (fl >= 0.f && fl + 1.f < fn) {
            const int i = static_cast<int>(fl);
            return p2_at(row, i, c) * (1.0f - frac) + p2_at(row, i + 1, c) * frac;
        } else if (fl >= 0.f && fl < fn) {
            return p2_at(row, static_cast<int>(fl), c);
        }
        return 0.f;
    }
    if (fl >= 0.f && fl + 1.f < fn) {
        const int i = static_cast<int>(fl);
        const float frac = src_x - fl;
        return p2_at(row, i, c) * (1.0f - frac) + p2_at(row, i + 1, c) * frac;
    } else if (fl >= 0.f && fl < fn) {
        return p2_at(row, static_cast<int>(fl), c);
    }
    return 0.f;
}

template <int N>
__device__ void filter_body_p2(const FilterParams& P, int blk) {
    extern __shared__ float2 sm[];
    constexpr int T = P2<N>::T;
    constexpr int PPW = 2048 / N;
    constexpr int RS = p2_row_slots(N);
    constexpr int LOGN = (N == 64) ? 6 : (N == 128) ? 7 : (N == 256) ? 8 : (N == 512) ? 9
                       : (N == 1024) ? 10 : 11;
    const int tid = threadIdx.x;
    const int q = tid / T;  // pair of this thread within the workgroup
    const int j = tid & (T - 1);
    const int pair0 = blk * PPW;
    const int total_pairs = (P.nrows + 1) / 2;
    const int pairs = min(PPW, total_pairs - pair0);

    // ---- load two rows per pair (row a -> re, row b -> im), zero past the end.
    // PPW * N = 2048: eight steps of 256 elements.  For N >= 256 every step
    // lies in one pair (IT = N / 256 consecutive steps per pair), so the f64
    // row sums (the per-view means' input, main.rs:615-623) are reduced over
    // the workgroup once per pair.
    constexpr int IT = N >= 256 ? N / 256 : 0;
    double sq = 0.0, sa = 0.0, sb = 0.0;
#pragma unroll
    for (int k = 0; k < PPW * N / kFilterThreads; ++k) {
        const int idx = tid + k * kFilterThreads;
        const int qq = idx >> LOGN;
        const int i = idx & (N - 1);
        float va = 0.f, vb = 0.f;
        if (qq < pairs) {
            const int ra = 2 * (pair0 + qq), rb = ra + 1;
            va = P.src[static_cast<long>(ra) * N + i];
            if (P.src_sub) va = va - P.src_sub[static_cast<long>(ra) * N + i];
            if (rb < P.nrows) {
                vb = P.src[static_cast<long>(rb) * N + i];
                if (P.src_sub) vb = vb - P.src_sub[static_cast<long>(rb) * N + i];
            }
        }
        if (P.sq_partial) sq += static_cast<double>(va) * va + static_cast<double>(vb) * vb;
        sm[qq * RS + p2_slot(i)] = make_float2(va, vb);
        if constexpr (IT > 0) {
            if (P.row_sum) {  // uniform
                sa += static_cast<double>(va);
                sb += static_cast<double>(vb);
                if ((k + 1) % IT == 0) {
                    const double2 t = wg_sum2(sa, sb);
                    const int pq = k / IT;
                    const int ra = 2 * (pair0 + pq);
                    if (tid == 0 && pq < pairs) {
                        P.row_sum[ra] = t.x;
                        if (ra + 1 < P.nrows) P.row_sum[ra + 1] = t.y;
                    }
                    sa = 0.0;
                    sb = 0.0;
                }
            }
        }
    }
    if (P.sq_partial) {
        const double t = wg_sum(sq);
        if (tid == 0) P.sq_partial[blk] = t;
    }
    __syncthreads();

    if (P.weights) {
        float2* row = sm + q * RS;
        // forward transform with the weight multiply fused into its last pass,
        // then the inverse transform of the conjugate with the 1/n scale fused
        p2_fft<N, 0, kP2Weight, true>(row, j, P.tw8, P.weights, 0.f);
        p2_fft<N, 0, kP2Scale, true>(row, j, P.tw8, nullptr, 1.0f / static_cast<float>(N));
    }
    // ---- cosine stretch (per row: its own view) and coalesced write.
    // For N >= 256 with the linear stretch, a step of 256 outputs lies in one
    // pair; when both rows of the pair belong to one view (always, for an even
    // row count per view) they share the source coordinate and the two
    // float2 reads, each component interpolated with the arithmetic of
    // p2_stretch_sample.
    if constexpr (IT > 0) {
        if (!P.cubic) {
            const float center = static_cast<float>(N) / 2.0f;
            const float fn = static_cast<float>(N);
            // not unrolled: the registers of eight steps in flight would cost
            // occupancy (89 instead of 63 VGPRs at n = 2048)
#pragma unroll 1
            for (int k = 0; k < PPW * N / kFilterThreads; ++k) {
                const int pq = k / IT;
                if (pq >= pairs) break;  // uniform
                const int ra = 2 * (pair0 + pq);
                const int rb = ra + 1;
                const int va_view = __builtin_amdgcn_readfirstlane(ra / P.rows_per_view);
                const bool hasb = rb < P.nrows;
                const bool same = !hasb || rb / P.rows_per_view == va_view;
                if (!same) {
                    // rows of two views: per row, as the general loop below
                    const int ox = (tid + k * kFilterThreads) & (N - 1);
                    const float2* row = sm + pq * RS;
#pragma unroll
                    for (int c = 0; c < 2; ++c) {
                        const int r = ra + c;
                        const float sv = P.stretch ? P.stretch[r / P.rows_per_view] : 0.f;
                        P.dst[static_cast<long>(r) * N + ox] =
                            sv != 0.f ? p2_stretch_sample(row, c, N, sv, 0, ox) : p2_at(row, ox, c);
                    }
                    continue;
                }
                const float sv = P.stretch ? P.stretch[va_view] : 0.f;
                const int ox = (tid + k * kFilterThreads) & (N - 1);
                const float2* row = sm + pq * RS;
                float oa, ob;
                if (sv != 0.f) {
                    const float dxs = static_cast<float>(ox) - center;
                    const float src_x = dxs * sv + center;
                    const float fl = floorf(src_x);
                    if (fl >= 0.f && fl + 1.f < fn) {
                        const int i = static_cast<int>(fl);
                        const float frac = src_x - fl;
                        const float2 v0 = row[p2_slot(i)], v1 = row[p2_slot(i + 1)];
                        oa = v0.x * (1.0f - frac) + v1.x * frac;
                        ob = v0.y * (1.0f - frac) + v1.y * frac;
                    } else if (fl >= 0.f && fl < fn) {
                        const float2 v = row[p2_slot(static_cast<int>(fl))];
                        oa = v.x;
                        ob = v.y;
                    } else {
                        oa = 0.f;
                        ob = 0.f;
                    }
                } else {
                    const float2 v = row[p2_slot(ox)];
                    oa = v.x;
                    ob = v.y;
                }
                P.dst[static_cast<long>(ra) * N + ox] = oa;
                if (hasb) P.dst[static_cast<long>(rb) * N + ox] = ob;
            }
            return;
        }
    }
    for (int idx = tid; idx < pairs * N; idx += kFilterThreads) {
        const int qq = idx >> LOGN;
        const int ox = idx & (N - 1);
        const float2* row = sm + qq * RS;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int r = 2 * (pair0 + qq) + c;
            if (r >= P.nrows) break;
            const float s = P.stretch ? P.stretch[r / P.rows_per_view] : 0.f;
            float v;
            if (s != 0.f)
                v = p2_stretch_sample(row, c, N, s, P.cubic, ox);
            else
                v = p2_at(row, ox, c);
            P.dst[static_cast<long>(r) * N + ox] = v;
        }
    }
}

// Workgroup `blk` of one filter launch (mixed-radix passes).
__device__ __forceinline__ void filter_body_mixed(const FilterParams& P, int blk) {
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
        const double t = wg_sum(sq);
        if (tid == 0) P.sq_partial[blk] = t;
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

// Workgroup `blk` of one filter launch: row pairs [blk*ppw, (blk+1)*ppw).
__device__ __forceinline__ void filter_body(const FilterParams& P, int blk) {
    if (P.tw8) {
        switch (P.log2n) {
            case 6: filter_body_p2<64>(P, blk); return;
            case 7: filter_body_p2<128>(P, blk); return;
            case 8: filter_body_p2<256>(P, blk); return;
            case 9: filter_body_p2<512>(P, blk); return;
            case 10: filter_body_p2<1024>(P, blk); return;
            case 11: filter_body_p2<2048>(P, blk); return;
            default: break;
        }
    }
    filter_body_mixed(P, blk);
}

__global__ void __launch_bounds__(256) filter_kernel(FilterParams P) { filter_body(P, blockIdx.x); }

// The power-of-two path as kernels of their own: their register allocation
// (63 VGPRs at n = 2048, 8 waves per SIMD) is not raised to the mixed-radix
// path's.
template <int N>
__global__ void __launch_bounds__(256) filter_p2_kernel(FilterParams P) {
    filter_body_p2<N>(P, blockIdx.x);
}

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

// Rows wider than 4096 on the mixed-radix path need more than the default 64 KiB
// of dynamic LDS per workgroup (two buffers of n complex values); a kernel opts
// in to the size a launch needs the first time it is needed.
static hipError_t allow_large_lds(const void* kernel, size_t lds) {
    return lds_opt_in(kernel, lds);
}

hipError_t lds_opt_in(const void* kernel, size_t lds) {
    if (lds <= 64 * 1024) return hipSuccess;
    if (lds > kFilterMaxLds) return hipErrorInvalidValue;
    static std::mutex mu;
    static std::vector<std::pair<const void*, size_t>> granted;
    std::lock_guard<std::mutex> lk(mu);
    for (auto& g : granted)
        if (g.first == kernel && g.second >= lds) return hipSuccess;
    const hipError_t e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                             static_cast<int>(lds));
    if (e != hipSuccess) return e;
    for (auto& g : granted)
        if (g.first == kernel) {
            g.second = lds;
            return hipSuccess;
        }
    granted.emplace_back(kernel, lds);
    return hipSuccess;
}

hipError_t launch_filter_batch(const FilterBatch& b, int total_blocks, size_t lds, hipStream_t st) {
    if (total_blocks <= 0) return hipSuccess;
    const hipError_t e = allow_large_lds(reinterpret_cast<const void*>(filter_batch_kernel), lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(filter_batch_kernel, dim3(total_blocks), dim3(256), lds, st, b);
    return hipGetLastError();
}

int filter_blocks(const FilterParams& p) {
    const int total_pairs = (p.nrows + 1) / 2;
    return (total_pairs + p.pairs_per_wg - 1) / p.pairs_per_wg;
}

size_t filter_lds_bytes(int n, int pairs_per_wg) {
    // one padded in-place buffer on the power-of-two path, two buffers otherwise
    if (pow2_twiddles(n, nullptr) > 0)
        return static_cast<size_t>(pairs_per_wg) * p2_row_slots(n) * sizeof(float2);
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

int pow2_twiddles(int n, float2* out, int max_n, int min_n) {
    if (n < min_n || n > max_n || (n & (n - 1)) != 0) return 0;
    int off = 0;
    for (int ns = 1; ns < n;) {
        const int R = n / ns >= 8 ? 8 : n / ns;
        for (int k = 0; k < ns; ++k)
            for (int r = 0; r < R; ++r) {
                const double a = -2.0 * M_PI * static_cast<double>(r) * k / (static_cast<double>(ns) * R);
                if (out) out[off + k * R + r] = make_float2(static_cast<float>(std::cos(a)),
                                                            static_cast<float>(std::sin(a)));
            }
        off += ns * R;
        ns *= R;
    }
    return off;
}

hipError_t launch_filter(const FilterParams& p, hipStream_t st) {
    const int total_pairs = (p.nrows + 1) / 2;
    const int nblk = (total_pairs + p.pairs_per_wg - 1) / p.pairs_per_wg;
    if (nblk <= 0) return hipSuccess;
    const size_t lds = filter_lds_bytes(p.n, p.pairs_per_wg);
    if (p.tw8) {
        void (*k)(FilterParams) = nullptr;
        switch (p.log2n) {
            case 6: k = filter_p2_kernel<64>; break;
            case 7: k = filter_p2_kernel<128>; break;
            case 8: k = filter_p2_kernel<256>; break;
            case 9: k = filter_p2_kernel<512>; break;
            case 10: k = filter_p2_kernel<1024>; break;
            case 11: k = filter_p2_kernel<2048>; break;
            default: break;
        }
        if (k) {
            hipLaunchKernelGGL(k, dim3(nblk), dim3(256), lds, st, p);
            return hipGetLastError();
        }
    }
    const hipError_t e = allow_large_lds(reinterpret_cast<const void*>(filter_kernel), lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(filter_kernel, dim3(nblk), dim3(256), lds, st, p);
    return hipGetLastError();
}

}  // namespace tiltgpu
