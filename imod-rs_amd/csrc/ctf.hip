// CTF phase flip of a tilt series for gfx950: the 2D-strip correction of
// imod-rs apps/ctfphaseflip/src/main.rs:353-705 (its CPU path) over a whole
// stack held in HBM.
//
// The strip geometry (main.rs:389-487) depends only on the image size, the
// tilt-axis angle and the strip width, so every section shares the strips'
// bounding boxes.  The stack is processed strip by strip, all sections of a
// chunk in one launch per stage:
//   1. rows:      bounding-box rows of the image as complex values, forward
//                 DFT along x                                     -> A [z][sh][sw]
//   2. transpose  A -> B [z][sw][sh]
//   3. columns:   forward DFT along y, the sign flip where the CTF is negative
//                 (decided in double precision as main.rs:563-630 does),
//                 inverse DFT along y, all in LDS                 -> B
//   4. transpose  B -> A
//   5. rows:      inverse DFT along x; the real part / (sw*sh) is added into
//                 the weighted sum and the weight total with the triangular
//                 strip weight (main.rs:660-679)
// and after the last strip out = wsum / wtot where wtot > 1e-10, else the
// input value (main.rs:682-688).  The membership test and the weights are
// evaluated with the reference's double-precision expressions (the file is
// compiled with -ffp-contract=off), so exactly the reference's pixels take part;
// the transforms are single precision with double-precision twiddles.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

#include "../../include/tiltgpu.h"
#include "fft_lds.h"
#include "session_internal.h"

namespace tiltgpu {
namespace {

constexpr int kCtfThreads = 256;
constexpr size_t kCtfMaxLds = 160 * 1024;    // LDS per workgroup on gfx950
constexpr long kCtfChunkElems = 64L << 20;   // complex elements per work buffer
constexpr double kPi = 3.14159265358979323846;

struct CtfAstig {
    double c2, s2;  // cos 2a, sin 2a of the astigmatism angle a
    double on;      // 1 where |defocus1 - defocus2| > 0.01 A (main.rs:385)
    double pad;
};

struct CtfStrip {
    const float* in;       // [zc][ny][nx] sections of the chunk
    float* wsum;           // [zc][ny][nx] weighted sums (the output stack)
    float* wtot;           // [zc][ny][nx] weight totals
    float2* a;             // [zc][sh][sw]
    float2* b;             // [zc][sw][sh]
    const double2* def;    // [zc] strip defocus 1 and 2 (Angstrom)
    const CtfAstig* astig; // [zc] cos 2a, sin 2a of the astigmatism angle a; on
    long plane;            // nx * ny
    int nx, zc, bx0, by0, sw, sh;
    FastDiv sw_div, sh_div;
    double cx, cy, px, py;     // image centre, direction across the tilt axis
    double scp, lo, hi, half;  // strip centre, membership [lo, hi), width / 2
    double den_x, den_y;       // sw * pixel, sh * pixel (Angstrom)
    double lambda, cs_a, alpha, cuton, plate, ca, sa;  // alpha = asin(amplitude contrast)
    int rotate;                // |tilt axis| > 1e-6 rad
    float scale;               // 1 / (sw * sh)
};

// main.rs:564-630 for the coefficient (fx, fy) of section z: true where the
// CTF (with amplitude contrast and phase plate) is negative and the spatial
// frequency is at or above the cut-on.  Evaluated in double precision with two
// identities instead of trigonometric calls:
//   -sin(c) sqrt(1 - w^2) + cos(c) w = -sin(c - asin w), negative exactly when
//     (c - asin w) / pi has an even floor and is not an integer;
//   def1 cos^2(t - a) + def2 sin^2(t - a) = (def1 + def2)/2 + (def1 - def2)/2 cos(2t - 2a),
//     with cos 2t = (sx^2 - sy^2)/s^2 and sin 2t = 2 sx sy / s^2 for t = atan2(sy, sx).
// The decision can differ from the reference's only within ~1e-14 (relative)
// of a CTF zero, the same as between two correctly rounded sin implementations.
__device__ bool ctf_negative(const CtfStrip& s, int fx, int fy, int z) {
    const double freq_y = fy <= s.sh / 2 ? static_cast<double>(fy)
                                         : static_cast<double>(fy) - static_cast<double>(s.sh);
    const double freq_x = fx <= s.sw / 2 ? static_cast<double>(fx)
                                         : static_cast<double>(fx) - static_cast<double>(s.sw);
    double rx = freq_x, ry = freq_y;
    if (s.rotate) {
        rx = freq_x * s.ca + freq_y * s.sa;
        ry = -freq_x * s.sa + freq_y * s.ca;
    }
    const double sx = rx / s.den_x;
    const double sy = ry / s.den_y;
    const double s2 = sx * sx + sy * sy;
    if (s.cuton > 0.0 && sqrt(s2) < s.cuton) return false;
    const double2 d = s.def[z];
    double def = d.x;
    const CtfAstig ag = s.astig[z];
    if (ag.on != 0.0 && s2 > 0.0) {
        const double c2t = (sx * sx - sy * sy) / s2, s2t = 2.0 * sx * sy / s2;
        def = 0.5 * (d.x + d.y) + 0.5 * (d.x - d.y) * (c2t * ag.c2 + s2t * ag.s2);
    }
    const double chi = kPi * s.lambda * s2 * (def - 0.5 * s.cs_a * s.lambda * s.lambda * s2);
    const double t = (chi + s.plate - s.alpha) * (1.0 / kPi);
    const double fl = floor(t);
    return t > fl && (static_cast<long long>(fl) & 1) == 0;
}

// Stage 1: rows [z][y] of the bounding box, forward DFT along x.
__global__ void __launch_bounds__(kCtfThreads) ctf_rows_fwd_kernel(CtfStrip s, FilterParams P) {
    extern __shared__ float2 sm[];
    const int n = P.n, ppw = P.pairs_per_wg;
    const int rows = s.zc * s.sh;
    const int r0 = blockIdx.x * ppw;
    const int cnt = min(ppw, rows - r0);
    float2* b0 = sm;
    float2* b1 = sm + ppw * n;
    for (int idx = threadIdx.x; idx < cnt * n; idx += kCtfThreads) {
        const int q = static_cast<int>(P.n_div.div(static_cast<uint32_t>(idx)));
        const int x = idx - q * n;
        const int r = r0 + q;
        const int z = static_cast<int>(s.sh_div.div(static_cast<uint32_t>(r)));
        const int y = r - z * s.sh;
        const float v = s.in[z * s.plane + static_cast<long>(s.by0 + y) * s.nx + s.bx0 + x];
        b0[idx] = make_float2(v, 0.f);
    }
    __syncthreads();
    const float2* X = fft_forward(b0, b1, P, cnt);
    float2* dst = s.a + static_cast<long>(r0) * n;
    for (int idx = threadIdx.x; idx < cnt * n; idx += kCtfThreads) dst[idx] = X[idx];
}

// Stage 3: rows [z][fx] of B, forward DFT along y, the flip, and the
// unnormalised inverse as conj(DFT(conj(.))).
__global__ void __launch_bounds__(kCtfThreads) ctf_cols_kernel(CtfStrip s, FilterParams P) {
    extern __shared__ float2 sm[];
    const int n = P.n, ppw = P.pairs_per_wg;
    const int rows = s.zc * s.sw;
    const int r0 = blockIdx.x * ppw;
    const int cnt = min(ppw, rows - r0);
    float2* b0 = sm;
    float2* b1 = sm + ppw * n;
    float2* row = s.b + static_cast<long>(r0) * n;
    for (int idx = threadIdx.x; idx < cnt * n; idx += kCtfThreads) b0[idx] = row[idx];
    __syncthreads();
    float2* X = fft_forward(b0, b1, P, cnt);
    for (int idx = threadIdx.x; idx < cnt * n; idx += kCtfThreads) {
        const int q = static_cast<int>(P.n_div.div(static_cast<uint32_t>(idx)));
        const int fy = idx - q * n;
        const int r = r0 + q;
        const int z = static_cast<int>(s.sw_div.div(static_cast<uint32_t>(r)));
        const int fx = r - z * s.sw;
        float2 v = X[idx];
        if (ctf_negative(s, fx, fy, z)) {
            v.x = -v.x;
            v.y = -v.y;
        }
        X[idx] = make_float2(v.x, -v.y);
    }
    __syncthreads();
    const float2* Y = fft_forward(X, X == b0 ? b1 : b0, P, cnt);
    for (int idx = threadIdx.x; idx < cnt * n; idx += kCtfThreads) {
        const float2 v = Y[idx];
        row[idx] = make_float2(v.x, -v.y);
    }
}

// Stages 2-4 in one, for mixed-radix columns: C adjacent columns of one
// section read straight from A [z][sh][sw] (C * 8-byte row segments), the
// forward DFT along y, the flip, the inverse, and the columns written back in
// place, so A needs no transposes.
template <int C>
__global__ void __launch_bounds__(kCtfThreads) ctf_cols_tile_kernel(CtfStrip s, FilterParams P) {
    extern __shared__ float2 sm[];
    __shared__ float2 tw_lds[512];  // the twiddle table (n <= 512 on this path)
    const int n = P.n;  // column length sh
    for (int i = threadIdx.x; i < n; i += kCtfThreads) tw_lds[i] = P.tw[i];
    const int ntile = (s.sw + C - 1) / C;
    // Neighbouring tiles share 128-byte L2 lines (a tile row is C * 8 bytes):
    // give each XCD a contiguous run of tiles so both halves of a line are
    // fetched once, into the one L2 that reads them both.
    const int t = xcd_remap(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x));
    const int z = t / ntile;
    const int x0 = (t - z * ntile) * C;
    float2* b0 = sm;
    float2* b1 = sm + C * n;
    float2* A = s.a + static_cast<long>(z) * s.sh * s.sw + x0;
    for (int idx = threadIdx.x; idx < C * n; idx += kCtfThreads) {
        const int y = idx / C, c = idx - y * C;
        b0[c * n + y] = x0 + c < s.sw ? A[static_cast<long>(y) * s.sw + c] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    float2* X = fft_forward_tw(b0, b1, P, C, tw_lds);
    for (int idx = threadIdx.x; idx < C * n; idx += kCtfThreads) {
        const int c = static_cast<int>(P.n_div.div(static_cast<uint32_t>(idx)));
        const int fy = idx - c * n;
        float2 v = X[idx];
        if (x0 + c < s.sw && ctf_negative(s, x0 + c, fy, z)) {
            v.x = -v.x;
            v.y = -v.y;
        }
        X[idx] = make_float2(v.x, -v.y);
    }
    __syncthreads();
    const float2* Y = fft_forward_tw(X, X == b0 ? b1 : b0, P, C, tw_lds);
    for (int idx = threadIdx.x; idx < C * n; idx += kCtfThreads) {
        const int y = idx / C, c = idx - y * C;
        if (x0 + c < s.sw) {
            const float2 v = Y[c * n + y];
            A[static_cast<long>(y) * s.sw + c] = make_float2(v.x, -v.y);
        }
    }
}

// Stage 5: rows [z][y] of A, inverse DFT along x (real part of DFT(conj(.))),
// blended into wsum / wtot where the pixel belongs to the strip.
__global__ void __launch_bounds__(kCtfThreads) ctf_rows_inv_kernel(CtfStrip s, FilterParams P) {
    extern __shared__ float2 sm[];
    const int n = P.n, ppw = P.pairs_per_wg;
    const int rows = s.zc * s.sh;
    const int r0 = blockIdx.x * ppw;
    const int cnt = min(ppw, rows - r0);
    float2* b0 = sm;
    float2* b1 = sm + ppw * n;
    const float2* src = s.a + static_cast<long>(r0) * n;
    for (int idx = threadIdx.x; idx < cnt * n; idx += kCtfThreads) {
        const float2 v = src[idx];
        b0[idx] = make_float2(v.x, -v.y);
    }
    __syncthreads();
    const float2* X = fft_forward(b0, b1, P, cnt);
    for (int idx = threadIdx.x; idx < cnt * n; idx += kCtfThreads) {
        const int q = static_cast<int>(P.n_div.div(static_cast<uint32_t>(idx)));
        const int x = idx - q * n;
        const int r = r0 + q;
        const int z = static_cast<int>(s.sh_div.div(static_cast<uint32_t>(r)));
        const int gx = s.bx0 + x, gy = s.by0 + r - z * s.sh;
        const double d = (static_cast<double>(gx) - s.cx) * s.px +
                         (static_cast<double>(gy) - s.cy) * s.py;
        if (d >= s.lo && d < s.hi) {
            const float dist = static_cast<float>(fabs((d - s.scp) / s.half));
            const float wt0 = 1.0f - dist;
            const float wt = wt0 > 0.0f ? wt0 : 0.0f;
            const float val = X[idx].x * s.scale;
            const long o = z * s.plane + static_cast<long>(gy) * s.nx + gx;
            s.wsum[o] += val * wt;
            s.wtot[o] += wt;
        }
    }
}

// ---- register transforms: rows of length n held in LDS rows of M = 2^L
// (16 <= M <= 4096) slots, max(1, 2048 / M) rows per workgroup, M / 8 threads
// per row.  n == M: the radix-8 passes directly.  n < M (Bluestein): the
// length-n DFT as the circular convolution of length M >= 2n - 1
//   X[k] = conj(w_k) * sum_j (x_j conj(w_j)) w_{k-j},  w_m = exp(i pi m^2 / n),
// i.e. two length-M transforms around the product with B = DFT_M(w) / M.
struct RegPlan {
    const float2* tw8;    // per-pass twiddles of the length-M transform
    const float2* chirp;  // [n] w_m, or null when n == M
    const float2* bspec;  // [M] DFT_M of the zero-padded symmetric chirp, / M
    int n;
};

template <int M>
struct RegShape {
    static constexpr int T = M / 8;                      // threads per row
    static constexpr int PPW = M >= 2048 ? 1 : 2048 / M;  // rows per workgroup
    static constexpr int TH = T * PPW;                    // threads per workgroup
    static constexpr int RS = p2_row_slots(M);            // LDS slots per row
};

__device__ __forceinline__ float2 cmul_conj(float2 a, float2 b) {  // a * conj(b)
    return make_float2(fmaf(a.x, b.x, a.y * b.y), fmaf(a.y, b.x, -(a.x * b.y)));
}

// Forward DFT of length R.n of every row of the workgroup (element i of row q
// at sm[q * RS + p2_slot(i)], i < n); all threads call it.
template <int M>
__device__ __forceinline__ void reg_dft(float2* sm, const RegPlan& R, const float2* twr) {
    using S = RegShape<M>;
    const int tid = threadIdx.x, q = tid / S::T, j = tid & (S::T - 1);
    if (!R.chirp) {
        p2_fft<M, 0, kP2None>(sm + q * S::RS, j, twr, nullptr, 0.f);
        return;
    }
    for (int idx = tid; idx < S::PPW * M; idx += S::TH) {
        const int qq = idx / M, i = idx & (M - 1);
        float2& e = sm[qq * S::RS + p2_slot(i)];
        e = i < R.n ? cmul_conj(e, R.chirp[i]) : make_float2(0.f, 0.f);
    }
    __syncthreads();
    p2_fft<M, 0, kP2None>(sm + q * S::RS, j, twr, nullptr, 0.f);
    for (int idx = tid; idx < S::PPW * M; idx += S::TH) {
        const int qq = idx / M, i = idx & (M - 1);
        float2& e = sm[qq * S::RS + p2_slot(i)];
        const float2 v = cmul(e, R.bspec[i]);
        e = make_float2(v.x, -v.y);
    }
    __syncthreads();
    p2_fft<M, 0, kP2None>(sm + q * S::RS, j, twr, nullptr, 0.f);
    for (int idx = tid; idx < S::PPW * M; idx += S::TH) {
        const int qq = idx / M, i = idx & (M - 1);
        if (i < R.n) {
            float2& e = sm[qq * S::RS + p2_slot(i)];
            const float2 v = cmul(e, R.chirp[i]);
            e = make_float2(v.x, -v.y);
        }
    }
    __syncthreads();
}

// Stage 1: two real rows packed as one complex row z = a + i b, one
// transform, the spectra separated as A[k] = (Z[k] + conj Z[n-k]) / 2,
// B[k] = (Z[k] - conj Z[n-k]) / (2i).
template <int M>
__global__ void __launch_bounds__(RegShape<M>::TH) ctf_rows_fwd_reg(CtfStrip s, RegPlan R) {
    using S = RegShape<M>;
    extern __shared__ float2 sm[];
    const int tid = threadIdx.x, j = tid & (S::T - 1);
    const int n = R.n;
    const int rows = s.zc * s.sh;
    const int pair0 = blockIdx.x * S::PPW;
    const int pairs = min(S::PPW, (rows + 1) / 2 - pair0);
    float2 twr[P2<M>::tw_total()];
    p2_load_tw<M, 0>(twr, j, R.tw8);
    for (int idx = tid; idx < S::PPW * M; idx += S::TH) {
        const int qq = idx / M, x = idx & (M - 1);
        float v[2] = {0.f, 0.f};
        if (qq < pairs && x < n) {
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const int r = 2 * (pair0 + qq) + c;
                if (r < rows) {
                    const int z = static_cast<int>(s.sh_div.div(static_cast<uint32_t>(r)));
                    const int y = r - z * s.sh;
                    v[c] = s.in[z * s.plane + static_cast<long>(s.by0 + y) * s.nx + s.bx0 + x];
                }
            }
        }
        sm[qq * S::RS + p2_slot(x)] = make_float2(v[0], v[1]);
    }
    __syncthreads();
    reg_dft<M>(sm, R, twr);
    for (int idx = tid; idx < pairs * M; idx += S::TH) {
        const int qq = idx / M, k = idx & (M - 1);
        if (k >= n) continue;
        const float2* row = sm + qq * S::RS;
        const float2 X = row[p2_slot(k)], Y = row[p2_slot(k == 0 ? 0 : n - k)];
        const int ra = 2 * (pair0 + qq);
        s.a[static_cast<long>(ra) * n + k] = make_float2(0.5f * (X.x + Y.x), 0.5f * (X.y - Y.y));
        if (ra + 1 < rows)
            s.a[static_cast<long>(ra + 1) * n + k] =
                make_float2(0.5f * (X.y + Y.y), 0.5f * (Y.x - X.x));
    }
}

// Stage 3: columns of length n, forward DFT, the flip, inverse as
// conj(DFT(conj(.))).
template <int M>
__global__ void __launch_bounds__(RegShape<M>::TH) ctf_cols_reg(CtfStrip s, RegPlan R) {
    using S = RegShape<M>;
    extern __shared__ float2 sm[];
    const int tid = threadIdx.x, j = tid & (S::T - 1);
    const int n = R.n;
    const int rows = s.zc * s.sw;
    const int r0 = blockIdx.x * S::PPW;
    const int cnt = min(S::PPW, rows - r0);
    float2 twr[P2<M>::tw_total()];
    p2_load_tw<M, 0>(twr, j, R.tw8);
    float2* base = s.b + static_cast<long>(r0) * n;
    for (int idx = tid; idx < S::PPW * M; idx += S::TH) {
        const int qq = idx / M, x = idx & (M - 1);
        sm[qq * S::RS + p2_slot(x)] =
            qq < cnt && x < n ? base[qq * n + x] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    reg_dft<M>(sm, R, twr);
    for (int idx = tid; idx < cnt * M; idx += S::TH) {
        const int qq = idx / M, fy = idx & (M - 1);
        if (fy >= n) continue;
        const int r = r0 + qq;
        const int z = static_cast<int>(s.sw_div.div(static_cast<uint32_t>(r)));
        const int fx = r - z * s.sw;
        float2& e = sm[qq * S::RS + p2_slot(fy)];
        float2 v = e;
        if (ctf_negative(s, fx, fy, z)) {
            v.x = -v.x;
            v.y = -v.y;
        }
        e = make_float2(v.x, -v.y);
    }
    __syncthreads();
    reg_dft<M>(sm, R, twr);
    for (int idx = tid; idx < cnt * M; idx += S::TH) {
        const int qq = idx / M, x = idx & (M - 1);
        if (x >= n) continue;
        const float2 v = sm[qq * S::RS + p2_slot(x)];
        base[qq * n + x] = make_float2(v.x, -v.y);
    }
}

// Stage 5, two rows per transform.  Only the real part of each row's inverse
// is kept, and Re(IDFT(A)) = IDFT(H) with H the Hermitian part
// H[k] = (A[k] + conj A[n-k]) / 2 (the flipped spectrum is not Hermitian where
// the CTF differs between +-f on the Nyquist lines, e.g. with a rotated tilt
// axis or astigmatism).  Z = H1 + i H2 then gives row 1 in the real part and
// row 2 in the imaginary part of conj(DFT(conj Z)).
template <int M>
__global__ void __launch_bounds__(RegShape<M>::TH) ctf_rows_inv_reg(CtfStrip s, RegPlan R) {
    using S = RegShape<M>;
    extern __shared__ float2 sm[];
    const int tid = threadIdx.x, j = tid & (S::T - 1);
    const int n = R.n;
    const int rows = s.zc * s.sh;
    const int pair0 = blockIdx.x * S::PPW;
    const int pairs = min(S::PPW, (rows + 1) / 2 - pair0);
    float2 twr[P2<M>::tw_total()];
    p2_load_tw<M, 0>(twr, j, R.tw8);
    for (int idx = tid; idx < S::PPW * M; idx += S::TH) {
        const int qq = idx / M, x = idx & (M - 1);
        float2 z = make_float2(0.f, 0.f);
        if (qq < pairs && x < n) {
            const int ra = 2 * (pair0 + qq);
            const int xm = x == 0 ? 0 : n - x;
            const float2* r1 = s.a + static_cast<long>(ra) * n;
            const float2 a1 = r1[x], b1 = r1[xm];
            float2 a2 = make_float2(0.f, 0.f), b2 = a2;
            if (ra + 1 < rows) {
                a2 = r1[n + x];
                b2 = r1[n + xm];
            }
            const float2 h1 = make_float2(0.5f * (a1.x + b1.x), 0.5f * (a1.y - b1.y));
            const float2 h2 = make_float2(0.5f * (a2.x + b2.x), 0.5f * (a2.y - b2.y));
            z = make_float2(h1.x - h2.y, -(h1.y + h2.x));
        }
        sm[qq * S::RS + p2_slot(x)] = z;
    }
    __syncthreads();
    reg_dft<M>(sm, R, twr);
    for (int idx = tid; idx < pairs * M; idx += S::TH) {
        const int qq = idx / M, x = idx & (M - 1);
        if (x >= n) continue;
        const float2 w = sm[qq * S::RS + p2_slot(x)];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int r = 2 * (pair0 + qq) + c;
            if (r >= rows) break;
            const int z = static_cast<int>(s.sh_div.div(static_cast<uint32_t>(r)));
            const int gx = s.bx0 + x, gy = s.by0 + r - z * s.sh;
            const double d = (static_cast<double>(gx) - s.cx) * s.px +
                             (static_cast<double>(gy) - s.cy) * s.py;
            if (d >= s.lo && d < s.hi) {
                const float dist = static_cast<float>(fabs((d - s.scp) / s.half));
                const float wt0 = 1.0f - dist;
                const float wt = wt0 > 0.0f ? wt0 : 0.0f;
                const float val = (c ? -w.y : w.x) * s.scale;
                const long o = z * s.plane + static_cast<long>(gy) * s.nx + gx;
                s.wsum[o] += val * wt;
                s.wtot[o] += wt;
            }
        }
    }
}

// src [batch][rows][cols] -> dst [batch][cols][rows] through a 32 x 32 LDS tile.
__global__ void __launch_bounds__(256) ctf_transpose_kernel(const float2* __restrict__ src,
                                                            float2* __restrict__ dst, int rows,
                                                            int cols) {
    __shared__ float2 tile[32][33];
    const long plane = static_cast<long>(rows) * cols;
    src += blockIdx.z * plane;
    dst += blockIdx.z * plane;
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
    for (int k = 0; k < 32; k += 8) {
        const int r = r0 + ty + k, c = c0 + tx;
        if (r < rows && c < cols) tile[ty + k][tx] = src[static_cast<long>(r) * cols + c];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 32; k += 8) {
        const int c = c0 + ty + k, r = r0 + tx;
        if (c < cols && r < rows) dst[static_cast<long>(c) * rows + r] = tile[tx][ty + k];
    }
}

// main.rs:682-688: out = wsum / wtot where wtot > 1e-10, else the input.
__global__ void __launch_bounds__(256) ctf_combine_kernel(const float* __restrict__ in,
                                                          float* __restrict__ out,
                                                          const float* __restrict__ wtot, long n) {
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += static_cast<long>(gridDim.x) * 256) {
        const float t = wtot[i];
        out[i] = t > 1e-10f ? out[i] / t : in[i];
    }
}

// ------------------------------------------------------------------ host side

struct StripBox {
    double scp, lo, hi;
    int bx0, bx1, by0, by1;
};

// (x - cx) * px + (y - cy) * py as main.rs:464 evaluates it (no contraction).
inline double perp_dist(int x, int y, double cx, double cy, double px, double py) {
    const double t1 = (static_cast<double>(x) - cx) * px;
    const double t2 = (static_cast<double>(y) - cy) * py;
    return t1 + t2;
}

// First x in [0, n) with pred(x) true for a predicate that is false then true
// along x (n when it never holds).
template <class Pred>
int first_true(int n, Pred pred) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (pred(mid))
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

// The strips of main.rs:389-487: the perpendicular range from the corners, the
// strip count, each strip's centre and membership interval, and the bounding
// box of the pixels with lo <= d < hi.  The distance is monotone along a row
// (its rounding is), so each row's member pixels are one interval found by
// bisection; the result equals the reference's full scan.
std::vector<StripBox> strip_layout(int nx, int ny, double axis_rad, int strip_width) {
    const double px = std::cos(axis_rad), py = std::sin(axis_rad);
    const double cx = nx / 2.0, cy = ny / 2.0;
    double min_perp = DBL_MAX, max_perp = -DBL_MAX;
    const double xs[4] = {0.0, static_cast<double>(nx), 0.0, static_cast<double>(nx)};
    const double ys[4] = {0.0, 0.0, static_cast<double>(ny), static_cast<double>(ny)};
    for (int k = 0; k < 4; ++k) {
        const double t1 = (xs[k] - cx) * px;
        const double t2 = (ys[k] - cy) * py;
        const double d = t1 + t2;
        if (d < min_perp) min_perp = d;
        if (d > max_perp) max_perp = d;
    }
    const double sw_d = static_cast<double>(strip_width);
    const int n_strips = static_cast<int>(std::ceil((max_perp - min_perp) / sw_d));
    std::vector<StripBox> out(std::max(0, n_strips));
    for (int si = 0; si < n_strips; ++si) {
        StripBox& b = out[si];
        b.scp = min_perp + (static_cast<double>(si) + 0.5) * sw_d;
        b.lo = b.scp - sw_d / 2.0;
        b.hi = b.scp + sw_d / 2.0;
        b.bx0 = nx;
        b.bx1 = 0;
        b.by0 = ny;
        b.by1 = 0;
        for (int y = 0; y < ny; ++y) {
            int a, e;
            if (px >= 0.0) {  // d non-decreasing in x
                a = first_true(nx, [&](int x) { return perp_dist(x, y, cx, cy, px, py) >= b.lo; });
                e = first_true(nx, [&](int x) { return perp_dist(x, y, cx, cy, px, py) >= b.hi; });
            } else {  // d non-increasing in x
                a = first_true(nx, [&](int x) { return perp_dist(x, y, cx, cy, px, py) < b.hi; });
                e = first_true(nx, [&](int x) { return perp_dist(x, y, cx, cy, px, py) < b.lo; });
            }
            if (a >= e) continue;
            b.bx0 = std::min(b.bx0, a);
            b.bx1 = std::max(b.bx1, e);
            b.by0 = std::min(b.by0, y);
            b.by1 = y + 1;
        }
    }
    return out;
}

// electron_wavelength (main.rs:709-719), Angstrom.
double wavelength_a(float voltage_kv) {
    const double v = static_cast<double>(voltage_kv) * 1000.0;
    const double m0 = 9.10938e-31, e = 1.60218e-19, c = 2.99792e8, h = 6.62607e-34;
    const double lam = h / std::sqrt(2.0 * m0 * e * v * (1.0 + e * v / (2.0 * m0 * c * c)));
    return lam * 1e10;
}

const double kF32Pi = static_cast<double>(static_cast<float>(kPi));  // std::f32::consts::PI

}  // namespace
}  // namespace tiltgpu

using namespace tiltgpu;
using tiltgpu_sess::fail;

// Transform plan of one length n: the register kernels (M > 0: length-M
// radix-8 passes, directly when n == M, by Bluestein otherwise) or the
// mixed-radix LDS passes (M == 0).
struct CtfPlan {
    FilterParams p{};
    int M = 0;
    float2* tw = nullptr;     // mixed-radix twiddles [n]
    float2* tw8 = nullptr;    // per-pass twiddles of the length-M transform
    float2* chirp = nullptr;  // Bluestein chirp [n]
    float2* bspec = nullptr;  // Bluestein kernel spectrum [M]
};

struct tiltgpu_ctf {
    int nx = 0, ny = 0, nz = 0, width = 0, zc = 0;
    tiltgpu_ctf_params prm{};
    std::vector<StripBox> strips;
    std::vector<double2> def_host;    // [n_strips][nz]
    std::vector<CtfAstig> astig_host;  // [nz]
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float* in = nullptr;
    float* out = nullptr;
    float* wtot = nullptr;
    float2* a = nullptr;
    float2* b = nullptr;
    double2* def = nullptr;
    CtfAstig* astig = nullptr;
    std::map<int, CtfPlan> plans;
    double lambda = 0, cs_a = 0, amp = 0, amp_phase = 0, cuton = 0, plate = 0, axis = 0;
    double last_ms = 0;
    int64_t launches = 0;
    std::vector<char> uploaded;  // [nz]: section uploaded
    bool done = false;           // a run has completed since the last upload
};

namespace {

void ctf_free_plan(CtfPlan& pl) {
    float2* ps[] = {pl.tw, pl.tw8, pl.chirp, pl.bspec};
    for (float2* q : ps)
        if (q) (void)hipFree(q);
}

void ctf_release(tiltgpu_ctf* c) {
    for (auto& kv : c->plans) ctf_free_plan(kv.second);
    void* bufs[] = {c->in, c->out, c->wtot, c->a, c->b, c->def, c->astig};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

template <class T>
int ctf_alloc(T** p, size_t count) {
    if (tiltgpu_sess::forced_alloc_failure()) {
        *p = nullptr;
        return fail(TILTGPU_ERR_BUDGET, "device buffer of %zu bytes refused (TILTGPU_FAIL_ALLOC_AT)",
                    count * sizeof(T));
    }
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(1, count) * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(e == hipErrorOutOfMemory ? TILTGPU_ERR_BUDGET : TILTGPU_ERR_HIP,
                    "device buffer of %zu bytes unavailable: %s", count * sizeof(T),
                    hipGetErrorString(e));
    }
    return TILTGPU_OK;
}

int largest_prime_factor(int n) {
    int best = 1;
    for (int p = 2; p * p <= n; ++p)
        while (n % p == 0) {
            best = p;
            n /= p;
        }
    return std::max(best, n);
}

// In-place radix-2 DFT of length m (a power of two) in double precision.
void host_fft(std::vector<double>& re, std::vector<double>& im) {
    const int m = static_cast<int>(re.size());
    for (int i = 1, j = 0; i < m; ++i) {
        int bit = m >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            std::swap(re[i], re[j]);
            std::swap(im[i], im[j]);
        }
    }
    for (int len = 2; len <= m; len <<= 1) {
        const int h = len / 2;
        for (int k = 0; k < h; ++k) {
            const double a = -2.0 * kPi * k / len, wr = std::cos(a), wi = std::sin(a);
            for (int i = 0; i < m; i += len) {
                const double xr = re[i + k + h] * wr - im[i + k + h] * wi;
                const double xi = re[i + k + h] * wi + im[i + k + h] * wr;
                re[i + k + h] = re[i + k] - xr;
                im[i + k + h] = im[i + k] - xi;
                re[i + k] += xr;
                im[i + k] += xi;
            }
        }
    }
}

template <class T>
int upload(T** dst, const std::vector<T>& v) {
    int rc = ctf_alloc(dst, v.size());
    if (rc) return rc;
    HIP_TRY(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return TILTGPU_OK;
}

// Transform plan of length n.  Powers of two from 16 to 4096 use the register
// radix-8 passes; other lengths up to 2048 with a prime factor above 13 (the
// mixed-radix passes cost O(n p) for a prime factor p) use Bluestein over the
// smallest power of two M >= max(64, 2n - 1); the rest use the mixed-radix LDS
// passes, max(1, 2048 / n) rows per workgroup.  TILTGPU_CTF_BLUESTEIN=all
// selects Bluestein for every other length up to 2048 as well.
int ctf_plan(tiltgpu_ctf* c, int n) {
    if (c->plans.count(n)) return TILTGPU_OK;
    CtfPlan pl;
    const char* env = std::getenv("TILTGPU_CTF_BLUESTEIN");
    const bool all = env && std::strcmp(env, "all") == 0;
    const bool pow2 = (n & (n - 1)) == 0;
    if (pow2 && n >= 16 && n <= 4096) {
        pl.M = n;
    } else if (n >= 2 && n <= 2048 && !pow2 && (all || largest_prime_factor(n) > 13)) {
        int m = 64;
        while (m < 2 * n - 1) m <<= 1;
        pl.M = m;
    }
    int rc;
    pl.p.n = n;
    if (pl.M > 0) {
        std::vector<float2> t8(pow2_twiddles(pl.M, nullptr, 4096, 16));
        pow2_twiddles(pl.M, t8.data(), 4096, 16);
        if ((rc = upload(&pl.tw8, t8))) return rc;
        if (pl.M != n) {
            // chirp w_m = exp(i pi m^2 / n), the phase reduced exactly (m^2 mod 2n)
            std::vector<float2> w(n);
            std::vector<double> br(pl.M, 0.0), bi(pl.M, 0.0);
            for (int k = 0; k < n; ++k) {
                const long long e = (static_cast<long long>(k) * k) % (2LL * n);
                const double a = kPi * static_cast<double>(e) / static_cast<double>(n);
                w[k] = make_float2(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a)));
                br[k] = std::cos(a);
                bi[k] = std::sin(a);
                if (k > 0) {
                    br[pl.M - k] = br[k];
                    bi[pl.M - k] = bi[k];
                }
            }
            host_fft(br, bi);
            std::vector<float2> bs(pl.M);
            for (int k = 0; k < pl.M; ++k)
                bs[k] = make_float2(static_cast<float>(br[k] / pl.M),
                                    static_cast<float>(bi[k] / pl.M));
            if ((rc = upload(&pl.chirp, w)) || (rc = upload(&pl.bspec, bs))) {
                ctf_free_plan(pl);
                return rc;
            }
        }
        c->plans[n] = pl;
        return TILTGPU_OK;
    }
    FilterParams& f = pl.p;
    f.pairs_per_wg = std::max(1, 2048 / n);
    if (n == 1) {
        f.nradix = 0;
    } else if (filter_plan(n, f.radix, &f.nradix) != 0) {
        return fail(TILTGPU_ERR_ARG, "no transform plan for length %d", n);
    }
    f.n_div = make_fastdiv(static_cast<uint32_t>(n));
    int ns = 1;
    for (int i = 0; i < f.nradix; ++i) {
        f.m_div[i] = make_fastdiv(static_cast<uint32_t>(n / f.radix[i]));
        f.ns_div[i] = make_fastdiv(static_cast<uint32_t>(ns));
        f.r_div[i] = make_fastdiv(static_cast<uint32_t>(f.radix[i]));
        ns *= f.radix[i];
    }
    std::vector<float2> tw(n);
    for (int t = 0; t < n; ++t) {
        const double ang = -2.0 * kPi * static_cast<double>(t) / static_cast<double>(n);
        tw[t] = make_float2(static_cast<float>(std::cos(ang)), static_cast<float>(std::sin(ang)));
    }
    if ((rc = upload(&pl.tw, tw))) return rc;
    f.tw = pl.tw;
    c->plans[n] = pl;
    return TILTGPU_OK;
}

using RegKernel = void (*)(CtfStrip, RegPlan);

// The register kernel of stage 1 (0), 3 (1) or 5 (2) for transform length M.
RegKernel reg_kernel(int stage, int M) {
#define TILTGPU_CTF_REG(N) \
    case N:                \
        return stage == 0 ? ctf_rows_fwd_reg<N> : stage == 1 ? ctf_cols_reg<N> : ctf_rows_inv_reg<N>;
    switch (M) {
        TILTGPU_CTF_REG(16)
        TILTGPU_CTF_REG(32)
        TILTGPU_CTF_REG(64)
        TILTGPU_CTF_REG(128)
        TILTGPU_CTF_REG(256)
        TILTGPU_CTF_REG(512)
        TILTGPU_CTF_REG(1024)
        TILTGPU_CTF_REG(2048)
        TILTGPU_CTF_REG(4096)
        default:
            return nullptr;
    }
#undef TILTGPU_CTF_REG
}

// Stage `stage` over `items` rows (stages 0 and 2: real rows, two per
// transform) with the register kernels.
hipError_t launch_reg(int stage, const CtfStrip& s, const CtfPlan& pl, int items, hipStream_t st) {
    const int M = pl.M, ppw = M >= 2048 ? 1 : 2048 / M, threads = (M / 8) * ppw;
    const int units = stage == 1 ? items : (items + 1) / 2;
    const int nblk = (units + ppw - 1) / ppw;
    CtfStrip a = s;
    RegPlan r{pl.tw8, pl.chirp, pl.bspec, pl.p.n};
    void* args[] = {&a, &r};
    const size_t lds = static_cast<size_t>(ppw) * p2_row_slots(M) * sizeof(float2);
    return hipLaunchKernel(reinterpret_cast<const void*>(reg_kernel(stage, M)), dim3(nblk),
                           dim3(threads), args, lds, st);
}

size_t plan_lds(const FilterParams& f) {  // two row buffers
    return static_cast<size_t>(2) * f.pairs_per_wg * f.n * sizeof(float2);
}

hipError_t launch_rows(void (*k)(CtfStrip, FilterParams), CtfStrip s, FilterParams f, int rows,
                       hipStream_t st) {
    const int nblk = (rows + f.pairs_per_wg - 1) / f.pairs_per_wg;
    void* args[] = {&s, &f};
    return hipLaunchKernel(reinterpret_cast<const void*>(k), dim3(nblk), dim3(kCtfThreads), args,
                           plan_lds(f), st);
}

// Columns per workgroup of the in-place column kernel for column length n (0:
// transpose instead): 8 columns (64-byte row segments) while the two LDS
// buffers stay within 64 KiB.  Measured at 2048x2048x121 with a tilt axis at
// 85 deg (195-point columns): 8 columns 118 ms, 16 columns 128 ms (half the
// workgroups per CU), 4 columns 154 ms (32-byte segments).
int cols_tile_width(int n) {
    return 2 * 8 * n * static_cast<int>(sizeof(float2)) <= 65536 ? 8 : 0;
}

hipError_t launch_cols_tile(int C, const CtfStrip& s, const FilterParams& f, int zc,
                            hipStream_t st) {
    const int ntile = (s.sw + C - 1) / C;
    CtfStrip a = s;
    FilterParams q = f;
    void* args[] = {&a, &q};
    if (C != 8) return hipErrorInvalidValue;
    const void* k = reinterpret_cast<const void*>(ctf_cols_tile_kernel<8>);
    const size_t lds = static_cast<size_t>(2) * C * f.n * sizeof(float2);
    return hipLaunchKernel(k, dim3(static_cast<unsigned>(zc) * ntile), dim3(kCtfThreads), args,
                           lds, st);
}

hipError_t launch_transpose(const float2* src, float2* dst, int rows, int cols, int batch,
                            hipStream_t st) {
    dim3 grid((cols + 31) / 32, (rows + 31) / 32, batch);
    hipLaunchKernelGGL(ctf_transpose_kernel, grid, dim3(256), 0, st, src, dst, rows, cols);
    return hipGetLastError();
}

int check_params(const tiltgpu_ctf_params* p) {
    if (!p) return fail(TILTGPU_ERR_ARG, "params is NULL");
    if (p->strip_width < 1) return fail(TILTGPU_ERR_ARG, "strip width %d < 1", p->strip_width);
    if (p->max_strip_width < 0)
        return fail(TILTGPU_ERR_ARG, "max strip width %d < 0", p->max_strip_width);
    if (!(p->pixel_a > 0.0f) || !std::isfinite(p->pixel_a))
        return fail(TILTGPU_ERR_ARG, "pixel size %g is not a positive number",
                    static_cast<double>(p->pixel_a));
    return TILTGPU_OK;
}

}  // namespace

extern "C" {

void tiltgpu_ctf_default_params(tiltgpu_ctf_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    p->voltage_kv = 300.0f;
    p->cs_mm = 2.7f;
    p->pixel_a = 1.0f;
    p->amplitude_contrast = 0.07f;
    p->defocus_nm = -3000.0f;
    p->strip_width = 256;
}

double tiltgpu_electron_wavelength(float voltage_kv) { return wavelength_a(voltage_kv); }

int tiltgpu_ctf_strip_width(const tiltgpu_ctf_params* p, int32_t nx, const float* vdef,
                            int32_t n_vdef) {
    int rc = check_params(p);
    if (rc) return rc;
    if (nx < 1) return fail(TILTGPU_ERR_ARG, "nx %d < 1", nx);
    if (vdef == nullptr && n_vdef > 0) return fail(TILTGPU_ERR_ARG, "vdef is NULL");
    long max_w;
    if (p->max_strip_width > 0) {
        max_w = p->max_strip_width;
    } else {
        double max_def = 0.0;
        if (vdef) {
            for (int i = 0; i < n_vdef; ++i) {
                const float m = std::max(std::fabs(vdef[3 * i]), std::fabs(vdef[3 * i + 1]));
                max_def = std::max(max_def, static_cast<double>(m) * 10.0);
            }
        } else {
            max_def = static_cast<double>(std::fabs(p->defocus_nm)) * 10.0;
        }
        const double lam = wavelength_a(p->voltage_kv);
        if (max_def > 1.0 && lam > 0.0) {
            const double per_px = max_def / static_cast<double>(nx);
            const double safe = static_cast<double>(p->pixel_a) / (lam * per_px);
            max_w = static_cast<long>(std::max(safe, 16.0));
        } else {
            max_w = p->strip_width;
        }
    }
    return static_cast<int>(std::min<long>(p->strip_width, max_w));
}

int tiltgpu_ctf_strip_layout(int32_t nx, int32_t ny, float tilt_axis_deg, int32_t strip_width,
                             int32_t* boxes, double* centres, int32_t cap) {
    if (nx < 1 || ny < 1 || strip_width < 1)
        return fail(TILTGPU_ERR_ARG, "invalid layout %d x %d, strip width %d", nx, ny, strip_width);
    const double axis = static_cast<double>(tilt_axis_deg) * kF32Pi / 180.0;
    const std::vector<StripBox> st = strip_layout(nx, ny, axis, strip_width);
    const int n = static_cast<int>(st.size());
    for (int i = 0; i < n && i < cap; ++i) {
        if (boxes) {
            boxes[4 * i] = st[i].bx0;
            boxes[4 * i + 1] = st[i].bx1;
            boxes[4 * i + 2] = st[i].by0;
            boxes[4 * i + 3] = st[i].by1;
        }
        if (centres) centres[i] = st[i].scp;
    }
    return n;
}

int tiltgpu_ctf_create(int32_t nx, int32_t ny, int32_t nz, const tiltgpu_ctf_params* p,
                       const float* tilt_deg, int32_t n_tilt, const float* vdef, int32_t n_vdef,
                       tiltgpu_ctf** out) {
    if (!out) return fail(TILTGPU_ERR_ARG, "out is NULL");
    *out = nullptr;
    int rc = check_params(p);
    if (rc) return rc;
    if (nx < 1 || ny < 1 || nz < 1)
        return fail(TILTGPU_ERR_ARG, "invalid stack %d x %d x %d", nx, ny, nz);
    if (n_tilt < 0 || (n_tilt > 0 && !tilt_deg) || n_vdef < 0 || (n_vdef > 0 && !vdef))
        return fail(TILTGPU_ERR_ARG, "tilt angles or defocus table missing");
    const int width = tiltgpu_ctf_strip_width(p, nx, vdef, n_vdef);
    if (width < 1) return width < 0 ? width : fail(TILTGPU_ERR_ARG, "strip width 0");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        return fail(TILTGPU_ERR_NODEV, "no GPU visible to HIP");

    auto* c = new tiltgpu_ctf();
    c->nx = nx;
    c->ny = ny;
    c->nz = nz;
    c->width = width;
    c->prm = *p;
    c->lambda = wavelength_a(p->voltage_kv);
    c->cs_a = static_cast<double>(p->cs_mm) * 1e7;
    c->amp = static_cast<double>(p->amplitude_contrast);
    c->amp_phase = std::sqrt(1.0 - c->amp * c->amp);
    c->cuton = static_cast<double>(p->cuton_freq);
    c->axis = static_cast<double>(p->tilt_axis_deg) * kF32Pi / 180.0;
    c->plate = static_cast<double>(p->plate_phase_deg) * kF32Pi / 180.0;
    c->strips = strip_layout(nx, ny, c->axis, width);

    // per-section strip defocus (main.rs:372-385, 445-450) and astigmatism
    const int ns = static_cast<int>(c->strips.size());
    c->def_host.resize(static_cast<size_t>(ns) * nz);
    c->astig_host.resize(nz);
    const double pixel = static_cast<double>(p->pixel_a);
    for (int z = 0; z < nz; ++z) {
        const float tdeg = z < n_tilt ? tilt_deg[z] : 0.0f;
        const double tilt_rad = static_cast<double>(tdeg) * kF32Pi / 180.0;
        float d1 = p->defocus_nm, d2 = p->defocus_nm, aa = 0.0f;
        if (vdef && z < n_vdef) {
            d1 = vdef[3 * z];
            d2 = vdef[3 * z + 1];
            aa = vdef[3 * z + 2];
        }
        const double def1 = static_cast<double>(d1) * 10.0, def2 = static_cast<double>(d2) * 10.0;
        const double arad = static_cast<double>(aa) * kF32Pi / 180.0;
        c->astig_host[z] = CtfAstig{std::cos(2.0 * arad), std::sin(2.0 * arad),
                                    std::fabs(def1 - def2) > 0.01 ? 1.0 : 0.0, 0.0};
        const double st = std::sin(tilt_rad);
        for (int si = 0; si < ns; ++si) {
            const double dx_a = c->strips[si].scp * pixel;
            c->def_host[static_cast<size_t>(si) * nz + z] =
                make_double2(def1 + dx_a * st, def2 + dx_a * st);
        }
    }

    long max_area = 1;
    for (const StripBox& b : c->strips) {
        if (b.bx1 <= b.bx0 || b.by1 <= b.by0) continue;
        const int sw = b.bx1 - b.bx0, sh = b.by1 - b.by0;
        max_area = std::max(max_area, static_cast<long>(sw) * sh);
    }
    c->zc = static_cast<int>(std::max(1L, std::min<long>(nz, kCtfChunkElems / max_area)));
    const size_t plane = static_cast<size_t>(nx) * ny;
    const size_t work = static_cast<size_t>(c->zc) * max_area;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        ctf_release(c);
        return fail(TILTGPU_ERR_HIP, "stream or event creation failed");
    }
    if ((rc = ctf_alloc(&c->in, plane * nz)) || (rc = ctf_alloc(&c->out, plane * nz)) ||
        (rc = ctf_alloc(&c->wtot, plane * c->zc)) || (rc = ctf_alloc(&c->a, work)) ||
        (rc = ctf_alloc(&c->b, work)) || (rc = ctf_alloc(&c->def, c->def_host.size())) ||
        (rc = ctf_alloc(&c->astig, c->astig_host.size()))) {
        ctf_release(c);
        return rc;
    }
    if (hipMemcpy(c->def, c->def_host.data(), c->def_host.size() * sizeof(double2),
                  hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->astig, c->astig_host.data(), c->astig_host.size() * sizeof(CtfAstig),
                  hipMemcpyHostToDevice) != hipSuccess) {
        ctf_release(c);
        return fail(TILTGPU_ERR_HIP, "defocus table upload failed");
    }
    size_t lds_max = 0;
    for (const StripBox& b : c->strips) {
        if (b.bx1 <= b.bx0 || b.by1 <= b.by0) continue;
        if ((rc = ctf_plan(c, b.bx1 - b.bx0)) || (rc = ctf_plan(c, b.by1 - b.by0))) {
            ctf_release(c);
            return rc;
        }
    }
    for (const auto& kv : c->plans)
        if (kv.second.M == 0) lds_max = std::max(lds_max, plan_lds(kv.second.p));
    if (lds_max > kCtfMaxLds) {  // mixed-radix rows beyond 10240 points
        ctf_release(c);
        return fail(TILTGPU_ERR_ARG,
                    "a strip bounding box side needs %zu bytes of LDS (more than %zu)", lds_max,
                    kCtfMaxLds);
    }
    if (lds_max > 65536) {  // rows longer than 4096: opt in to the larger LDS allocation
        const void* ks[] = {reinterpret_cast<const void*>(ctf_rows_fwd_kernel),
                            reinterpret_cast<const void*>(ctf_cols_kernel),
                            reinterpret_cast<const void*>(ctf_rows_inv_kernel)};
        for (const void* k : ks) {
            if (hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    static_cast<int>(lds_max)) != hipSuccess) {
                ctf_release(c);
                return fail(TILTGPU_ERR_HIP, "%zu bytes of LDS per workgroup unavailable", lds_max);
            }
        }
    }
    *out = c;
    return TILTGPU_OK;
}

int tiltgpu_ctf_strip_width_used(const tiltgpu_ctf* c) {
    if (!c) return fail(TILTGPU_ERR_ARG, "context is NULL");
    return c->width;
}

int tiltgpu_ctf_upload(tiltgpu_ctf* c, int32_t z0, int32_t nsec, const float* host) {
    if (!c || !host) return fail(TILTGPU_ERR_ARG, "context or data is NULL");
    if (z0 < 0 || nsec < 0 || z0 + nsec > c->nz)
        return fail(TILTGPU_ERR_ARG, "sections [%d, %d) outside [0, %d)", z0, z0 + nsec, c->nz);
    const size_t plane = static_cast<size_t>(c->nx) * c->ny;
    HIP_TRY(hipMemcpyAsync(c->in + plane * z0, host, plane * nsec * sizeof(float),
                           hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (c->uploaded.empty()) c->uploaded.assign(c->nz, 0);
    for (int z = z0; z < z0 + nsec; ++z) c->uploaded[z] = 1;
    c->done = false;
    return TILTGPU_OK;
}

int tiltgpu_ctf_run(tiltgpu_ctf* c) {
    if (!c) return fail(TILTGPU_ERR_ARG, "context is NULL");
    if (c->uploaded.empty() ||
        std::find(c->uploaded.begin(), c->uploaded.end(), 0) != c->uploaded.end())
        return fail(TILTGPU_ERR_STATE, "run before every section was uploaded");
    const long plane = static_cast<long>(c->nx) * c->ny;
    const hipStream_t st = c->stream;
    HIP_TRY(hipEventRecord(c->ev0, st));
    int64_t launches = 0;
    CtfStrip s{};
    s.plane = plane;
    s.nx = c->nx;
    s.cx = c->nx / 2.0;
    s.cy = c->ny / 2.0;
    s.px = std::cos(c->axis);
    s.py = std::sin(c->axis);
    s.half = static_cast<double>(c->width) / 2.0;
    s.lambda = c->lambda;
    s.cs_a = c->cs_a;
    s.alpha = std::atan2(c->amp, c->amp_phase);
    s.cuton = c->cuton;
    s.plate = c->plate;
    s.rotate = std::fabs(c->axis) > 1e-6 ? 1 : 0;
    s.ca = std::cos(c->axis);
    s.sa = std::sin(c->axis);
    s.a = c->a;
    s.b = c->b;
    s.wtot = c->wtot;
    const double pixel = static_cast<double>(c->prm.pixel_a);
    for (int z0 = 0; z0 < c->nz; z0 += c->zc) {
        const int zc = std::min(c->zc, c->nz - z0);
        s.zc = zc;
        s.in = c->in + plane * z0;
        s.wsum = c->out + plane * z0;
        s.astig = c->astig + z0;
        HIP_TRY(hipMemsetAsync(s.wsum, 0, plane * zc * sizeof(float), st));
        HIP_TRY(hipMemsetAsync(s.wtot, 0, plane * zc * sizeof(float), st));
        for (size_t si = 0; si < c->strips.size(); ++si) {
            const StripBox& b = c->strips[si];
            if (b.bx1 <= b.bx0 || b.by1 <= b.by0) continue;
            s.bx0 = b.bx0;
            s.by0 = b.by0;
            s.sw = b.bx1 - b.bx0;
            s.sh = b.by1 - b.by0;
            s.sw_div = make_fastdiv(static_cast<uint32_t>(s.sw));
            s.sh_div = make_fastdiv(static_cast<uint32_t>(s.sh));
            s.scp = b.scp;
            s.lo = b.lo;
            s.hi = b.hi;
            s.den_x = static_cast<double>(s.sw) * pixel;
            s.den_y = static_cast<double>(s.sh) * pixel;
            s.scale = 1.0f / static_cast<float>(s.sw * s.sh);
            s.def = c->def + si * c->nz + z0;
            const CtfPlan& px = c->plans.at(s.sw);
            const CtfPlan& py = c->plans.at(s.sh);
            if (px.M)
                HIP_TRY(launch_reg(0, s, px, zc * s.sh, st));
            else
                HIP_TRY(launch_rows(ctf_rows_fwd_kernel, s, px.p, zc * s.sh, st));
            const int tile = py.M ? 0 : cols_tile_width(s.sh);
            if (tile) {
                HIP_TRY(launch_cols_tile(tile, s, py.p, zc, st));
                launches -= 2;
            } else {
                HIP_TRY(launch_transpose(s.a, s.b, s.sh, s.sw, zc, st));
                if (py.M)
                    HIP_TRY(launch_reg(1, s, py, zc * s.sw, st));
                else
                    HIP_TRY(launch_rows(ctf_cols_kernel, s, py.p, zc * s.sw, st));
                HIP_TRY(launch_transpose(s.b, s.a, s.sw, s.sh, zc, st));
            }
            if (px.M)
                HIP_TRY(launch_reg(2, s, px, zc * s.sh, st));
            else
                HIP_TRY(launch_rows(ctf_rows_inv_kernel, s, px.p, zc * s.sh, st));
            launches += 5;
        }
        const long n = plane * zc;
        const int nblk = static_cast<int>(std::min<long>((n + 255) / 256, 65536));
        hipLaunchKernelGGL(ctf_combine_kernel, dim3(nblk), dim3(256), 0, st,
                           static_cast<const float*>(s.in), s.wsum,
                           static_cast<const float*>(s.wtot), n);
        HIP_TRY(hipGetLastError());
        launches += 1;
    }
    HIP_TRY(hipEventRecord(c->ev1, st));
    HIP_TRY(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->launches = launches;
    c->done = true;
    return TILTGPU_OK;
}

int tiltgpu_ctf_download(tiltgpu_ctf* c, int32_t z0, int32_t nsec, float* host) {
    if (!c || !host) return fail(TILTGPU_ERR_ARG, "context or data is NULL");
    if (!c->done) return fail(TILTGPU_ERR_STATE, "download before a run of the uploaded stack");
    if (z0 < 0 || nsec < 0 || z0 + nsec > c->nz)
        return fail(TILTGPU_ERR_ARG, "sections [%d, %d) outside [0, %d)", z0, z0 + nsec, c->nz);
    const size_t plane = static_cast<size_t>(c->nx) * c->ny;
    HIP_TRY(hipMemcpyAsync(host, c->out + plane * z0, plane * nsec * sizeof(float),
                           hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return TILTGPU_OK;
}

int tiltgpu_ctf_last_run(const tiltgpu_ctf* c, double* ms, int64_t* launches) {
    if (!c) return fail(TILTGPU_ERR_ARG, "context is NULL");
    if (ms) *ms = c->last_ms;
    if (launches) *launches = c->launches;
    return TILTGPU_OK;
}

void tiltgpu_ctf_destroy(tiltgpu_ctf* c) {
    if (!c) return;
    (void)hipStreamSynchronize(c->stream);
    ctf_release(c);
}

int tiltgpu_ctf_phaseflip(const float* in, float* out, int32_t nx, int32_t ny, int32_t nz,
                          const float* tilt_deg, int32_t n_tilt, const float* vdef,
                          int32_t n_vdef, const tiltgpu_ctf_params* p) {
    if (!in || !out) return fail(TILTGPU_ERR_ARG, "in or out is NULL");
    tiltgpu_ctf* c = nullptr;
    int rc = tiltgpu_ctf_create(nx, ny, nz, p, tilt_deg, n_tilt, vdef, n_vdef, &c);
    if (rc) return rc;
    if (!(rc = tiltgpu_ctf_upload(c, 0, nz, in)) && !(rc = tiltgpu_ctf_run(c)))
        rc = tiltgpu_ctf_download(c, 0, nz, out);
    tiltgpu_ctf_destroy(c);
    return rc;
}

}  // extern "C"
