/*
 * TEST INFRASTRUCTURE ONLY -- the oracle (checker and CPU baseline) for the
 * `tilt` reconstruction core.  Built into oracle/_build/liboracle.so by
 * oracle/Makefile with -O3 -ffp-contract=off (no -march=native: the reference
 * ships no .cargo/config.toml).  Never linked into the product.
 */
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------ utilities -- */

/* min_max_mean, crates/imod-math/src/stats.rs:844-861: f32 min/max, f64 sum. */
void or_min_max_mean(const float* d, long n, float* out3) {
    if (n == 0) {
        out3[0] = out3[1] = out3[2] = 0.0f;
        return;
    }
    float mn = FLT_MAX, mx = -FLT_MAX;
    double s = 0.0;
    for (long i = 0; i < n; i++) {
        float x = d[i];
        if (x < mn) mn = x;
        if (x > mx) mx = x;
        s += (double)x;
    }
    out3[0] = mn;
    out3[1] = mx;
    out3[2] = (float)(s / (double)n);
}

typedef void (*or_range_fn)(void* ctx, long i0, long i1);

typedef struct {
    or_range_fn fn;
    void* ctx;
    long n, grain;
    atomic_long next;
} or_pool_job;

static void* or_worker(void* arg) {
    or_pool_job* j = (or_pool_job*)arg;
    for (;;) {
        long i0 = atomic_fetch_add(&j->next, j->grain);
        if (i0 >= j->n) break;
        long i1 = i0 + j->grain < j->n ? i0 + j->grain : j->n;
        j->fn(j->ctx, i0, i1);
    }
    return NULL;
}

/* Dynamic parallel-for over [0, n) in pieces of `grain` (work stealing in the
 * spirit of rayon's par_iter; it never changes per-item arithmetic). */
void or_parallel_for(long n, long grain, int nthreads, or_range_fn fn, void* ctx) {
    if (n <= 0) return;
    if (nthreads <= 1 || n <= grain) {
        fn(ctx, 0, n);
        return;
    }
    or_pool_job j;
    j.fn = fn;
    j.ctx = ctx;
    j.n = n;
    j.grain = grain;
    atomic_init(&j.next, 0);
    int nt = nthreads;
    if ((long)nt > n) nt = (int)n;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nt);
    for (int t = 1; t < nt; t++) pthread_create(&th[t], NULL, or_worker, &j);
    or_worker(&j);
    for (int t = 1; t < nt; t++) pthread_join(th[t], NULL);
    free(th);
}

/* Options of one reconstruction (the `tilt` option surface, main.rs:18-120). */
typedef struct {
    int in_nx, in_ny, nz, out_nx, out_nz;
    float cutoff, falloff;
    int no_weight, exact_filter_size, hamming;
    int no_cosine_stretch, cosine_cubic;
    float max_cosine_tilt;
    int log_;
    int fill_mode;
    float densweight, outscale, outadd;
    int sirt;
    float sirt_relax;
    int fake_sirt;
    int nthreads;
} or_opts;

/* --------------------------------------------------- FP32 and FP64 forms -- */

#define REAL float
#define SUF f32
#define REAL_IS_DOUBLE 0
#include "tilt_oracle_impl.h"
#undef REAL
#undef SUF
#undef REAL_IS_DOUBLE

#define REAL double
#define SUF f64
#define REAL_IS_DOUBLE 1
#include "tilt_oracle_impl.h"
#undef REAL
#undef SUF
#undef REAL_IS_DOUBLE

/* imod-studio step_reconstruction, apps/imod-studio/src/main.rs:725-770: the
 * unweighted, unstretched backprojection of the studio's pipeline, in f32 in
 * its operation order: out_nz = in_nx; for each Y row, views pi < min(n, nz)
 * ascending, px = (ox - cx)*cos + (oz - cz)*sin + cx, a lerp of the row where
 * 0 <= p0 && p0 + 1 < in_nx (nothing otherwise), then * (1/nz).
 * out: [in_ny][in_nx][in_nx]. */
void or_studio_reconstruct(const float* P, int in_nx, int in_ny, int nz, const float* angles,
                           int n_angles, float* out) {
    const int out_nz = in_nx;
    const float cx = (float)in_nx / 2.0f, cz = (float)out_nz / 2.0f;
    const float pi_f = 3.14159265358979323846f;
    for (int iy = 0; iy < in_ny; iy++) {
        float* sl = out + (long)iy * out_nz * in_nx;
        memset(sl, 0, sizeof(float) * (size_t)out_nz * in_nx);
        for (int pi = 0; pi < n_angles && pi < nz; pi++) {
            const float r = angles[pi] * pi_f / 180.0f;
            const float cos_t = cosf(r), sin_t = sinf(r);
            const float* row = P + ((long)pi * in_ny + iy) * in_nx;
            for (int oz = 0; oz < out_nz; oz++) {
                const float bo = ((float)oz - cz) * sin_t;
                for (int ox = 0; ox < in_nx; ox++) {
                    const float px = ((float)ox - cx) * cos_t + bo + cx;
                    const float fl = floorf(px);
                    if (fl >= 0.0f && fl + 1.0f < (float)in_nx) {
                        const long p0 = (long)fl;
                        const float f = px - (float)p0;
                        sl[(long)oz * in_nx + ox] += row[p0] * (1.0f - f) + row[p0 + 1] * f;
                    }
                }
            }
        }
        const float inv = 1.0f / (float)nz;
        for (long i = 0; i < (long)out_nz * in_nx; i++) sl[i] *= inv;
    }
}
