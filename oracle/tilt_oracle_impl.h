/*
 * TEST INFRASTRUCTURE ONLY -- the oracle.  Nothing in the product library,
 * the drop-in host or the timed GPU path may include, link or call this code.
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it,
 * and only as the checker / the CPU baseline.
 *
 * A CPU restatement of the `tilt` reconstruction core of imod-rs
 * (apps/tilt/src/main.rs).  This header is included twice by tilt_oracle.c:
 *   REAL=float,  SUF=f32  -> the "faithful" FP32 form: the reference's operation
 *                            order, compiled with -ffp-contract=off so a*b+c is
 *                            never fused (Rust never contracts either);
 *   REAL=double, SUF=f64  -> the FP64 oracle: the same algorithm in double, with
 *                            geometry and filter tables also in double.
 *
 * The FFT inside the R-weighting is not rustfft (a third-party crate absent from
 * the reference tree, apps/tilt/Cargo.toml:17 "rustfft = 6", no lockfile): it is
 * a textbook mixed-radix decimation-in-time Cooley-Tukey DFT whose twiddles are
 * computed in double.  Its exact FP32 rounding is therefore "parity unpinned"
 * against rustfft; it is pinned against numpy.fft and closed forms in tests/.
 */

#define CAT_(a, b) a##_##b
#define CAT(a, b) CAT_(a, b)
#define F(name) CAT(name, SUF)

typedef struct {
    REAL re, im;
} F(cpx);

/* ------------------------------------------------------------------ FFT -- */

/* Smallest prime factor of n (n >= 2). */
static int F(spf)(int n) {
    if (n % 2 == 0) return 2;
    for (int p = 3; (long)p * p <= n; p += 2)
        if (n % p == 0) return p;
    return n;
}

/*
 * Recursive DIT DFT: out[k] = sum_j in[j*stride] * exp(sign*2*pi*i*j*k/n).
 * tw[t] = exp(sign*2*pi*i*t/N) for the top-level length N; level length n
 * divides N, so w_n^e = tw[(e*(N/n)) % N].
 */
static void F(dft_rec)(const F(cpx) * in, int stride, int n, F(cpx) * out,
                       F(cpx) * scratch, const F(cpx) * tw, int N) {
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    int p = F(spf)(n);
    int m = n / p;
    /* sub-DFTs of the p decimated sequences, stored contiguously in out */
    for (int r = 0; r < p; r++)
        F(dft_rec)(in + (long)r * stride, stride * p, m, out + (long)r * m,
                   scratch, tw, N);
    /* butterflies: X[k + q*m] = sum_r Y_r[k] * w_n^{r*(k+q*m)} */
    long step = N / n;
    for (int k = 0; k < m; k++) {
        for (int q = 0; q < p; q++) {
            int kk = k + q * m;
            REAL sr = 0, si = 0;
            for (int r = 0; r < p; r++) {
                long e = ((long)r * kk) % n;
                F(cpx) w = tw[(e * step) % N];
                F(cpx) y = out[(long)r * m + k];
                sr += y.re * w.re - y.im * w.im;
                si += y.re * w.im + y.im * w.re;
            }
            scratch[kk].re = sr;
            scratch[kk].im = si;
        }
    }
    for (int k = 0; k < n; k++) out[k] = scratch[k];
}

/* Unnormalised DFT of length n with sign -1 (forward) or +1 (inverse). */
static void F(dft)(F(cpx) * data, int n, int sign, F(cpx) * work3n) {
    F(cpx)* tw = work3n;           /* n */
    F(cpx)* res = work3n + n;      /* n */
    F(cpx)* scr = work3n + 2 * n;  /* n */
    for (int t = 0; t < n; t++) {
        double a = sign * 2.0 * M_PI * (double)t / (double)n;
        tw[t].re = (REAL)cos(a);
        tw[t].im = (REAL)sin(a);
    }
    F(dft_rec)(data, 1, n, res, scr, tw, n);
    for (int t = 0; t < n; t++) data[t] = res[t];
}

/* ------------------------------------------------------ radial weighting -- */

/*
 * Radial filter weight of one frequency bin.  Restates apps/tilt/src/main.rs:349-369:
 * freq_bin = i <= n/2 ? i : n-i;  w = freq_bin / filter_nxc;
 * Gaussian roll-off (default) or Hamming window.
 */
static REAL F(radial_weight)(int i, int n, int filter_nxc, REAL cutoff, REAL falloff,
                             int hamming) {
    int freq_bin = (i <= n / 2) ? i : n - i;
    REAL w = (REAL)freq_bin / (REAL)filter_nxc;
    if (hamming) {
        REAL h;
        if (cutoff > (REAL)1e-12 && w <= cutoff)
#if REAL_IS_DOUBLE
            h = 0.54 + 0.46 * cos(M_PI * w / cutoff);
#else
            h = 0.54f + 0.46f * cosf((float)M_PI * w / cutoff);
#endif
        else if (w > cutoff)
            h = 0;
        else
            h = 1;
        return w * h;
    }
    if (w <= cutoff) return w;
    REAL d = w - cutoff;
#if REAL_IS_DOUBLE
    return w * exp(-d * d / (2.0 * falloff * falloff));
#else
    return w * expf(-d * d / (2.0f * falloff * falloff));
#endif
}

void F(or_radial_weights)(int n, int exact_filter_size, float cutoff, float falloff,
                          int hamming, REAL* out) {
    int nxc = exact_filter_size > 0 ? exact_filter_size : n / 2 + 1;
    for (int i = 0; i < n; i++)
        out[i] = F(radial_weight)(i, n, nxc, (REAL)cutoff, (REAL)falloff, hamming);
}

typedef struct {
    REAL* data;
    long rows;
    int n;
    const REAL* wt;
} F(rw_ctx);

static void F(rw_rows)(void* vctx, long r0, long r1) {
    F(rw_ctx)* c = (F(rw_ctx)*)vctx;
    int n = c->n;
    F(cpx)* buf = (F(cpx)*)malloc(sizeof(F(cpx)) * 4 * (size_t)n);
    F(cpx)* work = buf + n;
    for (long r = r0; r < r1; r++) {
        REAL* row = c->data + r * (long)n;
        /* main.rs:344 -- real row into a complex buffer */
        for (int i = 0; i < n; i++) {
            buf[i].re = row[i];
            buf[i].im = 0;
        }
        F(dft)(buf, n, -1, work); /* main.rs:346 forward */
        for (int i = 0; i < n; i++) { /* main.rs:370-371 */
            buf[i].re *= c->wt[i];
            buf[i].im *= c->wt[i];
        }
        F(dft)(buf, n, +1, work); /* main.rs:374 inverse (unnormalised) */
        REAL inv_n = (REAL)1 / (REAL)n; /* main.rs:376-379 */
        for (int i = 0; i < n; i++) row[i] = buf[i].re * inv_n;
    }
    free(buf);
}

/* apply_rweighting, apps/tilt/src/main.rs:333-381.  data = rows x n, in place. */
void F(or_rweight)(REAL* data, long rows, int n, float cutoff, float falloff,
                   int exact_filter_size, int hamming, int nthreads) {
    REAL* wt = (REAL*)malloc(sizeof(REAL) * (size_t)n);
    F(or_radial_weights)(n, exact_filter_size, cutoff, falloff, hamming, wt);
    F(rw_ctx) c = {data, rows, n, wt};
    or_parallel_for(rows, 1, nthreads, F(rw_rows), &c);
    free(wt);
}

/* -------------------------------------------------------- cosine stretch -- */

/* cubic_interp_1d, main.rs:411-420 (Catmull-Rom) */
static REAL F(cubic_interp_1d)(REAL p0, REAL p1, REAL p2, REAL p3, REAL t) {
    REAL t2 = t * t;
    REAL t3 = t2 * t;
    REAL a = (REAL)-0.5 * p0 + (REAL)1.5 * p1 - (REAL)1.5 * p2 + (REAL)0.5 * p3;
    REAL b = p0 - (REAL)2.5 * p1 + (REAL)2.0 * p2 - (REAL)0.5 * p3;
    REAL c = (REAL)-0.5 * p0 + (REAL)0.5 * p2;
    REAL d = p1;
    return a * t3 + b * t2 + c * t + d;
}

/* sample_row_cubic, main.rs:425-441 */
static REAL F(sample_row_cubic)(const REAL* row, REAL src_x, int len) {
    long sx0 = (long)floor(src_x);
    REAL frac = src_x - (REAL)sx0;
    if (sx0 >= 1 && sx0 + 2 < len) {
        long i = sx0;
        return F(cubic_interp_1d)(row[i - 1], row[i], row[i + 1], row[i + 2], frac);
    } else if (sx0 >= 0 && sx0 + 1 < len) {
        return row[sx0] * ((REAL)1 - frac) + row[sx0 + 1] * frac;
    } else if (sx0 >= 0 && sx0 < len) {
        return row[sx0];
    }
    return 0;
}

/* Per-view stretch factor 1/cos(theta) or 0 when the view is not stretched
 * (main.rs:450-459). */
static REAL F(stretch_factor)(float deg, float max_tilt) {
    if (fabsf(deg) > max_tilt) return 0;
#if REAL_IS_DOUBLE
    double rad = (double)deg * M_PI / 180.0;
    double ct = cos(rad);
    if (fabs(ct) < 1e-6) return 0;
    return 1.0 / ct;
#else
    float rad = deg * (float)M_PI / 180.0f;
    float ct = cosf(rad);
    if (fabsf(ct) < 1e-6f) return 0;
    return 1.0f / ct;
#endif
}

/* apply_cosine_stretch, apps/tilt/src/main.rs:446-488.  `angles` must hold at
 * least nz entries (the reference indexes tilt_angles[p] for every p < nz). */
void F(or_cosine_stretch)(REAL* data, int in_nx, int in_ny, int nz, const float* angles,
                          int use_cubic, float max_tilt) {
    REAL center = (REAL)in_nx / (REAL)2;
    REAL* orig = (REAL*)malloc(sizeof(REAL) * (size_t)in_nx);
    for (int p = 0; p < nz; p++) {
        REAL stretch = F(stretch_factor)(angles[p], max_tilt);
        if (stretch == 0) continue;
        for (int iy = 0; iy < in_ny; iy++) {
            REAL* out_row = data + ((long)p * in_ny + iy) * in_nx;
            memcpy(orig, out_row, sizeof(REAL) * (size_t)in_nx);
            for (int ox = 0; ox < in_nx; ox++) {
                REAL dx = (REAL)ox - center;
                REAL src_x = dx * stretch + center;
                if (use_cubic) {
                    out_row[ox] = F(sample_row_cubic)(orig, src_x, in_nx);
                } else {
                    long sx0 = (long)floor(src_x);
                    if (sx0 >= 0 && sx0 + 1 < in_nx) {
                        REAL frac = src_x - (REAL)sx0;
                        out_row[ox] = orig[sx0] * ((REAL)1 - frac) + orig[sx0 + 1] * frac;
                    } else if (sx0 >= 0 && sx0 < in_nx) {
                        out_row[ox] = orig[sx0];
                    } else {
                        out_row[ox] = 0;
                    }
                }
            }
        }
    }
    free(orig);
}

/* -------------------------------------------------------------- geometry -- */

/*
 * ProjParams table, main.rs:683-700, for the first n = min(nz, n_angles) views.
 * xtilt / zfactor may be NULL (absent file); missing entries default to 0 deg / 1.0
 * (main.rs:689-691).  Output: 5 arrays of n REALs.
 */
void F(or_proj_params)(const float* angles, int n, const float* xtilt, int n_xtilt,
                       const float* zfac, int n_zfac, REAL* cos_t, REAL* sin_t,
                       REAL* cos_xt, REAL* sin_xt, REAL* zf) {
    for (int i = 0; i < n; i++) {
        float deg = angles[i];
        float xt_deg = (xtilt && i < n_xtilt) ? xtilt[i] : 0.0f;
        float z = (zfac && i < n_zfac) ? zfac[i] : 1.0f;
#if REAL_IS_DOUBLE
        double rad = (double)deg * M_PI / 180.0;
        double xrad = (double)xt_deg * M_PI / 180.0;
        cos_t[i] = cos(rad);
        sin_t[i] = sin(rad);
        cos_xt[i] = cos(xrad);
        sin_xt[i] = sin(xrad);
#else
        float rad = deg * (float)M_PI / 180.0f;
        float xrad = xt_deg * (float)M_PI / 180.0f;
        cos_t[i] = cosf(rad);
        sin_t[i] = sinf(rad);
        cos_xt[i] = cosf(xrad);
        sin_xt[i] = sinf(xrad);
#endif
        zf[i] = (REAL)z;
    }
}

/* ------------------------------------------------------- local alignment -- */

/* ViewPatches::find_bracket, main.rs:200-228 */
static void F(find_bracket)(const float* pos, int len, REAL val, int* idx, REAL* frac) {
    if (len == 1) {
        *idx = 0;
        *frac = 0;
        return;
    }
    int lo = 0, hi = len - 1;
    if (val <= (REAL)pos[lo]) {
        *idx = 0;
        *frac = 0;
        return;
    }
    if (val >= (REAL)pos[hi]) {
        *idx = hi;
        *frac = 0;
        return;
    }
    while (hi - lo > 1) {
        int mid = (lo + hi) / 2;
        if ((REAL)pos[mid] <= val)
            lo = mid;
        else
            hi = mid;
    }
    REAL span = (REAL)pos[hi] - (REAL)pos[lo];
    *idx = lo;
    *frac = span > (REAL)1e-12 ? (val - (REAL)pos[lo]) / span : 0;
}

/* ViewPatches::interpolate, main.rs:172-198.  nx, ny = grid size of the view. */
static void F(local_interp)(const float* xs, int nx, const float* ys, int ny,
                            const float* gdx, const float* gdy, REAL x, REAL y, REAL* odx,
                            REAL* ody) {
    if (nx == 0 || ny == 0) {
        *odx = 0;
        *ody = 0;
        return;
    }
    int ix0, iy0;
    REAL fx, fy;
    F(find_bracket)(xs, nx, x, &ix0, &fx);
    F(find_bracket)(ys, ny, y, &iy0, &fy);
    int ix1 = ix0 + 1 < nx - 1 ? ix0 + 1 : nx - 1;
    int iy1 = iy0 + 1 < ny - 1 ? iy0 + 1 : ny - 1;
    REAL d00x = gdx[iy0 * nx + ix0], d00y = gdy[iy0 * nx + ix0];
    REAL d10x = gdx[iy0 * nx + ix1], d10y = gdy[iy0 * nx + ix1];
    REAL d01x = gdx[iy1 * nx + ix0], d01y = gdy[iy1 * nx + ix0];
    REAL d11x = gdx[iy1 * nx + ix1], d11y = gdy[iy1 * nx + ix1];
#define LERP(a, b, t) ((a) + (t) * ((b) - (a)))
    *odx = LERP(LERP(d00x, d10x, fx), LERP(d01x, d11x, fx), fy);
    *ody = LERP(LERP(d00y, d10y, fx), LERP(d01y, d11y, fx), fy);
#undef LERP
}

/* ------------------------------------------------------------ projection -- */

/* The full geometry of one reconstruction, shared by BP and FP. */
typedef struct {
    int in_nx, in_ny, nz;   /* projection width, height, sections (views in file) */
    int nviews;             /* views iterated = min(nz, n_angles), main.rs:683-686 */
    int out_nx, out_nz;     /* output width, thickness */
    int fill_mode;          /* 0 mean, 1 zero, 2 edge */
    const REAL *cos_t, *sin_t, *cos_xt, *sin_xt, *zf; /* [nviews] */
    const REAL* means;      /* [nz] per-view means (fill value for Mean) */
    /* local alignment: per view v, grid nxs[v] x nys[v] at offsets */
    int has_local;
    const int *lnx, *lny;
    const long *lxoff, *lyoff, *lgoff;
    const float *lxs, *lys, *ldx, *ldy;
} F(or_geom);

/* sample_row_with_fill, main.rs:386-406 */
static inline REAL F(sample_fill)(const REAL* row, REAL proj_x, int row_len, int fill,
                                  REAL mean) {
    long px0 = (long)floor(proj_x);
    if (px0 >= 0 && px0 + 1 < row_len) {
        REAL frac = proj_x - (REAL)px0;
        return row[px0] * ((REAL)1 - frac) + row[px0 + 1] * frac;
    }
    if (fill == 1) return 0;
    if (fill == 0) return mean;
    if (px0 < 0) return row[0];
    return row_len > 0 ? row[row_len - 1] : 0;
}

/* Geometry of one voxel-view pair: proj_x, sample_y and the two-row flag.
 * main.rs:1015-1051 (BP) and :837-876 (FP) compute it identically. */
static inline void F(voxel_coords)(const F(or_geom) * g, int pi, int iy, int oz, int ox,
                                   REAL* proj_x, REAL* sample_y, int* two) {
    REAL center_x = (REAL)g->in_nx / (REAL)2;
    REAL center_z = (REAL)g->out_nz / (REAL)2;
    REAL out_center_x = (REAL)g->out_nx / (REAL)2;
    REAL dz = (REAL)oz - center_z;
    REAL z_contrib = dz * g->sin_t[pi] * g->zf[pi];
    REAL dx = (REAL)ox - out_center_x;
    REAL px = dx * g->cos_t[pi] * g->cos_xt[pi] + z_contrib + center_x;
    REAL y_off = dx * g->sin_xt[pi];
    REAL sy = (REAL)iy + y_off;
    if (g->has_local) {
        int nx = g->lnx[pi], ny = g->lny[pi];
        if (nx > 0) { /* "!vp.xs.is_empty()" */
            REAL ldx, ldy;
            F(local_interp)(g->lxs + g->lxoff[pi], nx, g->lys + g->lyoff[pi], ny,
                            g->ldx + g->lgoff[pi], g->ldy + g->lgoff[pi], (REAL)ox,
                            (REAL)iy, &ldx, &ldy);
            px += ldx;
            sy = sy + ldy;
        }
    }
    *proj_x = px;
    *sample_y = sy;
    *two = (fabs(y_off) > (REAL)1e-6) || g->has_local;
}

/* One output Y slice: the hot loop of main.rs:1011-1073 (== backproject_all,
 * main.rs:759-826).  slice = [out_nz][out_nx], multiplied by inv_n = 1/nz. */
static void F(bp_slice)(const F(or_geom) * g, const REAL* P, int iy, REAL* slice) {
    int in_nx = g->in_nx, in_ny = g->in_ny;
    long nxy = (long)in_nx * in_ny;
    memset(slice, 0, sizeof(REAL) * (size_t)g->out_nx * g->out_nz);
    for (int pi = 0; pi < g->nviews; pi++) {
        REAL pm = g->means[pi];
        for (int oz = 0; oz < g->out_nz; oz++) {
            REAL* srow = slice + (long)oz * g->out_nx;
            for (int ox = 0; ox < g->out_nx; ox++) {
                REAL px, sy;
                int two;
                F(voxel_coords)(g, pi, iy, oz, ox, &px, &sy, &two);
                long sy0 = (long)floor(sy);
                REAL fy = sy - (REAL)sy0;
                if (sy0 >= 0 && sy0 + 1 < in_ny && two) {
                    const REAL* r0 = P + pi * nxy + sy0 * in_nx;
                    const REAL* r1 = r0 + in_nx;
                    REAL v0 = F(sample_fill)(r0, px, in_nx, g->fill_mode, pm);
                    REAL v1 = F(sample_fill)(r1, px, in_nx, g->fill_mode, pm);
                    srow[ox] += v0 * ((REAL)1 - fy) + v1 * fy;
                } else if (sy0 >= 0 && sy0 < in_ny) {
                    const REAL* r0 = P + pi * nxy + sy0 * in_nx;
                    srow[ox] += F(sample_fill)(r0, px, in_nx, g->fill_mode, pm);
                }
            }
        }
    }
    REAL inv_n = (REAL)1 / (REAL)g->nz; /* main.rs:705 -- 1/nz even with fewer angles */
    long nv = (long)g->out_nx * g->out_nz;
    for (long i = 0; i < nv; i++) slice[i] *= inv_n;
}

typedef struct {
    const F(or_geom) * g;
    const REAL* P;
    int y0;
    REAL* out;
} F(bp_ctx);

static void F(bp_rows)(void* vctx, long r0, long r1) {
    F(bp_ctx)* c = (F(bp_ctx)*)vctx;
    long nv = (long)c->g->out_nx * c->g->out_nz;
    for (long r = r0; r < r1; r++) F(bp_slice)(c->g, c->P, c->y0 + (int)r, c->out + r * nv);
}

/* Backproject Y rows [y0, y1) into out[(iy-y0)][oz][ox]; rows are processed in
 * chunks of max(32, out_ny/20) and each chunk is split across threads, as in
 * main.rs:1004-1010 (rayon par_iter over the rows of a chunk). */
void F(or_backproject)(const F(or_geom) * g, const REAL* P, int y0, int y1, REAL* out,
                       int nthreads) {
    int out_ny = g->in_ny;
    int chunk = out_ny / 20 > 32 ? out_ny / 20 : 32;
    long nv = (long)g->out_nx * g->out_nz;
    for (int c0 = y0; c0 < y1; c0 += chunk) {
        int c1 = c0 + chunk < y1 ? c0 + chunk : y1;
        F(bp_ctx) c = {g, P, c0, out + (long)(c0 - y0) * nv};
        or_parallel_for(c1 - c0, 1, nthreads, F(bp_rows), &c);
    }
}

/* forward_project, main.rs:831-896: the adjoint scatter of the BP interpolation,
 * serial, loop order iy -> view -> oz -> ox.  V = [in_ny][out_nz][out_nx],
 * R = [nz][in_ny][in_nx] (zeroed here). */
typedef struct {
    const F(or_geom) * g;
    const REAL* V;
    REAL* R;
} F(fp_ctx);

/* Views [v0, v1) of forward_project.  A view writes only its own sinogram
 * section, and within a view the loop order is the reference's (iy, oz, ox), so
 * splitting the views over threads leaves every sum bitwise unchanged. */
static void F(fp_views)(void* vctx, long v0, long v1) {
    F(fp_ctx)* c = (F(fp_ctx)*)vctx;
    const F(or_geom)* g = c->g;
    int in_nx = g->in_nx, in_ny = g->in_ny;
    long nxy = (long)in_nx * in_ny;
    for (int iy = 0; iy < in_ny; iy++) {
        const REAL* slice = c->V + (long)iy * g->out_nz * g->out_nx;
        for (int pi = (int)v0; pi < (int)v1; pi++) {
            for (int oz = 0; oz < g->out_nz; oz++) {
                for (int ox = 0; ox < g->out_nx; ox++) {
                    REAL px, sy;
                    int two;
                    F(voxel_coords)(g, pi, iy, oz, ox, &px, &sy, &two);
                    long px0 = (long)floor(px);
                    if (!(px0 >= 0 && px0 + 1 < in_nx)) continue;
                    REAL frac = px - (REAL)px0;
                    REAL val = slice[(long)oz * g->out_nx + ox];
                    long sy0 = (long)floor(sy);
                    REAL fy = sy - (REAL)sy0;
                    if (sy0 >= 0 && sy0 + 1 < in_ny && two) {
                        REAL* r0 = c->R + pi * nxy + sy0 * in_nx;
                        REAL* r1 = r0 + in_nx;
                        REAL w0 = (REAL)1 - fy, w1 = fy;
                        r0[px0] += val * ((REAL)1 - frac) * w0;
                        r0[px0 + 1] += val * frac * w0;
                        r1[px0] += val * ((REAL)1 - frac) * w1;
                        r1[px0 + 1] += val * frac * w1;
                    } else if (sy0 >= 0 && sy0 < in_ny) {
                        REAL* r0 = c->R + pi * nxy + sy0 * in_nx;
                        r0[px0] += val * ((REAL)1 - frac);
                        r0[px0 + 1] += val * frac;
                    }
                }
            }
        }
    }
}

/* forward_project, main.rs:830-896 (loop order iy, view, oz, ox). */
void F(or_forward_project)(const F(or_geom) * g, const REAL* V, REAL* R, int nthreads) {
    memset(R, 0, sizeof(REAL) * (size_t)g->nz * g->in_nx * g->in_ny);
    F(fp_ctx) c = {g, V, R};
    or_parallel_for(g->nviews, 1, nthreads, F(fp_views), &c);
}

/* lowpass_3d_volume, main.rs:493-538: one separable 3x3x3 box pass (X, Z, Y)
 * with clamped edges.  vol = [ny][nz][nx], in place. */
void F(or_lowpass_3d)(REAL* vol, int ny, int nz, int nx) {
    if (ny == 0) return;
    long ns = (long)nz * nx;
    REAL* tmp = (REAL*)malloc(sizeof(REAL) * (size_t)ns);
    for (int iy = 0; iy < ny; iy++) {
        REAL* s = vol + iy * ns;
        for (int oz = 0; oz < nz; oz++)
            for (int ox = 0; ox < nx; ox++) {
                int x0 = ox > 0 ? ox - 1 : 0, x2 = ox + 1 < nx ? ox + 1 : nx - 1;
                long o = (long)oz * nx;
                tmp[o + ox] = (s[o + x0] + s[o + ox] + s[o + x2]) / (REAL)3;
            }
        memcpy(s, tmp, sizeof(REAL) * (size_t)ns);
    }
    for (int iy = 0; iy < ny; iy++) {
        REAL* s = vol + iy * ns;
        for (int oz = 0; oz < nz; oz++) {
            int z0 = oz > 0 ? oz - 1 : 0, z2 = oz + 1 < nz ? oz + 1 : nz - 1;
            for (int ox = 0; ox < nx; ox++)
                tmp[(long)oz * nx + ox] =
                    (s[(long)z0 * nx + ox] + s[(long)oz * nx + ox] + s[(long)z2 * nx + ox]) /
                    (REAL)3;
        }
        memcpy(s, tmp, sizeof(REAL) * (size_t)ns);
    }
    free(tmp);
    REAL* prev = (REAL*)malloc(sizeof(REAL) * (size_t)ny * ns);
    memcpy(prev, vol, sizeof(REAL) * (size_t)ny * ns);
    for (int iy = 0; iy < ny; iy++) {
        int y0 = iy > 0 ? iy - 1 : 0, y2 = iy + 1 < ny ? iy + 1 : ny - 1;
        for (long i = 0; i < ns; i++)
            vol[iy * ns + i] = (prev[y0 * ns + i] + prev[iy * ns + i] + prev[y2 * ns + i]) / (REAL)3;
    }
    free(prev);
}

/* The geometry of main.rs:682-705 (ProjParams for the first nviews views) plus
 * the --local-file grids (main.rs:232-307, as flattened per-view arrays; lnx
 * NULL = no file), filled into g.  tab: 5*nviews REALs; loff: 3*(nz+1) longs
 * (the per-view offsets of the flattened grids) when lnx is given.  g->means
 * is left NULL for the caller. */
static void F(geom_init)(F(or_geom) * g, int in_nx, int in_ny, int nz, int out_nx, int out_nz,
                         int fill_mode, const float* angles, int nviews, const float* xtilt,
                         int n_xtilt, const float* zfac, int n_zfac, const int* lnx,
                         const int* lny, const float* lxs, const float* lys, const float* ldx,
                         const float* ldy, REAL* tab, long* loff) {
    memset(g, 0, sizeof(*g));
    F(or_proj_params)(angles, nviews, xtilt, n_xtilt, zfac, n_zfac, tab, tab + nviews,
                      tab + 2 * nviews, tab + 3 * nviews, tab + 4 * nviews);
    g->in_nx = in_nx;
    g->in_ny = in_ny;
    g->nz = nz;
    g->nviews = nviews;
    g->out_nx = out_nx;
    g->out_nz = out_nz;
    g->fill_mode = fill_mode;
    g->cos_t = tab;
    g->sin_t = tab + nviews;
    g->cos_xt = tab + 2 * nviews;
    g->sin_xt = tab + 3 * nviews;
    g->zf = tab + 4 * nviews;
    if (lnx) {
        long *lxoff = loff, *lyoff = loff + nz + 1, *lgoff = loff + 2 * (nz + 1);
        lxoff[0] = lyoff[0] = lgoff[0] = 0;
        for (int v = 0; v < nz; v++) {
            lxoff[v + 1] = lxoff[v] + lnx[v];
            lyoff[v + 1] = lyoff[v] + lny[v];
            lgoff[v + 1] = lgoff[v] + (long)lnx[v] * lny[v];
        }
        g->has_local = 1;
        g->lnx = lnx;
        g->lny = lny;
        g->lxoff = lxoff;
        g->lyoff = lyoff;
        g->lgoff = lgoff;
        g->lxs = lxs;
        g->lys = lys;
        g->ldx = ldx;
        g->ldy = ldy;
    }
}

/* forward_project (main.rs:830-896) of a caller-given volume V [in_ny][out_nz]
 * [out_nx] with the full geometry: angles, x-axis tilts, z factors and local
 * grids (any may be absent) into R [nz][in_ny][in_nx]. */
void F(or_forward_project_full)(int in_nx, int in_ny, int nz, int out_nx, int out_nz,
                                const float* angles, int n_angles, const float* xtilt,
                                int n_xtilt, const float* zfac, int n_zfac, const int* lnx,
                                const int* lny, const float* lxs, const float* lys,
                                const float* ldx, const float* ldy, const float* V_in,
                                float* R_out, int nthreads) {
    int nviews = n_angles < nz ? n_angles : nz;
    long nvol = (long)in_ny * out_nx * out_nz;
    long total = (long)in_nx * in_ny * nz;
    REAL* tab = (REAL*)malloc(sizeof(REAL) * 5 * (size_t)(nviews > 0 ? nviews : 1));
    long* loff = lnx ? (long*)calloc(3 * ((size_t)nz + 1), sizeof(long)) : NULL;
    REAL zero = 0;
    F(or_geom) g;
    F(geom_init)(&g, in_nx, in_ny, nz, out_nx, out_nz, 1, angles, nviews, xtilt, n_xtilt,
                 zfac, n_zfac, lnx, lny, lxs, lys, ldx, ldy, tab, loff);
    g.means = &zero;
    REAL* V = (REAL*)malloc(sizeof(REAL) * (size_t)(nvol ? nvol : 1));
    for (long i = 0; i < nvol; i++) V[i] = (REAL)V_in[i];
    REAL* R = (REAL*)malloc(sizeof(REAL) * (size_t)(total ? total : 1));
    F(or_forward_project)(&g, V, R, nthreads);
    for (long i = 0; i < total; i++) R_out[i] = (float)R[i];
    free(R);
    free(V);
    free(loff);
    free(tab);
}

/* ------------------------------------------------------------- pipeline -- */

/*
 * The whole reconstruction of main.rs:601-1093 after file reading:
 *   log (:607-613), per-view means (:615-623), SIRT original copy (:625-630),
 *   R-weighting (:638-646), cosine stretch (:648-653), ProjParams (:682-700),
 *   then WBP (:1002-1090), SIRT (:919-975) or fake-SIRT (:976-1001),
 *   output transform and statistics (:1077-1093).
 * P is [nz][in_ny][in_nx] float (raw, modified in place).  vol receives
 * [in_ny][out_nz][out_nx] floats after the output transform.  stats[3] =
 * header amin, amax, amean.
 */

void F(or_tilt_run_rows)(const or_opts* o, float* P_in, const float* angles, int n_angles,
                         const float* xtilt, int n_xtilt, const float* zfac, int n_zfac,
                         const int* lnx, const int* lny, const float* lxs, const float* lys,
                         const float* ldx, const float* ldy, int y0, int y1, float* vol,
                         float* stats, const double* means_in);

void F(or_tilt_run)(const or_opts* o, float* P_in, const float* angles, int n_angles,
                    const float* xtilt, int n_xtilt, const float* zfac, int n_zfac,
                    const int* lnx, const int* lny, const float* lxs, const float* lys,
                    const float* ldx, const float* ldy, float* vol, float* stats) {
    F(or_tilt_run_rows)(o, P_in, angles, n_angles, xtilt, n_xtilt, zfac, n_zfac, lnx, lny,
                        lxs, lys, ldx, ldy, 0, o->in_ny, vol, stats, NULL);
}

/* Per-view means of the (logged) projections, main.rs:615-623 (min_max_mean's
 * f64 sum, rounded to f32 in the FP32 oracle). */
static REAL* F(view_means)(const or_opts* o, const float* P_in) {
    long nxy = (long)o->in_nx * o->in_ny;
    REAL* means = (REAL*)malloc(sizeof(REAL) * (size_t)(o->nz > 0 ? o->nz : 1));
    for (int p = 0; p < o->nz; p++) {
        float mmm[3];
        or_min_max_mean(P_in + p * nxy, nxy, mmm);
#if REAL_IS_DOUBLE
        double s = 0;
        for (long i = 0; i < nxy; i++) s += P_in[p * nxy + i];
        means[p] = nxy ? s / (double)nxy : 0;
#else
        means[p] = mmm[2];
#endif
    }
    return means;
}

/* The same for output rows [y0, y1) only (WBP: rows are independent); SIRT and
 * fake-SIRT always reconstruct every row first.  vol = [y1-y0][out_nz][out_nx]. */
void F(or_tilt_run_rows)(const or_opts* o, float* P_in, const float* angles, int n_angles,
                         const float* xtilt, int n_xtilt, const float* zfac, int n_zfac,
                         const int* lnx, const int* lny, const float* lxs, const float* lys,
                         const float* ldx, const float* ldy, int y0, int y1, float* vol,
                         float* stats, const double* means_in) {
    long nxy = (long)o->in_nx * o->in_ny;
    long total = nxy * o->nz;
    if (o->log_)
        for (long i = 0; i < total; i++) {
            float v = P_in[i] > 0.001f ? P_in[i] : 0.001f;
            P_in[i] = log10f(v);
        }
    /* means_in: the per-view means of the whole images when P_in holds only a
     * band of their rows (a row-subset run); NULL computes them from P_in */
    REAL* means = F(view_means)(o, P_in);
    if (means_in)
        for (int p = 0; p < o->nz; p++) means[p] = (REAL)means_in[p];
    REAL* P = (REAL*)malloc(sizeof(REAL) * (size_t)total);
    for (long i = 0; i < total; i++) P[i] = (REAL)P_in[i];
    REAL* O = NULL;
    if (o->sirt > 0) {
        O = (REAL*)malloc(sizeof(REAL) * (size_t)total);
        memcpy(O, P, sizeof(REAL) * (size_t)total);
    }
    if (!o->no_weight)
        F(or_rweight)(P, (long)o->nz * o->in_ny, o->in_nx, o->cutoff, o->falloff,
                      o->exact_filter_size, o->hamming, o->nthreads);
    if (!o->no_cosine_stretch)
        F(or_cosine_stretch)(P, o->in_nx, o->in_ny, o->nz, angles, o->cosine_cubic,
                             o->max_cosine_tilt);
    int nviews = n_angles < o->nz ? n_angles : o->nz;
    REAL* tab = (REAL*)malloc(sizeof(REAL) * 5 * (size_t)(nviews > 0 ? nviews : 1));
    long* loff = lnx ? (long*)calloc(3 * ((size_t)o->nz + 1), sizeof(long)) : NULL;
    F(or_geom) g;
    F(geom_init)(&g, o->in_nx, o->in_ny, o->nz, o->out_nx, o->out_nz, o->fill_mode, angles,
                 nviews, xtilt, n_xtilt, zfac, n_zfac, lnx, lny, lxs, lys, ldx, ldy, tab, loff);
    g.means = means;
    int out_ny = o->in_ny;
    long nv = (long)o->out_nx * o->out_nz;
    int whole = o->sirt > 0 || o->fake_sirt > 0;
    int ry0 = whole ? 0 : y0, ry1 = whole ? out_ny : y1;
    REAL* V = (REAL*)malloc(sizeof(REAL) * (size_t)((ry1 - ry0) * nv > 0 ? (ry1 - ry0) * nv : 1));
    F(or_backproject)(&g, P, ry0, ry1, V, o->nthreads);
    if (o->sirt > 0) {
        REAL* R = (REAL*)malloc(sizeof(REAL) * (size_t)total);
        REAL* D = (REAL*)malloc(sizeof(REAL) * (size_t)(out_ny * nv > 0 ? out_ny * nv : 1));
        for (int it = 0; it < o->sirt; it++) {
            F(or_forward_project)(&g, V, R, o->nthreads);
            for (long i = 0; i < total; i++) R[i] = O[i] - R[i];
            if (!o->no_weight)
                F(or_rweight)(R, (long)o->nz * o->in_ny, o->in_nx, o->cutoff, o->falloff,
                              o->exact_filter_size, o->hamming, o->nthreads);
            F(or_backproject)(&g, R, 0, out_ny, D, o->nthreads);
            REAL relax = (REAL)o->sirt_relax;
            for (long i = 0; i < out_ny * nv; i++) V[i] += relax * D[i];
        }
        free(R);
        free(D);
    }
    for (int k = 0; k < o->fake_sirt; k++) F(or_lowpass_3d)(V, out_ny, o->out_nz, o->out_nx);
    /* output transform + statistics, main.rs:1077-1093 */
    float gmin = FLT_MAX, gmax = -FLT_MAX;
    double gsum = 0;
    for (int iy = y0; iy < y1; iy++) {
        float* dst = vol + (long)(iy - y0) * nv;
        const REAL* src = V + (long)(iy - ry0) * nv;
        for (long i = 0; i < nv; i++) {
            REAL v = (src[i] * (REAL)o->densweight) * (REAL)o->outscale + (REAL)o->outadd;
            dst[i] = (float)v;
        }
        float mmm[3];
        or_min_max_mean(dst, nv, mmm);
        if (mmm[0] < gmin) gmin = mmm[0];
        if (mmm[1] > gmax) gmax = mmm[1];
        gsum += (double)mmm[2] * (double)nv;
    }
    stats[0] = gmin;
    stats[1] = gmax;
    stats[2] = (y1 > y0) ? (float)(gsum / (double)(nv * (y1 - y0))) : 0.0f;
    free(V);
    free(tab);
    free(means);
    free(P);
    free(O);
    free(loff);
}

/* SIRT (main.rs:919-953) on output rows [y0, y1) only, keeping the volume after
 * the initial WBP and after every iteration: trace = [sirt+1][y1-y0][out_nz]
 * [out_nx] (before the output transform), and resid[it] = the L2 norm of
 * O - FP(V) over those rows before iteration it's R-weighting.
 *
 * Only for the single-row geometry (no x-axis tilt, no local alignment): there
 * row iy of the volume projects into row iy of every view and back, the
 * R-weighting and the cosine stretch act along rows, so the rows evolve
 * independently and a row subset equals those rows of the whole run.  The
 * per-view means (fill values) come from the full images, as in the whole run.
 * P_in is the full [nz][in_ny][in_nx] series (logged in place with --log). */
void F(or_sirt_trace)(const or_opts* o, float* P_in, const float* angles, int n_angles,
                      const float* zfac, int n_zfac, int y0, int y1, float* trace,
                      double* resid) {
    long nxy = (long)o->in_nx * o->in_ny;
    long total = nxy * o->nz;
    if (o->log_)
        for (long i = 0; i < total; i++) {
            float v = P_in[i] > 0.001f ? P_in[i] : 0.001f;
            P_in[i] = log10f(v);
        }
    REAL* means = F(view_means)(o, P_in);
    const int rows = y1 - y0;
    const long sub_xy = (long)o->in_nx * rows;
    const long sub = sub_xy * o->nz;
    REAL* P = (REAL*)malloc(sizeof(REAL) * (size_t)(sub > 0 ? sub : 1));
    for (int p = 0; p < o->nz; p++)
        for (long i = 0; i < sub_xy; i++) P[p * sub_xy + i] = (REAL)P_in[p * nxy + y0 * (long)o->in_nx + i];
    REAL* O = (REAL*)malloc(sizeof(REAL) * (size_t)(sub > 0 ? sub : 1));
    memcpy(O, P, sizeof(REAL) * (size_t)sub);
    if (!o->no_weight)
        F(or_rweight)(P, (long)o->nz * rows, o->in_nx, o->cutoff, o->falloff,
                      o->exact_filter_size, o->hamming, o->nthreads);
    if (!o->no_cosine_stretch)
        F(or_cosine_stretch)(P, o->in_nx, rows, o->nz, angles, o->cosine_cubic,
                             o->max_cosine_tilt);
    int nviews = n_angles < o->nz ? n_angles : o->nz;
    REAL* tab = (REAL*)malloc(sizeof(REAL) * 5 * (size_t)(nviews > 0 ? nviews : 1));
    F(or_proj_params)(angles, nviews, NULL, 0, zfac, n_zfac, tab, tab + nviews,
                      tab + 2 * nviews, tab + 3 * nviews, tab + 4 * nviews);
    F(or_geom) g;
    memset(&g, 0, sizeof(g));
    g.in_nx = o->in_nx;
    g.in_ny = rows;
    g.nz = o->nz;
    g.nviews = nviews;
    g.out_nx = o->out_nx;
    g.out_nz = o->out_nz;
    g.fill_mode = o->fill_mode;
    g.cos_t = tab;
    g.sin_t = tab + nviews;
    g.cos_xt = tab + 2 * nviews;
    g.sin_xt = tab + 3 * nviews;
    g.zf = tab + 4 * nviews;
    g.means = means;
    const long nv = (long)o->out_nx * o->out_nz;
    const long nvol = nv * rows;
    REAL* V = (REAL*)malloc(sizeof(REAL) * (size_t)(nvol > 0 ? nvol : 1));
    REAL* R = (REAL*)malloc(sizeof(REAL) * (size_t)(sub > 0 ? sub : 1));
    REAL* D = (REAL*)malloc(sizeof(REAL) * (size_t)(nvol > 0 ? nvol : 1));
    F(or_backproject)(&g, P, 0, rows, V, o->nthreads);
    for (long i = 0; i < nvol; i++) trace[i] = (float)V[i];
    for (int it = 0; it < o->sirt; it++) {
        F(or_forward_project)(&g, V, R, o->nthreads);
        double ss = 0;
        for (long i = 0; i < sub; i++) {
            R[i] = O[i] - R[i];
            ss += (double)R[i] * (double)R[i];
        }
        if (resid) resid[it] = sqrt(ss);
        if (!o->no_weight)
            F(or_rweight)(R, (long)o->nz * rows, o->in_nx, o->cutoff, o->falloff,
                          o->exact_filter_size, o->hamming, o->nthreads);
        F(or_backproject)(&g, R, 0, rows, D, o->nthreads);
        REAL relax = (REAL)o->sirt_relax;
        for (long i = 0; i < nvol; i++) V[i] += relax * D[i];
        float* dst = trace + (long)(it + 1) * nvol;
        for (long i = 0; i < nvol; i++) dst[i] = (float)V[i];
    }
    free(D);
    free(R);
    free(V);
    free(tab);
    free(O);
    free(P);
    free(means);
}

/* Thin wrappers so tests can drive BP / FP on caller-given float arrays. */
void F(or_backproject_simple)(int in_nx, int in_ny, int nz, int nviews, int out_nx,
                              int out_nz, int fill_mode, const float* ptab /* 5*nviews */,
                              const float* means, const float* P_in, int y0, int y1,
                              float* out, int nthreads) {
    long total = (long)in_nx * in_ny * nz;
    REAL* P = (REAL*)malloc(sizeof(REAL) * (size_t)total);
    for (long i = 0; i < total; i++) P[i] = (REAL)P_in[i];
    REAL* tab = (REAL*)malloc(sizeof(REAL) * 5 * (size_t)(nviews ? nviews : 1));
    for (int i = 0; i < 5 * nviews; i++) tab[i] = (REAL)ptab[i];
    REAL* m = (REAL*)malloc(sizeof(REAL) * (size_t)(nz ? nz : 1));
    for (int i = 0; i < nz; i++) m[i] = (REAL)means[i];
    F(or_geom) g;
    memset(&g, 0, sizeof(g));
    g.in_nx = in_nx; g.in_ny = in_ny; g.nz = nz; g.nviews = nviews;
    g.out_nx = out_nx; g.out_nz = out_nz; g.fill_mode = fill_mode;
    g.cos_t = tab; g.sin_t = tab + nviews; g.cos_xt = tab + 2 * nviews;
    g.sin_xt = tab + 3 * nviews; g.zf = tab + 4 * nviews; g.means = m;
    long nv = (long)out_nx * out_nz;
    REAL* V = (REAL*)malloc(sizeof(REAL) * (size_t)((y1 - y0) * nv > 0 ? (y1 - y0) * nv : 1));
    F(or_backproject)(&g, P, y0, y1, V, nthreads);
    for (long i = 0; i < (y1 - y0) * nv; i++) out[i] = (float)V[i];
    free(V); free(m); free(tab); free(P);
}

void F(or_forward_project_simple)(int in_nx, int in_ny, int nz, int nviews, int out_nx,
                                  int out_nz, const float* ptab, const float* V_in,
                                  float* R_out, int nthreads) {
    long nvol = (long)in_ny * out_nx * out_nz;
    long total = (long)in_nx * in_ny * nz;
    REAL* V = (REAL*)malloc(sizeof(REAL) * (size_t)(nvol ? nvol : 1));
    for (long i = 0; i < nvol; i++) V[i] = (REAL)V_in[i];
    REAL* R = (REAL*)malloc(sizeof(REAL) * (size_t)(total ? total : 1));
    REAL* tab = (REAL*)malloc(sizeof(REAL) * 5 * (size_t)(nviews ? nviews : 1));
    for (int i = 0; i < 5 * nviews; i++) tab[i] = (REAL)ptab[i];
    REAL zero = 0;
    F(or_geom) g;
    memset(&g, 0, sizeof(g));
    g.in_nx = in_nx; g.in_ny = in_ny; g.nz = nz; g.nviews = nviews;
    g.out_nx = out_nx; g.out_nz = out_nz; g.fill_mode = 1;
    g.cos_t = tab; g.sin_t = tab + nviews; g.cos_xt = tab + 2 * nviews;
    g.sin_xt = tab + 3 * nviews; g.zf = tab + 4 * nviews; g.means = &zero;
    F(or_forward_project)(&g, V, R, nthreads);
    for (long i = 0; i < total; i++) R_out[i] = (float)R[i];
    free(tab); free(R); free(V);
}

/* R-weighting / stretch on float arrays for tests (computed in REAL). */
void F(or_rweight_float)(float* data, long rows, int n, float cutoff, float falloff,
                         int exact_filter_size, int hamming, int nthreads) {
    long total = rows * (long)n;
    REAL* d = (REAL*)malloc(sizeof(REAL) * (size_t)(total ? total : 1));
    for (long i = 0; i < total; i++) d[i] = (REAL)data[i];
    F(or_rweight)(d, rows, n, cutoff, falloff, exact_filter_size, hamming, nthreads);
    for (long i = 0; i < total; i++) data[i] = (float)d[i];
    free(d);
}

void F(or_cosine_stretch_float)(float* data, int in_nx, int in_ny, int nz,
                                const float* angles, int use_cubic, float max_tilt) {
    long total = (long)in_nx * in_ny * nz;
    REAL* d = (REAL*)malloc(sizeof(REAL) * (size_t)(total ? total : 1));
    for (long i = 0; i < total; i++) d[i] = (REAL)data[i];
    F(or_cosine_stretch)(d, in_nx, in_ny, nz, angles, use_cubic, max_tilt);
    for (long i = 0; i < total; i++) data[i] = (float)d[i];
    free(d);
}

void F(or_dft_float)(float* re, float* im, int n, int sign) {
    F(cpx)* b = (F(cpx)*)malloc(sizeof(F(cpx)) * 4 * (size_t)n);
    for (int i = 0; i < n; i++) { b[i].re = (REAL)re[i]; b[i].im = (REAL)im[i]; }
    F(dft)(b, n, sign, b + n);
    for (int i = 0; i < n; i++) { re[i] = (float)b[i].re; im[i] = (float)b[i].im; }
    free(b);
}

/* ----------------------------------------------- newstack -x section transforms
 * LinearTransform::inverse, crates/imod-transforms/src/xf.rs:95-109: det, the
 * four divisions and the shifted offsets, each in REAL in the reference's order. */
void F(or_xf_inverse)(const float* xf6, REAL* inv6) {
    const REAL a11 = xf6[0], a12 = xf6[1], a21 = xf6[2], a22 = xf6[3], dx = xf6[4], dy = xf6[5];
    const REAL det = a11 * a22 - a12 * a21;
    const REAL i11 = a22 / det, i12 = -a12 / det, i21 = -a21 / det, i22 = a11 / det;
    inv6[0] = i11;
    inv6[1] = i12;
    inv6[2] = i21;
    inv6[3] = i22;
    inv6[4] = -(i11 * dx + i12 * dy);
    inv6[5] = -(i21 * dx + i22 * dy);
}

/* Slice::get_clamped + interpolate_bilinear, crates/imod-slice/src/slice.rs:62-86.
 * x0 = floor(x) as isize and fx = x - x0 as f32: for every x whose floor fits an
 * isize, (x0 as f32) == floor(x), so fx = x - floor(x) and the tap bounds are
 * compared on floor(x) itself (a NaN position gives NaN either way). */
static REAL F(xf_bilinear)(const REAL* img, int nx, int ny, REAL x, REAL y, REAL fill) {
    const REAL flx = floor(x), fly = floor(y);
    const REAL fx = x - flx, fy = y - fly;
#define XF_IN(f, n) ((f) >= 0 && (f) < (REAL)(n))
#define XF_AT(fx0, fy0) ((XF_IN(fx0, nx) && XF_IN(fy0, ny)) ? img[(long)(fy0) * nx + (long)(fx0)] : fill)
    const REAL v00 = XF_AT(flx, fly), v10 = XF_AT(flx + 1, fly);
    const REAL v01 = XF_AT(flx, fly + 1), v11 = XF_AT(flx + 1, fly + 1);
#undef XF_AT
#undef XF_IN
    const REAL ofx = (REAL)1.0 - fx, ofy = (REAL)1.0 - fy;
    return v00 * ofx * ofy + v10 * fx * ofy + v01 * ofx * fy + v11 * fx * fy;
}

typedef struct {
    const float* in;
    float* out;
    int nx, ny, nxf;
    const float* xf;
    REAL fill;
} F(xf_ctx);

static void F(xf_sections)(void* vctx, long z0, long z1) {
    const F(xf_ctx)* c = (const F(xf_ctx)*)vctx;
    const long plane = (long)c->nx * c->ny;
    REAL* img = (REAL*)malloc(sizeof(REAL) * plane);
    const REAL xcen = (REAL)((float)c->nx / 2.0f), ycen = (REAL)((float)c->ny / 2.0f);
    for (long z = z0; z < z1; z++) {
        const float* src = c->in + z * plane;
        for (long i = 0; i < plane; i++) img[i] = (REAL)src[i];
        REAL inv[6];
        F(or_xf_inverse)(c->xf + 6 * (z < c->nxf - 1 ? z : c->nxf - 1), inv);
        float* dst = c->out + z * plane;
        for (int y = 0; y < c->ny; y++)
            for (int x = 0; x < c->nx; x++) {
                /* LinearTransform::apply, xf.rs:113-119 */
                const REAL xadj = (REAL)x - xcen, yadj = (REAL)y - ycen;
                const REAL sx = inv[0] * xadj + inv[1] * yadj + inv[4] + xcen;
                const REAL sy = inv[2] * xadj + inv[3] * yadj + inv[5] + ycen;
                dst[(long)y * c->nx + x] = (float)F(xf_bilinear)(img, c->nx, c->ny, sx, sy, c->fill);
            }
    }
    free(img);
}

/* newstack -x, apps/newstack/src/main.rs:1405-1462: section z takes transform
 * min(z, nxf-1) (:1281-1283 in the montage path, the same rule), every output
 * pixel samples the input at the inverse-transformed position about the centre
 * (nx/2, ny/2) (:1411-1412) with bilinear interpolation and fill outside. */
void F(or_apply_xf_stack)(const float* in, int nx, int ny, int nz, const float* xf, int nxf,
                          float fill, float* out, int nthreads) {
    F(xf_ctx) c = {in, out, nx, ny, nxf, xf, (REAL)fill};
    or_parallel_for(nz, 1, nthreads, F(xf_sections), &c);
}

/* --------------------------------------------------------- ctfphaseflip --
 * The CPU path of apps/ctfphaseflip/src/main.rs:369-712 for one section: strips
 * across the tilt axis, each strip's bounding box 2D-transformed (rows, then
 * columns), the Fourier coefficients negated where the CTF (evaluated in
 * double, :563-630) is negative, transformed back (columns, then rows) and
 * blended with the triangular strip weight; pixels with no weight keep their
 * value.  Transforms and pixel arithmetic in REAL. */
typedef struct {
    const float* in;
    float* out;
    int nx, ny, nz, strip_width;
    const float* tilt_deg;
    int n_tilt;
    const float* vdef;  /* [nz][3] defocus1 nm, defocus2 nm, astig deg, or NULL */
    float default_defocus_nm;
    double wavelength, cs_a, amp, cuton, plate_rad, tilt_axis_rad, pixel_a;
} F(ctf_ctx);

static void F(ctf_sections)(void* vctx, long z0, long z1) {
    const F(ctf_ctx)* c = (const F(ctf_ctx)*)vctx;
    const int nx = c->nx, ny = c->ny;
    const double pi = 3.14159265358979323846;
    const double perp_x = cos(c->tilt_axis_rad), perp_y = sin(c->tilt_axis_rad);
    const double center_x = nx / 2.0, center_y = ny / 2.0;
    REAL* wsum = (REAL*)malloc(sizeof(REAL) * (size_t)nx * ny);
    REAL* wtot = (REAL*)malloc(sizeof(REAL) * (size_t)nx * ny);
    F(cpx)* strip = (F(cpx)*)malloc(sizeof(F(cpx)) * (size_t)nx * ny);
    const int nmax = nx > ny ? nx : ny;
    F(cpx)* line = (F(cpx)*)malloc(sizeof(F(cpx)) * (size_t)nmax);
    F(cpx)* work = (F(cpx)*)malloc(sizeof(F(cpx)) * 3 * (size_t)nmax);
    const double w = c->amp, w_phase = sqrt(1.0 - w * w);
    for (long z = z0; z < z1; z++) {
        const float* data = c->in + z * (long)nx * ny;
        float* output = c->out + z * (long)nx * ny;
        const float tdeg = z < c->n_tilt ? c->tilt_deg[z] : 0.0f;
        const double tilt_rad = (double)tdeg * (double)(float)pi / 180.0;
        float d1 = c->default_defocus_nm, d2 = c->default_defocus_nm, aa = 0.0f;
        if (c->vdef) {
            d1 = c->vdef[3 * z];
            d2 = c->vdef[3 * z + 1];
            aa = c->vdef[3 * z + 2];
        }
        const double def1_a = (double)d1 * 10.0, def2_a = (double)d2 * 10.0;
        const double astig_rad = (double)aa * (double)(float)pi / 180.0;
        const int astig = fabs(def1_a - def2_a) > 0.01;
        for (long i = 0; i < (long)nx * ny; i++) {
            output[i] = data[i];
            wsum[i] = 0;
            wtot[i] = 0;
        }
        double min_perp = DBL_MAX, max_perp = -DBL_MAX;
        const double cxs[4] = {0.0, (double)nx, 0.0, (double)nx}, cys[4] = {0.0, 0.0, (double)ny, (double)ny};
        for (int k = 0; k < 4; k++) {
            const double d = (cxs[k] - center_x) * perp_x + (cys[k] - center_y) * perp_y;
            if (d < min_perp) min_perp = d;
            if (d > max_perp) max_perp = d;
        }
        const double sw_d = (double)c->strip_width;
        const int n_strips = (int)ceil((max_perp - min_perp) / sw_d);
        for (int si = 0; si < n_strips; si++) {
            const double scp = min_perp + ((double)si + 0.5) * sw_d;
            const double dx_a = scp * (double)(float)c->pixel_a;
            const double sbd = def1_a + dx_a * sin(tilt_rad);
            const double sbd2 = def2_a + dx_a * sin(tilt_rad);
            const double lo = scp - sw_d / 2.0, hi = scp + sw_d / 2.0;
            int bx0 = nx, bx1 = 0, by0 = ny, by1 = 0;
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++) {
                    const double d = ((double)x - center_x) * perp_x + ((double)y - center_y) * perp_y;
                    if (d >= lo && d < hi) {
                        if (x < bx0) bx0 = x;
                        if (x >= bx1) bx1 = x + 1;
                        if (y < by0) by0 = y;
                        if (y >= by1) by1 = y + 1;
                    }
                }
            if (bx1 <= bx0 || by1 <= by0) continue;
            const int sw = bx1 - bx0, sh = by1 - by0;
            for (int y = 0; y < sh; y++)
                for (int x = 0; x < sw; x++) {
                    strip[y * sw + x].re = (REAL)data[(long)(by0 + y) * nx + bx0 + x];
                    strip[y * sw + x].im = 0;
                }
            for (int y = 0; y < sh; y++) F(dft)(strip + (long)y * sw, sw, -1, work);
            for (int x = 0; x < sw; x++) {
                for (int y = 0; y < sh; y++) line[y] = strip[(long)y * sw + x];
                F(dft)(line, sh, -1, work);
                for (int y = 0; y < sh; y++) strip[(long)y * sw + x] = line[y];
            }
            for (int fy = 0; fy < sh; fy++) {
                const double freq_y = fy <= sh / 2 ? (double)fy : (double)fy - (double)sh;
                for (int fx = 0; fx < sw; fx++) {
                    const double freq_x = fx <= sw / 2 ? (double)fx : (double)fx - (double)sw;
                    double rx = freq_x, ry = freq_y;
                    if (fabs(c->tilt_axis_rad) > 1e-6) {
                        const double ca = cos(c->tilt_axis_rad), sa = sin(c->tilt_axis_rad);
                        rx = freq_x * ca + freq_y * sa;
                        ry = -freq_x * sa + freq_y * ca;
                    }
                    const double sx = rx / ((double)sw * (double)(float)c->pixel_a);
                    const double sy = ry / ((double)sh * (double)(float)c->pixel_a);
                    const double s2 = sx * sx + sy * sy;
                    const double sm = sqrt(s2);
                    if (sm < c->cuton) continue;
                    double def = sbd;
                    if (astig) {
                        const double da = atan2(sy, sx) - astig_rad;
                        const double cos2 = cos(da) * cos(da), sin2 = sin(da) * sin(da);
                        def = sbd * cos2 + sbd2 * sin2;
                    }
                    const double chi = pi * c->wavelength * s2 *
                                       (def - 0.5 * c->cs_a * c->wavelength * c->wavelength * s2);
                    const double chi_t = chi + c->plate_rad;
                    const double ctf = -sin(chi_t) * w_phase + cos(chi_t) * w;
                    if (ctf < 0.0) {
                        strip[(long)fy * sw + fx].re = -strip[(long)fy * sw + fx].re;
                        strip[(long)fy * sw + fx].im = -strip[(long)fy * sw + fx].im;
                    }
                }
            }
            for (int x = 0; x < sw; x++) {
                for (int y = 0; y < sh; y++) line[y] = strip[(long)y * sw + x];
                F(dft)(line, sh, +1, work);
                for (int y = 0; y < sh; y++) strip[(long)y * sw + x] = line[y];
            }
            for (int y = 0; y < sh; y++) F(dft)(strip + (long)y * sw, sw, +1, work);
            const REAL scale = (REAL)1 / (REAL)(sw * sh);
            for (int y = by0; y < by1; y++)
                for (int x = bx0; x < bx1; x++) {
                    const double d = ((double)x - center_x) * perp_x + ((double)y - center_y) * perp_y;
                    if (d >= lo && d < hi) {
                        const float dist = (float)fabs((d - scp) / (sw_d / 2.0));
                        const float wt0 = 1.0f - dist;
                        const REAL wt = (REAL)(wt0 > 0.0f ? wt0 : 0.0f);
                        const REAL val = strip[(long)(y - by0) * sw + (x - bx0)].re * scale;
                        wsum[(long)y * nx + x] += val * wt;
                        wtot[(long)y * nx + x] += wt;
                    }
                }
        }
        for (long i = 0; i < (long)nx * ny; i++)
            if (wtot[i] > (REAL)1e-10) output[i] = (float)(wsum[i] / wtot[i]);
    }
    free(wsum);
    free(wtot);
    free(strip);
    free(line);
    free(work);
}

void F(or_ctf_phaseflip)(const float* in, int nx, int ny, int nz, const float* tilt_deg,
                         int n_tilt, const float* vdef, float default_defocus_nm,
                         const double* phys /* wavelength, cs_a, amp, cuton, plate_rad,
                                               tilt_axis_rad, pixel_a */,
                         int strip_width, float* out, int nthreads) {
    F(ctf_ctx) c = {in, out, nx, ny, nz, strip_width, tilt_deg, n_tilt, vdef,
                    default_defocus_nm, phys[0], phys[1], phys[2], phys[3], phys[4], phys[5],
                    phys[6]};
    or_parallel_for(nz, 1, nthreads, F(ctf_sections), &c);
}

void F(or_local_interp_float)(const float* xs, int nx, const float* ys, int ny,
                              const float* gdx, const float* gdy, float x, float y,
                              float* out2) {
    REAL a, b;
    F(local_interp)(xs, nx, ys, ny, gdx, gdy, (REAL)x, (REAL)y, &a, &b);
    out2[0] = (float)a;
    out2[1] = (float)b;
}

#undef F
#undef CAT
#undef CAT_
