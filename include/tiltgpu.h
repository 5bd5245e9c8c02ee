/*
 * tiltgpu.h -- C ABI of libtiltgpu, an MI355X (gfx950) implementation of the
 * reconstruction core of the `tilt` app of imod-rs: R-weighted backprojection
 * (row FFT radial filter, cosine stretch, per-slice linear-interpolation
 * backprojection over views) and the SIRT reproject/backproject iteration.
 *
 * Plain C types only (pointers, sizes, fixed-width integers).  Every call
 * returns an int status: 0 = OK, negative = error; tiltgpu_last_error() gives
 * the message of the calling thread's last failure.  No exception or abort
 * crosses this boundary.  The caller owns every host buffer; a session owns its
 * device memory and its HIP stream.  One session is used by one host thread at
 * a time; the server API (tiltgpu_server_*) is thread-safe.
 *
 * Reference interfaces this library replaces (paths relative to imod-rs):
 *   apps/tilt/src/gpu.rs:26-269      GpuSession (the in-app GPU seam)
 *   apps/tilt/src/main.rs:601-1093   the CPU reconstruction pipeline
 *   apps/tilt/src/main.rs:18-120     the option surface it is driven by
 */
#ifndef TILTGPU_H
#define TILTGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status codes */
#define TILTGPU_OK 0
#define TILTGPU_ERR_ARG (-1)     /* invalid argument or unsupported shape      */
#define TILTGPU_ERR_HIP (-2)     /* a HIP runtime call failed                  */
#define TILTGPU_ERR_NODEV (-3)   /* no usable GPU                              */
#define TILTGPU_ERR_BUDGET (-4)  /* request exceeds the device-memory budget   */
#define TILTGPU_ERR_STATE (-5)   /* call out of order (e.g. run before upload) */
#define TILTGPU_ERR_IO (-6)      /* file could not be read or written          */
#define TILTGPU_ERR_KERNEL (-7)  /* a kernel reported a violated precondition  */

/* Fill value for out-of-range projection samples (main.rs:122-128,589-594). */
#define TILTGPU_FILL_MEAN 0
#define TILTGPU_FILL_ZERO 1
#define TILTGPU_FILL_EDGE 2

/* Storage precision of the filtered projections held in HBM.  Accumulation is
 * always FP32; F16/BF16 carry one FP32 scale per view. */
#define TILTGPU_STORE_F32 0
#define TILTGPU_STORE_F16 1
#define TILTGPU_STORE_BF16 2

/* Per-view geometry tables (cos/sin of the tilt and x-axis tilt, the z terms
 * (dz*sin t)*zfactor, the stretch factors 1/cos t): FP32 as the reference
 * computes them, or FP64 on the GPU rounded once to FP32 (precision form). */
#define TILTGPU_GEOMETRY_F32 0
#define TILTGPU_GEOMETRY_F64 1

/* Sizes of one reconstruction.  in_* come from the MRC header of the aligned
 * stack (main.rs:560-571): nz = number of sections; out_ny = in_ny. */
typedef struct {
    int32_t in_nx, in_ny, nz;
    int32_t out_nx;  /* --width, default in_nx                    */
    int32_t out_nz;  /* --thickness, default in_nx (when <= 0)    */
} tiltgpu_dims;

/* The `tilt` option surface (main.rs:18-120).  tiltgpu_default_options() fills
 * in the reference defaults. */
typedef struct {
    float radial_cutoff, radial_falloff; /* --radial "0.35,0.05"              */
    int32_t no_weight;                   /* --no-weight                        */
    int32_t exact_filter_size;           /* --exact-filter-size, 0 = n/2+1     */
    int32_t hamming;                     /* --hamming                          */
    int32_t no_cosine_stretch;           /* --no-cosine-stretch                */
    int32_t cosine_cubic;                /* --cosine-interp cubic              */
    float max_cosine_tilt;               /* --max-cosine-tilt 80               */
    int32_t log;                         /* --log                              */
    int32_t fill_mode;                   /* --fill-mode (TILTGPU_FILL_*)       */
    float densweight, outscale, outadd;  /* output transform                   */
    int32_t sirt;                        /* --sirt N                           */
    float sirt_relax;                    /* --sirt-relax                       */
    int32_t fake_sirt;                   /* --fake-sirt N                      */
    int32_t storage;                     /* TILTGPU_STORE_* (no reference flag) */
    int32_t geometry;                    /* TILTGPU_GEOMETRY_* (no ref. flag)   */
} tiltgpu_options;

/* Per-view geometry, the ProjParams of main.rs:130-140,682-700, in FP32 as the
 * reference computes it (the host fills it with tiltgpu_proj_params). */
typedef struct {
    float cos_t, sin_t, cos_xt, sin_xt, zfactor;
} tiltgpu_view_params;

/* Accumulated device time per kernel class since the last reset (ms), measured
 * with HIP events on the session's stream. */
typedef struct {
    double prep_ms;      /* log + per-view means                       */
    double filter_ms;    /* R-weighting + cosine stretch (+ storage)   */
    double bp_ms;        /* backprojection                             */
    double fp_ms;        /* forward projection (SIRT)                  */
    double other_ms;     /* SIRT update, lowpass, statistics           */
    int64_t bp_launches;
    int64_t filter_launches;
    int64_t fp_launches;
} tiltgpu_timing;

typedef struct tiltgpu_session tiltgpu_session;

/* ---------------------------------------------------------------- utility */
const char* tiltgpu_last_error(void);
const char* tiltgpu_version(void);
/* Number of visible GPUs (0 when none); never fails. */
int tiltgpu_device_count(void);
void tiltgpu_default_options(tiltgpu_options* opts);

/* ProjParams table for the first n views (main.rs:682-700): rad = deg*PI/180 in
 * FP32, cos/sin in FP32; xtilt / zfactor may be NULL (main.rs:689-691 defaults). */
int tiltgpu_proj_params(const float* angles_deg, int32_t n, const float* xtilt_deg,
                        int32_t n_xtilt, const float* zfactor, int32_t n_zfactor,
                        tiltgpu_view_params* out);

/* The per-view tables a session reconstructs with (after tilt angles are set):
 * views [nviews] (nviews = min(nz, angle count)), stretch [nz] (1/cos t, 0 for
 * skipped views) and zc [nviews][out_nz] (the z terms (dz*sin t)*zfactor,
 * dz = oz - out_nz/2).  Any pointer may be NULL.  FP32 tables follow
 * main.rs:449-459,688-697,1015-1016; TILTGPU_GEOMETRY_F64 tables are computed
 * in FP64 on the GPU and rounded once.  No counterpart in the reference:
 * exposes the tables for checking. */
int tiltgpu_geometry_tables(tiltgpu_session* s, tiltgpu_view_params* views, float* stretch,
                            float* zc);

/* Radial filter weights W[0..n) (main.rs:348-369). */
int tiltgpu_radial_weights(int32_t n, int32_t exact_filter_size, float cutoff,
                           float falloff, int32_t hamming, float* out);

/* ----------------------------------------------- full pipeline (drop-in) */

/* Create a session for one reconstruction; allocates the device buffers.
 * Replaces GpuSession::new (apps/tilt/src/gpu.rs:43-51), but carries the whole
 * option surface so the device path matches the CPU path of main.rs. */
int tiltgpu_create(const tiltgpu_dims* dims, const tiltgpu_options* opts,
                   tiltgpu_session** out);

/* Tilt angles in degrees as read by read_tilt_file
 * (crates/imod-transforms/src/tilt.rs:6-23).  Views iterated = min(nz, n);
 * cosine stretch needs n >= nz (main.rs:449-450 indexes every section). */
int tiltgpu_set_angles(tiltgpu_session* s, const float* angles_deg, int32_t n);
/* --xtilt-file and --zfactor-file contents (main.rs:664-680,689-691). */
int tiltgpu_set_xtilt(tiltgpu_session* s, const float* xtilt_deg, int32_t n);
int tiltgpu_set_zfactor(tiltgpu_session* s, const float* zfactor, int32_t n);
/* --local-file grids (main.rs:232-307), one per section v < nz: an nx[v] by
 * ny[v] grid with sorted positions xs/ys and row-major dx/dy [iy*nx+ix];
 * arrays are the per-view pieces concatenated in view order. */
int tiltgpu_set_local_alignment(tiltgpu_session* s, const int32_t* nx, const int32_t* ny,
                                const float* xs, const float* ys, const float* dx,
                                const float* dy);

/* Raw projections [nz][in_ny][in_nx] (main.rs:601-605), before --log. */
int tiltgpu_upload_projections(tiltgpu_session* s, const float* host, int64_t count);

/* Raw projections of sections [v0, v0+nv) ([nv][in_ny][in_nx]), for a stack
 * read in chunks (main.rs:601-605 reads it whole).  The session is uploaded
 * once every section has arrived; a later call replaces those sections.
 * Returns after the copy; from tiltgpu_host_alloc memory it runs at the
 * pinned-host rate.  No counterpart in the reference: added for streaming. */
int tiltgpu_upload_views(tiltgpu_session* s, int32_t v0, int32_t nv, const float* host);

/* Device buffers for callers without an allocator of their own (and for the
 * device-pointer entry points' tests): tiltgpu_device_alloc returns NULL on
 * failure (tiltgpu_last_error says why).  The copies are synchronous.  No
 * counterpart in the reference. */
void* tiltgpu_device_alloc(int64_t bytes);
void tiltgpu_device_free(void* p);
int tiltgpu_copy_to_device(void* dst, const void* src, int64_t bytes);
int tiltgpu_copy_to_host(void* dst, const void* src, int64_t bytes);

/* Page-locked host memory for upload and read-back buffers (NULL on failure,
 * tiltgpu_last_error says why); free with tiltgpu_host_free.  No counterpart
 * in the reference. */
void* tiltgpu_host_alloc(int64_t bytes);
void tiltgpu_host_free(void* p);

/* Run the reconstruction on the device: log, per-view means, R-weighting,
 * cosine stretch, then WBP, SIRT or fake-SIRT, and the output transform.  The
 * volume stays resident in HBM.  Replaces main.rs:607-1090 minus file I/O. */
int tiltgpu_run(tiltgpu_session* s);

/* WBP of output rows [y0, y1) only (then tiltgpu_read_rows(s, y0, y1, ...)): the
 * row-by-row form of the reference's GPU dispatch (main.rs:898-918 calls
 * GpuSession::reconstruct_row per row), with the filtering done on the device
 * once per upload and reused by later row ranges.  WBP only (sirt = fake_sirt = 0). */
int tiltgpu_run_rows(tiltgpu_session* s, int32_t y0, int32_t y1);

/* Copy output Y rows [y0, y1) to host as [y][out_nz][out_nx] (the MRC section
 * order of main.rs:1077-1085) with per-row statistics min, max, mean
 * (min_max_mean of each section; stats may be NULL). */
int tiltgpu_read_rows(tiltgpu_session* s, int32_t y0, int32_t y1, float* out,
                      float* stats);

/* SIRT residual norms ||O - FP(V)||_2 per iteration, reduced in FP64 on the
 * device (the reference computes none, main.rs:928-953). */
int tiltgpu_sirt_residuals(tiltgpu_session* s, double* out, int32_t n);

/* Launch shapes.  No counterpart in the reference (its CPU loops have none);
 * for imod-rs_amd/autotune.py, which measures them, and for callers pinning a
 * shape.  bp_tile_z: backprojection tile depth 32, 64 or 128; fp_bins_per_lane:
 * forward-projection bins per lane 4, 8 or 16; 0 = the tuned table's choice (or
 * the analytic rule where the table has no entry).  Results do not depend on
 * either (they change only the work split), except that the tile depth must fit
 * the geometry's LDS need. */
int tiltgpu_set_tuning(tiltgpu_session* s, int32_t bp_tile_z, int32_t fp_bins_per_lane);
/* out4 = {tile depth, its source, FP bins per lane, its source}; sources:
 * 0 analytic rule, 1 tuned table, 2 forced by tiltgpu_set_tuning. */
int tiltgpu_get_tuning(tiltgpu_session* s, int32_t* out4);
/* Diagnostic and tuning (no counterpart in the reference; for
 * imod-rs_amd/autotune.py and callers pinning a shape): the backprojection's
 * launch shape beyond its tile depth.  views_per_stage: views staged in LDS per
 * barrier (1 or 2); x_per_lane: output x positions per lane (1 or 2, tile x 64
 * or 128); 0 = the tuned table's choice.  Built shapes: (1, 1) at every depth,
 * (2, 1) at depths 32-128, (1, 2) at depths 32 and 64; a forced shape that is
 * not built or does not fit the geometry's LDS need fails the next run with
 * TILTGPU_ERR_ARG.  Results do not depend on the shape. */
int tiltgpu_set_bp_shape(tiltgpu_session* s, int32_t views_per_stage, int32_t x_per_lane);
/* out3 = {tile depth, views per stage, x per lane} the session runs with. */
int tiltgpu_get_bp_shape(tiltgpu_session* s, int32_t* out3);

/* Keep output rows [y0, y1) of the volume after the initial WBP and after every
 * SIRT iteration of the next tiltgpu_run (before the output transform), for
 * per-iteration monitoring of the iteration of apps/tilt/src/main.rs:927-953,
 * which reports only the iteration count (:951).  y0 == y1 turns it off.  Holds
 * (sirt+1) x (y1-y0) x out_nz x out_nx floats of device memory. */
int tiltgpu_set_sirt_trace(tiltgpu_session* s, int32_t y0, int32_t y1);
/* Copy trace k (0 = initial WBP, k = after iteration k) as [y1-y0][out_nz][out_nx]. */
int tiltgpu_read_sirt_trace(tiltgpu_session* s, int32_t k, float* out);

/* Forward projection of a caller-supplied volume [in_ny][out_nz][out_nx] with the
 * session's geometry into out [nz][in_ny][in_nx]: the reprojection step of the
 * SIRT loop, the `forward_project` closure of apps/tilt/src/main.rs:830-896.
 * Bit-identical to that closure's serial scatter for the single-row geometry
 * (no x-axis tilt, no local alignment); equal to rounding otherwise.  Replaces
 * the session's volume; sections beyond the tilt angles come out zero. */
int tiltgpu_reproject(tiltgpu_session* s, const float* vol, float* out);

/* Section transforms applied to the uploaded stack before the reconstruction:
 * `newstack -x` (apps/newstack/src/main.rs:1405-1462), the step that makes the
 * aligned stack `tilt` reads.  xf is [n][6] (a11 a12 a21 a22 dx dy, the .xf
 * order); section z takes transform min(z, n-1); output pixels sample the
 * section bilinearly at the inverse transform's position about (nx/2, ny/2)
 * (crates/imod-transforms/src/xf.rs:95-119, crates/imod-slice/src/slice.rs:
 * 62-86), `fill` outside.  n = 0 removes them.  Bit-identical to the
 * reference's FP32 arithmetic. */
int tiltgpu_set_section_transforms(tiltgpu_session* s, const float* xf, int32_t n, float fill);

/* Copy the stack the reconstruction reads back to the host, [nz][in_ny][in_nx]:
 * the uploaded projections, after the section transforms when they are set
 * (newstack's output).  No counterpart in the reference as an entry point. */
int tiltgpu_read_projections(tiltgpu_session* s, float* out);

/* Copy the filtered (and stretched) projections back, dequantised to FP32,
 * [nz][in_ny][in_nx]; for parity tests of the filter stage. */
int tiltgpu_read_filtered(tiltgpu_session* s, float* out);
/* Per-view means (main.rs:615-623) as computed on the device. */
int tiltgpu_read_means(tiltgpu_session* s, float* out);

int tiltgpu_synchronize(tiltgpu_session* s);
int tiltgpu_get_timing(tiltgpu_session* s, tiltgpu_timing* t);
int tiltgpu_reset_timing(tiltgpu_session* s);
/* Device bytes held by the session. */
int64_t tiltgpu_device_bytes(const tiltgpu_session* s);
void tiltgpu_destroy(tiltgpu_session* s);

/* ------------------------------------------- GpuSession-shaped seam entry */

/* Backprojection of already filtered and stretched projections, the shape of
 * GpuSession::new (apps/tilt/src/gpu.rs:43-51) with what the CPU path also
 * needs (per-view params incl. xtilt/zfactor, proj means, fill mode), so the
 * result equals main.rs:1011-1073 rather than the WGSL shader's. */
int tiltgpu_bp_session_new(const float* projections, const tiltgpu_view_params* params,
                           const float* proj_means, int32_t n_views, int32_t in_nx,
                           int32_t in_ny, int32_t nz, int32_t out_nx, int32_t out_nz,
                           int32_t fill_mode, tiltgpu_session** out);
/* GpuSession::reconstruct_row (gpu.rs:224): one [out_nz][out_nx] slice, already
 * multiplied by 1/nz, no output transform. */
int tiltgpu_reconstruct_row(tiltgpu_session* s, int32_t iy, float* out);
/* The same for rows [y0, y1) in one launch. */
int tiltgpu_reconstruct_rows(tiltgpu_session* s, int32_t y0, int32_t y1, float* out);

/* Allocation-free backprojection.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The computation of
 * tiltgpu_bp_session_new + tiltgpu_reconstruct_rows (bit-identical results) on
 * caller-owned device buffers: d_projections [nz][in_ny][in_nx] (filtered and
 * stretched) and d_out [y1-y0][out_nz][out_nx]; params [n_views] and proj_means
 * [nz] are small host tables.  stream is a hipStream_t (NULL: the default
 * stream); the call returns when the backprojection has finished.  The per-view
 * tables live in a static scratch array of TILTGPU_NOALLOC_SCRATCH_BYTES per
 * GPU, held until the process ends: geometries needing more (about n_views x
 * padded thickness x 4 bytes) are rejected with TILTGPU_ERR_ARG.  It calls no
 * hipMalloc or hipFree.  Calls on one GPU share that scratch: calls that overlap
 * on one GPU (from several threads or streams) race. */
#define TILTGPU_NOALLOC_SCRATCH_BYTES (8 << 20)
int tiltgpu_reconstruct_rows_noalloc(const float* d_projections,
                                     const tiltgpu_view_params* params, const float* proj_means,
                                     int32_t n_views, int32_t in_nx, int32_t in_ny, int32_t nz,
                                     int32_t out_nx, int32_t out_nz, int32_t fill_mode, int32_t y0,
                                     int32_t y1, float* d_out, void* stream);

/* ------------------------------------------------- continuous batching */
/* A server owns the device through one worker thread.  Requests from any host
 * thread are admitted in FIFO order while their device memory fits the budget;
 * every round batches the in-flight requests' filter, forward-projection and
 * backprojection work into shared launches (persistent per-XCD tile queues for
 * the backprojection).  A request's result is bit-identical alone and in any
 * mix.  No counterpart in the reference: added for serving. */
typedef struct tiltgpu_server tiltgpu_server;
typedef struct {
    int64_t completed, failed, rounds, bp_launches, bp_batched_requests;
    int32_t max_active, active, queued, pad;
    int64_t bytes_in_use, budget_bytes;
    double p50_ms, p99_ms, mean_ms;  /* submit-to-result latency of completed requests */
} tiltgpu_server_stats;

int tiltgpu_server_create(int64_t budget_bytes, tiltgpu_server** out);
/* A whole reconstruction (the pipeline of tiltgpu_run).  The projections are
 * read when the request is admitted, not copied at submit: they, and `out`
 * ([in_ny][out_nz][out_nx]), must stay valid and unchanged until the request
 * has completed (tiltgpu_server_wait returns, or tiltgpu_server_poll reports it
 * done or failed). */
int tiltgpu_server_submit(tiltgpu_server* s, const tiltgpu_dims* dims,
                          const tiltgpu_options* opts, const float* angles, int32_t n_angles,
                          const float* projections, float* out, int64_t* id);
/* Per-view geometry of a served reconstruction beyond the tilt angles: the
 * --xtilt-file and --zfactor-file contents (main.rs:664-680) and the
 * --local-file grids laid out as for tiltgpu_set_local_alignment (lnx, lny: one
 * entry per section).  Any part may be NULL / 0 (absent). */
typedef struct {
    const float* xtilt;
    int32_t n_xtilt;
    const float* zfactor;
    int32_t n_zfactor;
    const int32_t* lnx;
    const int32_t* lny;
    const float* lxs;
    const float* lys;
    const float* ldx;
    const float* ldy;
} tiltgpu_view_geometry;
/* tiltgpu_server_submit with the x-axis tilts, z factors and local grids of the
 * reconstruction (geom may be NULL): the served form of `tilt ... --xtilt-file
 * --zfactor-file --local-file`, bit-identical to a session given the same
 * tiltgpu_set_xtilt / set_zfactor / set_local_alignment.  The arrays are copied
 * at submit. */
int tiltgpu_server_submit_geometry(tiltgpu_server* s, const tiltgpu_dims* dims,
                                   const tiltgpu_options* opts, const float* angles,
                                   int32_t n_angles, const tiltgpu_view_geometry* geom,
                                   const float* projections, float* out, int64_t* id);
/* Keep a series resident (filtered once) for slab requests; returns when ready.
 * xtilt and the local grids (as tiltgpu_set_local_alignment) may be NULL. */
int tiltgpu_server_add_series(tiltgpu_server* s, const tiltgpu_dims* dims,
                              const tiltgpu_options* opts, const float* angles, int32_t n_angles,
                              const float* xtilt, int32_t n_xtilt, const int32_t* lnx,
                              const int32_t* lny, const float* lxs, const float* lys,
                              const float* ldx, const float* ldy, const float* projections,
                              int32_t* series_id);
/* Output rows [y0, y1) of a resident series: out = [y1-y0][out_nz][out_nx]. */
int tiltgpu_server_submit_slab(tiltgpu_server* s, int32_t series_id, int32_t y0, int32_t y1,
                               float* out, int64_t* id);
/* 1 done, 0 pending, < 0 the request's error. */
int tiltgpu_server_poll(tiltgpu_server* s, int64_t id);
/* Blocks until the request is done; its latency (submit to result in `out`). */
int tiltgpu_server_wait(tiltgpu_server* s, int64_t id, double* latency_ms);
int tiltgpu_server_get_stats(tiltgpu_server* s, tiltgpu_server_stats* st);
int tiltgpu_server_reset_stats(tiltgpu_server* s);
void tiltgpu_server_destroy(tiltgpu_server* s);

/* Diagnostic (no counterpart in the reference): `iters` launches of a plain
 * device-to-device copy kernel of `bytes` bytes; average ms per copy.  The known
 * byte count calibrates hardware traffic counters and gives the attainable
 * HBM bandwidth of the device. */
int tiltgpu_copy_calibration(int64_t bytes, int32_t iters, double* ms_per_copy);
/* Diagnostic (no counterpart in the reference): the attainable FP32 VALU rate of
 * the device with non-packed FMAs (the form every kernel here is built with)
 * whose operands are all in VGPRs, so that the VALU issues them on both of its
 * pipes, from `iters` iterations of eight independent FMA chains per lane at 8
 * waves per SIMD; TFLOP/s. */
int tiltgpu_fma_calibration(int32_t iters, double* tflops);
/* Diagnostic (no counterpart in the reference): time per wave instruction per
 * SIMD (ns) of an instruction mix at 8 waves per SIMD -- 0: FP32 FMA (VGPR
 * operands); 1: v_fract_f32 alone; 2: v_fract_f32 + FP32 FMA pairs; 3:
 * v_fract_f32 + integer add pairs; 4: the two-row backprojection voxel step
 * (10 instructions, no LDS read).  Which instructions share the VALU's second
 * pipe sets the rate a kernel's instruction mix can reach. */
int tiltgpu_issue_calibration(int32_t mix, int32_t iters, double* ns_per_instr);

/* ------------------------------------------------------------ ctfphaseflip
 * CTF phase flip of a tilt series by 2D strips: the CPU path of
 * apps/ctfphaseflip/src/main.rs:353-705 (and, for the flip itself, the work of
 * GpuCtfSession::correct_strip, apps/ctfphaseflip/src/gpu.rs:107), the whole
 * stack on the GPU: per strip the bounding box's 2D transform, the sign flip
 * where the CTF is negative (decided in double precision like the CPU path),
 * the inverse transform and the triangular-weight blend. */
typedef struct {
    float voltage_kv;          /* -v (default 300)                                   */
    float cs_mm;               /* -c (default 2.7)                                   */
    float pixel_a;             /* -p; the CLI defaults it to the header's xlen / mx  */
    float amplitude_contrast;  /* --amplitude-contrast (default 0.07)                */
    float cuton_freq;          /* --cuton-freq, 1/Angstrom (default 0)               */
    float plate_phase_deg;     /* --plate-phase (default 0)                          */
    float tilt_axis_deg;       /* --tilt-axis-angle (default 0)                      */
    float defocus_nm;          /* -d, used where no per-view defocus (default -3000) */
    int32_t strip_width;       /* -w, requested width (default 256)                  */
    int32_t max_strip_width;   /* --max-strip-width; 0 = automatic (main.rs:272-298)  */
} tiltgpu_ctf_params;
/* Args defaults, apps/ctfphaseflip/src/main.rs:20-89 (pixel_a = 1). */
void tiltgpu_ctf_default_params(tiltgpu_ctf_params* p);
/* electron_wavelength (main.rs:709-719), Angstrom. */
double tiltgpu_electron_wavelength(float voltage_kv);
/* The strip width actually used (main.rs:272-298): vdef = the interpolated
 * per-view defocus table [n_vdef][3] (defocus1 nm, defocus2 nm, astig deg) or
 * NULL when no defocus file is given.  Returns the width or an error. */
int tiltgpu_ctf_strip_width(const tiltgpu_ctf_params* p, int32_t nx, const float* vdef,
                            int32_t n_vdef);
/* The strips of main.rs:389-487 for an nx x ny image: returns the strip count n;
 * writes min(n, cap) bounding boxes [bx0, bx1, by0, by1] (empty when bx1 <= bx0)
 * and strip centres (perpendicular distance from the image centre). */
int tiltgpu_ctf_strip_layout(int32_t nx, int32_t ny, float tilt_axis_deg, int32_t strip_width,
                             int32_t* boxes, double* centres, int32_t cap);
/* A stack of nz sections of ny x nx resident on the device, with tilt angles
 * (read_tilt_file; missing angles are 0) and the per-view defocus table as for
 * tiltgpu_ctf_strip_width (sections beyond n_vdef use params->defocus_nm). */
typedef struct tiltgpu_ctf tiltgpu_ctf;
int tiltgpu_ctf_create(int32_t nx, int32_t ny, int32_t nz, const tiltgpu_ctf_params* p,
                       const float* tilt_deg, int32_t n_tilt, const float* vdef, int32_t n_vdef,
                       tiltgpu_ctf** out);
int tiltgpu_ctf_strip_width_used(const tiltgpu_ctf* c);
/* Sections [z0, z0 + nsec) from / to host memory ([nsec][ny][nx] floats). */
int tiltgpu_ctf_upload(tiltgpu_ctf* c, int32_t z0, int32_t nsec, const float* host);
int tiltgpu_ctf_download(tiltgpu_ctf* c, int32_t z0, int32_t nsec, float* host);
/* Corrects the uploaded stack (main.rs:353-691 for every section); blocks.
 * TILTGPU_ERR_STATE before every section has been uploaded; a download needs a
 * run after the last upload. */
int tiltgpu_ctf_run(tiltgpu_ctf* c);
/* Device time (ms) and kernel launches of the last tiltgpu_ctf_run. */
int tiltgpu_ctf_last_run(const tiltgpu_ctf* c, double* ms, int64_t* launches);
void tiltgpu_ctf_destroy(tiltgpu_ctf* c);
/* create + upload + run + download + destroy: host stack in, host stack out. */
int tiltgpu_ctf_phaseflip(const float* in, float* out, int32_t nx, int32_t ny, int32_t nz,
                          const float* tilt_deg, int32_t n_tilt, const float* vdef,
                          int32_t n_vdef, const tiltgpu_ctf_params* p);

/* -------------------------------------------------------- host utilities */

/* read_tilt_file (crates/imod-transforms/src/tilt.rs:6-23).  Writes at most cap
 * angles; returns the count (>= 0) or an error. */
int tiltgpu_read_tilt_file(const char* path, float* out, int32_t cap);
/* load_per_view_file (main.rs:310-326); a parse error is TILTGPU_ERR_IO. */
int tiltgpu_read_per_view_file(const char* path, float* out, int32_t cap);
/* read_xf_file (crates/imod-transforms/src/xf.rs:147-174): transforms of an .xf
 * file as [n][6] (a11 a12 a21 a22 dx dy) into out (up to cap transforms);
 * returns n or TILTGPU_ERR_IO. */
int tiltgpu_read_xf_file(const char* path, float* out, int32_t cap);

/* read_defocus_file (apps/ctfphaseflip/src/main.rs:103-157): entries of a
 * defocus file, views[k] and entries[k][3] = defocus1 nm, defocus2 nm, astig deg
 * (up to cap); returns the count, TILTGPU_ERR_IO when unreadable, or
 * -100 - line for the first malformed line. */
int tiltgpu_read_defocus_file(const char* path, int32_t* views, float* entries, int32_t cap);
/* interpolate_defocus (main.rs:164-198) into out [n_views][3]; returns n_views,
 * or 0 when n == 0 (no per-view table). */
int tiltgpu_interpolate_defocus(const int32_t* views, const float* entries, int32_t n,
                                int32_t n_views, float* out);

/* tiltalign's solved per-view alignment (ViewParams, apps/tiltalign/src/main.rs:
 * 147-154): shift, in-plane rotation (radians), magnification, tilt (radians). */
typedef struct {
    double dx, dy, rotation, mag, tilt;
} tiltgpu_align_view;
/* The local-alignment corrections tiltalign writes with --output-local
 * (compute_local_alignment, apps/tiltalign/src/main.rs:864-940, and the
 * --skip-edge-patches filter, :1374-1402): patches of patch_size pixels at a
 * step of patch_size * (1 - overlap) over the beads' bounding box; for every
 * patch and view, the mean residual (observed - projected) of the beads observed
 * inside the patch, where at least two are.  obs [n_beads][n_views][2]: centred
 * observed positions, NaN x for no observation; beads [n_beads][3]: solved
 * centred (x, y, z); cx, cy: the image centre; skip_edge 0 keeps every patch,
 * img_w / img_h <= 0 mean 2*cx / 2*cy.  Writes up to cap corrections (view,
 * patch index x, patch index y, dx, dy) in the reference's order and returns
 * their count.  The patch fields are grid INDICES, which `tilt --local-file`
 * (main.rs:232-307) reads as image coordinates -- kept as the reference does. */
int tiltgpu_local_corrections(const double* obs, const double* beads,
                              const tiltgpu_align_view* views, int32_t n_beads, int32_t n_views,
                              int32_t patch_size, float overlap, double cx, double cy,
                              int32_t skip_edge, double img_w, double img_h, int32_t cap,
                              int32_t* out_view, int32_t* out_px, int32_t* out_py,
                              double* out_dx, double* out_dy);
/* The --output-local file (main.rs:1404-1409): a comment line, then
 * "view patch_x patch_y dx dy" with dx, dy to 4 decimals. */
int tiltgpu_write_local_corrections(const char* path, int32_t n, const int32_t* view,
                                    const int32_t* px, const int32_t* py, const double* dx,
                                    const double* dy);

/* --local-file (load_local_alignment, main.rs:232-307) parsed into per-view
 * grids for tiltgpu_set_local_alignment.  sizes[3] receives the total lengths of
 * xs, ys and dx/dy; with nx == NULL only sizes are written.  nx, ny: nz ints. */
int tiltgpu_parse_local_file(const char* path, int32_t nz, int32_t* sizes, int32_t* nx,
                             int32_t* ny, float* xs, float* ys, float* dx, float* dy);

/* MRC 2014 files: the subset of crates/imod-mrc used by `tilt`. */
typedef struct {
    int32_t nx, ny, nz, mode, mx, my, mz, next, nlabl;
    int32_t swapped;  /* big-endian file (read.rs:36-52)   */
    int32_t is_imod;  /* IMOD stamp present (header.rs:137) */
    float xlen, ylen, zlen, amin, amax, amean;
    char label0[81];
} tiltgpu_mrc_info;
/* MrcReader::open (crates/imod-mrc/src/read.rs:28-93) header fields. */
int tiltgpu_mrc_info_read(const char* path, tiltgpu_mrc_info* out);
/* read_slice_f32 (read.rs:127-231) of sections [z0, z1) into out (cap floats). */
int tiltgpu_mrc_read_sections(const char* path, int32_t z0, int32_t z1, float* out,
                              int64_t cap);
/* MrcHeader::new + add_label + MrcWriter create / write_slice_f32 / finish
 * (header.rs:181-258, write.rs:20-153) of an nx x ny x nz float volume, with the
 * header statistics computed as main.rs:1080-1093 does. */
int tiltgpu_mrc_write_float(const char* path, const float* data, int32_t nx, int32_t ny,
                            int32_t nz, float xlen, float ylen, float zlen, const char* label);

/* Synthetic tilt series: the seeded ellipsoid + bead phantom projected
 * analytically with the tilt geometry (per-view angle, x-axis tilt, in-plane
 * shift), plus Gaussian noise of sigma = noise_rel * std of the noiseless
 * stack.  out = [n_views][ny][nx].  Runs on host threads. */
typedef struct {
    int32_t vol_nx, vol_ny, vol_nz;  /* phantom box (x, y, thickness)      */
    int32_t n_ellipsoids, n_beads;
    float axis_min, axis_max;        /* ellipsoid semi-axes (px)           */
    float bead_radius, bead_density;
    float noise_rel;
    uint64_t seed;
} tiltgpu_phantom;
int tiltgpu_synth_series(const tiltgpu_phantom* ph, int32_t nx, int32_t ny,
                         int32_t n_views, const float* angles_deg, const float* xtilt_deg,
                         const float* shift_x, const float* shift_y, int32_t nthreads,
                         float* out);

#ifdef __cplusplus
}
#endif
#endif /* TILTGPU_H */
