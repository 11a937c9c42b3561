/* fyg.h -- the thin C ABI between the C host code (the files under host/, which
 * implement the fysics.h entry points) and the HIP side (the .hip files: device
 * runtime and kernels).
 *
 * Everything here is plain C: device buffers are plain device addresses and
 * streams are opaque handles.  The host code never includes HIP
 * headers; the HIP files never see fysics.h structs.  Every int-returning
 * function returns 0 on success and nonzero on any HIP error.
 */
#ifndef FYG_H
#define FYG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ runtime */
typedef struct fyg_stream fyg_stream;

/* the calling host thread's stream (created on first use, default priority) */
fyg_stream *fyg_thread_stream(void);
/* wrap an external hipStream_t (NULL -> the thread stream) */
fyg_stream *fyg_stream_wrap(void *hip_stream);
/* frees the stream object of hip_stream (NULL: the calling thread's own
 * stream) after waiting for it; the next call makes a new one */
int   fyg_stream_release(void *hip_stream);
void *fyg_stream_handle(fyg_stream *s);          /* the hipStream_t */
/* a private stream object on hip_stream (NULL: the thread stream's), with its
 * own workspaces and flags and tables from the process-wide cache, for work
 * that is recorded into a graph */
fyg_stream *fyg_stream_private(void *hip_stream);
void  fyg_stream_private_free(fyg_stream *s);
int   fyg_stream_persistent(fyg_stream *s);
/* waits for s, then frees its workspaces (the next fyg_ws regrows them) */
int   fyg_stream_trim(fyg_stream *s);
/* Fork / join: helper i (< FYG_SUBS) of s is a stream object of its own
 * (workspaces, default priority; persistent when s is), created on first use
 * and freed with s; an eager helper has a hardware queue of its own (a
 * full-CU-mask stream, up to 12 per process) so it never queues behind s.  fyg_fork makes helpers 0..n-1 wait for the work queued
 * on s so far; fyg_join makes s wait for everything queued on them.  Both are
 * event dependencies, so they also hold inside a graph capture of s. */
#define FYG_SUBS 3
fyg_stream *fyg_stream_sub(fyg_stream *s, int i);
int   fyg_fork(fyg_stream *s, int n);
int   fyg_join(fyg_stream *s, int n);
/* an owned stream object on a new non-blocking hipStream_t; fyg_stream_free
 * waits for it, frees its buffers and destroys the stream */
fyg_stream *fyg_stream_new(void);
void  fyg_stream_free(fyg_stream *s);
/* work queued on s after this call waits for ev (recorded on any stream) */
int   fyg_stream_wait(fyg_stream *s, void *ev);
/* pinned host memory */
int   fyg_host_alloc(void **p, size_t bytes);
void  fyg_host_free(void *p);
/* graph capture of the launches queued on s between begin and end */
int   fyg_capture_begin(fyg_stream *s);
int   fyg_capture_end(fyg_stream *s, void **exec);
int   fyg_graph_launch(void *exec, fyg_stream *s);
void  fyg_graph_free(void *exec);
int   fyg_sync(fyg_stream *s);

/* per-stream grow-only device workspace: slot in [0, FYG_WS_SLOTS) */
#define FYG_WS_SLOTS 9
void *fyg_ws(fyg_stream *s, int slot, size_t bytes);
/* per-stream grow-only pinned host staging buffer */
void *fyg_pinned(fyg_stream *s, int slot, size_t bytes);

int  fyg_malloc(void **p, size_t bytes);
void fyg_free(void *p);
int  fyg_h2d(void *dst, const void *src, size_t bytes, fyg_stream *s);
int  fyg_d2h(void *dst, const void *src, size_t bytes, fyg_stream *s);
int  fyg_d2d(void *dst, const void *src, size_t bytes, fyg_stream *s);
int  fyg_zero(void *dst, size_t bytes, fyg_stream *s);    /* fill kernel, graph-safe */
int  fyg_last_error(void);                                /* sticky launch error check */
/* per-stream all-air flags: enqueues the step of a device generation word g
 * (>= 1, a one-thread kernel, so a captured graph steps it on every replay) and
 * returns the flags and g's address; kernels that see a nonzero input voxel
 * store g, consumers test (*flag == g), so flags are never cleared between
 * calls.  Returns 0 on error. */
int  fyg_flag(fyg_stream *s, int **d_flag, const int **d_gen);
/* the same with n flags (one per cube of a batch), all tested against one g */
int  fyg_flags(fyg_stream *s, int n, int **d_flags, const int **d_gen);
/* fyg_flags without the step: the caller steps g (nflag flags follow it) with
 * fyg_gen_step or in the first kernel of its run, before anything reads g */
int  fyg_flags_reserve(fyg_stream *s, int n, int **d_flags, int **d_gen, int *nflag);
int  fyg_gen_step(int *d_gen, int nflag, fyg_stream *s);
const char *fyg_error_string(void);
void fyg_set_error(const char *msg);            /* host-side failures (I/O, arguments) */

/* process-wide cache of small read-only device tables (twiddles, transfer
 * LUTs, per-volume blobs).  Returns the device copy for key, uploading data on
 * a miss.  Thread safe.  Twiddles ('T' keys) live until process end; the other
 * tables form a bounded cache that evicts the least recently used entry that
 * no hold covers, and frees its memory only after the events of all its
 * releases have completed. */
const void *fyg_table(const void *key, size_t keylen, const void *data, size_t bytes);
const void *fyg_table_find(const void *key, size_t keylen);
/* Holds: every table the calling thread finds or makes between fyg_hold_begin
 * and fyg_hold_end (or release) is held -- never evicted -- until
 * fyg_hold_release, which records an event on s after the work queued so far
 * (s NULL or capturing: no event) and frees the hold.  Holds nest (LIFO); each
 * open hold holds the tables.  A captured graph keeps its hold until it is
 * destroyed. */
typedef struct fyg_hold fyg_hold;
fyg_hold *fyg_hold_begin(void);
void fyg_hold_end(fyg_hold *h);
int  fyg_hold_release(fyg_hold *h, fyg_stream *s);
/* st[0..6]: entries, evictable entries, held entries, evicted entries not yet
 * freed, evictions so far, frees so far, the cache limit */
void fyg_table_stats(long *st);
/* sets the limit on evictable entries (n >= 1; 0: unchanged); returns the old one */
size_t fyg_table_limit(size_t n);
/* per-stream table slots for tables that change from call to call (per-request
 * transfer LUTs): on a miss the host data is staged into pinned memory and
 * copied on the stream, so no call waits on the device.  FYG_ARENA_SLOTS slots
 * of FYG_ARENA_SLOT bytes; the least recently used one is recycled, which is
 * safe in stream order as long as a caller holds fewer than FYG_ARENA_SLOTS/2
 * tables between planning and launching.  data == NULL: lookup only.  Returns
 * NULL if bytes exceeds a slot. */
#define FYG_ARENA_SLOTS 256
#define FYG_ARENA_SLOT ((size_t)128 << 10)
const void *fyg_stream_table(fyg_stream *s, const void *key, size_t keylen, const void *data, size_t bytes);

/* completion events on a stream */
int  fyg_event_record(fyg_stream *s, void **ev);     /* creates *ev on first use */
int  fyg_event_done(void *ev);                       /* 1 done, 0 pending, -1 error */
int  fyg_event_wait(void *ev);
void fyg_event_destroy(void *ev);

/* fp32 twiddles cos/sin(2 pi m / L), m < L, evaluated in double (float2[L]) */
const void *fyg_twiddles(int L);
/* the same in fp64 (double2[L]) */
const void *fyg_twiddles64(int L);

/* per-kernel timing with HIP events recorded on the launch stream */
void fyg_prof_enable(int on);
int  fyg_prof_collect(char *names, int names_len, double *ms, long *count, int max);
void fyg_prof_reset(void);

/* Cubes of one shape run as one launch per stage: up to FYG_BATCH cubes'
 * descriptors travel in the kernel arguments (no descriptor upload, so the
 * launches can be captured), and larger groups take several launches. */
#define FYG_BATCH 16

/* Cubes of mixed shapes run as one launch per stage for all of them: the
 * launch's blocks are cut into one segment per cube (at most FYG_MIX_MAX),
 * and the segments and descriptors travel in the kernel arguments. */
#define FYG_MIX_MAX 128

/* ------------------------------------------------------------------ FFT / deconv */
int fyg_fft_supported(int L);                   /* 2^k or 3*2^k, L <= FYG_FFT_MAX */
#define FYG_FFT_MAX 4096

/* tile prologue applied to u8 input voxels: f = u8/255, ring subtraction,
 * z-drift factor, dewindow (pipeline.c:492-494,266-268) */
typedef struct {
    int dering; const float *ring; int nr, nslab, slab_z; double cy, cx; float scale;
    int ring_fast;                   /* axis on the (half-)integer grid and radii < 512: float radius */
    int zdrift; const float *zfactor;
    long gz0, gy0, gx0;              /* global origin of the buffer */
    int window; float wlo, wspan;
    /* deconvolution input only: workspace of fyg_ring_image_bytes() for the
     * per-slab ring correction image scale*ring[slab][radius(y,x)] */
    float *cimg;
} fyg_prologue;

size_t fyg_ring_image_bytes(const fyg_prologue *p, int nz, int ny, int nx);

/* epilogue applied to every deconvolved voxel before it is stored
 * (pipeline.c:273-283): rewindow, dec-range rescale, clamp to [0,1] */
typedef struct {
    int window; float wlo, winv;
    int dec; float dlo, dinv;
    int clamp01;
} fyg_epilogue;

/* one real-input Wiener deconvolution on device buffers */
typedef struct {
    int nz, ny, nx;                  /* logical (cropped) volume */
    int pz, py, px, hx, hxp;         /* padded FFT sizes, half-spectrum width and row pitch */
    const void *lut;                 /* 4097 weights over u = f^2: fp32, or fp64 when f64 */
    double us;                       /* LUT index scale: t = u*us */
    const double *fz2, *fy2, *fx2;   /* squared frequencies (fz2 pre-scaled by sigma_z ratio) */
    const void *in; int in_u8;       /* fp32 or u8 input (nz*ny*nx) */
    fyg_prologue pro;                /* u8 input only */
    void *out;                       /* fp32 nz*ny*nx (double when f64), may alias a fp32 in */
    fyg_epilogue epi;
    void *spec;                      /* complex workspace pz*py*hxp float2 (double2 when f64) */
    int *nonzero_flag;               /* optional: receives *nonzero_gen if any input u8 is nonzero */
    const int *nonzero_gen;          /* device generation word of the call */
    int f64;                         /* all-FP64 variant: spectrum, weights and out in double */
    int *step_gen;                   /* optional (a[0] of a batch): generation word to step before
                                        anything reads it -- folded into the first ring-image launch */
    int step_nflag;                  /* flags that follow *step_gen */
    const void *twx, *twy, *twz;     /* optional twiddle tables for px, py, pz (NULL: the cache's) */
    int no_tune;                     /* never time radix plans here (callers that may not allocate):
                                        lengths not decided yet take plan 0 */
} fyg_deconv_args;
int fyg_deconv_run(const fyg_deconv_args *a, fyg_stream *s);
/* n deconvolutions of one shape (same nz..hxp, in_u8 and pro.dering) */
int fyg_deconv_run_batch(const fyg_deconv_args *a, int n, fyg_stream *s);
size_t fyg_deconv_spec_bytes(int pz, int py, int px);
size_t fyg_deconv_spec_bytes64(int pz, int py, int px);
int fyg_deconv_run_batch_f64(const fyg_deconv_args *a, int n, fyg_stream *s);
int fyg_deconv_hxp(int px);

/* the deconvolution with nothing allocated: spectrum, transfer tables and
 * twiddles live in a file-scope device array (fyg_noalloc.hip) of
 * FYG_NOALLOC_SPEC_MAX bytes of spectrum plus a table area, shared by every
 * call on a device; the tables are computed on the device each call.  phys8
 * is fy_physics as 8 doubles; us the LUT index scale.  Wiener weight of the
 * recon transfer, fz2 scale 1. */
#define FYG_NOALLOC_SPEC_MAX ((size_t)80 << 20)
int fyg_noalloc_deconv(const float *in, float *out, int nz, int ny, int nx, int pz, int py, int px,
                       const double *phys8, double reg, double cap, double us, fyg_stream *s);

/* pass structure of the deconvolution for cubes whose padded X and Y axes are
 * one length of 64 or 96: 1 (default) fused X+Y plane kernels, 0 the separate
 * X and Y passes (the same transforms; fp32 rounding may differ in the last
 * bits); returns the previous mode */
int fyg_fft_plane_mode(void);
int fyg_fft_plane_set(int mode);

/* Radix plans of the fp32 deconvolution passes.  Lengths 64..512 with two
 * plans (radix-16 stages, or radix-8 stages) take one per pass kind (X, Y, Z,
 * fused XY planes), decided once per process and length: mode 0 plan 0
 * everywhere, 2 plan 1 everywhere, 1 (default) time both on a synthetic
 * problem of the length the first time a deconvolution that may tune meets it
 * (not under capture, not no_tune) or take the tuning file's choice.
 * Returns the previous mode. */
int fyg_fft_radix_mode(int mode);
void fyg_fft_radix_clear(void);
/* the plan variant of pass kind (FYG_RK_*) at length L for a deconvolution on
 * s, timed now when may_tune and not decided yet */
enum { FYG_RK_X, FYG_RK_Y, FYG_RK_Z, FYG_RK_XY };
int fyg_radix_plan(int kind, int L, fyg_stream *s, int may_tune);
/* "fft\t<kind>\t<length>\t<plan>" lines of every decided (kind, length) */
int fyg_fft_radix_lines(char *buf, int len);
/* one such line (without "fft\t"): 0 ok, 1 malformed */
int fyg_fft_radix_load_line(const char *line);
/* human-readable decisions with the timings they came from */
int fyg_fft_radix_report(char *buf, int len);

/* complex FFT of split re/im device arrays (fy_fft1d/fy_fft3d) */
int fyg_fft3d_run(float *d_re, float *d_im, int nz, int ny, int nx, int sign,
                  void *work /* 8*nz*ny*nx bytes */, fyg_stream *s);
int fyg_scale_run(float *d, size_t n, float k, fyg_stream *s);
/* in-place complex 3-D FFT of interleaved (re, im) fp32 device data */
int fyg_cfft3d_run(void *data, int nz, int ny, int nx, int sign, fyg_stream *s);

/* ------------------------------------------------------------------ phase correlation */
typedef struct {
    float best;                     /* surface maximum (first in index order) */
    float maxmag;                   /* largest cross-power magnitude */
    long long index;                /* its padded index */
    float c[7];                     /* surface at the peak, then -z, +z, -y, +y, -x, +x (circular) */
} fyg_pc_result;
size_t fyg_pc_ws_bytes(int pz, int py, int px);
/* phasecorr.c:44-145 on device volumes: w* are the per-axis windows (device) */
int fyg_phase_correlate_run(const float *fixed, const float *moving, int nz, int ny, int nx, int pz, int py,
                            int px, const float *wz, const float *wy, const float *wx, void *ws,
                            fyg_pc_result *out, fyg_stream *s);

/* ------------------------------------------------------------------ guided / box */
typedef struct {
    int quant;                       /* 0: fp32 out over the whole volume; 1: u8 of the inner tile */
    float *out;                      /* fp32 nz*ny*nx (quant 0); may alias in */
    unsigned char *q;                /* u8 tz*ty*tx (quant 1) */
    long iz0, iy0, ix0, tz, ty, tx;  /* inner tile, local coordinates */
    long gz0, gy0, gx0;              /* global origin of the buffer (dither coordinates) */
    int no_dither;
    int recenter; float rmid, rk;    /* pipeline.c:413-421 */
    const int *nonzero_flag;         /* optional: unless it holds *nonzero_gen the tile is all air -> zeros */
    const int *nonzero_gen;
} fyg_emit;

typedef struct {
    int nz, ny, nx, r, s;            /* s <= 1: exact path */
    double eps;
    const float *in;
    fyg_emit em;
    void *ws;                        /* device workspace, fyg_guided_ws_bytes() */
    int f64;                         /* all-FP64 variant (fast path, s >= 2): guide in64, ws in */
    const double *in64;              /* doubles of fyg_guided_ws_bytes64(); emitted in fp32 */
} fyg_guided_args;
size_t fyg_guided_ws_bytes(int nz, int ny, int nx, int s);
size_t fyg_guided_ws_bytes64(int nz, int ny, int nx, int s);
int fyg_guided_run(const fyg_guided_args *a, fyg_stream *s);
/* tile-shape tuner of the fast guided filter: mode 0 first candidates, 1 tune
 * on first use (default), 2 + k force candidate k; returns the previous mode */
int fyg_tune_mode(int mode);
/* one line per tuned shape; returns the report's length */
int fyg_tune_report(char *buf, int len);
/* the tuned shapes to / from a text file; clear forgets them all */
int fyg_tune_save(const char *path);
int fyg_tune_load(const char *path);
void fyg_tune_clear(void);
/* n guided filters; cubes of one shape on the fast path share launches */
int fyg_guided_run_batch(const fyg_guided_args *a, int n, fyg_stream *s);
/* 1 when a padded FFT length runs in mixed launches (64, 96, 128, 192, 256) */
int fyg_mix_len_ok(int L);
/* n deconvolutions (fp32, u8 input, all with or all without rings, lengths
 * that fyg_mix_len_ok accepts): ring images (or the generation step), then
 * X forward, Y forward, Z, Y inverse, X inverse, one launch each */
int fyg_mix_deconv_run(const fyg_deconv_args *a, int n, fyg_stream *s);
/* the fast guided filter of a cube can run in a mixed launch; n filters in
 * one coefficient launch and one upsample launch */
int fyg_mix_guided_ok(const fyg_guided_args *a);
int fyg_mix_guided_run(const fyg_guided_args *a, int n, fyg_stream *s);
/* batches of several shapes where every cube allows it: 1 (default) mixed
 * launches over parts of the batch whose spectra fit the last-level cache,
 * on two streams; 2 mixed launches over the whole batch; 0 one launch per
 * stage per shape group.  Returns the previous mode. */
int fyg_mix_mode(void);
int fyg_mix_set(int mode);
/* prologue-only path (deconv disabled): u8 -> f32 through the tile prologue */
int fyg_prologue_run(const unsigned char *in, float *out, int nz, int ny, int nx,
                     const fyg_prologue *pro, int *nonzero_flag, const int *gen, fyg_stream *s);
/* emit without guided filtering: recenter + quantise the inner tile */
int fyg_emit_run(const float *f, int nz, int ny, int nx, const fyg_emit *em, fyg_stream *s);
int fyg_box_mean_run(const float *in, float *out, int nz, int ny, int nx, int r,
                     void *ws, fyg_stream *s);

/* ------------------------------------------------------------------ elementwise */
int fyg_u8_to_f32_run(const unsigned char *in, float *out, size_t n, float lo, float scale, fyg_stream *s);
/* 2x LOD downscale, method 0 box / 1 contrast-maintaining box (downscale.c) */
int fyg_downscale2x_run(const unsigned char *in, int nx, int ny, int nz, unsigned char *out, int method,
                        float alpha, fyg_stream *s);
int fyg_f32_to_u8_run(const float *in, unsigned char *out, size_t n, float lo, float k, fyg_stream *s);
int fyg_zdrift_run(float *f, int nz, int ny, int nx, const float *factor_at_z0, fyg_stream *s);
int fyg_dering_apply_run(float *f, long dz, long dy, long dx, const fyg_prologue *pro, fyg_stream *s);
/* accumulate (sector,radius) sums/counts from a u8 region with a per-(y,x) bin map */
int fyg_dering_accum_run(const unsigned char *buf, long dz, long dy, long dx, int ss,
                         const int *map, int nbins, unsigned long long *sum,
                         unsigned int *cnt, fyg_stream *s);

/* ------------------------------------------------------------------ pass-1 calibration */
/* hist[b] += count of byte value b in d[0..n) (any alignment); part is scratch
 * of fyg_hist_part_bytes(n) bytes */
size_t fyg_hist_part_bytes(size_t n);
int fyg_hist_u8_run(const unsigned char *d, size_t n, unsigned int *part, unsigned long long *hist,
                    fyg_stream *s);
/* acc[0..255] += tile[0..255], acc[256] += n; skip_zero: not when tile[0] == n.
 * Clears tile. */
int fyg_hist_merge_run(unsigned long long *tile, unsigned long long n, unsigned long long *acc,
                       int skip_zero, fyg_stream *s);
/* per slice z < dz: zs[z] += sum of u8/255 > 0.10 over (y, x) multiples of ss in
 * units of 2^-27, zc[z] += count (integer atomics) */
int fyg_zdrift_u8_run(const unsigned char *d, long dz, long dy, long dx, int ss, unsigned long long *zs,
                      unsigned long long *zc, fyg_stream *s);
/* per slice z < nz: sums[z] += sum of values > thresh (double), cnts[z] += count */
int fyg_zdrift_f32_run(const float *d, int nz, int ny, int nx, float thresh, double *sums,
                       long long *cnts, fyg_stream *s);
/* std of each blk^3 block with mean > 0.1, -1 for the others (blocks x fastest) */
int fyg_flat_blocks_run(const float *v, int nz, int ny, int nx, int blk, double *std_out, fyg_stream *s);
/* out = clamp((in - lo) / (hi - lo), 0, 1) in fp32 (1/(hi-lo) -> 1 when hi <= lo) */
int fyg_norm_u8_run(const unsigned char *in, float *out, size_t n, unsigned char lo, unsigned char hi,
                    fyg_stream *s);

/* hist[b] += sum of part[r*256 + b] over rows r < rows */
int fyg_hist_reduce_run(const unsigned int *part, int rows, unsigned long long *hist, fyg_stream *s);

/* ------------------------------------------------------------------ air-zero */
typedef struct {
    int scratch_passes;             /* <= 0 -> 5 */
    int cut_u8, band;               /* cut_u8 < 0: dark-peak rule */
    int min_component;              /* <= 0: no despeckle */
    double thresh, aggr;
} fyg_air_args;
/* valley depth of a 256-bin histogram (stream.c:138-170) */
double fyg_valley_depth(const long *hist, int *dark, int *light, int *valley);
size_t fyg_air_ws_bytes(int nz, int ny, int nx);
/* stage (d) on f in place; orig = the corrected tile before the deconvolution */
int fyg_air_zero_run(float *f, const float *orig, int nz, int ny, int nx, const fyg_air_args *a, void *ws,
                     fyg_stream *s);
/* the union-find's broken-invariant flag of the last run (waits for the stream) */
int fyg_air_err_read(void *ws, int nz, int ny, int nx, int *out, fyg_stream *s);

/* ------------------------------------------------------------------ noise model */
/* local (2r+1)^3 box mean m and variance v (double) with sliding sums per line */
int fyg_noise_local(const float *d_in, int nz, int ny, int nx, int r, double wc, double *m, double *v,
                    double *t0, double *t1, double *t2, double *t3, fyg_stream *s);
size_t fyg_sort_pairs_temp_bytes(size_t n);
int fyg_sort_pairs_f64(const double *kin, double *kout, const double *vin, double *vout, size_t n,
                       void *temp, size_t temp_bytes, fyg_stream *s);
size_t fyg_segsort_temp_bytes(size_t n, int nseg);
int fyg_segsort_f64(const double *in, double *out, size_t n, int nseg, const int *d_begin,
                    const int *d_end, void *temp, size_t temp_bytes, fyg_stream *s);
int fyg_lower_bounds_f64(const double *sorted, size_t n, const double *d_q, int nq, long *d_idx,
                         fyg_stream *s);
int fyg_gather_f64(const double *src, const long *d_idx, int n, double *d_out, fyg_stream *s);
int fyg_aniso_blocks(const float *v, int nz, int ny, int nx, float *bvar, int *ok, fyg_stream *s);
int fyg_aniso_lag(const float *v, int nz, int ny, int nx, float thresh, double *part, int *used,
                  fyg_stream *s);

#ifdef __cplusplus
}
#endif
#endif
