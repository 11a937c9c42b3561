/* fy_core.c -- FFT sizes and transforms, transfer functions, u8/float and
 * physical-window conversions, z-drift apply, kernel halo.  Array work runs on
 * the GPU through the fyg ABI; only scalar formulas are evaluated here.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fyh.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ---------------------------------------------------------------- sizes (fft.c:25-38) */
int fy_is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

int fy_next_pow2(int n) {
    int p = 1;
    while (p < n) p <<= 1;
    return p;
}

/* smallest 2^k or 3*2^k >= n */
int fy_next_fft_size(int n) {
    int p2 = fy_next_pow2(n);
    int p3 = 3 * fy_next_pow2((n + 2) / 3);
    return (p3 >= n && p3 < p2) ? p3 : p2;
}

/* ---------------------------------------------------------------- complex FFTs */
/* n must be 2^k or 3*2^k and at most FYG_FFT_MAX; other lengths leave the data
 * untouched (the reference's transform is undefined for them). */
static void fft3d_host(float *re, float *im, int nz, int ny, int nx, int sign) {
    if (nz <= 0 || ny <= 0 || nx <= 0) return;
    if (!fyg_fft_supported(nx) || !fyg_fft_supported(ny) || !fyg_fft_supported(nz)) return;
    fyg_stream *s = fyg_thread_stream();
    if (!s) return;
    size_t n = (size_t)nz * ny * nx;
    float *d = (float *)fyg_ws(s, WS_IN, sizeof(float) * 2 * n);
    void *work = fyg_ws(s, WS_MISC, sizeof(float) * 2 * n);
    if (!d || !work) return;
    if (fyg_h2d(d, re, sizeof(float) * n, s) || fyg_h2d(d + n, im, sizeof(float) * n, s)) return;
    if (fyg_fft3d_run(d, d + n, nz, ny, nx, sign, work, s)) return;
    if (fyg_d2h(re, d, sizeof(float) * n, s) || fyg_d2h(im, d + n, sizeof(float) * n, s)) return;
    fyg_sync(s);
}

void fy_fft1d(float *re, float *im, int n, int sign) { fft3d_host(re, im, 1, 1, n, sign); }

void fy_fft3d(float *re, float *im, int nz, int ny, int nx, int sign) {
    fft3d_host(re, im, nz, ny, nx, sign);
}

void fy_fft3d_normalize(float *re, float *im, int nz, int ny, int nx) {
    size_t n = (size_t)nz * ny * nx;
    if (!n) return;
    fyg_stream *s = fyg_thread_stream();
    if (!s) return;
    float *d = (float *)fyg_ws(s, WS_IN, sizeof(float) * 2 * n);
    if (!d) return;
    float k = 1.0f / (float)n;                                  /* fft.c:226 */
    if (fyg_h2d(d, re, sizeof(float) * n, s) || fyg_h2d(d + n, im, sizeof(float) * n, s)) return;
    if (fyg_scale_run(d, 2 * n, k, s)) return;
    if (fyg_d2h(re, d, sizeof(float) * n, s) || fyg_d2h(im, d + n, sizeof(float) * n, s)) return;
    fyg_sync(s);
}

/* ---------------------------------------------------------------- conversions (stream.c:27-63) */
static void u8_to_f32_host(const unsigned char *in, float *out, size_t n, float lo, float scale) {
    if (!n) return;
    fyg_stream *s = fyg_thread_stream();
    if (!s) return;
    unsigned char *du = (unsigned char *)fyg_ws(s, WS_U8, n);
    float *df = (float *)fyg_ws(s, WS_IN, sizeof(float) * n);
    if (!du || !df) return;
    if (fyg_h2d(du, in, n, s) || fyg_u8_to_f32_run(du, df, n, lo, scale, s)) return;
    if (fyg_d2h(out, df, sizeof(float) * n, s)) return;
    fyg_sync(s);
}

static void f32_to_u8_host(const float *in, unsigned char *out, size_t n, float lo, float k) {
    if (!n) return;
    fyg_stream *s = fyg_thread_stream();
    if (!s) return;
    unsigned char *du = (unsigned char *)fyg_ws(s, WS_U8, n);
    float *df = (float *)fyg_ws(s, WS_IN, sizeof(float) * n);
    if (!du || !df) return;
    if (fyg_h2d(df, in, sizeof(float) * n, s) || fyg_f32_to_u8_run(df, du, n, lo, k, s)) return;
    if (fyg_d2h(out, du, n, s)) return;
    fyg_sync(s);
}

/* out = in * (1/255): written as 0 + in*(1/255) on the device, which is exact */
void fy_u8_to_float(const unsigned char *in, float *out, size_t n) {
    u8_to_f32_host(in, out, n, 0.0f, 1.0f / 255.0f);
}

void fy_float_to_u8(const float *in, unsigned char *out, size_t n) {
    f32_to_u8_host(in, out, n, 0.0f, 255.0f);
}

void fy_u8_to_phys(const unsigned char *in, float *out, size_t n, double f32_min, double f32_max) {
    u8_to_f32_host(in, out, n, (float)f32_min, (float)((f32_max - f32_min) / 255.0));
}

void fy_phys_to_u8(const float *in, unsigned char *out, size_t n, double f32_min, double f32_max) {
    double span = f32_max - f32_min;
    float k = (span != 0.0) ? (float)(255.0 / span) : 0.0f;
    f32_to_u8_host(in, out, n, (float)f32_min, k);
}

float fy_phys_to_u8_level(double phys, double f32_min, double f32_max) {
    double span = f32_max - f32_min;
    if (span == 0.0) return 0.0f;
    return (float)((phys - f32_min) / span * 255.0);
}

/* ---------------------------------------------------------------- z-drift apply (zdrift.c:64-71) */
void fy_zdrift_apply(float *chunk, int nz_slab, int ny, int nx, int z0, const float *factor) {
    size_t n = (size_t)nz_slab * ny * nx;
    if (!n) return;
    fyg_stream *s = fyg_thread_stream();
    if (!s) return;
    float *d = (float *)fyg_ws(s, WS_IN, sizeof(float) * n);
    float *df = (float *)fyg_ws(s, WS_MISC, sizeof(float) * (size_t)nz_slab);
    if (!d || !df) return;
    if (fyg_h2d(d, chunk, sizeof(float) * n, s) || fyg_h2d(df, factor + z0, sizeof(float) * nz_slab, s)) return;
    if (fyg_zdrift_run(d, nz_slab, ny, nx, df, s)) return;
    if (fyg_d2h(chunk, d, sizeof(float) * n, s)) return;
    fyg_sync(s);
}

/* ---------------------------------------------------------------- transfers (paganin.c:15-91) */
#define HC_KEV_UM 1.23984199e-3   /* h*c in keV*micron */

double fy_paganin_transfer(double f_cyc_per_voxel, const fy_physics *p) {
    double lam = HC_KEV_UM / p->energy_kev;         /* micron */
    double dist = p->distance_mm * 1000.0;          /* micron */
    double px = p->pixel_um;
    if (px > 0.0 && px < 0.1) px *= 1000.0;         /* a millimetre value was passed */
    double fu = f_cyc_per_voxel / px;               /* cycles / micron */
    return 1.0 / (1.0 + p->delta_beta * lam * dist * M_PI * (fu * fu));
}

double fy_unsharp_transfer(double f_cyc_per_voxel, const fy_physics *p) {
    if (p->unsharp_coeff == 0.0 || p->unsharp_sigma == 0.0) return 1.0;
    double sg = p->unsharp_sigma;
    double g = exp(-2.0 * M_PI * M_PI * sg * sg * f_cyc_per_voxel * f_cyc_per_voxel);
    return 1.0 + p->unsharp_coeff * (1.0 - g);
}

double fy_recon_transfer(double f_cyc_per_voxel, const fy_physics *p) {
    double h = fy_paganin_transfer(f_cyc_per_voxel, p) * fy_unsharp_transfer(f_cyc_per_voxel, p);
    return h < 1e-4 ? 1e-4 : h;
}

double fy_auto_reg(const fy_physics *p, double base) {
    (void)p;
    return base > 0.0 ? base : 0.015;
}

/* partial delta/beta inversion from the Paganin low-pass strength at Nyquist */
double fy_auto_deltabeta_scale(const fy_physics *p) {
    const double lo = 0.0025, hi = 0.0070, s_lo = 0.35;
    double hn = fy_paganin_transfer(0.5, p);
    if (!isfinite(hn) || hn <= 0) return 1.0;
    if (hn <= lo) return s_lo;
    if (hn >= hi) return 1.0;
    return s_lo + (hn - lo) / (hi - lo) * (1.0 - s_lo);
}

/* spatial half-extent of H/(H^2+0.05) along one axis: inverse FFT of the
 * 256-sample response on the GPU, then the last lag above 1% of the peak,
 * clamped to [8, 96] (paganin.c:93-118) */
int fy_kernel_halo(const fy_physics *p) {
    enum { N = 256 };
    float re[N], im[N];
    for (int i = 0; i < N; i++) {
        double f = (i <= N / 2) ? (double)i / N : (double)(i - N) / N;
        double h = fy_recon_transfer(fabs(f), p);
        re[i] = (float)(h / (h * h + 0.05));
        im[i] = 0.0f;
    }
    fy_fft1d(re, im, N, +1);
    float peak = 0.0f;
    for (int i = 0; i < N; i++) peak = fabsf(re[i]) > peak ? fabsf(re[i]) : peak;
    int half = 8;
    for (int d = 1; d < N / 2; d++)
        if (fabsf(re[d]) > 0.01f * peak) half = d;
    if (half < 8) half = 8;
    if (half > 96) half = 96;
    return half;
}

/* ---------------------------------------------------------------- device hooks */
/* ---------------------------------------------------------------- phase correlation (phasecorr.c) */
/* Hann weights 0.5 - 0.5 cos(2 pi i / (n-1)) in fp32 with the C library's cosf,
 * as phasecorr.c:36-40 evaluates them; all ones without the window */
static void pc_window(float *w, int n, int window) {
    if (!window || n == 1) {
        for (int i = 0; i < n; i++) w[i] = 1.0f;
        return;
    }
    for (int i = 0; i < n; i++) w[i] = 0.5f - 0.5f * cosf(2.0f * (float)M_PI * i / (float)(n - 1));
}

/* Foroosh's estimator on one axis (phasecorr.c:150-176) */
static double pc_foroosh(double c0, double cm, double cp) {
    double c1 = cp >= cm ? cp : cm, sg = cp >= cm ? 1.0 : -1.0, d = 0.0;
    if (fabs(c1 - c0) > 1e-20 && fabs(c1 + c0) > 1e-20) {
        double d1 = c1 / (c1 - c0), d2 = c1 / (c1 + c0);
        d = sg * (fabs(d1) <= fabs(d2) ? d1 : d2);
    }
    return d > 0.5 ? 0.5 : (d < -0.5 ? -0.5 : d);
}

static int pc_dev(const float *df, const float *dm, int nz, int ny, int nx, double *shift, double *peak, int window,
                  fyg_stream *s) {
    int pz = fy_next_pow2(nz), py = fy_next_pow2(ny), px = fy_next_pow2(nx);
    size_t wb = sizeof(float) * (size_t)(nz + ny + nx);
    float *hw = (float *)malloc(wb);
    char *ws = (char *)fyg_ws(s, WS_SPEC, fyg_pc_ws_bytes(pz, py, px) + ((wb + 255) & ~(size_t)255));
    if (!hw || !ws) { free(hw); return 1; }
    pc_window(hw, nz, window);
    pc_window(hw + nz, ny, window);
    pc_window(hw + nz + ny, nx, window);
    float *dw = (float *)(ws + fyg_pc_ws_bytes(pz, py, px));
    fyg_pc_result r;
    int rc = fyg_h2d(dw, hw, wb, s) ||
             fyg_phase_correlate_run(df, dm, nz, ny, nx, pz, py, px, dw, dw + nz, dw + nz + ny, ws, &r, s);
    free(hw);
    if (rc) return 1;
    if (r.maxmag < 1e-20) {                                     /* a flat input: no phase */
        shift[0] = shift[1] = shift[2] = 0.0;
        if (peak) *peak = 0.0;
        return 1;
    }
    long long pk = r.index;
    int kz = (int)(pk / ((long long)py * px)), ky = (int)((pk / px) % py), kx = (int)(pk % px);
    double c0 = r.c[0];
    shift[0] = (kz <= pz / 2 ? kz : kz - pz) + pc_foroosh(c0, r.c[1], r.c[2]);
    shift[1] = (ky <= py / 2 ? ky : ky - py) + pc_foroosh(c0, r.c[3], r.c[4]);
    shift[2] = (kx <= px / 2 ? kx : kx - px) + pc_foroosh(c0, r.c[5], r.c[6]);
    if (peak) *peak = (double)r.best / (double)((size_t)pz * py * px);
    return 0;
}

int fy_phase_correlate(const float *fixed, const float *moving, int nz, int ny, int nx, double *shift, double *peak,
                       int window) {                            /* phasecorr.c:44-186 */
    if (!fixed || !moving || !shift || nz < 2 || ny < 2 || nx < 2) return 1;
    fyg_stream *s = fyg_thread_stream();
    size_t n = (size_t)nz * ny * nx;
    float *d = s ? (float *)fyg_ws(s, WS_IN, 2 * n * sizeof(float)) : NULL;
    if (!d || fyg_h2d(d, fixed, n * sizeof(float), s) || fyg_h2d(d + n, moving, n * sizeof(float), s)) return 1;
    return pc_dev(d, d + n, nz, ny, nx, shift, peak, window, s);
}

int fy_gpu_phase_correlate(const float *d_fixed, const float *d_moving, int nz, int ny, int nx, double *shift,
                           double *peak, int window, void *stream) {
    if (!d_fixed || !d_moving || !shift || nz < 2 || ny < 2 || nx < 2) return 1;
    fyg_stream *s = fyg_stream_wrap(stream);
    return !s || pc_dev(d_fixed, d_moving, nz, ny, nx, shift, peak, window, s);
}

/* ---------------------------------------------------------------- 2x LOD (downscale.c) */
/* input bytes per slab of fy_downscale2x (FYSICS_DS_SLAB_BYTES overrides: a
 * positive byte count, at least two planes are taken) */
static size_t ds_slab_bytes(void) {
    const char *e = getenv("FYSICS_DS_SLAB_BYTES");
    long long v = e ? atoll(e) : 0;
    return v > 0 ? (size_t)v : ((size_t)256 << 20);
}

void fy_downscale2x(const unsigned char *in, int nx, int ny, int nz, unsigned char *out, int *oxp, int *oyp,
                    int *ozp, fy_ds_method method, float alpha) {
    int ox = (nx + 1) / 2, oy = (ny + 1) / 2, oz = (nz + 1) / 2;
    if (ox < 1) ox = 1;
    if (oy < 1) oy = 1;
    if (oz < 1) oz = 1;
    if (oxp) *oxp = ox;
    if (oyp) *oyp = oy;
    if (ozp) *ozp = oz;
    if (nx < 1 || ny < 1 || nz < 1) return;
    /* Slabs of output planes through a bounded device workspace (each output
     * plane reads its two parent planes, so slabs are independent).  The
     * reference cannot fail here (downscale.c:303-320 returns void); if the
     * device does, out is zeroed and fy_gpu_last_error() says why. */
    const size_t plane_in = (size_t)ny * nx, plane_out = (size_t)oy * ox;
    long per = (long)(ds_slab_bytes() / (2 * plane_in));
    if (per < 1) per = 1;
    if (per > oz) per = oz;
    fyg_stream *s = fyg_thread_stream();
    unsigned char *d = s ? (unsigned char *)fyg_ws(s, WS_U8, 2 * plane_in * (size_t)per) : NULL;
    unsigned char *o = s ? (unsigned char *)fyg_ws(s, WS_OUT, plane_out * (size_t)per) : NULL;
    int rc = !d || !o;
    for (long z0 = 0; z0 < oz && !rc; z0 += per) {
        long nzs = oz - z0 < per ? oz - z0 : per;
        long zin = 2 * z0, nin = 2 * nzs < nz - zin ? 2 * nzs : nz - zin;
        rc = fyg_h2d(d, in + (size_t)zin * plane_in, (size_t)nin * plane_in, s) ||
             fyg_downscale2x_run(d, nx, ny, (int)nin, o, (int)method, alpha, s) ||
             fyg_d2h(out + (size_t)z0 * plane_out, o, (size_t)nzs * plane_out, s);
    }
    /* copies already queued must end before out is handed back (or zeroed) */
    if (s && fyg_sync(s)) rc = 1;
    if (rc) {
        const char *why = fyg_error_string();
        char msg[200];
        snprintf(msg, sizeof msg, "fy_downscale2x: output zeroed after a device failure (%s)",
                 why && *why ? why : "no device");
        memset(out, 0, (size_t)ox * oy * oz);
        fyg_set_error(msg);
    }
}

int fy_gpu_downscale2x(const unsigned char *d_in, int nx, int ny, int nz, unsigned char *d_out, fy_ds_method method,
                       float alpha, void *stream) {
    fyg_stream *s = fyg_stream_wrap(stream);
    return !s || fyg_downscale2x_run(d_in, nx, ny, nz, d_out, (int)method, alpha, s);
}

int fy_gpu_tuning(int mode) { return fyg_tune_mode(mode); }
int fy_gpu_tuning_report(char *buf, int len) { return fyg_tune_report(buf, len); }
int fy_gpu_tuning_save(const char *path) { return path ? fyg_tune_save(path) : 1; }
int fy_gpu_tuning_load(const char *path) { return path ? fyg_tune_load(path) : 1; }
void fy_gpu_tuning_clear(void) { fyg_tune_clear(); }

int fy_gpu_sync(void *stream) {
    fyg_stream *s = fyg_stream_wrap(stream);
    return s ? fyg_sync(s) : 1;
}
int fy_gpu_stream_release(void *stream) { return fyg_stream_release(stream); }
void fy_gpu_profile(int on) { fyg_prof_enable(on); }
int fy_gpu_profile_read(char *names, int names_len, double *ms, long *count, int max) {
    return fyg_prof_collect(names, names_len, ms, count, max);
}
void fy_gpu_profile_reset(void) { fyg_prof_reset(); }
const char *fy_gpu_last_error(void) { return fyg_error_string(); }

long fy_gpu_table_cache(long *stats, long limit) {
    size_t old = fyg_table_limit(limit > 0 ? (size_t)limit : 0);
    if (stats) fyg_table_stats(stats);
    return (long)old;
}

int fy_gpu_fft_planes(int mode) { return fyg_fft_plane_set(mode); }
int fy_gpu_batch_mix(int mode) { return fyg_mix_set(mode); }
int fy_gpu_fft_radix(int mode) { return fyg_fft_radix_mode(mode); }
