/* oracle.c -- CPU restatement of the fysics hot path (TEST INFRASTRUCTURE ONLY).
 * See oracle.h for what this is and what pins it.  Builds liboracle.so
 * (oracle/Makefile); symbols are prefixed orc_ so it never collides with the
 * reference or the product.
 */
#define _GNU_SOURCE
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define ORC_LUT_N 4096          /* FY_DECONV_LUT_N, paganin.c:138 */
#define ORC_MAX_GAIN 8.0        /* FY_MAX_DECONV_GAIN, paganin.c:50 */
#define ORC_HC_KEV_UM 1.23984199e-3

/* ------------------------------------------------------------------ sizes (fft.c:25-38) */
static int orc_is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
static int orc_next_pow2(int n) { int p = 1; while (p < n) p *= 2; return p; }
int orc_next_fft_size(int n) {
    int p2 = orc_next_pow2(n);
    int p3 = 3 * orc_next_pow2((n + 2) / 3);
    return (p3 >= n && p3 < p2) ? p3 : p2;
}

/* reflect-pad index (paganin.c:121-128) */
static int orc_reflect(int i, int n) {
    if (n == 1) return 0;
    while (i < 0 || i >= n) {
        if (i < 0) i = -i - 1;
        if (i >= n) i = 2 * n - i - 1;
    }
    return i;
}

/* ------------------------------------------------------------------ FFT, both precisions */
#define REAL float
#define SFX(x) x##_f32
#include "orc_fft_impl.h"
#include "orc_deconv_impl.h"
#include "orc_guided_impl.h"
#undef REAL
#undef SFX
#define REAL double
#define SFX(x) x##_f64
#include "orc_fft_impl.h"
#include "orc_deconv_impl.h"
#include "orc_guided_impl.h"
#undef REAL
#undef SFX

void orc_fft3d_normalize_f32(float *re, float *im, int nz, int ny, int nx) {
    size_t n = (size_t)nz * ny * nx;
    float s = 1.0f / (float)n;
    for (size_t i = 0; i < n; i++) { re[i] *= s; im[i] *= s; }
}

/* ------------------------------------------------------------------ transfers (paganin.c:17-91) */
double orc_paganin_transfer(double f, const fy_physics *p) {
    double lam = ORC_HC_KEV_UM / p->energy_kev;
    double dist = p->distance_mm * 1000.0;
    double px = p->pixel_um;
    if (px > 0.0 && px < 0.1) px *= 1000.0;
    double fu = f / px;
    return 1.0 / (1.0 + p->delta_beta * lam * dist * M_PI * (fu * fu));
}
double orc_unsharp_transfer(double f, const fy_physics *p) {
    if (p->unsharp_coeff == 0.0 || p->unsharp_sigma == 0.0) return 1.0;
    double sg = p->unsharp_sigma;
    double g = exp(-2.0 * M_PI * M_PI * sg * sg * f * f);
    return 1.0 + p->unsharp_coeff * (1.0 - g);
}
double orc_recon_transfer(double f, const fy_physics *p) {
    double h = orc_paganin_transfer(f, p) * orc_unsharp_transfer(f, p);
    return h < 1e-4 ? 1e-4 : h;
}
double orc_auto_reg(const fy_physics *p, double base) { (void)p; return base > 0.0 ? base : 0.015; }
double orc_auto_deltabeta_scale(const fy_physics *p) {
    double h = orc_paganin_transfer(0.5, p);
    if (!isfinite(h) || h <= 0) return 1.0;
    if (h <= 0.0025) return 0.35;
    if (h >= 0.0070) return 1.0;
    return 0.35 + (h - 0.0025) / (0.0070 - 0.0025) * (1.0 - 0.35);
}
/* impulse-response extent of H/(H^2+0.05) along one axis (paganin.c:93-118) */
int orc_kernel_halo(const fy_physics *p) {
    const int n = 256;
    float re[256], im[256];
    for (int i = 0; i < n; i++) {
        double f = (i <= n / 2) ? (double)i / n : (double)(i - n) / n;
        double h = orc_recon_transfer(fabs(f), p);
        re[i] = (float)(h / (h * h + 0.05)); im[i] = 0.0f;
    }
    orc_fft1d_f32(re, im, n, +1);
    float peak = 0.0f;
    for (int i = 0; i < n; i++) if (fabsf(re[i]) > peak) peak = fabsf(re[i]);
    int half = 8;
    for (int d = 1; d < n / 2; d++) if (fabsf(re[d]) > 0.01f * peak) half = d;
    return half < 8 ? 8 : (half > 96 ? 96 : half);
}

/* ------------------------------------------------------------------ deconv LUT (paganin.c:212-230,296-328) */
static double orc_weight(double fr, const fy_physics *p, double param, int kind) {
    if (kind == 0) {                                   /* wiener_weight */
        double h = orc_recon_transfer(fr, p);
        return h / (h * h + param);
    }
    double sg = p->psf_sigma_vox > 0 ? p->psf_sigma_vox : 1.0;   /* matched_transfer */
    double g = exp(-2.0 * M_PI * M_PI * sg * sg * (fr * fr));
    double hf = g * orc_unsharp_transfer(fr, p);
    return hf / (hf * hf + param);
}
static void orc_freq_tables(int pz, int py, int px, double fz2_scale,
                            double *fz2, double *fy2, double *fx2, double *umax) {
    int hx = px / 2 + 1;
    double mz = 0, my = 0, mx = 0;
    if (fz2_scale <= 0) fz2_scale = 1.0;
    for (int i = 0; i < pz; i++) {
        double f = (i <= pz / 2) ? (double)i / pz : (double)(i - pz) / pz;
        fz2[i] = f * f * fz2_scale; if (fz2[i] > mz) mz = fz2[i];
    }
    for (int i = 0; i < py; i++) {
        double f = (i <= py / 2) ? (double)i / py : (double)(i - py) / py;
        fy2[i] = f * f; if (fy2[i] > my) my = fy2[i];
    }
    for (int i = 0; i < hx; i++) { double f = (double)i / px; fx2[i] = f * f; if (fx2[i] > mx) mx = fx2[i]; }
    *umax = mz + my + mx;
}
static void orc_lut_d(int pz, int py, int px, const fy_physics *p, double param, int kind,
                      double cap, double umax, double *lutd, double *us) {
    *us = ORC_LUT_N / umax;
    for (int i = 0; i <= ORC_LUT_N; i++) {
        double w = orc_weight(sqrt((double)i / *us), p, param, kind);
        if (cap > 0 && w > cap) w = cap;
        lutd[i] = w;
    }
    (void)pz; (void)py; (void)px;
}
void orc_deconv_lut(int pz, int py, int px, const fy_physics *p, double param, int kind,
                    double cap, double fz2_scale, float *lut, double *us) {
    double *fz2 = malloc(sizeof(double) * pz), *fy2 = malloc(sizeof(double) * py);
    double *fx2 = malloc(sizeof(double) * (px / 2 + 1)), umax, lutd[ORC_LUT_N + 1];
    orc_freq_tables(pz, py, px, fz2_scale, fz2, fy2, fx2, &umax);
    orc_lut_d(pz, py, px, p, param, kind, cap, umax, lutd, us);
    for (int i = 0; i <= ORC_LUT_N; i++) lut[i] = (float)lutd[i];
    free(fz2); free(fy2); free(fx2);
}

static int orc_deconv_radial_any(const float *in, void *out, int f64, int nz, int ny, int nx,
                                 const fy_physics *p, double param, int kind, double cap,
                                 double fz2_scale) {
    int pz = orc_next_fft_size(nz), py = orc_next_fft_size(ny), px = orc_next_fft_size(nx);
    double *fz2 = malloc(sizeof(double) * pz), *fy2 = malloc(sizeof(double) * py);
    double *fx2 = malloc(sizeof(double) * (px / 2 + 1)), umax, us, lutd[ORC_LUT_N + 1];
    if (!fz2 || !fy2 || !fx2) { free(fz2); free(fy2); free(fx2); return 1; }
    orc_freq_tables(pz, py, px, fz2_scale, fz2, fy2, fx2, &umax);
    orc_lut_d(pz, py, px, p, param, kind, cap, umax, lutd, &us);
    int rc = f64 ? orc_deconv_core_f64(in, (double *)out, nz, ny, nx, lutd, us, fz2, fy2, fx2)
                 : orc_deconv_core_f32(in, (float *)out, nz, ny, nx, lutd, us, fz2, fy2, fx2);
    free(fz2); free(fy2); free(fx2);
    return rc;
}
int orc_deconv_radial_f32(const float *in, float *out, int nz, int ny, int nx, const fy_physics *p,
                          double param, int kind, double cap, double fz2_scale) {
    size_t n = (size_t)nz * ny * nx;
    float *tmp = malloc(sizeof(float) * n);      /* out may alias in */
    if (!tmp) return 1;
    int rc = orc_deconv_radial_any(in, tmp, 0, nz, ny, nx, p, param, kind, cap, fz2_scale);
    if (!rc) memcpy(out, tmp, sizeof(float) * n);
    free(tmp);
    return rc;
}
int orc_deconv_radial_f64(const float *in, double *out, int nz, int ny, int nx, const fy_physics *p,
                          double param, int kind, double cap, double fz2_scale) {
    return orc_deconv_radial_any(in, out, 1, nz, ny, nx, p, param, kind, cap, fz2_scale);
}
/* fy_deconvolve (paganin.c:301-308) */
int orc_deconvolve(const float *in, float *out, int nz, int ny, int nx, const fy_physics *p, double reg) {
    if (reg <= 0.0) reg = orc_auto_reg(p, 0.015);
    return orc_deconv_radial_f32(in, out, nz, ny, nx, p, reg, 0, ORC_MAX_GAIN, 1.0);
}
/* fy_deconvolve_matched (paganin.c:330-342) */
double orc_matched_fz2_scale(const fy_physics *p) {
    double sxy = p->psf_sigma_vox > 0 ? p->psf_sigma_vox : 1.0;
    if (p->psf_sigma_z_vox > 0 && sxy > 0) { double q = p->psf_sigma_z_vox / sxy; return q * q; }
    return 1.0;
}
int orc_deconvolve_matched(const float *in, float *out, int nz, int ny, int nx, const fy_physics *p,
                           double tikhonov) {
    if (tikhonov <= 0) tikhonov = 0.05;
    return orc_deconv_radial_f32(in, out, nz, ny, nx, p, tikhonov, 1, 0.0, orc_matched_fz2_scale(p));
}

/* ------------------------------------------------------------------ conversions (stream.c:27-63) */
void orc_u8_to_float(const unsigned char *in, float *out, size_t n) {
    const float k = 1.0f / 255.0f;
    for (size_t i = 0; i < n; i++) out[i] = in[i] * k;
}
static unsigned char orc_clamp_u8(float v) { return v < 0 ? 0 : (v > 255 ? 255 : (unsigned char)v); }
void orc_float_to_u8(const float *in, unsigned char *out, size_t n) {
    for (size_t i = 0; i < n; i++) out[i] = orc_clamp_u8(in[i] * 255.0f + 0.5f);
}
void orc_u8_to_phys(const unsigned char *in, float *out, size_t n, double lo, double hi) {
    float l = (float)lo, sc = (float)((hi - lo) / 255.0);
    for (size_t i = 0; i < n; i++) out[i] = l + in[i] * sc;
}
void orc_phys_to_u8(const float *in, unsigned char *out, size_t n, double lo, double hi) {
    double span = hi - lo;
    float k = span != 0.0 ? (float)(255.0 / span) : 0.0f, l = (float)lo;
    for (size_t i = 0; i < n; i++) out[i] = orc_clamp_u8((in[i] - l) * k + 0.5f);
}
float orc_phys_to_u8_level(double phys, double lo, double hi) {
    double span = hi - lo;
    return span == 0.0 ? 0.0f : (float)((phys - lo) / span * 255.0);
}

/* ------------------------------------------------------------------ quantiser (pipeline.c:53-65) */
float orc_dither01(long z, long y, long x) {
    unsigned int h = (unsigned int)(z * 73856093L) ^ (unsigned int)(y * 19349663L) ^ (unsigned int)(x * 83492791L);
    h ^= h >> 16; h *= 0x7feb352dU;
    h ^= h >> 15; h *= 0x846ca68bU;
    h ^= h >> 16;
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}
unsigned char orc_quant_u8(float v, long gz, long gy, long gx, int no_dither) {
    float t = v * 255.0f + (no_dither ? 0.5f : orc_dither01(gz, gy, gx));
    if (t <= 0.0f) return 0;
    int q = (int)t;
    return q > 255 ? 255 : (unsigned char)q;
}

/* ------------------------------------------------------------------ dering (dering.c) */
int orc_dering_init(fy_dering *d, long Z, long Y, long X, double cy, double cx, int slab_z, int ns) {
    memset(d, 0, sizeof(*d));
    d->Z = Z; d->Y = Y; d->X = X;
    d->cy = cy >= 0 ? cy : (Y - 1) / 2.0;
    d->cx = cx >= 0 ? cx : (X - 1) / 2.0;
    d->slab_z = slab_z > 0 ? slab_z : 512;
    d->nslab = (int)((Z + d->slab_z - 1) / d->slab_z);
    d->ns = ns > 0 ? ns : 8;
    double ry = d->cy > Y - 1 - d->cy ? d->cy : Y - 1 - d->cy;
    double rx = d->cx > X - 1 - d->cx ? d->cx : X - 1 - d->cx;
    d->nr = (int)sqrt(ry * ry + rx * rx) + 2;
    size_t nb = (size_t)d->nslab * d->ns * d->nr;
    d->sum = calloc(nb, sizeof(float));
    d->cnt = calloc(nb, sizeof(unsigned int));
    d->ring = calloc((size_t)d->nslab * d->nr, sizeof(float));
    if (!d->sum || !d->cnt || !d->ring) { orc_dering_free(d); return 1; }
    return 0;
}
void orc_dering_free(fy_dering *d) {
    free(d->sum); free(d->cnt); free(d->ring);
    d->sum = NULL; d->cnt = NULL; d->ring = NULL;
}
int orc_dering_tile_init(fy_dering *t, const fy_dering *d) {
    *t = *d;
    t->nslab = 1;
    t->sum = calloc((size_t)t->ns * t->nr, sizeof(float));
    t->cnt = calloc((size_t)t->ns * t->nr, sizeof(unsigned int));
    t->ring = NULL;
    if (!t->sum || !t->cnt) { free(t->sum); free(t->cnt); return 1; }
    return 0;
}
/* bin of (y,x): sector*nr + radius, -1 when the radius is out of range (dering.c:89-99) */
static int orc_ring_bin(const fy_dering *t, long gy_i, long gx_i) {
    double gy = (double)gy_i - t->cy, gx = (double)gx_i - t->cx;
    int ri = (int)(sqrt(gy * gy + gx * gx) + 0.5);
    if (ri >= t->nr) return -1;
    int q = (int)((atan2(gy, gx) + M_PI) * (t->ns / (2.0 * M_PI)));
    if (q >= t->ns) q = t->ns - 1;
    return q * t->nr + ri;
}
void orc_dering_accumulate_u8(fy_dering *t, const unsigned char *buf,
                              long y0, long x0, long dz, long dy, long dx, int ss) {
    if (ss < 1) ss = 1;
    for (long z = 0; z < dz; z++)
        for (long y = 0; y < dy; y += ss)
            for (long x = 0; x < dx; x += ss) {
                unsigned char v = buf[((size_t)z * dy + y) * dx + x];
                if (!v) continue;
                int b = orc_ring_bin(t, y0 + y, x0 + x);
                if (b < 0) continue;
                t->sum[b] += v;
                t->cnt[b]++;
            }
}
void orc_dering_merge_tile(fy_dering *d, int slab, const fy_dering *t) {
    if (slab < 0 || slab >= d->nslab) return;
    size_t nb = (size_t)d->ns * d->nr, off = (size_t)slab * nb;
    for (size_t i = 0; i < nb; i++) { d->sum[off + i] += t->sum[i]; d->cnt[off + i] += t->cnt[i]; }
}
static int orc_cmpf(const void *a, const void *b) {
    float x = *(const float *)a, y = *(const float *)b;
    return (x > y) - (x < y);
}
/* per slab: sector mean profiles, high-pass against the valid-bin box mean,
 * all-sectors-agree sign vote, median, floor/clamp (dering.c:136-194) */
long orc_dering_finalize(fy_dering *d, int hp_win, unsigned int min_cnt, double min_amp, double max_amp) {
    if (hp_win < 3) hp_win = 15;
    if (min_amp <= 0) min_amp = 0.5;
    if (max_amp <= 0) max_amp = 6.0;
    int ns = d->ns, nr = d->nr, hw = hp_win / 2;
    float *prof = malloc(sizeof(float) * (size_t)ns * nr), *hp = malloc(sizeof(float) * (size_t)ns * nr);
    unsigned char *ok = malloc((size_t)ns * nr);
    float vals[64];
    long found = 0;
    for (int s = 0; s < d->nslab; s++) {
        size_t so = (size_t)s * ns * nr;
        for (int q = 0; q < ns; q++) {
            float *pq = prof + (size_t)q * nr, *hq = hp + (size_t)q * nr;
            unsigned char *vq = ok + (size_t)q * nr;
            for (int r = 0; r < nr; r++) {
                unsigned int c = d->cnt[so + (size_t)q * nr + r];
                vq[r] = c >= min_cnt;
                pq[r] = vq[r] ? d->sum[so + (size_t)q * nr + r] / c : 0.0f;
            }
            for (int r = 0; r < nr; r++) {
                if (!vq[r]) { hq[r] = 0.0f; continue; }
                double acc = 0; int c = 0;
                int lo = r - hw < 0 ? 0 : r - hw, hi = r + hw >= nr ? nr - 1 : r + hw;
                for (int j = lo; j <= hi; j++) if (vq[j]) { acc += pq[j]; c++; }
                if (c >= hw) hq[r] = pq[r] - (float)(acc / c);
                else { hq[r] = 0.0f; vq[r] = 0; }
            }
        }
        float *ring = d->ring + (size_t)s * nr;
        for (int r = 0; r < nr; r++) {
            int nv = 0, pos = 0, neg = 0;
            for (int q = 0; q < ns; q++) {
                if (!ok[(size_t)q * nr + r]) continue;
                float v = hp[(size_t)q * nr + r];
                vals[nv++] = v;
                pos += v > 0; neg += v < 0;
            }
            if (nv < ns - 2 || (pos && neg)) { ring[r] = 0; continue; }
            qsort(vals, nv, sizeof(float), orc_cmpf);
            float m = (nv & 1) ? vals[nv / 2] : 0.5f * (vals[nv / 2 - 1] + vals[nv / 2]);
            if (fabsf(m) < (float)min_amp) { ring[r] = 0; continue; }
            if (m > (float)max_amp) m = (float)max_amp;
            if (m < (float)-max_amp) m = (float)-max_amp;
            ring[r] = m;
            found++;
        }
    }
    free(prof); free(hp); free(ok);
    d->have_rings = found >= 2;
    return found;
}
void orc_dering_apply(const fy_dering *d, float *f, long z0, long y0, long x0,
                      long dz, long dy, long dx, double scale) {
    if (!d->have_rings) return;
    float sc = (float)scale;
    for (long z = 0; z < dz; z++) {
        int slab = (int)((z0 + z) / d->slab_z);
        if (slab < 0) slab = 0;
        if (slab >= d->nslab) slab = d->nslab - 1;
        const float *ring = d->ring + (size_t)slab * d->nr;
        for (long y = 0; y < dy; y++) {
            double gy = (double)(y0 + y) - d->cy;
            for (long x = 0; x < dx; x++) {
                float *v = f + ((size_t)z * dy + y) * dx + x;
                double gx = (double)(x0 + x) - d->cx;
                int ri = (int)(sqrt(gy * gy + gx * gx) + 0.5);
                if (ri >= d->nr || *v <= 0.0f) continue;
                float w = *v - sc * ring[ri];
                *v = w > 0.0f ? w : 0.0f;
            }
        }
    }
}

/* ------------------------------------------------------------------ noise (noise.c) */
static void orc_box_sum_d(const float *in, double *out, double *tmp, int nz, int ny, int nx, int r) {
    size_t plane = (size_t)ny * nx, n = plane * nz;
    for (size_t row = 0; row < (size_t)nz * ny; row++) {      /* X (noise.c:24-34) */
        const float *ir = in + row * nx; double *orow = out + row * nx, s = 0;
        for (int x = 0; x <= r && x < nx; x++) s += ir[x];
        for (int x = 0; x < nx; x++) {
            orow[x] = s;
            if (x + r + 1 < nx) s += ir[x + r + 1];
            if (x - r >= 0) s -= ir[x - r];
        }
    }
    for (int a = 1; a < 3; a++) {                              /* Y then Z (noise.c:36-58) */
        memcpy(tmp, out, sizeof(double) * n);
        int len = a == 1 ? ny : nz;
        size_t st = a == 1 ? (size_t)nx : plane;
        for (size_t line = 0; line < n / (size_t)len; line++) {
            size_t b = a == 1 ? (line / nx) * plane + (line % nx) : line;
            double s = 0;
            for (int i = 0; i <= r && i < len; i++) s += tmp[b + (size_t)i * st];
            for (int i = 0; i < len; i++) {
                out[b + (size_t)i * st] = s;
                if (i + r + 1 < len) s += tmp[b + (size_t)(i + r + 1) * st];
                if (i - r >= 0) s -= tmp[b + (size_t)(i - r) * st];
            }
        }
    }
}
static int orc_cmpd(const void *a, const void *b) {
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}
static double orc_pct(double *v, int n, double pct) {      /* noise.c:66-73 (sorts v) */
    if (n <= 0) return 0.0;
    qsort(v, n, sizeof(double), orc_cmpd);
    double idx = pct / 100.0 * (n - 1);
    int lo = (int)idx; double fr = idx - lo;
    return lo + 1 < n ? v[lo] * (1 - fr) + v[lo + 1] * fr : v[lo];
}
int orc_estimate_noise(const float *in, int nz, int ny, int nx, int win, double flat_pct,
                       double ref_intensity, fy_noise_model *out) {
    if (!out) return 1;
    if (win < 3) win = 5;
    if (flat_pct <= 0) flat_pct = 10.0;
    int r = win / 2;
    size_t n = (size_t)nz * ny * nx;
    double *sm = malloc(sizeof(double) * n), *sq = malloc(sizeof(double) * n), *tmp = malloc(sizeof(double) * n);
    float *i2 = malloc(sizeof(float) * n);
    double *cp = malloc(sizeof(double) * n);
    if (!sm || !sq || !tmp || !i2 || !cp) { free(sm); free(sq); free(tmp); free(i2); free(cp); return 1; }
    for (size_t i = 0; i < n; i++) i2[i] = in[i] * in[i];
    orc_box_sum_d(in, sm, tmp, nz, ny, nx, r);
    orc_box_sum_d(i2, sq, tmp, nz, ny, nx, r);
    double wc = (double)win * win * win;
    for (size_t i = 0; i < n; i++) {          /* tmp := local mean, sq := local var */
        double m = sm[i] / wc, v = sq[i] / wc - m * m;
        tmp[i] = m; sq[i] = v > 0 ? v : 0;
    }
    memcpy(cp, tmp, sizeof(double) * n);
    double lo = orc_pct(cp, (int)n, 1.0);
    double hi = orc_pct(cp, (int)n, 99.0);    /* cp is sorted; re-sorting is a no-op */
    if (hi <= lo) hi = lo + 1e-6;
    enum { NB = 24 };
    double bi[NB], bf[NB]; long bc[NB];
    int used = 0;
    for (int k = 0; k < NB; k++) {            /* per-bin variance floor (noise.c:117-134) */
        double b0 = lo + (hi - lo) * k / NB, b1 = lo + (hi - lo) * (k + 1) / NB;
        long c = 0;
        for (size_t i = 0; i < n; i++) if (tmp[i] >= b0 && tmp[i] < b1) cp[c++] = sq[i];
        if (c < 200) continue;
        bf[used] = orc_pct(cp, (int)c, flat_pct);
        bi[used] = 0.5 * (b0 + b1);
        bc[used] = c;
        used++;
    }
    out->n_bins_used = used;
    double g = 0, b = 0;
    if (used >= 3) {                           /* IRLS (noise.c:139-161) */
        double w[NB];
        for (int i = 0; i < used; i++) w[i] = sqrt((double)bc[i]);
        for (int it = 0; it < 4; it++) {
            double s0 = 0, s1 = 0, sy = 0, s2 = 0, sxy = 0;
            for (int i = 0; i < used; i++) {
                s0 += w[i]; s1 += w[i] * bi[i]; sy += w[i] * bf[i];
                s2 += w[i] * bi[i] * bi[i]; sxy += w[i] * bi[i] * bf[i];
            }
            double det = s0 * s2 - s1 * s1;
            if (fabs(det) < 1e-20) break;
            g = (s0 * sxy - s1 * sy) / det;
            b = (s2 * sy - s1 * sxy) / det;
            double res[NB], rs[NB];
            for (int i = 0; i < used; i++) { res[i] = fabs(bf[i] - (g * bi[i] + b)); rs[i] = res[i]; }
            double med = orc_pct(rs, used, 50.0) + 1e-9;
            for (int i = 0; i < used; i++) {
                double q = res[i] / (2.0 * med);
                w[i] = sqrt((double)bc[i]) / (1.0 + q * q);
            }
        }
    } else if (used > 0) {
        double s = 0; for (int i = 0; i < used; i++) s += bf[i];
        b = s / used;
    } else {
        memcpy(cp, sq, sizeof(double) * n);
        b = orc_pct(cp, (int)n, flat_pct);
    }
    if (!isfinite(g)) g = 0;
    if (!isfinite(b)) b = 0;
    out->g = g; out->b = b;
    double vr = g * ref_intensity + b;
    if (vr < 0) vr = b > 0 ? b : 0;
    out->noise_ref = sqrt(vr) * 1.15;
    out->ref_intensity = ref_intensity;
    free(sm); free(sq); free(tmp); free(i2); free(cp);
    return 0;
}
double orc_guided_eps_for_noise(double noise_ref) {     /* noise.c:197-202 */
    double e = (3.0 * noise_ref) * (3.0 * noise_ref);
    return e > 1e-8 ? e : 1e-8;
}
/* lag-1 residual autocorrelation in flat 8^3 blocks (noise.c:210-269) */
int orc_noise_aniso(const float *v, int nz, int ny, int nx, double *rho_z, double *rho_xy) {
    const int B = 8, B3 = 512;
    int bz = nz / B, by = ny / B, bx = nx / B;
    long nb = (long)bz * by * bx;
    if (nb < 64) return 1;
    float *bv = malloc(sizeof(float) * nb);
    if (!bv) return 1;
    long nv = 0;
    float blk[512];
    for (int z = 0; z < bz; z++) for (int y = 0; y < by; y++) for (int x = 0; x < bx; x++) {
        double s = 0, s2 = 0; int pos = 1;
        for (int i = 0; i < B3; i++) {
            float t = v[((size_t)(z * B + i / 64) * ny + (y * B + (i / 8) % 8)) * nx + (size_t)x * B + i % 8];
            if (t <= 0) pos = 0;
            s += t; s2 += (double)t * t;
        }
        if (!pos) continue;
        double m = s / B3;
        bv[nv++] = (float)(s2 / B3 - m * m);
    }
    if (nv < 64) { free(bv); return 1; }
    /* the reference insertion-sorts bv ascending and takes bv[nv/5] */
    qsort(bv, nv, sizeof(float), orc_cmpf);
    float th = bv[nv / 5];
    double nzs = 0, nys = 0, nxs = 0, den = 0;
    long used = 0;
    for (int z = 0; z < bz; z++) for (int y = 0; y < by; y++) for (int x = 0; x < bx; x++) {
        double s = 0, s2 = 0; int pos = 1;
        for (int i = 0; i < B3; i++) {
            float t = v[((size_t)(z * B + i / 64) * ny + (y * B + (i / 8) % 8)) * nx + (size_t)x * B + i % 8];
            if (t <= 0) pos = 0;
            blk[i] = t; s += t; s2 += (double)t * t;
        }
        if (!pos) continue;
        double m = s / B3;
        if (s2 / B3 - m * m > th) continue;
        for (int i = 0; i < B3; i++) blk[i] -= (float)m;
        for (int i = 0; i < B3; i++) {
            int dz = i / 64, dy = (i / 8) % 8, dx = i % 8;
            double c = blk[i];
            den += c * c;
            if (dz + 1 < B) nzs += c * blk[i + 64];
            if (dy + 1 < B) nys += c * blk[i + 8];
            if (dx + 1 < B) nxs += c * blk[i + 1];
        }
        used++;
    }
    free(bv);
    if (used < 32 || den <= 0) return 1;
    double k = (double)B / (B - 1);
    *rho_z = k * nzs / den;
    *rho_xy = k * 0.5 * (nys + nxs) / den;
    return 0;
}

/* ------------------------------------------------------------------ pass-1 calibration */
void orc_hist_accumulate_u8(fy_hist_state *s, const unsigned char *chunk, size_t n) {
    for (size_t i = 0; i < n; i++) s->hist[chunk[i]]++;           /* stream.c:76-86 */
    s->total += (long)n;
}
int orc_hist_percentile_u8(const fy_hist_state *s, double pct) {  /* stream.c:96-101 */
    long target = (long)(pct / 100.0 * (double)s->total), acc = 0;
    for (int i = 0; i < 256; i++) {
        acc += s->hist[i];
        if (acc >= target) return i;
    }
    return 255;
}
int orc_hist_percentile_u8_inner(const fy_hist_state *s, double pct) {  /* stream.c:105-112 */
    long total = 0, acc = 0;
    for (int i = 1; i <= 254; i++) total += s->hist[i];
    if (total <= 0) return 0;
    long target = (long)(pct / 100.0 * (double)total);
    for (int i = 1; i <= 254; i++) {
        acc += s->hist[i];
        if (acc >= target) return i;
    }
    return 254;
}
void orc_norm_apply_u8(const unsigned char *in, float *out, size_t n,
                       unsigned char lo, unsigned char hi) {      /* stream.c:122-129 */
    float flo = lo, inv = hi > lo ? 1.0f / (float)(hi - lo) : 1.0f;
    for (size_t i = 0; i < n; i++) {
        float v = ((float)in[i] - flo) * inv;
        out[i] = v < 0 ? 0 : (v > 1 ? 1 : v);
    }
}
void orc_zdrift_accumulate(const float *chunk, int nz, int ny, int nx, int z0,
                           double *sums, long *counts, float thresh) {  /* zdrift.c:23-35 */
    size_t plane = (size_t)ny * nx;
    for (int z = 0; z < nz; z++) {
        double sm = 0;
        long c = 0;
        for (size_t i = 0; i < plane; i++) {
            float v = chunk[(size_t)z * plane + i];
            if (v > thresh) { sm += v; c++; }
        }
        sums[z0 + z] += sm;
        counts[z0 + z] += c;
    }
}
void orc_zdrift_finalize(const double *sums, const long *counts, int nz, float *factor) {
    double *mean = (double *)malloc(sizeof(double) * (nz > 0 ? nz : 1));  /* zdrift.c:41-62 */
    double gsum = 0;
    long gcount = 0;
    for (int z = 0; z < nz; z++) {
        mean[z] = counts[z] ? sums[z] / counts[z] : 0.0;
        gsum += sums[z];
        gcount += counts[z];
    }
    double gmean = gcount ? gsum / gcount : 1.0;
    int win = nz / 20 + 1;
    for (int z = 0; z < nz; z++) {
        double sm = 0;
        int c = 0;
        for (int d = -win; d <= win; d++) {
            int zz = z + d;
            if (zz >= 0 && zz < nz && mean[zz] > 0) { sm += mean[zz]; c++; }
        }
        double m = c ? sm / c : gmean;
        factor[z] = m > 1e-6 ? (float)(gmean / m) : 1.0f;
    }
    free(mean);
}
/* the pass-1 z-drift sums of one u8 ptile: 4x4 (y, x) subsample, v = u8/255 in
 * fp32, papyrus where v > 0.10f (pipeline.c:686-701) */
void orc_zdrift_u8_sub(const unsigned char *pb, long dz, long dy, long dx, long z0,
                       double *sums, long *counts) {
    const int ss = 4;
    for (long z = 0; z < dz; z++) {
        const unsigned char *sl = pb + (size_t)z * dy * dx;
        double sm = 0;
        long c = 0;
        for (long y = 0; y < dy; y += ss)
            for (long x = 0; x < dx; x += ss) {
                float v = sl[(size_t)y * dx + x] * (1.0f / 255.0f);
                if (v > 0.10f) { sm += v; c++; }
            }
        sums[z0 + z] += sm;
        counts[z0 + z] += c;
    }
}
double orc_flat_noise(const float *v, int nz, int ny, int nx, int blk) {  /* stream.c:174-200 */
    if (blk < 4) blk = 8;
    long nb = (long)(nz / blk) * (ny / blk) * (nx / blk);
    if (nb <= 0) return 0;
    double *stds = (double *)malloc(sizeof(double) * nb);
    long m = 0;
    for (int z = 0; z + blk <= nz; z += blk)
        for (int y = 0; y + blk <= ny; y += blk)
            for (int x = 0; x + blk <= nx; x += blk) {
                double sm = 0, s2 = 0;
                int c = 0;
                for (int dz = 0; dz < blk; dz++)
                    for (int dy = 0; dy < blk; dy++)
                        for (int dx = 0; dx < blk; dx++) {
                            float val = v[((size_t)(z + dz) * ny + (y + dy)) * nx + (x + dx)];
                            float sq = val * val;                 /* fp32 square */
                            sm += val;
                            s2 += sq;
                            c++;
                        }
                double mean = sm / c, var = s2 / c - mean * mean;
                if (mean > 0.1) stds[m++] = sqrt(var > 0 ? var : 0);
            }
    double nf = 0;
    if (m > 0) {
        qsort(stds, m, sizeof(double), orc_cmpd);
        nf = stds[m / 10];
    }
    free(stds);
    return nf;
}

/* ------------------------------------------------------------------ air-zero */
double orc_valley_depth(const long hist[256], int *dark_out, int *light_out, int *valley_out) {
    double hf[256];                                             /* stream.c:138-170 */
    for (int i = 0; i < 256; i++) {
        double sm = 0;
        int cnt = 0;
        for (int k = -2; k <= 2; k++) {
            int j = i + k;
            if (j >= 0 && j < 256) { sm += hist[j]; cnt++; }
        }
        hf[i] = sm / cnt;
    }
    hf[0] = hf[254] = hf[255] = 0;
    double mx = 0;
    for (int i = 0; i < 256; i++) if (hf[i] > mx) mx = hf[i];
    if (mx <= 0) return -1;
    int peaks[256], np = 0;
    for (int u = 3; u < 253; u++)
        if (hf[u] >= hf[u - 1] && hf[u] >= hf[u + 1] && hf[u] > 0.05 * mx) {
            if (np && u - peaks[np - 1] <= 10) {
                if (hf[u] > hf[peaks[np - 1]]) peaks[np - 1] = u;
            } else {
                peaks[np++] = u;
            }
        }
    if (np < 2) return -1;
    int p1 = peaks[0], p2 = peaks[1];
    for (int i = 0; i < np; i++) {
        if (hf[peaks[i]] > hf[p1]) { p2 = p1; p1 = peaks[i]; }
        else if (hf[peaks[i]] > hf[p2] && peaks[i] != p1) p2 = peaks[i];
    }
    int a = p1 < p2 ? p1 : p2, b = p1 < p2 ? p2 : p1;
    int v = a;
    double vmin = hf[a];
    for (int u = a; u <= b; u++) if (hf[u] < vmin) { vmin = hf[u]; v = u; }
    if (dark_out) *dark_out = a;
    if (light_out) *light_out = b;
    if (valley_out) *valley_out = v;
    double mn = hf[a] < hf[b] ? hf[a] : hf[b];
    return 1.0 - vmin / (mn + 1e-12);
}

/* 6-connected components of f > 0: zero those that touch no face of the region
 * and hold fewer than min_size voxels (pipeline.c:227-250) */
void orc_air_cc_despeckle(float *f, int nz, int ny, int nx, int min_size) {
    size_t n = (size_t)nz * ny * nx;
    int *lab = (int *)malloc(sizeof(int) * (n ? n : 1)), *q = (int *)malloc(sizeof(int) * (n ? n : 1));
    if (!lab || !q) { free(lab); free(q); return; }
    for (size_t i = 0; i < n; i++) lab[i] = f[i] > 0.0f ? -1 : 0;
    int plane = ny * nx;
    for (size_t s0 = 0; s0 < n; s0++) {
        if (lab[s0] != -1) continue;
        int head = 0, tail = 0, edge = 0;
        q[tail++] = (int)s0;
        lab[s0] = 1;
        while (head < tail) {
            int p = q[head++];
            int z = p / plane, r = p - z * plane, y = r / nx, x = r - y * nx;
            if (z == 0 || z == nz - 1 || y == 0 || y == ny - 1 || x == 0 || x == nx - 1) edge = 1;
            int nb[6] = {z > 0 ? p - plane : -1, z < nz - 1 ? p + plane : -1, y > 0 ? p - nx : -1,
                         y < ny - 1 ? p + nx : -1, x > 0 ? p - 1 : -1, x < nx - 1 ? p + 1 : -1};
            for (int k = 0; k < 6; k++)
                if (nb[k] >= 0 && lab[nb[k]] == -1) { lab[nb[k]] = 1; q[tail++] = nb[k]; }
        }
        if (!edge && tail < min_size)
            for (int k = 0; k < tail; k++) f[q[k]] = 0.0f;
    }
    free(lab);
    free(q);
}

/* process_tile stage (d) on f (after the guided filter) with orig = the
 * corrected tile before the deconvolution (pipeline.c:302-399) */
void orc_air_zero(float *f, const float *orig, int nz, int ny, int nx, const fy_cube_params *P) {
    size_t n = (size_t)nz * ny * nx;
    int sr = P->scratch_passes > 0 ? P->scratch_passes : 5;
    long hist[256];
    memset(hist, 0, sizeof(hist));
    int lz = (nz + 1) / 2, ly = (ny + 1) / 2, lx = (nx + 1) / 2;
    size_t nl = (size_t)lz * ly * lx;
    int lowres = lz >= 2 && ly >= 2 && lx >= 2;
    float *sm = (float *)malloc(sizeof(float) * (lowres ? nl : n));
    float *low = lowres ? (float *)malloc(sizeof(float) * nl) : NULL;
    if (!sm || (lowres && !low)) { free(sm); free(low); return; }
    if (lowres) {
        for (int z = 0; z < lz; z++)
            for (int y = 0; y < ly; y++)
                for (int x = 0; x < lx; x++)
                    low[((size_t)z * ly + y) * lx + x] = orig[((size_t)(2 * z) * ny + 2 * y) * nx + 2 * x];
        int rl = (sr + 1) / 2;
        if (rl < 1) rl = 1;
        orc_box_mean_f32(low, sm, lz, ly, lx, rl);
    } else {
        orc_box_mean_f32(orig, sm, nz, ny, nx, sr);
    }
    size_t nh = lowres ? nl : n;
    for (size_t i = 0; i < nh; i++) {
        int b = (int)(sm[i] * 255.0f + 0.5f);
        if (b < 0) b = 0;
        if (b > 255) b = 255;
        hist[b]++;
    }
    int dark = 0, light = 0, valley = 0;
    double d = orc_valley_depth(hist, &dark, &light, &valley);
    int cut;
    if (P->air_cut_u8 >= 0) {
        int gg = P->air_cut_u8, band = P->air_cut_band;
        if (d >= 0) {
            int pf = P->air_thresh > 0 ? (int)(P->air_thresh * 255 + 0.5) : 39;
            double a = P->air_cut_aggr;
            if (a < 0) a = 0;
            if (a > 2) a = 2;
            int local;
            if (valley <= pf) local = pf;
            else if (a <= 1.0) local = (int)(pf + a * (valley - pf) + 0.5);
            else local = (int)(valley + (a - 1.0) * (light > valley ? light - valley : 0) + 0.5);
            cut = local;
            if (cut < gg - band) cut = gg - band;
            if (cut > gg + band) cut = gg + band;
        } else {
            cut = gg;
        }
    } else if (d >= 0) {
        int mid = (dark + valley) / 2;
        cut = dark + 8 < mid ? dark + 8 : mid;
    } else {
        cut = (int)((P->air_thresh > 0 ? P->air_thresh : 0.05) * 255);
    }
    int pf2 = P->air_thresh > 0 ? (int)(P->air_thresh * 255 + 0.5) : 39;
    int spread = valley > pf2 ? (valley - pf2) / 2 : 8;
    int cu = cut + spread;
    if (valley > cut && cu > valley) cu = valley;
    int ct = cut - spread;
    if (ct < pf2) ct = pf2;
    float cutf_uni = cu / 255.0f, cutf_tex = ct / 255.0f;
    const float AIR_TEX = 0.03f;
    if (lowres) {
        float *row = (float *)malloc(sizeof(float) * lx);
        for (int z = 0; z < nz; z++) {
            int z0 = z >> 1;
            if (z0 > lz - 2) z0 = lz - 2;
            float zf = 0.5f * z - z0;
            if (zf > 1.0f) zf = 1.0f;
            for (int y = 0; y < ny; y++) {
                int y0 = y >> 1;
                if (y0 > ly - 2) y0 = ly - 2;
                float yf = 0.5f * y - y0;
                if (yf > 1.0f) yf = 1.0f;
                size_t r00 = ((size_t)z0 * ly + y0) * lx, r01 = r00 + lx;
                size_t r10 = r00 + (size_t)ly * lx, r11 = r10 + lx;
                float w00 = (1 - zf) * (1 - yf), w01 = (1 - zf) * yf, w10 = zf * (1 - yf), w11 = zf * yf;
                for (int k = 0; k < lx; k++)
                    row[k] = w00 * sm[r00 + k] + w01 * sm[r01 + k] + w10 * sm[r10 + k] + w11 * sm[r11 + k];
                float *frow = f + ((size_t)z * ny + y) * nx;
                const float *orow = orig + ((size_t)z * ny + y) * nx;
                for (int x = 0; x < nx; x++) {
                    int x0 = x >> 1;
                    if (x0 > lx - 2) x0 = lx - 2;
                    float fx = 0.5f * x - x0;
                    if (fx > 1.0f) fx = 1.0f;
                    float v = row[x0] + (row[x0 + 1] - row[x0]) * fx;
                    float tex = orow[x] - v;
                    if (tex < 0) tex = -tex;
                    if (v < (tex < AIR_TEX ? cutf_uni : cutf_tex)) frow[x] = 0.0f;
                }
            }
        }
        free(row);
    } else {
        for (size_t i = 0; i < n; i++) {
            float tex = orig[i] - sm[i];
            if (tex < 0) tex = -tex;
            if (sm[i] < (tex < AIR_TEX ? cutf_uni : cutf_tex)) f[i] = 0.0f;
        }
    }
    free(sm);
    free(low);
    if (P->air_min_component > 0) orc_air_cc_despeckle(f, nz, ny, nx, P->air_min_component);
}

/* ------------------------------------------------------------------ LOD downscale */
static unsigned char orc_ds_u8(float v) { return (unsigned char)(v < 0 ? 0 : v > 255 ? 255 : (v + 0.5f)); }
void orc_downscale2x(const unsigned char *in, int nx, int ny, int nz, unsigned char *out,
                     int *oxp, int *oyp, int *ozp, int method, float alpha) {   /* downscale.c:15-45 */
    int ox = (nx + 1) / 2, oy = (ny + 1) / 2, oz = (nz + 1) / 2;
    if (ox < 1) ox = 1;
    if (oy < 1) oy = 1;
    if (oz < 1) oz = 1;
    if (oxp) *oxp = ox;
    if (oyp) *oyp = oy;
    if (ozp) *ozp = oz;
    for (int z = 0; z < oz; z++)
        for (int y = 0; y < oy; y++)
            for (int x = 0; x < ox; x++) {
                unsigned char c[8];
                int k = 0;
                unsigned acc = 0;
                for (int dz = 0; dz < 2; dz++)
                    for (int dy = 0; dy < 2; dy++)
                        for (int dx = 0; dx < 2; dx++) {
                            int z2 = 2 * z + dz, y2 = 2 * y + dy, x2 = 2 * x + dx;
                            if (z2 >= nz) z2 = nz - 1;
                            if (y2 >= ny) y2 = ny - 1;
                            if (x2 >= nx) x2 = nx - 1;
                            c[k] = in[((size_t)z2 * ny + y2) * nx + x2];
                            acc |= c[k++];
                        }
                unsigned char r = 0;
                if (acc) {
                    float mean = (c[0] + c[1] + c[2] + c[3] + c[4] + c[5] + c[6] + c[7]) * 0.125f;
                    if (method == 0) {
                        r = orc_ds_u8(mean);
                    } else {
                        float best = mean, bestdev = -1.0f;
                        for (int i = 0; i < 8; i++) {
                            float d = (float)c[i] - mean, ad = d < 0 ? -d : d;
                            if (ad > bestdev) { bestdev = ad; best = c[i]; }
                        }
                        r = orc_ds_u8(mean + alpha * (best - mean));
                    }
                }
                out[((size_t)z * oy + y) * ox + x] = r;
            }
}

/* ------------------------------------------------------------------ phase correlation */
static void orc_hann(float *w, int n) {                        /* phasecorr.c:36-40 */
    if (n == 1) { w[0] = 1.0f; return; }
    for (int i = 0; i < n; i++) w[i] = 0.5f - 0.5f * cosf(2.0f * (float)M_PI * i / (float)(n - 1));
}

/* Foroosh's estimator along one axis from the peak c0 and its circular
 * neighbours cm, cp (phasecorr.c:150-176) */
static double orc_foroosh(double c0, double cm, double cp) {
    double c1 = cp >= cm ? cp : cm, sg = cp >= cm ? 1.0 : -1.0, d = 0.0;
    if (fabs(c1 - c0) > 1e-20 && fabs(c1 + c0) > 1e-20) {
        double d1 = c1 / (c1 - c0), d2 = c1 / (c1 + c0);
        d = sg * (fabs(d1) <= fabs(d2) ? d1 : d2);
    }
    if (d > 0.5) d = 0.5;
    if (d < -0.5) d = -0.5;
    return d;
}

int orc_phase_correlate(const float *fixed, const float *moving, int nz, int ny, int nx,
                        double *shift, double *peak, int window) {   /* phasecorr.c:44-186 */
    if (!fixed || !moving || !shift || nz < 2 || ny < 2 || nx < 2) return 1;
    int pz = orc_next_pow2(nz), py = orc_next_pow2(ny), px = orc_next_pow2(nx);
    size_t pn = (size_t)pz * py * px, n = (size_t)nz * ny * nx;
    float *fr = calloc(pn, sizeof(float)), *fi = calloc(pn, sizeof(float));
    float *mr = calloc(pn, sizeof(float)), *mi = calloc(pn, sizeof(float));
    float *wz = malloc(sizeof(float) * nz), *wy = malloc(sizeof(float) * ny), *wx = malloc(sizeof(float) * nx);
    if (window) { orc_hann(wz, nz); orc_hann(wy, ny); orc_hann(wx, nx); }
    else {
        for (int i = 0; i < nz; i++) wz[i] = 1;
        for (int i = 0; i < ny; i++) wy[i] = 1;
        for (int i = 0; i < nx; i++) wx[i] = 1;
    }
    double sf = 0, sm = 0;
    for (size_t i = 0; i < n; i++) { sf += fixed[i]; sm += moving[i]; }
    float mf = (float)(sf / n), mm = (float)(sm / n);
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++) {
            float wzy = wz[z] * wy[y];
            for (int x = 0; x < nx; x++) {
                float w = wzy * wx[x];
                size_t i = ((size_t)z * ny + y) * nx + x, o = ((size_t)z * py + y) * px + x;
                fr[o] = (fixed[i] - mf) * w;
                mr[o] = (moving[i] - mm) * w;
            }
        }
    free(wz); free(wy); free(wx);
    orc_fft3d_f32(fr, fi, pz, py, px, -1);
    orc_fft3d_f32(mr, mi, pz, py, px, -1);
    double maxmag = 0.0;
    for (size_t i = 0; i < pn; i++) {
        float cr = fr[i] * mr[i] + fi[i] * mi[i], ci = fr[i] * mi[i] - fi[i] * mr[i];
        float mag = sqrtf(cr * cr + ci * ci);
        if (mag > maxmag) maxmag = mag;
        if (mag > 1e-12f) { fr[i] = cr / mag; fi[i] = ci / mag; }
        else { fr[i] = 0.0f; fi[i] = 0.0f; }
    }
    free(mr); free(mi);
    if (maxmag < 1e-20) {
        free(fr); free(fi);
        shift[0] = shift[1] = shift[2] = 0.0;
        if (peak) *peak = 0.0;
        return 1;
    }
    orc_fft3d_f32(fr, fi, pz, py, px, 1);
    size_t pk = 0;
    float best = fr[0];
    for (size_t i = 1; i < pn; i++) if (fr[i] > best) { best = fr[i]; pk = i; }
    int kz = (int)(pk / ((size_t)py * px)), ky = (int)((pk / px) % py), kx = (int)(pk % px);
#define ORC_S(Z, Y, X) fr[((size_t)(Z) * py + (Y)) * px + (X)]
    double c0 = ORC_S(kz, ky, kx);
    double dz = orc_foroosh(c0, ORC_S((kz - 1 + pz) % pz, ky, kx), ORC_S((kz + 1) % pz, ky, kx));
    double dy = orc_foroosh(c0, ORC_S(kz, (ky - 1 + py) % py, kx), ORC_S(kz, (ky + 1) % py, kx));
    double dx = orc_foroosh(c0, ORC_S(kz, ky, (kx - 1 + px) % px), ORC_S(kz, ky, (kx + 1) % px));
#undef ORC_S
    shift[0] = (kz <= pz / 2 ? kz : kz - pz) + dz;
    shift[1] = (ky <= py / 2 ? ky : ky - py) + dy;
    shift[2] = (kx <= px / 2 ? kx : kx - px) + dx;
    if (peak) *peak = (double)best / (double)pn;
    free(fr); free(fi);
    return 0;
}

/* ------------------------------------------------------------------ chain pieces */
void orc_zdrift_apply(float *chunk, int nz, int ny, int nx, int z0, const float *factor) {
    size_t plane = (size_t)ny * nx;                              /* zdrift.c:64-71 */
    for (int z = 0; z < nz; z++) for (size_t i = 0; i < plane; i++) chunk[(size_t)z * plane + i] *= factor[z0 + z];
}
double orc_eps_for_tile(const fy_pipeline_cfg *cfg, double ecy, double ecx,
                        long y0, long x0, long ty, long tx) {  /* pipeline.c:447-456 */
    if (!cfg->have_eps_r || cfg->guided_eps <= 0 || cfg->eps_fn_med <= 0) return 0.0;
    double ry = y0 + ty / 2.0 - ecy, rx = x0 + tx / 2.0 - ecx;
    double fn = cfg->eps_fn_a + cfg->eps_fn_b * sqrt(ry * ry + rx * rx);
    double s = fn / cfg->eps_fn_med, s2 = s * s;
    if (s2 < 0.4) s2 = 0.4;
    if (s2 > 8.0) s2 = 8.0;
    return cfg->guided_eps * s2;
}
static double orc_scaled_db(const fy_pipeline_cfg *c, fy_physics *p) {   /* pipeline.c:89-104 */
    p->delta_beta = c->delta_beta; p->energy_kev = c->energy_kev; p->distance_mm = c->distance_mm;
    p->pixel_um = c->pixel_um; p->unsharp_sigma = c->unsharp_sigma; p->unsharp_coeff = c->unsharp_coeff;
    p->psf_sigma_vox = c->psf_sigma_vox;
    p->psf_sigma_z_vox = (c->psf_z_ratio > 1.0 && c->psf_sigma_vox > 0) ? c->psf_z_ratio * c->psf_sigma_vox : 0.0;
    p->delta_beta *= orc_auto_deltabeta_scale(p);
    return p->delta_beta;
}

/* The tile chain with explicit parameters (the fy_cube_params of
 * fysics_gpu.h): u8->f01, rings, z-drift, dewindow, deconv_radial with the
 * given weight/param/cap/fz2_scale, rewindow, dec-range, clamp, guided filter,
 * recenter, dithered quantisation of the inner tile (pipeline.c:252-297,
 * 413-421, 467-510).  ftile (optional) receives the float tile after the
 * guided filter, before recenter and quantisation. */
int orc_chain_cube(const fy_cube_params *P, const unsigned char *u8buf,
                   long rz0, long ry0, long rx0, long hz, long hy, long hx,
                   long z0, long y0, long x0, long tz, long ty, long tx,
                   unsigned char *out, int *all_air, float *ftile) {
    size_t hn = (size_t)hz * hy * hx;
    int any = 0;
    for (size_t i = 0; i < hn && !any; i++) any = u8buf[i] != 0;
    if (!any) {
        memset(out, 0, (size_t)tz * ty * tx);
        if (ftile) memset(ftile, 0, sizeof(float) * hn);
        if (all_air) *all_air = 1;
        return 0;
    }
    float *f = malloc(sizeof(float) * hn), *b1 = malloc(sizeof(float) * hn);
    if (!f || !b1) { free(f); free(b1); return 1; }
    orc_u8_to_float(u8buf, f, hn);
    if (P->dering && P->dering->have_rings) orc_dering_apply(P->dering, f, rz0, ry0, rx0, hz, hy, hx, 1.0 / 255.0);
    if (P->zdrift_factor) orc_zdrift_apply(f, (int)hz, (int)hy, (int)hx, (int)rz0, P->zdrift_factor);
    int nz = (int)hz, ny = (int)hy, nx = (int)hx;
    float *orig = NULL;                                        /* air-blend baseline */
    if (P->air_zero) {
        orig = malloc(sizeof(float) * hn);
        if (!orig) { free(f); free(b1); return 1; }
        memcpy(orig, f, sizeof(float) * hn);
    }
    if (P->deconv) {                                           /* (b) */
        int win = P->window_hi > P->window_lo;
        float wlo = (float)P->window_lo, wspan = (float)(P->window_hi - P->window_lo);
        if (win) for (size_t i = 0; i < hn; i++) f[i] = wlo + f[i] * wspan;
        int matched = P->deconv == 2;
        double param = P->deconv_param;
        if (param <= 0) param = matched ? 0.05 : 0.015;
        int rc = orc_deconv_radial_f32(f, b1, nz, ny, nx, &P->phys, param, matched, P->deconv_cap,
                                       P->fz2_scale > 0 ? P->fz2_scale : 1.0);
        if (rc == 0) {
            if (win) { float k = 1.0f / wspan; for (size_t i = 0; i < hn; i++) b1[i] = (b1[i] - wlo) * k; }
            if (P->have_dec_range && P->dec_hi - P->dec_lo > 1e-6) {
                float lo = (float)P->dec_lo, k = 1.0f / (float)(P->dec_hi - P->dec_lo);
                for (size_t i = 0; i < hn; i++) b1[i] = (b1[i] - lo) * k;
            }
            for (size_t i = 0; i < hn; i++) { float v = b1[i]; f[i] = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); }
        } else if (win) {
            float k = 1.0f / wspan;
            for (size_t i = 0; i < hn; i++) f[i] = (f[i] - wlo) * k;
        }
    }
    if (P->guided_eps > 0) {                                   /* (c) */
        int gs = P->guided_s == 0 ? 2 : P->guided_s;
        int r = P->guided_radius < 1 ? 2 : P->guided_radius;
        if (orc_guided_fast_f32(f, b1, nz, ny, nx, r, P->guided_eps, gs) == 0) memcpy(f, b1, sizeof(float) * hn);
    }
    if (orig) {                                                /* (d) */
        orc_air_zero(f, orig, nz, ny, nx, P);
        free(orig);
    }
    if (ftile) memcpy(ftile, f, sizeof(float) * hn);
    if (P->recenter && P->norm_hi > P->norm_lo) {              /* (d.5) */
        float mid = 0.5f * (float)(P->norm_lo + P->norm_hi) / 255.0f;
        float k = (2.0f * 0.375f) * 255.0f / (float)(P->norm_hi - P->norm_lo);
        for (size_t i = 0; i < hn; i++) {
            if (f[i] <= 0.0f) { f[i] = 0.0f; continue; }
            float v = 0.5f + (f[i] - mid) * k;
            f[i] = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
        }
    }
    long iz = z0 - rz0, iy = y0 - ry0, ix = x0 - rx0, o = 0;
    for (long z = iz; z < iz + tz; z++)
        for (long y = iy; y < iy + ty; y++)
            for (long x = ix; x < ix + tx; x++)
                out[o++] = orc_quant_u8(f[((size_t)z * hy + y) * hx + x], rz0 + z, ry0 + y, rx0 + x, P->no_dither);
    if (all_air) *all_air = 0;
    free(f); free(b1);
    return 0;
}

/* orc_chain_cube with the deconvolution and the guided filter evaluated in
 * fp64 (the same LUT semantics; orc_deconv_radial_f64, orc_guided_*_f64): the
 * reference to hold the fp32 chain's float tile to relative L2 1e-5.  The u8
 * prologue (u8/255, ring subtraction, z-drift, dewindow) stays fp32, as the
 * reference's input to the deconvolution is fp32.  ftile gets the double tile
 * after the guided filter (before recenter); out the inner tile quantised from
 * it (recenter and dither as orc_chain_cube).  Air-zero is not covered (1). */
int orc_chain_cube_f64(const fy_cube_params *P, const unsigned char *u8buf,
                       long rz0, long ry0, long rx0, long hz, long hy, long hx,
                       long z0, long y0, long x0, long tz, long ty, long tx,
                       unsigned char *out, double *ftile) {
    if (P->air_zero) return 1;
    size_t hn = (size_t)hz * hy * hx;
    int any = 0;
    for (size_t i = 0; i < hn && !any; i++) any = u8buf[i] != 0;
    if (!any) {
        memset(out, 0, (size_t)tz * ty * tx);
        if (ftile) memset(ftile, 0, sizeof(double) * hn);
        return 0;
    }
    float *f = malloc(sizeof(float) * hn);
    double *d = malloc(sizeof(double) * hn), *d1 = malloc(sizeof(double) * hn);
    if (!f || !d || !d1) { free(f); free(d); free(d1); return 1; }
    orc_u8_to_float(u8buf, f, hn);
    if (P->dering && P->dering->have_rings) orc_dering_apply(P->dering, f, rz0, ry0, rx0, hz, hy, hx, 1.0 / 255.0);
    if (P->zdrift_factor) orc_zdrift_apply(f, (int)hz, (int)hy, (int)hx, (int)rz0, P->zdrift_factor);
    int nz = (int)hz, ny = (int)hy, nx = (int)hx, rc = 0;
    if (P->deconv) {
        int win = P->window_hi > P->window_lo;
        float wlo = (float)P->window_lo, wspan = (float)(P->window_hi - P->window_lo);
        if (win) for (size_t i = 0; i < hn; i++) f[i] = wlo + f[i] * wspan;
        int matched = P->deconv == 2;
        double param = P->deconv_param;
        if (param <= 0) param = matched ? 0.05 : 0.015;
        rc = orc_deconv_radial_f64(f, d, nz, ny, nx, &P->phys, param, matched, P->deconv_cap,
                                   P->fz2_scale > 0 ? P->fz2_scale : 1.0);
        if (!rc) {
            double lo = win ? P->window_lo : 0.0, k = win ? 1.0 / (P->window_hi - P->window_lo) : 1.0;
            int dec = P->have_dec_range && P->dec_hi - P->dec_lo > 1e-6;
            for (size_t i = 0; i < hn; i++) {
                double v = (d[i] - lo) * k;
                if (dec) v = (v - P->dec_lo) / (P->dec_hi - P->dec_lo);
                d[i] = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
            }
        }
    } else {
        for (size_t i = 0; i < hn; i++) d[i] = f[i];
    }
    if (!rc && P->guided_eps > 0) {
        int gs = P->guided_s == 0 ? 2 : P->guided_s;
        int r = P->guided_radius < 1 ? 2 : P->guided_radius;
        rc = gs <= 1 ? orc_guided_exact_f64(d, d1, nz, ny, nx, r, P->guided_eps)
                     : orc_guided_fast_f64(d, d1, nz, ny, nx, r, P->guided_eps, gs);
        if (!rc) memcpy(d, d1, sizeof(double) * hn);
    }
    if (!rc) {
        if (ftile) memcpy(ftile, d, sizeof(double) * hn);
        float mid = 0.5f * (float)(P->norm_lo + P->norm_hi) / 255.0f;
        float k = P->norm_hi > P->norm_lo ? (2.0f * 0.375f) * 255.0f / (float)(P->norm_hi - P->norm_lo) : 0.0f;
        int rec = P->recenter && P->norm_hi > P->norm_lo;
        long iz = z0 - rz0, iy = y0 - ry0, ix = x0 - rx0, o = 0;
        for (long z = iz; z < iz + tz; z++)
            for (long y = iy; y < iy + ty; y++)
                for (long x = ix; x < ix + tx; x++) {
                    float v = (float)d[((size_t)z * hy + y) * hx + x];
                    if (rec) {
                        if (v <= 0.0f) v = 0.0f;
                        else {
                            v = 0.5f + (v - mid) * k;
                            v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
                        }
                    }
                    out[o++] = orc_quant_u8(v, rz0 + z, ry0 + y, rx0 + x, P->no_dither);
                }
    }
    free(f); free(d); free(d1);
    return rc;
}

/* fy_cube_params derived from a pipeline cfg as fy_process_buffer does
 * (pipeline.c:89-104,264-297,447-456) */
int orc_cube_params_from_cfg(const fy_pipeline_cfg *cfg, long y0, long x0, long ty, long tx,
                             long vol_y, long vol_x, fy_cube_params *o) {
    memset(o, 0, sizeof *o);
    if (cfg->do_musica) return 1;
    orc_scaled_db(cfg, &o->phys);
    if (cfg->do_deconv) {
        int matched = cfg->use_matched_deconv && cfg->psf_sigma_vox > 0;
        o->deconv = matched ? 2 : 1;
        o->deconv_param = matched ? cfg->deconv_tikhonov : 0.015;
        o->deconv_cap = matched ? 0.0 : ORC_MAX_GAIN;
        o->fz2_scale = matched ? orc_matched_fz2_scale(&o->phys) : 1.0;
        o->window_lo = cfg->window_lo; o->window_hi = cfg->window_hi;
        o->have_dec_range = cfg->have_dec_range; o->dec_lo = cfg->dec_lo; o->dec_hi = cfg->dec_hi;
    }
    double ecy = cfg->dering_cy > 0 ? cfg->dering_cy : (vol_y - 1) / 2.0;
    double ecx = cfg->dering_cx > 0 ? cfg->dering_cx : (vol_x - 1) / 2.0;
    double eps_t = orc_eps_for_tile(cfg, ecy, ecx, y0, x0, ty, tx);
    if (cfg->guided_eps > 0 || eps_t > 0) {
        o->guided_eps = eps_t > 0 ? eps_t : cfg->guided_eps;
        o->guided_radius = 2;
        o->guided_s = cfg->guided_subsample == 0 ? 2 : cfg->guided_subsample;
    }
    o->no_dither = cfg->no_dither;
    if (cfg->have_norm && cfg->norm_hi > cfg->norm_lo) { o->recenter = 1; o->norm_lo = cfg->norm_lo; o->norm_hi = cfg->norm_hi; }
    if (cfg->have_dering && cfg->dering) o->dering = cfg->dering;
    if (cfg->do_air_zero) {                                    /* pipeline.c:302-399 */
        o->air_zero = 1;
        o->air_cut_u8 = cfg->air_cut_u8; o->air_cut_band = cfg->air_cut_band;
        o->air_thresh = cfg->air_thresh; o->air_cut_aggr = cfg->air_cut_aggr;
        o->air_min_component = cfg->air_min_component; o->scratch_passes = cfg->scratch_passes;
    }
    if (cfg->have_zdrift && cfg->zdrift_factor) { o->zdrift_factor = cfg->zdrift_factor; o->vol_z = cfg->vol_z; }
    return 0;
}

/* fy_process_buffer restricted to the hot-path stages (pipeline.c:467-510).
 * Returns 1 for a cfg that enables the air-zero or MUSICA stages. */
int orc_process_buffer_f(const fy_pipeline_cfg *cfg, const unsigned char *u8buf,
                         long rz0, long ry0, long rx0, long hz, long hy, long hx,
                         long z0, long y0, long x0, long tz, long ty, long tx,
                         long vol_y, long vol_x, unsigned char *out, int *all_air, float *ftile) {
    fy_cube_params P;
    if (orc_cube_params_from_cfg(cfg, y0, x0, ty, tx, vol_y, vol_x, &P)) return 1;
    return orc_chain_cube(&P, u8buf, rz0, ry0, rx0, hz, hy, hx, z0, y0, x0, tz, ty, tx, out, all_air, ftile);
}
int orc_process_buffer(const fy_pipeline_cfg *cfg, const unsigned char *u8buf,
                       long rz0, long ry0, long rx0, long hz, long hy, long hx,
                       long z0, long y0, long x0, long tz, long ty, long tx,
                       long vol_y, long vol_x, unsigned char *out, int *all_air) {
    return orc_process_buffer_f(cfg, u8buf, rz0, ry0, rx0, hz, hy, hx, z0, y0, x0, tz, ty, tx,
                                vol_y, vol_x, out, all_air, NULL);
}

/* ------------------------------------------------------------------ pass-1 commits */
/* median of a copy, insertion-sorted (pipeline.c:74-78 sorts in place with qsort) */
static double orc_median(const double *v, int n) {
    if (n <= 0) return 0.0;
    double *t = malloc(sizeof(double) * (size_t)n);
    for (int i = 0; i < n; i++) {
        int j = i;
        while (j > 0 && t[j - 1] > v[i]) { t[j] = t[j - 1]; j--; }
        t[j] = v[i];
    }
    double m = (n & 1) ? t[n / 2] : 0.5 * (t[n / 2 - 1] + t[n / 2]);
    free(t);
    return m;
}

void orc_calib_commit_eps(fy_pipeline_cfg *cfg, const double *rads, const double *floors, int n) {
    cfg->have_eps_r = 0;                                       /* pipeline.c:856-877 */
    if (n >= 8) {
        double sr = 0, sf2 = 0, srr = 0, srf = 0, sff = 0;
        for (int i = 0; i < n; i++) {
            sr += rads[i]; sf2 += floors[i];
            srr += rads[i] * rads[i]; srf += rads[i] * floors[i]; sff += floors[i] * floors[i];
        }
        double nn = n, den = nn * srr - sr * sr;
        if (den > 1e-9) {
            double b = (nn * srf - sr * sf2) / den;
            double a = (sf2 - b * sr) / nn;
            double cden = sqrt((nn * srr - sr * sr) * (nn * sff - sf2 * sf2));
            double corr = cden > 1e-12 ? (nn * srf - sr * sf2) / cden : 0;
            double rmax = 0;
            for (int i = 0; i < n; i++) if (rads[i] > rmax) rmax = rads[i];
            double f0 = a, f1 = a + b * rmax;
            if (b > 0 && corr >= 0.5 && f0 > 0 && f1 / f0 >= 1.3) {
                cfg->eps_fn_a = a; cfg->eps_fn_b = b; cfg->have_eps_r = 1;
            }
        }
    }
    double med = n ? orc_median(floors, n) : 0.015;            /* pipeline.c:878-891 */
    cfg->eps_fn_med = med;
    cfg->noise_floor = med * 255.0;
    if (cfg->guided_eps <= 0 && cfg->denoise_k >= 0) {
        double k = cfg->denoise_k > 0 ? cfg->denoise_k : 4.2;
        cfg->guided_eps = (k * med) * (k * med);
    }
}

double orc_aniso_ratio(double arz, double arxy) {              /* pipeline.c:736-739 */
    if (arz > 0.05 && arz < 0.95 && arxy > 0.05 && arxy < 0.95) {
        double ratio = sqrt(log(1.0 / arxy) / log(1.0 / arz));
        if (ratio > 0.5 && ratio < 3.0) return ratio;
    }
    return -1.0;
}

void orc_calib_commit_aniso(fy_pipeline_cfg *cfg, const double *ratios, int n) {   /* pipeline.c:894-900 */
    if (n >= 5) {
        double ratio = orc_median(ratios, n);
        if (ratio >= 1.05) cfg->psf_z_ratio = ratio < 1.6 ? ratio : 1.6;
    }
}
