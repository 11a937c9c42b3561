/* fy_calib.c -- pass-1 calibration statistics (stream.c:65-200, zdrift.c,
 * pipeline.c:576-701).
 *
 * The per-voxel work (histogram, z-drift sums, flat-block stds, normalisation
 * apply) runs on the GPU; what remains on the host is O(256) or O(slices) logic
 * on the small results: percentiles, the z-drift smoothing, the 10th-percentile
 * selection of the block stds.
 *
 * The host-buffer entry points keep the reference signatures (void functions
 * included).  When the device fails they leave their outputs untouched and
 * fy_gpu_last_error() says why; fy_flat_noise then returns -1, a value no
 * successful call returns.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "fyh.h"

static size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

/* ------------------------------------------------------------------ histogram */
void fy_hist_init(fy_hist_state *s) { memset(s, 0, sizeof(*s)); }      /* stream.c:70-73 */

void fy_hist_accumulate_u8(fy_hist_state *st, const unsigned char *chunk, size_t n) {
    if (!n) return;                                                     /* stream.c:76-86 */
    fyg_stream *s = fyg_thread_stream();
    unsigned long long hh[256];
    unsigned char *d = s ? (unsigned char *)fyg_ws(s, WS_U8, n) : NULL;
    char *m = s ? (char *)fyg_ws(s, WS_MISC, al256(sizeof(hh)) + fyg_hist_part_bytes(n)) : NULL;
    if (!d || !m) return;
    unsigned long long *dh = (unsigned long long *)m;
    unsigned int *part = (unsigned int *)(m + al256(sizeof(hh)));
    if (fyg_h2d(d, chunk, n, s) || fyg_zero(dh, sizeof(hh), s) || fyg_hist_u8_run(d, n, part, dh, s) ||
        fyg_d2h(hh, dh, sizeof(hh), s) || fyg_sync(s))
        return;
    for (int b = 0; b < 256; b++) st->hist[b] += (long)hh[b];
    st->total += (long)n;
}

void fy_hist_merge(fy_hist_state *dst, const fy_hist_state *src) {     /* stream.c:89-92 */
    for (int b = 0; b < 256; b++) dst->hist[b] += src->hist[b];
    dst->total += src->total;
}

/* first bin whose running count reaches (long)(pct/100 * total) (stream.c:95-101) */
int fy_hist_percentile_u8(const fy_hist_state *s, double pct) {
    long target = (long)(pct / 100.0 * (double)s->total), acc = 0;
    for (int b = 0; b < 256; b++) {
        acc += s->hist[b];
        if (acc >= target) return b;
    }
    return 255;
}

/* the same over bins 1..254 (stream.c:105-112) */
int fy_hist_percentile_u8_inner(const fy_hist_state *s, double pct) {
    long total = 0, acc = 0;
    for (int b = 1; b <= 254; b++) total += s->hist[b];
    if (total <= 0) return 0;
    long target = (long)(pct / 100.0 * (double)total);
    for (int b = 1; b <= 254; b++) {
        acc += s->hist[b];
        if (acc >= target) return b;
    }
    return 254;
}

void fy_norm_finalize(const fy_hist_state *s, double lo_pct, double hi_pct, unsigned char *lo_out,
                      unsigned char *hi_out) {                          /* stream.c:116-120 */
    *lo_out = (unsigned char)fy_hist_percentile_u8(s, lo_pct);
    *hi_out = (unsigned char)fy_hist_percentile_u8(s, hi_pct);
}

void fy_norm_apply_u8(const unsigned char *in, float *out, size_t n, unsigned char lo, unsigned char hi) {
    if (!n) return;                                                     /* stream.c:122-129 */
    fyg_stream *s = fyg_thread_stream();
    unsigned char *d = s ? (unsigned char *)fyg_ws(s, WS_U8, n) : NULL;
    float *f = s ? (float *)fyg_ws(s, WS_TILE, n * sizeof(float)) : NULL;
    if (!d || !f) return;
    if (fyg_h2d(d, in, n, s) || fyg_norm_u8_run(d, f, n, lo, hi, s) || fyg_d2h(out, f, n * sizeof(float), s))
        return;
    fyg_sync(s);
}

/* bimodal valley depth (stream.c:138-170): the same routine the device air-zero
 * cut runs, compiled for the host */
double fy_valley_depth(const long hist[256], int *dark_out, int *light_out, int *valley_out) {
    return fyg_valley_depth(hist, dark_out, light_out, valley_out);
}

/* ------------------------------------------------------------------ flat-block noise */
static int cmp_double(const void *a, const void *b) {
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

/* median of n values (the mean of the middle two for even n) of a sorted copy */
static double median_of(const double *v, int n) {
    if (n <= 0) return 0.0;
    double *t = (double *)malloc(sizeof(double) * (size_t)n);
    if (!t) return 0.0;
    memcpy(t, v, sizeof(double) * (size_t)n);
    qsort(t, (size_t)n, sizeof(double), cmp_double);
    double m = (n & 1) ? t[n / 2] : 0.5 * (t[n / 2 - 1] + t[n / 2]);
    free(t);
    return m;
}

/* pipeline.c:851-891: radial eps fit, median floor, guided eps commit */
void fy_calib_commit_eps(fy_pipeline_cfg *cfg, const double *rads, const double *floors, int n) {
    cfg->have_eps_r = 0;
    if (n >= 8) {
        double sr = 0, sf = 0, srr = 0, srf = 0, sff = 0;
        for (int i = 0; i < n; i++) {
            sr += rads[i];
            sf += floors[i];
            srr += rads[i] * rads[i];
            srf += rads[i] * floors[i];
            sff += floors[i] * floors[i];
        }
        const double nn = n, den = nn * srr - sr * sr;
        if (den > 1e-9) {
            const double b = (nn * srf - sr * sf) / den;
            const double a = (sf - b * sr) / nn;
            const double cden = sqrt((nn * srr - sr * sr) * (nn * sff - sf * sf));
            const double corr = cden > 1e-12 ? (nn * srf - sr * sf) / cden : 0;
            double rmax = 0;
            for (int i = 0; i < n; i++)
                if (rads[i] > rmax) rmax = rads[i];
            const double f0 = a, f1 = a + b * rmax;
            if (b > 0 && corr >= 0.5 && f0 > 0 && f1 / f0 >= 1.3) {
                cfg->eps_fn_a = a;
                cfg->eps_fn_b = b;
                cfg->have_eps_r = 1;
            }
        }
    }
    const double med = n > 0 ? median_of(floors, n) : 0.015;
    cfg->eps_fn_med = med;
    cfg->noise_floor = med * 255.0;
    if (cfg->guided_eps <= 0 && cfg->denoise_k >= 0) {
        const double k = cfg->denoise_k > 0 ? cfg->denoise_k : 4.2;
        cfg->guided_eps = (k * med) * (k * med);
    }
}

/* pipeline.c:736-739 */
double fy_aniso_ratio(double arz, double arxy) {
    if (!(arz > 0.05 && arz < 0.95 && arxy > 0.05 && arxy < 0.95)) return -1.0;
    const double ratio = sqrt(log(1.0 / arxy) / log(1.0 / arz));
    return (ratio > 0.5 && ratio < 3.0) ? ratio : -1.0;
}

/* pipeline.c:894-900 */
void fy_calib_commit_aniso(fy_pipeline_cfg *cfg, const double *ratios, int n) {
    if (n < 5) return;
    const double r = median_of(ratios, n);
    if (r >= 1.05) cfg->psf_z_ratio = r < 1.6 ? r : 1.6;
}

/* the 10th percentile (element m/10 of the ascending order) of the kept stds */
static double flat_select(double *std, long nb) {
    long m = 0;
    for (long i = 0; i < nb; i++)
        if (std[i] >= 0.0) std[m++] = std[i];
    if (!m) return 0.0;
    qsort(std, (size_t)m, sizeof(double), cmp_double);
    return std[m / 10];
}

static int flat_noise_dev(const float *d_v, int nz, int ny, int nx, int blk, double *out, fyg_stream *s) {
    if (blk < 4) blk = 8;                                               /* stream.c:174-200 */
    long nb = (long)(nz / blk) * (ny / blk) * (nx / blk);
    if (nz <= 0 || ny <= 0 || nx <= 0 || nb <= 0) { *out = 0.0; return 0; }
    double *d_std = (double *)fyg_ws(s, WS_MISC, sizeof(double) * (size_t)nb);
    double *h = (double *)malloc(sizeof(double) * (size_t)nb);
    int rc = !d_std || !h || fyg_flat_blocks_run(d_v, nz, ny, nx, blk, d_std, s) ||
             fyg_d2h(h, d_std, sizeof(double) * (size_t)nb, s) || fyg_sync(s);
    if (!rc) *out = flat_select(h, nb);
    free(h);
    return rc;
}

double fy_flat_noise(const float *v, int nz, int ny, int nx, int blk) {
    fyg_stream *s = fyg_thread_stream();
    if (!s) return -1.0;
    size_t n = (nz > 0 && ny > 0 && nx > 0) ? (size_t)nz * ny * nx : 0;
    float *d = n ? (float *)fyg_ws(s, WS_IN, n * sizeof(float)) : NULL;
    if (n && (!d || fyg_h2d(d, v, n * sizeof(float), s))) return -1.0;
    double r;
    return flat_noise_dev(d, nz, ny, nx, blk, &r, s) ? -1.0 : r;
}

int fy_gpu_flat_noise(const float *d_v, int nz, int ny, int nx, int blk, double *out, void *stream) {
    fyg_stream *s = fyg_stream_wrap(stream);
    return !s || !out || flat_noise_dev(d_v, nz, ny, nx, blk, out, s);
}

/* ------------------------------------------------------------------ z-drift */
static int zdrift_sums_dev(const float *d, int nz, int ny, int nx, float thresh, double *hs, long long *hc,
                           fyg_stream *s) {
    char *m = (char *)fyg_ws(s, WS_MISC, 2 * al256(sizeof(double) * (size_t)nz));
    if (!m) return 1;
    double *ds = (double *)m;
    long long *dc = (long long *)(m + al256(sizeof(double) * (size_t)nz));
    return fyg_zero(m, 2 * al256(sizeof(double) * (size_t)nz), s) ||
           fyg_zdrift_f32_run(d, nz, ny, nx, thresh, ds, dc, s) ||
           fyg_d2h(hs, ds, sizeof(double) * (size_t)nz, s) || fyg_d2h(hc, dc, sizeof(long long) * (size_t)nz, s) ||
           fyg_sync(s);
}

void fy_zdrift_accumulate(const float *chunk, int nz, int ny, int nx, int z0, double *sums, long *counts,
                          float thresh) {                               /* zdrift.c:23-35 */
    if (nz <= 0 || ny <= 0 || nx <= 0) return;
    fyg_stream *s = fyg_thread_stream();
    size_t n = (size_t)nz * ny * nx;
    float *d = s ? (float *)fyg_ws(s, WS_IN, n * sizeof(float)) : NULL;
    double *hs = (double *)malloc(sizeof(double) * (size_t)nz);
    long long *hc = (long long *)malloc(sizeof(long long) * (size_t)nz);
    if (d && hs && hc && !fyg_h2d(d, chunk, n * sizeof(float), s) && !zdrift_sums_dev(d, nz, ny, nx, thresh, hs, hc, s))
        for (int z = 0; z < nz; z++) {
            sums[z0 + z] += hs[z];
            counts[z0 + z] += (long)hc[z];
        }
    free(hs);
    free(hc);
}

/* per-slice factor = global papyrus mean / mean of the slice means within
 * nz/20+1 slices (slices without papyrus skipped), 1 where that mean is tiny
 * (zdrift.c:41-62) */
void fy_zdrift_finalize(const double *sums, const long *counts, int nz, float *factor) {
    if (nz <= 0) return;
    double *mean = (double *)malloc(sizeof(double) * (size_t)nz);
    if (!mean) return;
    double gsum = 0.0;
    long gcount = 0;
    for (int z = 0; z < nz; z++) {
        mean[z] = counts[z] ? sums[z] / counts[z] : 0.0;
        gsum += sums[z];
        gcount += counts[z];
    }
    double gmean = gcount ? gsum / gcount : 1.0;
    int win = nz / 20 + 1;
    for (int z = 0; z < nz; z++) {
        double acc = 0.0;
        int c = 0;
        for (int d = -win; d <= win; d++) {
            int zz = z + d;
            if (zz >= 0 && zz < nz && mean[zz] > 0) { acc += mean[zz]; c++; }
        }
        double sm = c ? acc / c : gmean;
        factor[z] = sm > 1e-6 ? (float)(gmean / sm) : 1.0f;
    }
    free(mean);
}

int fy_correct_zdrift(float *vol, int nz, int ny, int nx, float thresh) {   /* zdrift.c:74-85 */
    if (nz <= 0 || ny <= 0 || nx <= 0) return 0;
    fyg_stream *s = fyg_thread_stream();
    size_t n = (size_t)nz * ny * nx;
    float *d = s ? (float *)fyg_ws(s, WS_IN, n * sizeof(float)) : NULL;
    float *dfac = s ? (float *)fyg_ws(s, WS_FLAG, sizeof(float) * (size_t)nz) : NULL;
    double *hs = (double *)malloc(sizeof(double) * (size_t)nz);
    long long *hc = (long long *)malloc(sizeof(long long) * (size_t)nz);
    long *lc = (long *)malloc(sizeof(long) * (size_t)nz);
    float *fac = (float *)malloc(sizeof(float) * (size_t)nz);
    int rc = !d || !dfac || !hs || !hc || !lc || !fac || fyg_h2d(d, vol, n * sizeof(float), s) ||
             zdrift_sums_dev(d, nz, ny, nx, thresh, hs, hc, s);
    if (!rc) {
        for (int z = 0; z < nz; z++) lc[z] = (long)hc[z];
        fy_zdrift_finalize(hs, lc, nz, fac);
        rc = fyg_h2d(dfac, fac, sizeof(float) * (size_t)nz, s) || fyg_zdrift_run(d, nz, ny, nx, dfac, s) ||
             fyg_d2h(vol, d, n * sizeof(float), s) || fyg_sync(s);
    }
    free(hs); free(hc); free(lc); free(fac);
    return rc ? 1 : 0;
}

/* ------------------------------------------------------------------ device sweep */
struct fy_calib {
    long Z;
    fyg_stream *s;
    void *mem;
    size_t bytes;
    unsigned long long *acc;    /* 256 bins + total */
    unsigned long long *tile;   /* one tile's bins (cleared by the merge) */
    unsigned long long *zs;     /* Z slice sums, units of 2^-27 */
    unsigned long long *zc;     /* Z slice counts */
    unsigned int *part;         /* per-workgroup bins of the histogram kernel */
};

fy_calib *fy_calib_create(long Z, void *stream) {
    if (Z < 0) return NULL;
    fy_calib *c = (fy_calib *)calloc(1, sizeof(*c));
    if (!c) return NULL;
    c->Z = Z;
    c->s = fyg_stream_wrap(stream);
    size_t zb = al256(sizeof(unsigned long long) * (size_t)(Z ? Z : 1));
    size_t o1 = al256(257 * sizeof(unsigned long long)), o2 = o1 + al256(256 * sizeof(unsigned long long));
    size_t o3 = o2 + zb, o4 = o3 + zb;
    c->bytes = o4 + fyg_hist_part_bytes(0);
    if (!c->s || fyg_malloc(&c->mem, c->bytes)) { free(c); return NULL; }
    char *m = (char *)c->mem;
    c->acc = (unsigned long long *)m;
    c->tile = (unsigned long long *)(m + o1);
    c->zs = (unsigned long long *)(m + o2);
    c->zc = (unsigned long long *)(m + o3);
    c->part = (unsigned int *)(m + o4);
    if (fyg_zero(c->mem, c->bytes, c->s)) { fy_calib_destroy(c); return NULL; }
    return c;
}

int fy_calib_reset(fy_calib *c) { return !c || fyg_zero(c->mem, c->bytes, c->s); }

/* pipeline.c:676-701: the histogram of a tile that is not all zero, and the
 * per-slice papyrus sums of its 4x4 subsample */
int fy_calib_accumulate_u8(fy_calib *c, const unsigned char *d_tile, long z0, long dz, long dy, long dx,
                           int flags) {
    if (!c || dz < 0 || dy < 0 || dx < 0) return 1;
    size_t n = (size_t)dz * dy * dx;
    if (n && !d_tile) return 1;
    if (flags & FY_CALIB_HIST) {
        if ((n && fyg_hist_u8_run(d_tile, n, c->part, c->tile, c->s)) ||
            fyg_hist_merge_run(c->tile, (unsigned long long)n, c->acc, 1, c->s))
            return 1;
    }
    if ((flags & FY_CALIB_ZDRIFT) && n) {
        if (z0 < 0 || z0 + dz > c->Z) return 1;
        if (fyg_zdrift_u8_run(d_tile, dz, dy, dx, 4, c->zs + z0, c->zc + z0, c->s)) return 1;
    }
    return 0;
}

/* slice sums leave the device as integers in units of 2^-27: exact in double
 * while a slice's papyrus total stays below 2^26 */
int fy_calib_read(fy_calib *c, fy_hist_state *hist, double *zsums, long *zcnts) {
    if (!c) return 1;
    unsigned long long acc[257];
    size_t nz = (size_t)c->Z;
    unsigned long long *hs = (unsigned long long *)malloc(sizeof(unsigned long long) * (nz ? nz : 1));
    unsigned long long *hc = (unsigned long long *)malloc(sizeof(unsigned long long) * (nz ? nz : 1));
    int rc = !hs || !hc || fyg_d2h(acc, c->acc, sizeof(acc), c->s) ||
             (nz && fyg_d2h(hs, c->zs, sizeof(unsigned long long) * nz, c->s)) ||
             (nz && fyg_d2h(hc, c->zc, sizeof(unsigned long long) * nz, c->s)) || fyg_sync(c->s);
    if (!rc) {
        if (hist) {
            for (int b = 0; b < 256; b++) hist->hist[b] += (long)acc[b];
            hist->total += (long)acc[256];
        }
        for (size_t z = 0; z < nz; z++) {
            if (zsums) zsums[z] += (double)hs[z] * (1.0 / 134217728.0);
            if (zcnts) zcnts[z] += (long)hc[z];
        }
    }
    free(hs);
    free(hc);
    return rc ? 1 : 0;
}

void fy_calib_destroy(fy_calib *c) {
    if (!c) return;
    if (c->s) fyg_sync(c->s);
    fyg_free(c->mem);
    free(c);
}
