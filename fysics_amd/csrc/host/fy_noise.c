/* fy_noise.c -- per-volume noise model and noise anisotropy (noise.c:75-269).
 *
 * The voxel-level work (box sums, local moments, the sorts behind every
 * percentile, the per-block variances and lag products) runs on the GPU
 * (fyg_noise.hip); this file picks order statistics out of the sorted device
 * arrays and runs the small fp64 fits: the 4-iteration IRLS line var = g*I + b
 * over at most 24 intensity bins, and the final correlation ratios.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "fyh.h"

enum { NB = 24 };

/* host-side percentile over a small array (noise.c:66-73; sorts v) */
static int cmp_double(const void *a, const void *b) {
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}
static double pct_host(double *v, int n, double pct) {
    if (n <= 0) return 0.0;
    qsort(v, n, sizeof(double), cmp_double);
    double idx = pct / 100.0 * (n - 1);
    int lo = (int)idx;
    double f = idx - lo;
    return lo + 1 < n ? v[lo] * (1 - f) + v[lo + 1] * f : v[lo];
}

/* indices of the order statistics the percentile of a sorted run of length n needs */
static void pct_index(long n, double pct, long *i0, long *i1, double *f) {
    double idx = pct / 100.0 * (n - 1);
    long lo = (long)idx;
    *f = idx - lo;
    *i0 = lo;
    *i1 = lo + 1 < n ? lo + 1 : -1;
}
static double pct_value(double v0, double v1, long i1, double f) {
    return i1 >= 0 ? v0 * (1 - f) + v1 * f : v0;
}

/* gather sorted[idx[k]] for k < m (device -> host) */
static int gather(const double *d_sorted, const long *idx, int m, double *out, long *d_idx, double *d_out,
                  fyg_stream *s) {
    if (m <= 0) return 0;
    return fyg_h2d(d_idx, idx, sizeof(long) * m, s) || fyg_gather_f64(d_sorted, d_idx, m, d_out, s) ||
           fyg_d2h(out, d_out, sizeof(double) * m, s) || fyg_sync(s);
}

int fy_estimate_noise(const float *in, int nz, int ny, int nx, int win, double flat_pct, double ref_intensity,
                      fy_noise_model *out) {
    if (!out) return 1;
    if (win < 3) win = 5;
    if (flat_pct <= 0) flat_pct = 10.0;
    int r = win / 2;
    size_t n = (size_t)nz * ny * nx;
    if (nz <= 0 || ny <= 0 || nx <= 0) n = 0;
    double g = 0, b = 0;
    int used = 0;
    double bin_i[NB], bin_f[NB];
    long bin_c[NB];
    if (n > 0) {
        fyg_stream *s = fyg_thread_stream();
        if (!s) return 1;
        size_t tmpb = fyg_sort_pairs_temp_bytes(n);
        size_t segb = fyg_segsort_temp_bytes(n, NB);
        if (segb > tmpb) tmpb = segb;
        float *d_in = (float *)fyg_ws(s, WS_IN, sizeof(float) * n);
        double *w = (double *)fyg_ws(s, WS_MISC, sizeof(double) * 8 * n);
        char *small = (char *)fyg_ws(s, WS_FLAG, 4096);
        void *temp = fyg_ws(s, WS_GUIDED, tmpb);
        if (!d_in || !w || !small || !temp) return 1;
        double *m = w, *v = w + n, *ms = w + 2 * n, *vp = w + 3 * n, *vs = w + 4 * n;
        double *t0 = w + 5 * n, *t1 = w + 6 * n, *t2 = w + 7 * n, *t3 = w + 4 * n;   /* t3 reuses vs */
        long *d_idx = (long *)small;
        double *d_val = (double *)(small + 1024);
        double *d_q = (double *)(small + 2048);
        int *d_seg = (int *)(small + 3072);
        double wc = (double)win * win * win;        /* fixed count, as in the reference */
        if (fyg_h2d(d_in, in, sizeof(float) * n, s) ||
            fyg_noise_local(d_in, nz, ny, nx, r, wc, m, v, t0, t1, t2, t3, s) ||
            fyg_sort_pairs_f64(m, ms, v, vp, n, temp, tmpb, s))
            return 1;
        /* 1st / 99th percentile of the local means */
        long ix[4];
        double fx[2], val[4];
        pct_index((long)n, 1.0, &ix[0], &ix[1], &fx[0]);
        pct_index((long)n, 99.0, &ix[2], &ix[3], &fx[1]);
        long gi[4];
        for (int k = 0; k < 4; k++) gi[k] = ix[k] >= 0 ? ix[k] : 0;
        if (gather(ms, gi, 4, val, d_idx, d_val, s)) return 1;
        double lo = pct_value(val[0], val[1], ix[1], fx[0]);
        double hi = pct_value(val[2], val[3], ix[3], fx[1]);
        if (hi <= lo) hi = lo + 1e-6;
        /* bin edges and their positions in the sorted means */
        double q[NB + 1];
        for (int k = 0; k <= NB; k++) q[k] = lo + (hi - lo) * k / NB;
        long edge[NB + 1];
        if (fyg_h2d(d_q, q, sizeof q, s) || fyg_lower_bounds_f64(ms, n, d_q, NB + 1, d_idx, s) ||
            fyg_d2h(edge, d_idx, sizeof edge, s) || fyg_sync(s))
            return 1;
        /* per-bin variance floor: segmented sort of the variances of each bin */
        int beg[NB], end[NB], nseg = 0, segbin[NB];
        for (int k = 0; k < NB; k++) {
            /* the upper edge of bin k is the lower edge of bin k+1: the same double */
            long c = edge[k + 1] - edge[k];
            if (c < 200) continue;
            beg[nseg] = (int)edge[k];
            end[nseg] = (int)edge[k + 1];
            segbin[nseg] = k;
            nseg++;
        }
        if (nseg > 0) {
            if (fyg_h2d(d_seg, beg, sizeof(int) * nseg, s) || fyg_h2d(d_seg + NB, end, sizeof(int) * nseg, s) ||
                fyg_segsort_f64(vp, vs, n, nseg, d_seg, d_seg + NB, temp, tmpb, s))
                return 1;
            long pi[2 * NB], pidx[2 * NB];
            double pf[NB], pv[2 * NB];
            for (int j = 0; j < nseg; j++) {
                long c = end[j] - beg[j];
                pct_index(c, flat_pct, &pi[2 * j], &pi[2 * j + 1], &pf[j]);
                pidx[2 * j] = beg[j] + pi[2 * j];
                pidx[2 * j + 1] = beg[j] + (pi[2 * j + 1] >= 0 ? pi[2 * j + 1] : pi[2 * j]);
            }
            if (gather(vs, pidx, 2 * nseg, pv, d_idx, d_val, s)) return 1;
            for (int j = 0; j < nseg; j++) {
                int k = segbin[j];
                double b0 = lo + (hi - lo) * k / NB, b1 = lo + (hi - lo) * (k + 1) / NB;
                bin_f[used] = pct_value(pv[2 * j], pv[2 * j + 1], pi[2 * j + 1], pf[j]);
                bin_i[used] = 0.5 * (b0 + b1);
                bin_c[used] = end[j] - beg[j];
                used++;
            }
        }
        if (used == 0) {                            /* global low percentile of the variance */
            if (fyg_sort_pairs_f64(v, vs, m, ms, n, temp, tmpb, s)) return 1;
            long a0, a1;
            double f;
            pct_index((long)n, flat_pct, &a0, &a1, &f);
            long gi2[2] = {a0, a1 >= 0 ? a1 : a0};
            double vv[2];
            if (gather(vs, gi2, 2, vv, d_idx, d_val, s)) return 1;
            b = pct_value(vv[0], vv[1], a1, f);
        }
    }
    out->n_bins_used = used;
    if (used >= 3) {                                /* IRLS line fit (noise.c:139-161) */
        double wt[NB];
        for (int i = 0; i < used; i++) wt[i] = sqrt((double)bin_c[i]);
        for (int it = 0; it < 4; it++) {
            double sw = 0, swx = 0, swy = 0, swxx = 0, swxy = 0;
            for (int i = 0; i < used; i++) {
                double wi = wt[i], xi = bin_i[i], yi = bin_f[i];
                sw += wi; swx += wi * xi; swy += wi * yi; swxx += wi * xi * xi; swxy += wi * xi * yi;
            }
            double det = sw * swxx - swx * swx;
            if (fabs(det) < 1e-20) break;
            g = (sw * swxy - swx * swy) / det;
            b = (swxx * swy - swx * swxy) / det;
            double res[NB], rs[NB];
            for (int i = 0; i < used; i++) { res[i] = fabs(bin_f[i] - (g * bin_i[i] + b)); rs[i] = res[i]; }
            double med = pct_host(rs, used, 50.0) + 1e-9;
            for (int i = 0; i < used; i++) {
                double rr = res[i] / (2.0 * med);
                wt[i] = sqrt((double)bin_c[i]) / (1.0 + rr * rr);
            }
        }
    } else if (used > 0) {                          /* flat model */
        double sm = 0;
        for (int i = 0; i < used; i++) sm += bin_f[i];
        b = sm / used;
        g = 0;
    }
    if (!isfinite(g)) g = 0;
    if (!isfinite(b)) b = 0;
    out->g = g;
    out->b = b;
    double var_ref = g * ref_intensity + b;
    if (var_ref < 0) var_ref = b > 0 ? b : 0;
    out->noise_ref = sqrt(var_ref) * 1.15;        /* window-averaging correction of the reference */
    out->ref_intensity = ref_intensity;
    return 0;
}

static int cmp_float(const void *a, const void *b) {
    float x = *(const float *)a, y = *(const float *)b;
    return (x > y) - (x < y);
}

int fy_noise_aniso(const float *v, int nz, int ny, int nx, double *rho_z, double *rho_xy) {
    int bz = nz / 8, by = ny / 8, bx = nx / 8;
    long nb = (long)bz * by * bx;
    if (nb < 64) return 1;
    fyg_stream *s = fyg_thread_stream();
    if (!s) return 1;
    size_t n = (size_t)nz * ny * nx;
    float *d_v = (float *)fyg_ws(s, WS_IN, sizeof(float) * n);
    char *w = (char *)fyg_ws(s, WS_MISC, (size_t)nb * (sizeof(float) + 2 * sizeof(int) + 4 * sizeof(double)) + 64);
    float *bvar = (float *)malloc(sizeof(float) * nb);
    int *ok = (int *)malloc(sizeof(int) * nb);
    double *part = (double *)malloc(sizeof(double) * 4 * nb);
    int rc = 1;
    if (!d_v || !w || !bvar || !ok || !part) goto done;
    double *d_part = (double *)w;
    float *d_bvar = (float *)(d_part + 4 * nb);
    int *d_ok = (int *)(d_bvar + nb);
    int *d_used = d_ok + nb;
    if (fyg_h2d(d_v, v, sizeof(float) * n, s) || fyg_aniso_blocks(d_v, nz, ny, nx, d_bvar, d_ok, s) ||
        fyg_d2h(bvar, d_bvar, sizeof(float) * nb, s) || fyg_d2h(ok, d_ok, sizeof(int) * nb, s) || fyg_sync(s))
        goto done;
    long nv = 0;
    for (long i = 0; i < nb; i++)
        if (ok[i]) bvar[nv++] = bvar[i];
    if (nv < 64) goto done;
    qsort(bvar, (size_t)nv, sizeof(float), cmp_float);     /* 20th-percentile flat threshold */
    float thresh = bvar[nv / 5];
    if (fyg_aniso_lag(d_v, nz, ny, nx, thresh, d_part, d_used, s) ||
        fyg_d2h(part, d_part, sizeof(double) * 4 * nb, s) || fyg_d2h(ok, d_used, sizeof(int) * nb, s) ||
        fyg_sync(s))
        goto done;
    double den = 0, az = 0, ay = 0, ax = 0;
    long used = 0;
    for (long i = 0; i < nb; i++) {
        if (!ok[i]) continue;
        den += part[4 * i]; az += part[4 * i + 1]; ay += part[4 * i + 2]; ax += part[4 * i + 3];
        used++;
    }
    if (used < 32 || den <= 0) goto done;
    double corr = 8.0 / 7.0;                        /* lag pairs cover 7/8 of the energy terms */
    *rho_z = corr * az / den;
    *rho_xy = corr * 0.5 * (ay + ax) / den;
    rc = 0;
done:
    free(bvar); free(ok); free(part);
    return rc;
}
