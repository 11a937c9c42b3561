/* fy_dering.c -- detect-then-subtract ring removal (dering.c).
 *
 * State lives in host memory exactly as in the reference struct (callers read
 * d->ring, d->have_rings).  Accumulation and subtraction run on the GPU; the
 * finalize step (sector vote over nslab*ns*nr bins) is host logic on the small
 * profile arrays.  The (sector, radius) bin of each (y, x) sample is computed
 * on the host with the C library's sqrt/atan2 -- the same libm calls the
 * reference makes -- and shipped to the GPU as a map, so sector boundaries that
 * fall exactly on a sample (diagonals, axes) resolve identically; the device
 * then only adds integers, which makes sums and counts exact and independent
 * of summation order.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "fyh.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* The ring radius floor(sqrt(dy^2 + dx^2) + 0.5) may be evaluated in fp32 when
 * the axis sits on the half-integer grid (so every dy, dx and the squared radius
 * d are exact multiples of 1/4) and every radius of the region is below 512.
 * Then d either equals a boundary (k+1/2)^2 -- a perfect square, for which sqrt
 * and the +0.5 are exact in both precisions -- or differs from every boundary by
 * at least 1/4, i.e. sqrt(d) is at least 1/4/(2*512+1) away from k+1/2, far more
 * than the fp32 rounding of sqrt(d) + 0.5 for d < 2^18. */
int fyh_ring_fast(double cy, double cx, long y0, long x0, long dy, long dx) {
    if (!(isfinite(cy) && isfinite(cx))) return 0;
    if (2.0 * cy != floor(2.0 * cy) || 2.0 * cx != floor(2.0 * cx)) return 0;
    double ay = fmax(fabs((double)y0 - cy), fabs((double)(y0 + dy - 1) - cy));
    double ax = fmax(fabs((double)x0 - cx), fabs((double)(x0 + dx - 1) - cx));
    return ay * ay + ax * ax < 512.0 * 512.0;
}

int fy_dering_init(fy_dering *d, long Z, long Y, long X, double cy, double cx, int slab_z, int ns) {
    memset(d, 0, sizeof(*d));
    d->Z = Z; d->Y = Y; d->X = X;
    d->cy = cy >= 0 ? cy : (Y - 1) / 2.0;
    d->cx = cx >= 0 ? cx : (X - 1) / 2.0;
    d->slab_z = slab_z > 0 ? slab_z : 512;
    d->nslab = (int)((Z + d->slab_z - 1) / d->slab_z);
    d->ns = ns > 0 ? ns : 8;
    double ry = d->cy > Y - 1 - d->cy ? d->cy : Y - 1 - d->cy;
    double rx = d->cx > X - 1 - d->cx ? d->cx : X - 1 - d->cx;
    d->nr = (int)(sqrt(ry * ry + rx * rx)) + 2;
    size_t nb = (size_t)d->nslab * d->ns * d->nr;
    d->sum = (float *)calloc(nb, sizeof(float));
    d->cnt = (unsigned int *)calloc(nb, sizeof(unsigned int));
    d->ring = (float *)calloc((size_t)d->nslab * d->nr, sizeof(float));
    if (!d->sum || !d->cnt || !d->ring) { fy_dering_free(d); return 1; }
    return 0;
}

void fy_dering_free(fy_dering *d) {
    free(d->sum); free(d->cnt); free(d->ring);
    d->sum = NULL; d->cnt = NULL; d->ring = NULL;
}

int fy_dering_tile_init(fy_dering *t, const fy_dering *d) {
    *t = *d;
    t->nslab = 1;
    size_t nb = (size_t)t->ns * t->nr;
    t->sum = (float *)calloc(nb, sizeof(float));
    t->cnt = (unsigned int *)calloc(nb, sizeof(unsigned int));
    t->ring = NULL;
    if (!t->sum || !t->cnt) { free(t->sum); free(t->cnt); t->sum = NULL; t->cnt = NULL; return 1; }
    return 0;
}

void fy_dering_tile_reset(fy_dering *t) {
    size_t nb = (size_t)t->ns * t->nr;
    memset(t->sum, 0, nb * sizeof(float));
    memset(t->cnt, 0, nb * sizeof(unsigned int));
}

void fy_dering_accumulate_u8(fy_dering *t, const unsigned char *buf, long y0, long x0, long dz, long dy, long dx,
                             int ss) {
    if (ss < 1) ss = 1;
    if (dz <= 0 || dy <= 0 || dx <= 0) return;
    int ns = t->ns, nr = t->nr, nb = ns * nr;
    long sy = (dy + ss - 1) / ss, sx = (dx + ss - 1) / ss;
    int *map = (int *)malloc(sizeof(int) * (size_t)sy * sx);
    unsigned long long *hs = (unsigned long long *)malloc(sizeof(unsigned long long) * (size_t)nb);
    unsigned int *hc = (unsigned int *)malloc(sizeof(unsigned int) * (size_t)nb);
    if (!map || !hs || !hc) { free(map); free(hs); free(hc); return; }
    double inv_sect = ns / (2.0 * M_PI);
    for (long yi = 0; yi < sy; yi++) {                       /* dering.c:89-99 */
        double gy = (double)(y0 + yi * ss) - t->cy, gy2 = gy * gy;
        for (long xi = 0; xi < sx; xi++) {
            double gx = (double)(x0 + xi * ss) - t->cx;
            int ri = (int)(sqrt(gy2 + gx * gx) + 0.5);
            int b = -1;
            if (ri < nr) {
                int q = (int)((atan2(gy, gx) + M_PI) * inv_sect);
                if (q >= ns) q = ns - 1;
                b = q * nr + ri;
            }
            map[yi * sx + xi] = b;
        }
    }
    fyg_stream *s = fyg_thread_stream();
    size_t nbuf = (size_t)dz * dy * dx;
    unsigned char *dbuf = s ? (unsigned char *)fyg_ws(s, WS_U8, nbuf) : NULL;
    char *dm = s ? (char *)fyg_ws(s, WS_MISC, sizeof(int) * (size_t)sy * sx + 12 * (size_t)nb + 64) : NULL;
    if (dbuf && dm) {
        int *dmap = (int *)dm;
        unsigned long long *dsum = (unsigned long long *)(dm + ((sizeof(int) * (size_t)sy * sx + 15) & ~(size_t)15));
        unsigned int *dcnt = (unsigned int *)(dsum + nb);
        int rc = fyg_h2d(dbuf, buf, nbuf, s) || fyg_h2d(dmap, map, sizeof(int) * (size_t)sy * sx, s) ||
                 fyg_zero(dsum, 12 * (size_t)nb, s) ||
                 fyg_dering_accum_run(dbuf, dz, dy, dx, ss, dmap, nb, dsum, dcnt, s) ||
                 fyg_d2h(hs, dsum, sizeof(unsigned long long) * nb, s) ||
                 fyg_d2h(hc, dcnt, sizeof(unsigned int) * nb, s) || fyg_sync(s);
        if (!rc)
            for (int b = 0; b < nb; b++) {
                if (!hc[b]) continue;
                t->sum[b] += (float)hs[b];
                t->cnt[b] += hc[b];
            }
    }
    free(map); free(hs); free(hc);
}

void fy_dering_merge_tile(fy_dering *d, int slab, const fy_dering *t) {
    if (slab < 0 || slab >= d->nslab) return;
    size_t nb = (size_t)d->ns * d->nr, off = (size_t)slab * nb;
    for (size_t i = 0; i < nb; i++) {
        d->sum[off + i] += t->sum[i];
        d->cnt[off + i] += t->cnt[i];
    }
}

static int cmp_float(const void *a, const void *b) {
    float x = *(const float *)a, y = *(const float *)b;
    return (x > y) - (x < y);
}

/* per slab: sector mean profiles, minus their box mean over valid radii, then a
 * radius is a ring only if >= ns-2 sectors are valid and all agree in sign; the
 * ring value is the sector median, floored at min_amp and clamped at max_amp
 * (dering.c:136-194) */
long fy_dering_finalize(fy_dering *d, int hp_win, unsigned int min_cnt, double min_amp, double max_amp) {
    if (hp_win < 3) hp_win = 15;
    if (min_amp <= 0) min_amp = 0.5;
    if (max_amp <= 0) max_amp = 6.0;
    int ns = d->ns, nr = d->nr, hw = hp_win / 2;
    float *prof = (float *)malloc(sizeof(float) * (size_t)ns * nr);
    float *hp = (float *)malloc(sizeof(float) * (size_t)ns * nr);
    unsigned char *ok = (unsigned char *)malloc((size_t)ns * nr);
    float vals[64];
    long found = 0;
    if (!prof || !hp || !ok || ns > 64) { free(prof); free(hp); free(ok); return 0; }
    for (int sl = 0; sl < d->nslab; sl++) {
        size_t so = (size_t)sl * ns * nr;
        for (int q = 0; q < ns; q++) {
            const float *sm = d->sum + so + (size_t)q * nr;
            const unsigned int *cn = d->cnt + so + (size_t)q * nr;
            float *pq = prof + (size_t)q * nr, *hq = hp + (size_t)q * nr;
            unsigned char *vq = ok + (size_t)q * nr;
            for (int r = 0; r < nr; r++) {
                vq[r] = cn[r] >= min_cnt;
                pq[r] = vq[r] ? sm[r] / cn[r] : 0.0f;
            }
            for (int r = 0; r < nr; r++) {
                if (!vq[r]) { hq[r] = 0.0f; continue; }
                double acc = 0;
                int c = 0;
                int lo = r - hw < 0 ? 0 : r - hw, hi = r + hw >= nr ? nr - 1 : r + hw;
                for (int j = lo; j <= hi; j++)
                    if (vq[j]) { acc += pq[j]; c++; }
                if (c >= hw) hq[r] = pq[r] - (float)(acc / c);
                else { hq[r] = 0.0f; vq[r] = 0; }
            }
        }
        float *ring = d->ring + (size_t)sl * nr;
        for (int r = 0; r < nr; r++) {
            int nv = 0, pos = 0, neg = 0;
            for (int q = 0; q < ns; q++) {
                if (!ok[(size_t)q * nr + r]) continue;
                float v = hp[(size_t)q * nr + r];
                vals[nv++] = v;
                if (v > 0) pos++;
                else if (v < 0) neg++;
            }
            if (nv < ns - 2 || (pos && neg)) { ring[r] = 0; continue; }
            qsort(vals, nv, sizeof(float), cmp_float);
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

void fy_dering_apply(const fy_dering *d, float *f, long z0, long y0, long x0, long dz, long dy, long dx,
                     double scale) {
    if (!d->have_rings) return;
    size_t n = (size_t)dz * dy * dx;
    if (!n) return;
    fyg_stream *s = fyg_thread_stream();
    if (!s) return;
    fyg_hold *hold = fyg_hold_begin();
    const float *ring = (const float *)fyh_const_blob('R', d->ring, sizeof(float) * (size_t)d->nslab * d->nr);
    float *df = (float *)fyg_ws(s, WS_IN, sizeof(float) * n);
    if (!ring || !df) {
        fyg_hold_release(hold, s);
        return;
    }
    fyg_prologue pro;
    memset(&pro, 0, sizeof pro);
    pro.dering = 1; pro.ring = ring; pro.nr = d->nr; pro.nslab = d->nslab; pro.slab_z = d->slab_z;
    pro.cy = d->cy; pro.cx = d->cx; pro.scale = (float)scale;
    pro.gz0 = z0; pro.gy0 = y0; pro.gx0 = x0;
    pro.ring_fast = fyh_ring_fast(d->cy, d->cx, y0, x0, dy, dx);
    int rc = fyg_h2d(df, f, sizeof(float) * n, s) || fyg_dering_apply_run(df, dz, dy, dx, &pro, s) ||
             fyg_d2h(f, df, sizeof(float) * n, s);
    fyg_hold_release(hold, s);
    if (!rc) fyg_sync(s);
}
