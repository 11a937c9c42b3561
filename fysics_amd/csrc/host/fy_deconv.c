/* fy_deconv.c -- Wiener deconvolution entry points (paganin.c:152-342).
 *
 * The host side builds, once per (weight, physics, padded shape), the transfer
 * tables deconv_radial uses: the squared-frequency axes (fftfreq convention,
 * fz^2 scaled by (sigma_z/sigma_xy)^2) and the 4097-entry fp32 LUT of the radial
 * weight over u = f^2, evaluated in double (paganin.c:212-230).  They live in
 * the calling stream's table slots (uploaded asynchronously on a miss), or in
 * the process-wide cache when too large for a slot; the GPU passes in
 * fyg_fft.hip do the rest.
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fyh.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* H/(H^2+reg) of the recon transfer (paganin.c:296-299) */
static double wiener_weight(double fr, const fy_physics *p, double reg) {
    double h = fy_recon_transfer(fr, p);
    return h / (h * h + reg);
}

/* G*U/((G*U)^2+gamma), G the Gaussian system PSF (paganin.c:321-328) */
static double matched_weight(double fr, const fy_physics *p, double gamma) {
    double sg = p->psf_sigma_vox > 0 ? p->psf_sigma_vox : 1.0;
    double g = exp(-2.0 * M_PI * M_PI * sg * sg * (fr * fr));
    double hf = g * fy_unsharp_transfer(fr, p);
    return hf / (hf * hf + gamma);
}

double fyh_matched_fz2_scale(const fy_physics *p) {
    double sxy = p->psf_sigma_vox > 0 ? p->psf_sigma_vox : 1.0;
    if (p->psf_sigma_z_vox > 0 && sxy > 0) {
        double q = p->psf_sigma_z_vox / sxy;
        return q * q;
    }
    return 1.0;
}

typedef struct {
    char tag[4];
    int kind, pz, py, px;
    double phys[8], param, cap, fz2_scale;
} table_key;

/* padded shape, frequency axes (fftfreq convention, fz^2 scaled), LUT scale
 * and cache key of one table; *axes holds fz2, fy2, fx2 (caller frees) */
static int table_shape(const fyh_table_req *r, fyh_tables *t, double **axes, table_key *key, size_t *lut_bytes,
                       size_t *bytes) {
    int pz = fy_next_fft_size(r->nz), py = fy_next_fft_size(r->ny), px = fy_next_fft_size(r->nx);
    int hx = px / 2 + 1;
    double fz2_scale = r->fz2_scale > 0 ? r->fz2_scale : 1.0;
    t->pz = pz; t->py = py; t->px = px; t->hx = hx; t->hxp = fyg_deconv_hxp(px);
    /* layout of the device block: lut[4097] (padded to 4100 entries; fp32, or the
     * same weights in fp64 for the all-FP64 variant), fz2, fy2, fx2 */
    *lut_bytes = (r->f64 ? sizeof(double) : sizeof(float)) * 4100;
    *bytes = *lut_bytes + sizeof(double) * (size_t)(pz + py + hx);
    double *fz2 = (double *)malloc(sizeof(double) * (size_t)(pz + py + hx));
    if (!fz2) return 1;
    double *fy2 = fz2 + pz, *fx2 = fy2 + py;
    double mz = 0, my = 0, mx = 0;
    for (int i = 0; i < pz; i++) {
        double f = (i <= pz / 2) ? (double)i / pz : (double)(i - pz) / pz;
        fz2[i] = f * f * fz2_scale;
        if (fz2[i] > mz) mz = fz2[i];
    }
    for (int i = 0; i < py; i++) {
        double f = (i <= py / 2) ? (double)i / py : (double)(i - py) / py;
        fy2[i] = f * f;
        if (fy2[i] > my) my = fy2[i];
    }
    for (int i = 0; i < hx; i++) {
        double f = (double)i / px;
        fx2[i] = f * f;
        if (fx2[i] > mx) mx = fx2[i];
    }
    t->us = FYH_LUT_N / (mz + my + mx);
    memset(key, 0, sizeof *key);
    memcpy(key->tag, r->f64 ? "L64" : "LUT", 4);
    key->kind = r->kind; key->pz = pz; key->py = py; key->px = px;
    memcpy(key->phys, r->p, sizeof(double) * 8);
    key->param = r->param; key->cap = r->cap; key->fz2_scale = fz2_scale;
    *axes = fz2;
    return 0;
}

/* the LUT index scale of a padded shape: 4096 over the largest
 * fz^2*scale + fy^2 + fx^2 (the maxima table_shape's axes reach) */
static double lut_us(int pz, int py, int px, double fz2_scale) {
    int hx = px / 2 + 1;
    double mz = 0, my = 0, mx = 0;
    for (int i = 0; i < pz; i++) {
        double f = (i <= pz / 2) ? (double)i / pz : (double)(i - pz) / pz;
        double v = f * f * fz2_scale;
        if (v > mz) mz = v;
    }
    for (int i = 0; i < py; i++) {
        double f = (i <= py / 2) ? (double)i / py : (double)(i - py) / py;
        double v = f * f;
        if (v > my) my = v;
    }
    for (int i = 0; i < hx; i++) {
        double f = (double)i / px;
        double v = f * f;
        if (v > mx) mx = v;
    }
    return FYH_LUT_N / (mz + my + mx);
}

/* per-stream slots first (no device wait on a miss), the process-wide cache
 * for tables too large for a slot; data == NULL: lookup only */
static const void *table_put(const table_key *key, const void *data, size_t bytes, fyg_stream *s) {
    int slot = bytes <= FYG_ARENA_SLOT && !fyg_stream_persistent(s);
    if (slot) return fyg_stream_table(s, key, sizeof *key, data, bytes);
    return data ? fyg_table(key, sizeof *key, data, bytes) : fyg_table_find(key, sizeof *key);
}

/* the 4097 weights over u = i/us, evaluated in double (paganin.c:212-230),
 * plus three zero pad entries */
static void lut_fill(void *lut, const fy_physics *p, double param, int kind, double cap, int f64, double us) {
    float *l32 = (float *)lut;
    double *l64 = (double *)lut;
    for (int i = 0; i <= FYH_LUT_N; i++) {
        double fr = sqrt((double)i / us);
        double w = kind == 0 ? wiener_weight(fr, p, param) : matched_weight(fr, p, param);
        if (cap > 0 && w > cap) w = cap;
        if (f64) l64[i] = w;
        else l32[i] = (float)w;
    }
    for (int i = FYH_LUT_N + 1; i < 4100; i++) {
        if (f64) l64[i] = 0.0;
        else l32[i] = 0.0f;
    }
}

/* weights built ahead of a planning pass */
typedef struct {
    table_key key;
    fy_physics phys;
    double us;
    int f64;
    void *lut;
} memo_ent;

struct fyh_lut_memo {
    int n, next;
    memo_ent *e;
    pthread_mutex_t mu;
};

static void *memo_worker(void *arg) {
    fyh_lut_memo *m = (fyh_lut_memo *)arg;
    for (;;) {
        pthread_mutex_lock(&m->mu);
        int i = m->next < m->n ? m->next++ : -1;
        pthread_mutex_unlock(&m->mu);
        if (i < 0) return NULL;
        memo_ent *e = &m->e[i];
        lut_fill(e->lut, &e->phys, e->key.param, e->key.kind, e->key.cap, e->f64, e->us);
    }
}

/* Each table costs 4097 transcendental weight evaluations on the host; a batch
 * of cubes that each bring their own physics (a serving wave) would otherwise
 * evaluate them one table after another on the calling thread.  The tables
 * not cached yet are built here, one per worker at a time. */
fyh_lut_memo *fyh_tables_prefill(const fyh_table_req *req, int n, fyg_stream *s) {
    fyh_lut_memo *m = (fyh_lut_memo *)calloc(1, sizeof *m);
    if (!m) return NULL;
    m->e = (memo_ent *)calloc((size_t)(n > 0 ? n : 1), sizeof *m->e);
    if (!m->e) { free(m); return NULL; }
    pthread_mutex_init(&m->mu, NULL);
    for (int i = 0; i < n; i++) {
        fyh_tables t;
        double *axes = NULL;
        table_key key;
        size_t lut_bytes, bytes;
        if (table_shape(&req[i], &t, &axes, &key, &lut_bytes, &bytes)) continue;
        free(axes);
        if (table_put(&key, NULL, bytes, s)) continue;               /* cached */
        int dup = 0;
        for (int j = 0; j < m->n && !dup; j++) dup = !memcmp(&m->e[j].key, &key, sizeof key);
        if (dup) continue;
        memo_ent *e = &m->e[m->n];
        e->lut = malloc(lut_bytes);
        if (!e->lut) continue;
        e->key = key; e->phys = *req[i].p; e->us = t.us; e->f64 = req[i].f64;
        m->n++;
    }
    enum { MAXW = 8 };
    pthread_t th[MAXW];
    int nw = m->n / 2 < MAXW ? m->n / 2 : MAXW, started = 0;
    for (int w = 0; w < nw; w++)
        if (!pthread_create(&th[started], NULL, memo_worker, m)) started++;
    memo_worker(m);
    for (int w = 0; w < started; w++) pthread_join(th[w], NULL);
    return m;
}

void fyh_lut_memo_free(fyh_lut_memo *m) {
    if (!m) return;
    for (int i = 0; i < m->n; i++) free(m->e[i].lut);
    pthread_mutex_destroy(&m->mu);
    free(m->e);
    free(m);
}

int fyh_tables_get_m(const fyh_table_req *r, const fyh_lut_memo *m, fyg_stream *s, fyh_tables *t) {
    double *fz2 = NULL;
    table_key key;
    size_t lut_bytes, bytes;
    if (table_shape(r, t, &fz2, &key, &lut_bytes, &bytes)) return 1;
    const int pz = t->pz, py = t->py, hx = t->hx;
    const unsigned char *dev = (const unsigned char *)table_put(&key, NULL, bytes, s);
    if (!dev) {
        unsigned char *blk = (unsigned char *)malloc(bytes);
        if (!blk) { free(fz2); return 1; }
        const memo_ent *e = NULL;
        for (int j = 0; m && j < m->n && !e; j++)
            if (!memcmp(&m->e[j].key, &key, sizeof key)) e = &m->e[j];
        if (e) memcpy(blk, e->lut, lut_bytes);
        else lut_fill(blk, r->p, r->param, r->kind, r->cap, r->f64, t->us);
        memcpy(blk + lut_bytes, fz2, sizeof(double) * (size_t)(pz + py + hx));
        dev = (const unsigned char *)table_put(&key, blk, bytes, s);
        free(blk);
    }
    free(fz2);
    if (!dev) return 1;
    t->lut = dev;
    t->fz2 = (const double *)(dev + lut_bytes);
    t->fy2 = t->fz2 + pz;
    t->fx2 = t->fy2 + py;
    return 0;
}

int fyh_tables_get(int nz, int ny, int nx, const fy_physics *p, double param, int kind,
                   double cap, double fz2_scale, int f64, fyg_stream *s, fyh_tables *t) {
    const fyh_table_req r = {nz, ny, nx, kind, f64, p, param, cap, fz2_scale};
    return fyh_tables_get_m(&r, NULL, s, t);
}

int fyh_deconv_dev(const float *d_in, float *d_out, int nz, int ny, int nx, const fy_physics *p,
                   double param, int kind, double cap, double fz2_scale, fyg_stream *s) {
    if (nz <= 0 || ny <= 0 || nx <= 0) return 0;
    fyh_tables t;
    fyg_hold *hold = fyg_hold_begin();          /* tables of the process-wide cache stay until queued */
    int rc = fyh_tables_get(nz, ny, nx, p, param, kind, cap, fz2_scale, 0, s, &t) ||
             !fyg_fft_supported(t.pz) || !fyg_fft_supported(t.py) || !fyg_fft_supported(t.px);
    void *spec = rc ? NULL : fyg_ws(s, WS_SPEC, fyg_deconv_spec_bytes(t.pz, t.py, t.px));
    if (!spec) {
        fyg_hold_release(hold, s);
        return 1;
    }
    fyg_deconv_args a;
    memset(&a, 0, sizeof a);
    a.nz = nz; a.ny = ny; a.nx = nx;
    a.pz = t.pz; a.py = t.py; a.px = t.px; a.hx = t.hx; a.hxp = t.hxp;
    a.lut = t.lut; a.us = t.us; a.fz2 = t.fz2; a.fy2 = t.fy2; a.fx2 = t.fx2;
    a.in = d_in; a.in_u8 = 0;
    a.out = d_out;
    a.spec = spec;
    rc = fyg_deconv_run(&a, s);
    if (fyg_hold_release(hold, s)) rc = 1;
    return rc;
}

static int deconv_host(const float *in, float *out, int nz, int ny, int nx, const fy_physics *p,
                       double param, int kind, double cap, double fz2_scale) {
    if (nz <= 0 || ny <= 0 || nx <= 0) return 0;
    fyg_stream *s = fyg_thread_stream();
    if (!s) return 1;
    size_t n = (size_t)nz * ny * nx;
    float *d = (float *)fyg_ws(s, WS_IN, sizeof(float) * n);
    if (!d) return 1;
    if (fyg_h2d(d, in, sizeof(float) * n, s)) return 1;
    if (fyh_deconv_dev(d, d, nz, ny, nx, p, param, kind, cap, fz2_scale, s)) return 1;
    if (fyg_d2h(out, d, sizeof(float) * n, s)) return 1;
    return fyg_sync(s);
}

int fy_deconvolve(const float *in, float *out, int nz, int ny, int nx, const fy_physics *p, double reg) {
    if (reg <= 0.0) reg = fy_auto_reg(p, 0.015);
    return deconv_host(in, out, nz, ny, nx, p, reg, 0, FYH_MAX_GAIN, 1.0);
}

int fy_deconvolve_matched(const float *in, float *out, int nz, int ny, int nx, const fy_physics *p,
                          double tikhonov) {
    if (tikhonov <= 0) tikhonov = 0.05;
    return deconv_host(in, out, nz, ny, nx, p, tikhonov, 1, 0.0, fyh_matched_fz2_scale(p));
}

int fy_gpu_deconvolve(const float *d_in, float *d_out, int nz, int ny, int nx, const fy_physics *p,
                      double reg, void *stream) {
    fyg_stream *s = fyg_stream_wrap(stream);
    if (!s) return 1;
    if (reg <= 0.0) reg = fy_auto_reg(p, 0.015);
    return fyh_deconv_dev(d_in, d_out, nz, ny, nx, p, reg, 0, FYH_MAX_GAIN, 1.0, s);
}

int fy_gpu_deconvolve_matched(const float *d_in, float *d_out, int nz, int ny, int nx, const fy_physics *p,
                              double tikhonov, void *stream) {
    fyg_stream *s = fyg_stream_wrap(stream);
    if (!s) return 1;
    if (tikhonov <= 0) tikhonov = 0.05;
    return fyh_deconv_dev(d_in, d_out, nz, ny, nx, p, tikhonov, 1, 0.0, fyh_matched_fz2_scale(p), s);
}

int fy_gpu_deconv_radial(const float *d_in, float *d_out, int nz, int ny, int nx, const fy_physics *p,
                         double param, int kind, double cap, double fz2_scale, void *stream) {
    fyg_stream *s = fyg_stream_wrap(stream);
    if (!s) return 1;
    return fyh_deconv_dev(d_in, d_out, nz, ny, nx, p, param, kind, cap, fz2_scale, s);
}

/* the all-FP64 variant of deconv_radial: fp32 input, fp64 spectrum, weights and
 * output (the speed reference of the precision study) */
int fy_gpu_deconv_radial_f64(const float *d_in, double *d_out, int nz, int ny, int nx, const fy_physics *p,
                             double param, int kind, double cap, double fz2_scale, void *stream) {
    if (nz <= 0 || ny <= 0 || nx <= 0) return 0;
    fyg_stream *s = fyg_stream_wrap(stream);
    if (!s) return 1;
    fyh_tables t;
    fyg_hold *hold = fyg_hold_begin();
    int rc = fyh_tables_get(nz, ny, nx, p, param, kind, cap, fz2_scale, 1, s, &t) ||
             !fyg_fft_supported(t.pz) || !fyg_fft_supported(t.py) || !fyg_fft_supported(t.px);
    void *spec = rc ? NULL : fyg_ws(s, WS_SPEC, fyg_deconv_spec_bytes64(t.pz, t.py, t.px));
    if (!spec) {
        fyg_hold_release(hold, s);
        return 1;
    }
    fyg_deconv_args a;
    memset(&a, 0, sizeof a);
    a.nz = nz; a.ny = ny; a.nx = nx;
    a.pz = t.pz; a.py = t.py; a.px = t.px; a.hx = t.hx; a.hxp = t.hxp;
    a.lut = t.lut; a.us = t.us; a.fz2 = t.fz2; a.fy2 = t.fy2; a.fx2 = t.fx2;
    a.in = d_in; a.in_u8 = 0;
    a.out = d_out;
    a.spec = spec;
    a.f64 = 1;
    rc = fyg_deconv_run(&a, s);
    if (fyg_hold_release(hold, s)) rc = 1;
    return rc;
}

/* fy_gpu_deconvolve with nothing allocated: the spectrum and the tables live
 * in the library's per-GPU scratch array (fyg_noalloc.hip); sizes whose
 * spectrum exceeds it are refused */
int fy_gpu_deconvolve_noalloc(const float *d_in, float *d_out, int nz, int ny, int nx, const fy_physics *p,
                              double reg, void *stream) {
    if (nz <= 0 || ny <= 0 || nx <= 0) return 0;
    if (!p || !d_in || !d_out) {
        fyg_set_error("fy_gpu_deconvolve_noalloc: NULL argument");
        return 1;
    }
    int pz = fy_next_fft_size(nz), py = fy_next_fft_size(ny), px = fy_next_fft_size(nx);
    if (!fyg_fft_supported(pz) || !fyg_fft_supported(py) || !fyg_fft_supported(px)) {
        fyg_set_error("fy_gpu_deconvolve_noalloc: padded axis longer than 4096");
        return 1;
    }
    size_t need = fyg_deconv_spec_bytes(pz, py, px);
    if (need > FYG_NOALLOC_SPEC_MAX) {
        char msg[256];
        snprintf(msg, sizeof msg,
                 "fy_gpu_deconvolve_noalloc: %dx%dx%d pads to %dx%dx%d, whose spectrum needs %zu bytes; "
                 "the scratch holds %zu (use fy_gpu_deconvolve)",
                 nz, ny, nx, pz, py, px, need, (size_t)FYG_NOALLOC_SPEC_MAX);
        fyg_set_error(msg);
        return 1;
    }
    fyg_stream *s = fyg_stream_wrap(stream);
    if (!s) return 1;
    if (reg <= 0.0) reg = fy_auto_reg(p, 0.015);
    const double phys8[8] = {p->delta_beta, p->energy_kev, p->distance_mm, p->pixel_um,
                             p->unsharp_sigma, p->unsharp_coeff, p->psf_sigma_vox, p->psf_sigma_z_vox};
    return fyg_noalloc_deconv(d_in, d_out, nz, ny, nx, pz, py, px, phys8, reg, FYH_MAX_GAIN,
                              lut_us(pz, py, px, 1.0), s);
}
