/* fy_chain.c -- the per-tile chain of fy_process_buffer (pipeline.c:252-297,
 * 413-421, 447-510) scheduled as device work:
 *
 *   X-forward FFT kernel with the tile prologue fused into its gather
 *       (u8 -> f01, ring subtraction, z-drift, dewindow)
 *   Y forward, Z forward*weight*inverse, Y inverse
 *   X-inverse kernel with rewindow / dec-range rescale / clamp fused into its store
 *   guided filter (fast s=2 by default) whose last kernel emits the inner tile:
 *       recenter, dithered quantisation at global coordinates, all-air handling
 *
 * With deconvolution off, one prologue kernel produces the float tile instead.
 * The all-zero test of the reference (pipeline.c:473-477) becomes a device flag
 * set by whichever kernel reads the u8 input; the emit stage writes zeros when
 * it is clear, so no host round trip sits between the stages.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "fyh.h"

/* FNV-1a over the content: cache key for small per-volume arrays */
static unsigned long long fnv1a(const void *p, size_t n) {
    const unsigned char *b = (const unsigned char *)p;
    unsigned long long h = 1469598103934665603ULL;
    for (size_t i = 0; i < n; i++) { h ^= b[i]; h *= 1099511628211ULL; }
    return h;
}

const void *fyh_const_blob(char tag, const void *data, size_t bytes) {
    struct { char tag[8]; unsigned long long h; size_t n; } key;
    memset(&key, 0, sizeof key);
    key.tag[0] = 'B'; key.tag[1] = tag;
    key.h = fnv1a(data, bytes);
    key.n = bytes;
    const void *d = fyg_table_find(&key, sizeof key);
    return d ? d : fyg_table(&key, sizeof key, data, bytes);
}

void fyh_scaled_physics(const fy_pipeline_cfg *c, fy_physics *p) {
    p->delta_beta = c->delta_beta; p->energy_kev = c->energy_kev; p->distance_mm = c->distance_mm;
    p->pixel_um = c->pixel_um; p->unsharp_sigma = c->unsharp_sigma; p->unsharp_coeff = c->unsharp_coeff;
    p->psf_sigma_vox = c->psf_sigma_vox;
    p->psf_sigma_z_vox = (c->psf_z_ratio > 1.0 && c->psf_sigma_vox > 0) ? c->psf_z_ratio * c->psf_sigma_vox : 0.0;
    p->delta_beta *= fy_auto_deltabeta_scale(p);
}

/* per-tile guided eps from the radial noise fit (pipeline.c:447-456) */
static double eps_for_tile(const fy_pipeline_cfg *cfg, double ecy, double ecx, long y0, long x0, long ty,
                           long tx) {
    if (!cfg->have_eps_r || cfg->guided_eps <= 0 || cfg->eps_fn_med <= 0) return 0.0;
    double ry = y0 + ty / 2.0 - ecy, rx = x0 + tx / 2.0 - ecx;
    double fn = cfg->eps_fn_a + cfg->eps_fn_b * sqrt(ry * ry + rx * rx);
    double sc = fn / cfg->eps_fn_med, s2 = sc * sc;
    if (s2 < 0.4) s2 = 0.4;
    if (s2 > 8.0) s2 = 8.0;
    return cfg->guided_eps * s2;
}

int fy_cube_params_from_cfg(const fy_pipeline_cfg *cfg, long y0, long x0, long ty, long tx, long vol_y,
                            long vol_x, fy_cube_params *o) {
    memset(o, 0, sizeof *o);
    if (cfg->do_musica) return 1;                              /* stage not provided */
    fyh_scaled_physics(cfg, &o->phys);
    if (cfg->do_deconv) {
        int matched = cfg->use_matched_deconv && cfg->psf_sigma_vox > 0;
        o->deconv = matched ? 2 : 1;
        o->deconv_param = matched ? cfg->deconv_tikhonov : -1.0;
        o->deconv_cap = matched ? 0.0 : FYH_MAX_GAIN;
        o->fz2_scale = matched ? fyh_matched_fz2_scale(&o->phys) : 1.0;
        o->window_lo = cfg->window_lo;
        o->window_hi = cfg->window_hi;
        o->have_dec_range = cfg->have_dec_range;
        o->dec_lo = cfg->dec_lo;
        o->dec_hi = cfg->dec_hi;
    }
    double ecy = cfg->dering_cy > 0 ? cfg->dering_cy : (vol_y - 1) / 2.0;
    double ecx = cfg->dering_cx > 0 ? cfg->dering_cx : (vol_x - 1) / 2.0;
    double eps_t = eps_for_tile(cfg, ecy, ecx, y0, x0, ty, tx);
    if (cfg->guided_eps > 0 || eps_t > 0) {
        o->guided_eps = eps_t > 0 ? eps_t : cfg->guided_eps;
        o->guided_radius = 2;
        o->guided_s = cfg->guided_subsample == 0 ? 2 : cfg->guided_subsample;
    }
    o->no_dither = cfg->no_dither;
    if (cfg->have_norm && cfg->norm_hi > cfg->norm_lo) {
        o->recenter = 1;
        o->norm_lo = cfg->norm_lo;
        o->norm_hi = cfg->norm_hi;
    }
    if (cfg->have_dering && cfg->dering) o->dering = cfg->dering;
    if (cfg->have_zdrift && cfg->zdrift_factor) {
        o->zdrift_factor = cfg->zdrift_factor;
        o->vol_z = cfg->vol_z;
    }
    if (cfg->do_air_zero) {                                     /* pipeline.c:302-399 */
        o->air_zero = 1;
        o->air_cut_u8 = cfg->air_cut_u8;
        o->air_cut_band = cfg->air_cut_band;
        o->air_thresh = cfg->air_thresh;
        o->air_cut_aggr = cfg->air_cut_aggr;
        o->air_min_component = cfg->air_min_component;
        o->scratch_passes = cfg->scratch_passes;
    }
    return 0;
}

/* One cube's chain, planned on the host: every argument of its device stages
 * except the workspace addresses, which run_group() hands out per group. */
typedef struct {
    const fy_cube *c;
    int skip;                   /* empty region: nothing to do */
    size_t hn;                  /* voxels of the halo'd region */
    int deconv, guided, f64;
    size_t tile_bytes;          /* the deconvolved tile: fp32, or fp64 for the all-FP64 variant */
    fyg_prologue pro;           /* u8 prologue (both paths) */
    fyg_deconv_args d;          /* deconv = 1 */
    fyg_guided_args g;          /* guided = 1 */
    fyg_emit em;
    int air;                    /* stage (d): air-zero + despeckle */
    fyg_air_args air_args;
    size_t spec_bytes, img_bytes, gws_bytes, air_bytes;
} cube_plan;

static size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

/* stage (d) workspace: the pre-deconvolution tile, the float tile it zeroes
 * (when the caller gives none) and the kernels' scratch */
static size_t air_bytes(long hz, long hy, long hx) {
    size_t hn = (size_t)hz * hy * hx;
    return 2 * al256(sizeof(float) * hn) + al256(fyg_air_ws_bytes((int)hz, (int)hy, (int)hx));
}

/* the transfer table a cube's deconvolution uses; 0: the cube has none */
static int deconv_req(const fy_cube *c, fyh_table_req *r) {
    const fy_cube_params *P = c->params;
    if (!P || !P->deconv || c->hz <= 0 || c->hy <= 0 || c->hx <= 0) return 0;
    int matched = P->deconv == 2;
    double param = P->deconv_param;
    if (param <= 0) param = matched ? 0.05 : fy_auto_reg(&P->phys, 0.015);
    r->nz = (int)c->hz; r->ny = (int)c->hy; r->nx = (int)c->hx;
    r->kind = matched; r->f64 = P->fp64 != 0; r->p = &P->phys;
    r->param = param; r->cap = P->deconv_cap; r->fz2_scale = P->fz2_scale > 0 ? P->fz2_scale : 1.0;
    return 1;
}

static int plan_cube(const fy_cube *c, cube_plan *pl, const fyh_lut_memo *memo, fyg_stream *s) {
    memset(pl, 0, sizeof *pl);
    pl->c = c;
    const fy_cube_params *P = c->params;
    long rz0 = c->rz0, ry0 = c->ry0, rx0 = c->rx0, hz = c->hz, hy = c->hy, hx = c->hx;
    if (hz <= 0 || hy <= 0 || hx <= 0 || c->tz <= 0 || c->ty <= 0 || c->tx <= 0) {
        pl->skip = 1;
        return 0;
    }
    if (!P || c->z0 < rz0 || c->y0 < ry0 || c->x0 < rx0 || c->z0 + c->tz > rz0 + hz ||
        c->y0 + c->ty > ry0 + hy || c->x0 + c->tx > rx0 + hx)
        return 1;                                               /* inner tile must lie in the region */
    pl->hn = (size_t)hz * hy * hx;
    pl->f64 = P->fp64 != 0;
    /* the all-FP64 variant covers deconvolution + fast guided filtering */
    if (pl->f64 && (!P->deconv || P->guided_eps <= 0 || P->guided_s == 1)) return 1;
    pl->tile_bytes = (pl->f64 ? sizeof(double) : sizeof(float)) * pl->hn;

    /* prologue (pipeline.c:492-494 and the dewindow of :266-268) */
    fyg_prologue *pro = &pl->pro;
    pro->gz0 = rz0; pro->gy0 = ry0; pro->gx0 = rx0;
    if (P->dering && P->dering->have_rings) {
        const fy_dering *d = P->dering;
        pro->ring = (const float *)fyh_const_blob('R', d->ring, sizeof(float) * (size_t)d->nslab * d->nr);
        if (!pro->ring) return 1;
        pro->dering = 1; pro->nr = d->nr; pro->nslab = d->nslab; pro->slab_z = d->slab_z;
        pro->cy = d->cy; pro->cx = d->cx; pro->scale = (float)(1.0 / 255.0);
        pro->ring_fast = fyh_ring_fast(d->cy, d->cx, ry0, rx0, hy, hx);
    }
    if (P->zdrift_factor) {
        if (P->vol_z < rz0 + hz || rz0 < 0) return 1;           /* factor table must cover the region */
        pro->zfactor = (const float *)fyh_const_blob('Z', P->zdrift_factor, sizeof(float) * (size_t)P->vol_z);
        if (!pro->zfactor) return 1;
        pro->zdrift = 1;
    }
    int windowed = P->window_hi > P->window_lo;
    float wlo = (float)P->window_lo, wspan = (float)(P->window_hi - P->window_lo);

    if (P->deconv) {                                            /* (b) */
        pl->deconv = 1;
        fyh_table_req rq;
        fyh_tables t;
        if (!deconv_req(c, &rq) || fyh_tables_get_m(&rq, memo, s, &t)) return 1;
        if (!fyg_fft_supported(t.pz) || !fyg_fft_supported(t.py) || !fyg_fft_supported(t.px)) return 1;
        fyg_deconv_args *a = &pl->d;
        a->nz = (int)hz; a->ny = (int)hy; a->nx = (int)hx;
        a->pz = t.pz; a->py = t.py; a->px = t.px; a->hx = t.hx; a->hxp = t.hxp;
        a->lut = t.lut; a->us = t.us; a->fz2 = t.fz2; a->fy2 = t.fy2; a->fx2 = t.fx2;
        a->in = c->d_in; a->in_u8 = 1;
        a->pro = *pro;
        if (windowed) { a->pro.window = 1; a->pro.wlo = wlo; a->pro.wspan = wspan; }
        if (windowed) { a->epi.window = 1; a->epi.wlo = wlo; a->epi.winv = 1.0f / wspan; }
        if (P->have_dec_range && P->dec_hi - P->dec_lo > 1e-6) {
            a->epi.dec = 1; a->epi.dlo = (float)P->dec_lo; a->epi.dinv = 1.0f / (float)(P->dec_hi - P->dec_lo);
        }
        a->epi.clamp01 = 1;
        a->f64 = pl->f64;
        pl->spec_bytes = pl->f64 ? fyg_deconv_spec_bytes64(t.pz, t.py, t.px) : fyg_deconv_spec_bytes(t.pz, t.py, t.px);
        if (pro->dering) pl->img_bytes = fyg_ring_image_bytes(pro, (int)hz, (int)hy, (int)hx);
    }

    fyg_emit *em = &pl->em;
    em->quant = 1;
    em->q = c->d_out;
    em->iz0 = c->z0 - rz0; em->iy0 = c->y0 - ry0; em->ix0 = c->x0 - rx0;
    em->tz = c->tz; em->ty = c->ty; em->tx = c->tx;
    em->gz0 = rz0; em->gy0 = ry0; em->gx0 = rx0;
    em->no_dither = P->no_dither;
    if (P->recenter && P->norm_hi > P->norm_lo) {               /* (d.5) */
        em->recenter = 1;
        em->rmid = 0.5f * (float)(P->norm_lo + P->norm_hi) / 255.0f;
        em->rk = (2.0f * 0.375f) * 255.0f / (float)(P->norm_hi - P->norm_lo);
    }

    if (P->guided_eps > 0) {                                    /* (c) */
        pl->guided = 1;
        fyg_guided_args *g = &pl->g;
        g->nz = (int)hz; g->ny = (int)hy; g->nx = (int)hx;
        g->r = P->guided_radius < 1 ? 2 : P->guided_radius;
        g->s = P->guided_s == 0 ? 2 : P->guided_s;
        g->eps = P->guided_eps;
        g->f64 = pl->f64;
        pl->gws_bytes = pl->f64 ? fyg_guided_ws_bytes64(g->nz, g->ny, g->nx, g->s)
                                : fyg_guided_ws_bytes(g->nz, g->ny, g->nx, g->s);
        if (pl->f64 && (g->s < 2 || g->r / g->s > 2 || (hz + g->s - 1) / g->s < 2 || (hy + g->s - 1) / g->s < 2 ||
                        (hx + g->s - 1) / g->s < 2))
            return 1;
    }
    if (P->air_zero) {                                          /* (d) */
        if (pl->f64 && !pl->guided) return 1;
        pl->air = 1;
        fyg_air_args *aa = &pl->air_args;
        aa->scratch_passes = P->scratch_passes;
        aa->cut_u8 = P->air_cut_u8;
        aa->band = P->air_cut_band;
        aa->min_component = P->air_min_component;
        aa->thresh = P->air_thresh;
        aa->aggr = P->air_cut_aggr;
        pl->air_bytes = air_bytes(hz, hy, hx);
    }
    return 0;
}

/* cubes that can share every launch of the chain */
static int same_group(const cube_plan *a, const cube_plan *b) {
    const fy_cube *x = a->c, *y = b->c;
    if (x->hz != y->hz || x->hy != y->hy || x->hx != y->hx || x->tz != y->tz || x->ty != y->ty || x->tx != y->tx)
        return 0;
    if (a->deconv != b->deconv || a->guided != b->guided || a->pro.dering != b->pro.dering || a->f64 != b->f64 ||
        a->air != b->air)
        return 0;
    if ((x->d_float != NULL) != (y->d_float != NULL)) return 0;
    if (a->guided && (a->g.s != b->g.s || a->g.r != b->g.r)) return 0;
    return 1;
}

/* the generation word of a run, still to be stepped before its first kernel
 * reads it (NULL once stepped), and the number of flags that follow it */
typedef struct {
    int *gen;
    int nflag;
} gen_step;

/* the chain of n planned cubes of one group; flags[i] is cube i's all-air flag.
 * The run's first group steps the generation: inside its first ring-image
 * launch when it has one, else with a launch of its own. */
static int run_group(cube_plan *pl, int n, int **flags, const int *gen, gen_step *st, fyg_stream *s) {
    const cube_plan *h = &pl[0];
    const int fold = st->gen && h->deconv && h->pro.dering;
    if (st->gen && !fold) {
        if (fyg_gen_step(st->gen, st->nflag, s)) return 1;
        st->gen = NULL;
    }
    size_t tile = al256(h->tile_bytes), spec = al256(h->spec_bytes), img = 0, gws = al256(h->gws_bytes);
    for (int i = 0; i < n; i++)                                 /* ring images may span more slabs */
        if (al256(pl[i].img_bytes) > img) img = al256(pl[i].img_bytes);
    char *T = (char *)fyg_ws(s, WS_TILE, tile * n);
    char *S = spec ? (char *)fyg_ws(s, WS_SPEC, spec * n) : NULL;
    char *I = img ? (char *)fyg_ws(s, WS_MISC, img * n) : NULL;
    char *G = gws ? (char *)fyg_ws(s, WS_GUIDED, gws * n) : NULL;
    size_t air = al256(h->air_bytes);
    char *A = air ? (char *)fyg_ws(s, WS_AIR, air * n) : NULL;
    if (!T || (spec && !S) || (img && !I) || (gws && !G) || (air && !A)) return 1;

    if (h->deconv) {
        fyg_deconv_args *da = (fyg_deconv_args *)malloc(sizeof *da * n);
        if (!da) return 1;
        for (int i = 0; i < n; i++) {
            da[i] = pl[i].d;
            da[i].out = T + tile * i;
            da[i].spec = S + spec * i;
            if (pl[i].pro.dering) da[i].pro.cimg = (float *)(I + img * i);
            da[i].nonzero_flag = flags[i];
            da[i].nonzero_gen = gen;
        }
        if (fold) {
            da[0].step_gen = st->gen;
            da[0].step_nflag = st->nflag;
            st->gen = NULL;
        }
        int rc = fyg_deconv_run_batch(da, n, s);
        free(da);
        if (rc) return 1;
    } else {
        for (int i = 0; i < n; i++)
            if (fyg_prologue_run(pl[i].c->d_in, (float *)(T + tile * i), (int)pl[i].c->hz, (int)pl[i].c->hy,
                                 (int)pl[i].c->hx, &pl[i].pro, flags[i], gen, s))
                return 1;
    }

    for (int i = 0; i < n; i++) {
        pl[i].em.nonzero_flag = flags[i];
        pl[i].em.nonzero_gen = gen;
    }
    if (h->guided && !h->c->d_float && !h->air) {               /* filter + emit in one pass */
        fyg_guided_args *ga = (fyg_guided_args *)malloc(sizeof *ga * n);
        if (!ga) return 1;
        for (int i = 0; i < n; i++) {
            ga[i] = pl[i].g;
            if (pl[i].f64) ga[i].in64 = (const double *)(T + tile * i);
            else ga[i].in = (const float *)(T + tile * i);
            ga[i].ws = G + gws * i;
            ga[i].em = pl[i].em;
        }
        int rc = fyg_guided_run_batch(ga, n, s);
        free(ga);
        return rc;
    }
    for (int i = 0; i < n; i++) {               /* fp32 copy requested, air-zero, or no filter */
        const fy_cube *c = pl[i].c;
        const float *src = (const float *)(T + tile * i);
        size_t fb = al256(sizeof(float) * pl[i].hn);
        float *orig = h->air ? (float *)(A + air * i) : NULL;
        float *F = c->d_float ? c->d_float : (h->air ? (float *)(A + air * i + fb) : NULL);
        if (h->guided) {
            fyg_guided_args g = pl[i].g;
            if (pl[i].f64) g.in64 = (const double *)(T + tile * i);
            else g.in = src;
            g.ws = G + gws * i;
            g.em = pl[i].em;
            g.em.quant = 0;
            g.em.out = F;
            if (fyg_guided_run(&g, s)) return 1;
            src = F;
        } else if (F) {
            if (fyg_d2d(F, src, sizeof(float) * pl[i].hn, s)) return 1;
            src = F;
        }
        if (h->air) {                                           /* (d), pipeline.c:302-399 */
            if (fyg_prologue_run(c->d_in, orig, (int)c->hz, (int)c->hy, (int)c->hx, &pl[i].pro, flags[i], gen, s) ||
                fyg_air_zero_run(F, orig, (int)c->hz, (int)c->hy, (int)c->hx, &pl[i].air_args, A + air * i + 2 * fb,
                                 s))
                return 1;
        }
        if (fyg_emit_run(src, (int)c->hz, (int)c->hy, (int)c->hx, &pl[i].em, s)) return 1;
    }
    return 0;
}

/* set on a thread whose caller runs several streams itself (the served form's
 * lanes): its groups then run one after another on the one stream */
static __thread int no_fork;

/* device workspace one group may hold at once; larger groups run in parts */
#define FYH_GROUP_BYTES ((size_t)8 << 30)

/* the cube can share mixed launches: the u8 chain of deconvolution and fast
 * guided filter emitting its inner tile, in fp32, at lengths those launches
 * cover */
static int mix_ok(const cube_plan *p) {
    if (!p->deconv || !p->guided || p->f64 || p->air || p->c->d_float) return 0;
    const fyg_deconv_args *d = &p->d;
    return d->in_u8 && fyg_mix_len_ok(d->px) && fyg_mix_len_ok(d->py) && fyg_mix_len_ok(d->pz) &&
           fyg_mix_guided_ok(&p->g);
}

/* n planned cubes of several shapes: every stage one launch for all of them
 * (fyg_mix_*), each cube with its own workspace slices and flag, as in
 * run_group */
static int run_mix(cube_plan *pl, int n, int **flags, const int *gen, gen_step *st, fyg_stream *s) {
    size_t tile = 0, spec = 0, img = 0, gws = 0;
    for (int i = 0; i < n; i++) {
        tile += al256(pl[i].tile_bytes);
        spec += al256(pl[i].spec_bytes);
        img += al256(pl[i].img_bytes);
        gws += al256(pl[i].gws_bytes);
    }
    char *T = (char *)fyg_ws(s, WS_TILE, tile);
    char *S = (char *)fyg_ws(s, WS_SPEC, spec);
    char *I = img ? (char *)fyg_ws(s, WS_MISC, img) : NULL;
    char *G = (char *)fyg_ws(s, WS_GUIDED, gws);
    if (!T || !S || (img && !I) || !G) return 1;
    fyg_deconv_args *da = (fyg_deconv_args *)malloc(sizeof *da * (size_t)n);
    fyg_guided_args *ga = (fyg_guided_args *)malloc(sizeof *ga * (size_t)n);
    int rc = !da || !ga;
    size_t ot = 0, os = 0, oi = 0, og = 0;
    for (int i = 0; i < n && !rc; i++) {
        da[i] = pl[i].d;
        da[i].out = T + ot;
        da[i].spec = S + os;
        if (pl[i].pro.dering) da[i].pro.cimg = (float *)(I + oi);
        da[i].nonzero_flag = flags[i];
        da[i].nonzero_gen = gen;
        ga[i] = pl[i].g;
        ga[i].in = (const float *)(T + ot);
        ga[i].ws = G + og;
        ga[i].em = pl[i].em;
        ga[i].em.nonzero_flag = flags[i];
        ga[i].em.nonzero_gen = gen;
        ot += al256(pl[i].tile_bytes);
        os += al256(pl[i].spec_bytes);
        oi += al256(pl[i].img_bytes);
        og += al256(pl[i].gws_bytes);
    }
    if (!rc && st->gen) {                       /* the first launch steps the generation */
        da[0].step_gen = st->gen;
        da[0].step_nflag = st->nflag;
        st->gen = NULL;
    }
    if (!rc) rc = fyg_mix_deconv_run(da, n, s) || fyg_mix_guided_run(ga, n, s);
    free(da);
    free(ga);
    return rc;
}

/* cubes planned before their launches are queued: bounds the per-stream
 * transfer tables held at once (FYG_ARENA_SLOTS/2) */
#define FYH_PLAN_CHUNK (FYG_ARENA_SLOTS / 2)

/* Mixed launches over n cubes.  Mode 2: every stage one launch for all of
 * them.  Mode 1: the cubes are packed (largest spectrum first, first fit)
 * into parts whose half spectra fit FYH_MIX_PART_BYTES, each part runs as one
 * launch per stage, and the parts alternate between two lanes (the stream and
 * a helper stream).  Between a cube's passes only the spectra of the two
 * parts in flight (at most 192 MB) stream through the chip, which the 256 MB
 * last-level cache (MALL) holds, instead of every cube's; the second lane
 * fills the first one's stage tails.  Measured on the 96-cube mix (MI355X):
 * whole batch 2.52 ms, parts of 224 MB on one lane 2.39, parts of 96 MB on two
 * lanes 2.18, grouped launches 2.29. */
#define FYH_MIX_PART_BYTES ((size_t)96 << 20)
#define FYH_MIX_LANES 2
static int run_mixed(cube_plan *pl, int n, int **flags, const int *gen, gen_step *st, fyg_stream *s) {
    if (fyg_mix_mode() == 2 || n < 2) return run_mix(pl, n, flags, gen, st, s);
    int *ord = (int *)malloc(sizeof(int) * (size_t)n), *part = (int *)malloc(sizeof(int) * (size_t)n);
    size_t *fill = (size_t *)calloc((size_t)n, sizeof(size_t));
    cube_plan *pp = (cube_plan *)malloc(sizeof *pp * (size_t)n);
    int **pf = (int **)malloc(sizeof(int *) * (size_t)n);
    int rc = !ord || !part || !fill || !pp || !pf;
    int nparts = 0;
    for (int i = 0; i < n && !rc; i++) ord[i] = i;
    for (int i = 1; i < n && !rc; i++)                  /* insertion sort, largest spectrum first (stable) */
        for (int j = i; j > 0 && pl[ord[j]].spec_bytes > pl[ord[j - 1]].spec_bytes; j--) {
            int t = ord[j]; ord[j] = ord[j - 1]; ord[j - 1] = t;
        }
    for (int k = 0; k < n && !rc; k++) {
        const int i = ord[k];
        const size_t b = al256(pl[i].spec_bytes);
        int p = 0;
        while (p < nparts && fill[p] + b > FYH_MIX_PART_BYTES) p++;
        if (p == nparts) nparts++;
        fill[p] += b;
        part[i] = p;
    }
    const int lanes = nparts < FYH_MIX_LANES ? nparts : FYH_MIX_LANES;
    if (!rc && lanes > 1) {
        if (st->gen) {
            rc = fyg_gen_step(st->gen, st->nflag, s);
            st->gen = NULL;
        }
        if (!rc) rc = fyg_fork(s, lanes - 1);
    }
    int m = 0;
    for (int p = 0; p < nparts && !rc; p++) {
        const int m0 = m;
        for (int i = 0; i < n; i++)
            if (part[i] == p) {
                pp[m] = pl[i];
                pf[m] = flags[i];
                m++;
            }
        fyg_stream *ls = (lanes > 1 && p % lanes) ? fyg_stream_sub(s, p % lanes - 1) : s;
        rc = !ls || run_mix(pp + m0, m - m0, pf + m0, gen, st, ls);
    }
    if (lanes > 1 && fyg_join(s, lanes - 1)) rc = 1;
    free(ord); free(part); free(fill); free(pp); free(pf);
    return rc;
}

/* Plan every cube of a chunk, then run the chunk group by group (one shape and
 * stage set per group, in order of first appearance), each group in as few
 * launches per stage as FYG_BATCH and FYH_GROUP_BYTES allow.  Results do not
 * depend on the grouping: every cube has its own workspace slice and flag. */
static int run_chunk(const fy_cube *cubes, int n, int *d_flags, const int *gen, gen_step *st, fyg_stream *s) {
    /* the cached tables the plans point at (ring and z-drift blobs, transfer
     * tables of the process-wide cache) stay held until the chunk's launches
     * are queued and their completion is tracked by an event */
    fyg_hold *hold = fyg_hold_begin();
    cube_plan *pl = (cube_plan *)malloc(sizeof *pl * (size_t)n);
    cube_plan *grp = (cube_plan *)malloc(sizeof *grp * (size_t)n);
    int *gi = (int *)malloc(sizeof(int) * (size_t)n);
    int **gf = (int **)malloc(sizeof(int *) * (size_t)n);
    int rc = (!pl || !grp || !gi || !gf) ? 1 : 0;
    /* transfer tables of cubes with their own physics: built ahead, in parallel */
    fyh_lut_memo *memo = NULL;
    fyh_table_req *rq = rc || n < 2 ? NULL : (fyh_table_req *)malloc(sizeof *rq * (size_t)n);
    if (rq) {
        int nr = 0;
        for (int i = 0; i < n; i++) nr += deconv_req(&cubes[i], &rq[nr]);
        if (nr > 1) memo = fyh_tables_prefill(rq, nr, s);
        free(rq);
    }
    for (int i = 0; i < n && !rc; i++) {
        rc = plan_cube(&cubes[i], &pl[i], memo, s);
        gi[i] = -1;
    }
    fyh_lut_memo_free(memo);
    /* the launch groups: cubes of one shape and stage set, in parts that fit
     * FYH_GROUP_BYTES; grp is filled group by group */
    typedef struct { int first, m; double work; int lane; } part_t;
    part_t *parts = rc ? NULL : (part_t *)malloc(sizeof(part_t) * (size_t)n);
    int np = 0, ng = 0;
    if (!rc && !parts) rc = 1;
    for (int i = 0; i < n && !rc; i++) {
        if (gi[i] >= 0 || pl[i].skip) continue;
        int m = 0;
        for (int j = i; j < n; j++) {
            if (gi[j] < 0 && !pl[j].skip && same_group(&pl[i], &pl[j])) {
                gi[j] = i;
                grp[ng + m] = pl[j];
                gf[ng + m] = d_flags + j;
                m++;
            }
        }
        size_t per = al256(pl[i].tile_bytes) + al256(pl[i].spec_bytes) + al256(pl[i].gws_bytes) +
                     2 * al256(pl[i].img_bytes) + al256(pl[i].air_bytes);
        int part = (int)(FYH_GROUP_BYTES / (per ? per : 1));
        if (part < 1) part = 1;
        for (int k = 0; k < m; k += part) {
            int mk = m - k < part ? m - k : part;
            parts[np].first = ng + k;
            parts[np].m = mk;
            parts[np].work = (double)mk * (double)(per ? per : 1);
            parts[np].lane = 0;
            np++;
        }
        ng += m;
    }
    /* Groups of several shapes: mixed launches when each cube's chain allows
     * it and their workspaces fit one group's bound.  Not on a served lane,
     * whose waves already overlap on the device: there the grouped launches
     * measured faster (capacity 86.7 against 77.6 Gvox/s). */
    if (!rc && np > 1 && fyg_mix_mode() && !no_fork) {
        int ok = ng <= FYG_MIX_MAX;
        size_t per = 0;
        for (int i = 0; i < ng && ok; i++) {
            ok = mix_ok(&grp[i]) && grp[i].pro.dering == grp[0].pro.dering;
            per += al256(grp[i].tile_bytes) + al256(grp[i].spec_bytes) + al256(grp[i].gws_bytes) +
                   al256(grp[i].img_bytes);
        }
        if (ok && per <= FYH_GROUP_BYTES) {
            rc = run_mixed(grp, ng, gf, gen, st, s);
            free(parts);
            free(pl); free(grp); free(gi); free(gf);
            if (fyg_hold_release(hold, s)) rc = 1;
            return rc;
        }
    }
    /* Several groups run side by side on helper streams: the small cubes' short
     * grids then share the chip with the large ones instead of leaving it
     * half idle one group after another.  Each lane has its own workspaces;
     * results do not depend on the lane. */
    int lanes = no_fork ? 1 : (np < 1 + FYG_SUBS ? np : 1 + FYG_SUBS);
    if (!rc && lanes > 1) {
        double load[1 + FYG_SUBS] = {0};
        for (int a = 0; a < np; a++) {                  /* largest part first, to the least loaded lane */
            int best = -1;
            for (int b = 0; b < np; b++)
                if (parts[b].lane == 0 && parts[b].work >= 0 && (best < 0 || parts[b].work > parts[best].work))
                    best = b;
            int l = 0;
            for (int k = 1; k < lanes; k++)
                if (load[k] < load[l]) l = k;
            load[l] += parts[best].work;
            parts[best].lane = l + 1;                   /* 1-based while assigning */
            parts[best].work = -1.0;
        }
        for (int a = 0; a < np; a++) parts[a].lane -= 1;
        /* the generation word is stepped before any lane reads it */
        if (st->gen) {
            rc = fyg_gen_step(st->gen, st->nflag, s);
            st->gen = NULL;
        }
        if (!rc) rc = fyg_fork(s, lanes - 1);
    }
    for (int l = 0; l < (lanes > 1 ? lanes : 1) && !rc; l++) {
        fyg_stream *ls = l == 0 ? s : fyg_stream_sub(s, l - 1);
        if (!ls) rc = 1;
        for (int a = 0; a < np && !rc; a++)
            if (parts[a].lane == l) rc = run_group(grp + parts[a].first, parts[a].m, gf + parts[a].first, gen, st, ls);
    }
    if (lanes > 1 && fyg_join(s, lanes - 1)) rc = 1;
    free(parts);
    free(pl); free(grp); free(gi); free(gf);
    if (fyg_hold_release(hold, s)) rc = 1;
    return rc;
}

/* n cubes on the stream's flags (reserved, not yet stepped) */
static int run_cubes(const fy_cube *cubes, int n, int *d_flags, int *gen, int nflag, fyg_stream *s) {
    gen_step st = {gen, nflag};
    for (int i = 0; i < n; i += FYH_PLAN_CHUNK)
        if (run_chunk(cubes + i, n - i < FYH_PLAN_CHUNK ? n - i : FYH_PLAN_CHUNK, d_flags + i, gen, &st, s))
            return 1;
    return st.gen ? fyg_gen_step(st.gen, st.nflag, s) : 0;    /* nothing ran: step anyway */
}

static int chain_cube(const fy_cube *c, int *d_flag, int *gen, int nflag, fyg_stream *s) {
    return run_cubes(c, 1, d_flag, gen, nflag, s);
}

size_t fy_cube_workspace_bytes(const fy_cube *c) {
    if (!c || !c->params || c->hz <= 0 || c->hy <= 0 || c->hx <= 0) return 0;
    const fy_cube_params *P = c->params;
    size_t hn = (size_t)c->hz * c->hy * c->hx;
    int f64 = P->fp64 != 0;
    size_t b = al256((f64 ? sizeof(double) : sizeof(float)) * hn);
    if (P->deconv) {
        int pz = fy_next_fft_size((int)c->hz), py = fy_next_fft_size((int)c->hy), px = fy_next_fft_size((int)c->hx);
        b += al256(f64 ? fyg_deconv_spec_bytes64(pz, py, px) : fyg_deconv_spec_bytes(pz, py, px));
        if (P->dering && P->dering->have_rings) {
            fyg_prologue pro;
            memset(&pro, 0, sizeof pro);
            pro.dering = 1; pro.nslab = P->dering->nslab; pro.slab_z = P->dering->slab_z; pro.gz0 = c->rz0;
            b += al256(fyg_ring_image_bytes(&pro, (int)c->hz, (int)c->hy, (int)c->hx));
        }
    }
    if (P->guided_eps > 0) {
        int sd = P->guided_s == 0 ? 2 : P->guided_s;
        b += al256(f64 ? fyg_guided_ws_bytes64((int)c->hz, (int)c->hy, (int)c->hx, sd)
                       : fyg_guided_ws_bytes((int)c->hz, (int)c->hy, (int)c->hx, sd));
    }
    if (P->air_zero) b += al256(air_bytes(c->hz, c->hy, c->hx));
    return b;
}

int fyh_process_cubes(const fy_cube *cubes, int n, fyg_stream *s) {
    if (n <= 0) return 0;
    if (!s) return 1;
    int *dflags = NULL, *dgen = NULL, nflag = 0;
    if (!fyg_flags_reserve(s, n, &dflags, &dgen, &nflag)) return 1;
    return run_cubes(cubes, n, dflags, dgen, nflag, s);
}

int fyh_process_cubes_serial(const fy_cube *cubes, int n, fyg_stream *s) {
    int old = no_fork;
    no_fork = 1;
    int rc = fyh_process_cubes(cubes, n, s);
    no_fork = old;
    return rc;
}

int fy_gpu_process_cubes(const fy_cube *cubes, int n, void *stream) {
    if (n <= 0) return 0;
    return fyh_process_cubes(cubes, n, fyg_stream_wrap(stream));
}

/* ---- captured batched form: the launches of one fy_gpu_process_cubes call
 * recorded once into a HIP graph and replayed.  The graph owns a private
 * stream object, so its workspaces and flags are never regrown or shared, and
 * its tables live in the process-wide cache. */
struct fy_graph {
    fyg_stream *s;
    void *exec;
    fyg_hold *hold;             /* the tables the recorded launches read */
};

fy_graph *fy_gpu_graph_cubes(const fy_cube *cubes, int n, void *stream) {
    if (n <= 0 || !cubes) return NULL;
    fy_graph *g = (fy_graph *)calloc(1, sizeof *g);
    if (!g) return NULL;
    g->s = fyg_stream_private(stream);
    int *dflags = NULL, *dgen = NULL, nflag = 0;
    /* an uncaptured call first: sizes every workspace and makes every table resident */
    int rc = !g->s || !fyg_flags_reserve(g->s, n, &dflags, &dgen, &nflag) ||
             run_cubes(cubes, n, dflags, dgen, nflag, g->s) || fyg_sync(g->s);
    if (!rc) {
        g->hold = fyg_hold_begin();
        rc = fyg_capture_begin(g->s);
        if (!rc) {
            int r1 = !fyg_flags_reserve(g->s, n, &dflags, &dgen, &nflag) ||
                     run_cubes(cubes, n, dflags, dgen, nflag, g->s);
            int r2 = fyg_capture_end(g->s, &g->exec);
            rc = r1 || r2;
        }
        fyg_hold_end(g->hold);
    }
    if (rc) {
        fy_graph_destroy(g);
        return NULL;
    }
    return g;
}

int fy_graph_launch(fy_graph *g) { return !g || fyg_graph_launch(g->exec, g->s); }

void fy_graph_destroy(fy_graph *g) {
    if (!g) return;
    fyg_graph_free(g->exec);
    /* replays still queued on the stream read the tables: the release's event
     * follows them */
    (void)fyg_hold_release(g->hold, g->s);
    fyg_stream_private_free(g->s);
    free(g);
}

int fyh_chain(const fy_pipeline_cfg *cfg, const unsigned char *d_u8, long rz0, long ry0, long rx0, long hz,
              long hy, long hx, long z0, long y0, long x0, long tz, long ty, long tx, long vol_y, long vol_x,
              unsigned char *d_out, int *d_flag, int *gen, int nflag, fyg_stream *s) {
    fy_cube_params P;
    if (fy_cube_params_from_cfg(cfg, y0, x0, ty, tx, vol_y, vol_x, &P)) return 1;
    fy_cube c = {d_u8, rz0, ry0, rx0, hz, hy, hx, z0, y0, x0, tz, ty, tx, d_out, NULL, &P};
    return chain_cube(&c, d_flag, gen, nflag, s);
}

int fy_process_buffer(const fy_pipeline_cfg *cfg, const unsigned char *u8buf, long rz0, long ry0, long rx0,
                      long hz, long hy, long hx, long z0, long y0, long x0, long tz, long ty, long tx,
                      long vol_y, long vol_x, unsigned char *out, int *all_air) {
    size_t hn = (size_t)hz * hy * hx, tn = (size_t)tz * ty * tx;
    if (!hn || !tn) {
        if (all_air) *all_air = 1;
        return 0;
    }
    fyg_stream *s = fyg_thread_stream();
    if (!s) return 1;
    unsigned char *du = (unsigned char *)fyg_ws(s, WS_U8, hn);
    unsigned char *dq = (unsigned char *)fyg_ws(s, WS_OUT, tn);
    int *dflag = NULL, *dgen = NULL, nflag = 0;
    if (!du || !dq || !fyg_flags_reserve(s, 1, &dflag, &dgen, &nflag)) return 1;
    if (fyg_h2d(du, u8buf, hn, s)) return 1;
    if (fyh_chain(cfg, du, rz0, ry0, rx0, hz, hy, hx, z0, y0, x0, tz, ty, tx, vol_y, vol_x, dq, dflag, dgen, nflag,
                  s))
        return 1;
    int flag = 0, gen = 0;
    if (fyg_d2h(out, dq, tn, s) || fyg_d2h(&flag, dflag, sizeof(int), s) || fyg_d2h(&gen, dgen, sizeof(int), s) ||
        fyg_sync(s))
        return 1;
    if (all_air) *all_air = flag != gen;
    return 0;
}

int fy_gpu_process_buffer(const fy_pipeline_cfg *cfg, const unsigned char *d_u8, long rz0, long ry0, long rx0,
                          long hz, long hy, long hx, long z0, long y0, long x0, long tz, long ty, long tx,
                          long vol_y, long vol_x, unsigned char *d_out, int *all_air, void *stream) {
    fyg_stream *s = fyg_stream_wrap(stream);
    if (!s) return 1;
    int *dflag = NULL, *dgen = NULL, nflag = 0;
    if (!fyg_flags_reserve(s, 1, &dflag, &dgen, &nflag)) return 1;
    if (fyh_chain(cfg, d_u8, rz0, ry0, rx0, hz, hy, hx, z0, y0, x0, tz, ty, tx, vol_y, vol_x, d_out, dflag, dgen,
                  nflag, s))
        return 1;
    if (all_air) {
        int flag = 0, gen = 0;
        if (fyg_d2h(&flag, dflag, sizeof(int), s) || fyg_d2h(&gen, dgen, sizeof(int), s) || fyg_sync(s)) return 1;
        *all_air = flag != gen;
    }
    return 0;
}
