/* fy_export.c -- feeding zarr tiles through the GPU chain.
 *
 * fy_gpu_process_chunk (fy_process_chunk, pipeline.c:518-538): one tile, synchronous, its halo'd
 * region read into the thread's pinned staging buffer.
 *
 * fy_gpu_export: the pass-2 tile loop of fy_run_pipeline (pipeline.c:1000-1070)
 * as a pipeline.  The reference gives each OpenMP thread one tile at a time
 * (read, chain, quantise, write).  Here host threads only do I/O -- they read
 * halo'd regions straight into pinned slots and write finished chunks -- while
 * the calling thread moves slots through three GPU stages on three streams.
 * The reads share one mapping of each input chunk (fy_gpu_zarr_cache): the up
 * to 27 chunks a halo'd tile touches are each mapped once per export, not once per
 * neighbouring tile.
 *
 *   reads (threads) -> H2D (copy stream) -> fy_gpu_process_cubes (the caller's
 *   stream, shape-grouped launches) -> D2H (second copy stream) -> writes (threads)
 *
 * Six slots of `batch` tiles circulate, so the reads of the next slots, the
 * kernels of slot b and the writes of the previous ones overlap, and both PCIe
 * directions run beside the kernels.  Events order the stages; a slot is refilled only after
 * its chunks are written, which is after every stage that touched its buffers.
 * Tiles whose region is all zero are skipped and their chunks stay absent
 * (read back as fill 0), as in the reference; failed reads and writes are
 * counted, and the call then returns 1.
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "fyh.h"

static long lmin(long a, long b) { return a < b ? a : b; }
static long lmax(long a, long b) { return a > b ? a : b; }

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static int any_nonzero(const unsigned char *p, size_t n) {
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
        unsigned long long w;
        memcpy(&w, p + i, 8);
        if (w) return 1;
    }
    for (; i < n; i++)
        if (p[i]) return 1;
    return 0;
}

/* tile t of the pass-2 grid (pipeline.c:1021-1028): inner origin and size,
 * halo'd region origin and size, all clamped to the volume */
typedef struct {
    long iz, iy, ix;
    long z0, y0, x0, tz, ty, tx;
    long rz0, ry0, rx0, hz, hy, hx;
} tile_geom;

static void geom(long Z, long Y, long X, long nty, long ntx, int tile, int halo, long t, tile_geom *g) {
    g->iz = t / (nty * ntx);
    long r = t % (nty * ntx);
    g->iy = r / ntx;
    g->ix = r % ntx;
    g->z0 = g->iz * tile;
    g->y0 = g->iy * tile;
    g->x0 = g->ix * tile;
    g->tz = lmin(tile, Z - g->z0);
    g->ty = lmin(tile, Y - g->y0);
    g->tx = lmin(tile, X - g->x0);
    g->rz0 = lmax(0, g->z0 - halo);
    g->ry0 = lmax(0, g->y0 - halo);
    g->rx0 = lmax(0, g->x0 - halo);
    g->hz = lmin(Z, g->z0 + g->tz + halo) - g->rz0;
    g->hy = lmin(Y, g->y0 + g->ty + halo) - g->ry0;
    g->hx = lmin(X, g->x0 + g->tx + halo) - g->rx0;
}

int fy_gpu_process_chunk(const fy_zarr *zin, const fy_pipeline_cfg *cfg, long z0, long y0, long x0, int tile,
                         unsigned char *out, int *out_tz, int *out_ty, int *out_tx, int *all_air) {
    const long Z = zin->shape[0], Y = zin->shape[1], X = zin->shape[2];
    const int halo = cfg->halo > 0 ? cfg->halo : 8;
    const long tz = lmin(tile, Z - z0), ty = lmin(tile, Y - y0), tx = lmin(tile, X - x0);
    if (out_tz) *out_tz = (int)tz;
    if (out_ty) *out_ty = (int)ty;
    if (out_tx) *out_tx = (int)tx;
    if (all_air) *all_air = 1;
    if (tile < 1 || z0 < 0 || y0 < 0 || x0 < 0 || tz <= 0 || ty <= 0 || tx <= 0) {
        fyg_set_error("fy_gpu_process_chunk: tile outside the volume");
        return 1;
    }
    const long rz0 = lmax(0, z0 - halo), ry0 = lmax(0, y0 - halo), rx0 = lmax(0, x0 - halo);
    const long hz = lmin(Z, z0 + tz + halo) - rz0, hy = lmin(Y, y0 + ty + halo) - ry0,
               hx = lmin(X, x0 + tx + halo) - rx0;
    fyg_stream *s = fyg_thread_stream();
    if (!s) return 1;
    unsigned char *u8 = (unsigned char *)fyg_pinned(s, 0, (size_t)hz * hy * hx);
    if (!u8) return 1;
    if (fy_gpu_zarr_read(zin, rz0, ry0, rx0, hz, hy, hx, u8)) return 1;
    return fy_process_buffer(cfg, u8, rz0, ry0, rx0, hz, hy, hx, z0, y0, x0, tz, ty, tx, Y, X, out, all_air);
}

/* ------------------------------------------------------------------ export */
#define NSLOT 6
enum { S_FREE = 0, S_READING, S_SUBMITTED };

typedef struct {
    int state;
    long t0;
    int n;                      /* tiles t0 .. t0+n-1 */
    int reads_left, writes_left;
    unsigned char *hin, *hout;  /* pinned: batch regions, batch inner tiles */
    unsigned char *din, *dout;  /* device */
    tile_geom *g;
    int *air, *fail, *cube_of;
    fy_cube *cubes;
    fy_cube_params *prm;
    void *e_in, *e_comp, *e_out;
} slot_t;

typedef struct {
    int kind;                   /* 0 read, 1 write, 2 stop, 3 await a slot then queue its writes */
    int slot, k;
} task_t;

typedef struct {
    const fy_zarr *zin, *zout;
    const fy_pipeline_cfg *cfg;
    int tile, halo, batch;
    long Z, Y, X, nty, ntx;
    size_t cap, tcap;
    slot_t sl[NSLOT];
    fy_gpu_zarr_cache *zc;             /* input chunk mappings shared by the reads */
    task_t *q;
    int qcap, qhead, qn;
    pthread_mutex_t mu;
    pthread_cond_t cv_task, cv_main;
    long written, air, io_fail;
    double read_s, write_s;
    int dev_err;
    char err[256];
} exporter;

static void push(exporter *e, task_t t) {   /* mu held */
    e->q[(e->qhead + e->qn) % e->qcap] = t;
    e->qn++;
    pthread_cond_signal(&e->cv_task);
}

static void note_err(exporter *e, const char *msg) {   /* mu held */
    if (!e->err[0]) snprintf(e->err, sizeof e->err, "%s", msg);
}

static void do_read(exporter *e, slot_t *s, int k) {
    const tile_geom *g = &s->g[k];
    unsigned char *dst = s->hin + (size_t)k * e->cap;
    const size_t hn = (size_t)g->hz * g->hy * g->hx;
    double t = now_s();
    int bad = fy_gpu_zarr_read_cached(e->zin, e->zc, g->rz0, g->ry0, g->rx0, g->hz, g->hy, g->hx, dst);
    int air = bad ? 1 : !any_nonzero(dst, hn);
    t = now_s() - t;
    pthread_mutex_lock(&e->mu);
    s->fail[k] = bad;
    s->air[k] = air;
    if (bad) {
        e->io_fail++;
        note_err(e, fyg_error_string());
    } else if (air) {
        e->air++;
    }
    e->read_s += t;
    if (--s->reads_left == 0) pthread_cond_broadcast(&e->cv_main);
    pthread_mutex_unlock(&e->mu);
}

static void do_write(exporter *e, slot_t *s, int k) {
    const tile_geom *g = &s->g[k];
    int bad = fyg_event_wait(s->e_out) != 0;
    double t = now_s();
    if (!bad)
        bad = fy_gpu_zarr_write_chunk(e->zout, g->iz, g->iy, g->ix, s->hout + (size_t)k * e->tcap, g->tz, g->ty,
                                  g->tx) != 0;
    t = now_s() - t;
    pthread_mutex_lock(&e->mu);
    if (bad) {
        e->io_fail++;
        note_err(e, fyg_error_string());
    } else {
        e->written++;
    }
    e->write_s += t;
    if (--s->writes_left == 0) {
        s->state = S_FREE;
        pthread_cond_broadcast(&e->cv_main);
    }
    pthread_mutex_unlock(&e->mu);
}

static void *worker(void *arg) {
    exporter *e = (exporter *)arg;
    for (;;) {
        pthread_mutex_lock(&e->mu);
        while (!e->qn) pthread_cond_wait(&e->cv_task, &e->mu);
        task_t t = e->q[e->qhead];
        e->qhead = (e->qhead + 1) % e->qcap;
        e->qn--;
        pthread_mutex_unlock(&e->mu);
        if (t.kind == 2) break;
        if (t.kind == 0) {
            do_read(e, &e->sl[t.slot], t.k);
        } else if (t.kind == 3) {
            /* one thread waits for a slot's copies back, then queues its chunk
             * writes, so the other threads keep reading meanwhile */
            slot_t *sl = &e->sl[t.slot];
            (void)fyg_event_wait(sl->e_out);
            pthread_mutex_lock(&e->mu);
            for (int k = 0; k < sl->n; k++)
                if (sl->cube_of[k] >= 0) push(e, (task_t){1, t.slot, k});
            pthread_mutex_unlock(&e->mu);
        } else {
            do_write(e, &e->sl[t.slot], t.k);
        }
    }
    return NULL;
}

/* queue stage 2-4 of a slot whose reads are done; returns the number of
 * chunks to write (0: the slot is free again) or -1 on a device error */
static int submit(exporter *e, slot_t *s, fyg_stream *cin, fyg_stream *comp, fyg_stream *cout, void *stream) {
    int nc = 0;
    for (int k = 0; k < s->n; k++) {
        s->cube_of[k] = -1;
        if (s->air[k] || s->fail[k]) continue;
        const tile_geom *g = &s->g[k];
        if (fy_cube_params_from_cfg(e->cfg, g->y0, g->x0, g->ty, g->tx, e->Y, e->X, &s->prm[nc])) return -1;
        fy_cube *c = &s->cubes[nc];
        c->d_in = s->din + (size_t)k * e->cap;
        c->rz0 = g->rz0; c->ry0 = g->ry0; c->rx0 = g->rx0;
        c->hz = g->hz; c->hy = g->hy; c->hx = g->hx;
        c->z0 = g->z0; c->y0 = g->y0; c->x0 = g->x0;
        c->tz = g->tz; c->ty = g->ty; c->tx = g->tx;
        c->d_out = s->dout + (size_t)k * e->tcap;
        c->d_float = NULL;
        c->params = &s->prm[nc];
        s->cube_of[k] = nc++;
    }
    if (!nc) return 0;
    for (int k = 0; k < s->n; k++) {
        if (s->cube_of[k] < 0) continue;
        const tile_geom *g = &s->g[k];
        if (fyg_h2d(s->din + (size_t)k * e->cap, s->hin + (size_t)k * e->cap, (size_t)g->hz * g->hy * g->hx, cin))
            return -1;
    }
    if (fyg_event_record(cin, &s->e_in) || fyg_stream_wait(comp, s->e_in)) return -1;
    if (fy_gpu_process_cubes(s->cubes, nc, stream)) return -1;
    if (fyg_event_record(comp, &s->e_comp) || fyg_stream_wait(cout, s->e_comp)) return -1;
    for (int k = 0; k < s->n; k++) {
        if (s->cube_of[k] < 0) continue;
        const tile_geom *g = &s->g[k];
        if (fyg_d2h(s->hout + (size_t)k * e->tcap, s->dout + (size_t)k * e->tcap, (size_t)g->tz * g->ty * g->tx,
                    cout))
            return -1;
    }
    if (fyg_event_record(cout, &s->e_out)) return -1;
    return nc;
}

int fy_gpu_export(const fy_zarr *zin, const fy_zarr *zout, const fy_pipeline_cfg *cfg, int tile, long t_begin,
                  long t_end, int threads, int batch, fy_export_stats *stats, void *stream) {
    double t_start = now_s();
    if (stats) memset(stats, 0, sizeof *stats);
    if (!zin || !zout || !cfg || tile < 1) {
        fyg_set_error("fy_gpu_export: bad arguments");
        return 1;
    }
    exporter E;
    memset(&E, 0, sizeof E);
    exporter *e = &E;
    e->zin = zin;
    e->zout = zout;
    e->cfg = cfg;
    e->tile = tile;
    e->halo = cfg->halo > 0 ? cfg->halo : 8;
    e->Z = zin->shape[0];
    e->Y = zin->shape[1];
    e->X = zin->shape[2];
    /* tile == output chunk (pipeline.c:1049-1050) */
    for (int i = 0; i < 3; i++)
        if (zout->shape[i] != zin->shape[i] || zout->chunk[i] != tile) {
            fyg_set_error("fy_gpu_export: the output zarr must have the input's shape and tile^3 chunks");
            return 1;
        }
    const long ntz = (e->Z + tile - 1) / tile;
    e->nty = (e->Y + tile - 1) / tile;
    e->ntx = (e->X + tile - 1) / tile;
    const long ntiles = ntz * e->nty * e->ntx;
    if (t_end <= 0 || t_end > ntiles) t_end = ntiles;
    if (t_begin < 0) t_begin = 0;
    if (t_begin >= t_end) {
        if (stats) stats->wall_s = now_s() - t_start;
        return 0;
    }
    if (batch < 1) batch = 8;
    if (threads < 1) {
        long nc = sysconf(_SC_NPROCESSORS_ONLN);
        threads = nc < 2 ? 2 : (nc > 16 ? 16 : (int)nc);
    }
    e->batch = batch;
    const long ext = tile + 2L * e->halo;
    e->cap = (size_t)lmin(ext, e->Z) * lmin(ext, e->Y) * lmin(ext, e->X);
    e->tcap = (size_t)tile * tile * tile;
    e->qcap = NSLOT * batch * 2 + NSLOT + threads + 4;
    e->q = (task_t *)calloc((size_t)e->qcap, sizeof(task_t));
    e->zc = fy_gpu_zarr_cache_new((size_t)1 << 31);
    int rc = e->q == NULL || e->zc == NULL;
    for (int i = 0; i < NSLOT && !rc; i++) {
        slot_t *s = &e->sl[i];
        rc |= fyg_host_alloc((void **)&s->hin, e->cap * batch);
        rc |= fyg_host_alloc((void **)&s->hout, e->tcap * batch);
        rc |= fyg_malloc((void **)&s->din, e->cap * batch);
        rc |= fyg_malloc((void **)&s->dout, e->tcap * batch);
        s->g = (tile_geom *)calloc((size_t)batch, sizeof(tile_geom));
        s->air = (int *)calloc((size_t)batch, sizeof(int));
        s->fail = (int *)calloc((size_t)batch, sizeof(int));
        s->cube_of = (int *)calloc((size_t)batch, sizeof(int));
        s->cubes = (fy_cube *)calloc((size_t)batch, sizeof(fy_cube));
        s->prm = (fy_cube_params *)calloc((size_t)batch, sizeof(fy_cube_params));
        rc |= !s->g || !s->air || !s->fail || !s->cube_of || !s->cubes || !s->prm;
    }
    fyg_stream *comp = fyg_stream_wrap(stream);
    fyg_stream *cin = rc ? NULL : fyg_stream_new();
    fyg_stream *cout = rc ? NULL : fyg_stream_new();
    rc |= !comp || !cin || !cout;
    pthread_t *th = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    int nth = 0;
    if (!rc && th) {
        pthread_mutex_init(&e->mu, NULL);
        pthread_cond_init(&e->cv_task, NULL);
        pthread_cond_init(&e->cv_main, NULL);
        for (; nth < threads; nth++)
            if (pthread_create(&th[nth], NULL, worker, e)) break;
        rc |= nth == 0;
    } else {
        rc = 1;
    }

    const long nb = (t_end - t_begin + batch - 1) / batch;
    long next = 0;
    for (long b = 0; b < nb && !rc; b++) {
        pthread_mutex_lock(&e->mu);
        /* start reading the batches whose slots are free, up to NSLOT ahead,
         * without waiting; wait for a slot only when batch b itself has not
         * started reading (its slot still holds batch b - NSLOT) */
        while (next < nb && next < b + NSLOT && (e->sl[next % NSLOT].state == S_FREE || next <= b)) {
            slot_t *s = &e->sl[next % NSLOT];
            while (s->state != S_FREE) pthread_cond_wait(&e->cv_main, &e->mu);
            s->t0 = t_begin + next * batch;
            s->n = (int)lmin(batch, t_end - s->t0);
            s->reads_left = s->n;
            s->state = S_READING;
            for (int k = 0; k < s->n; k++) {
                geom(e->Z, e->Y, e->X, e->nty, e->ntx, tile, e->halo, s->t0 + k, &s->g[k]);
                push(e, (task_t){0, (int)(next % NSLOT), k});
            }
            next++;
        }
        slot_t *s = &e->sl[b % NSLOT];
        while (s->reads_left > 0) pthread_cond_wait(&e->cv_main, &e->mu);
        pthread_mutex_unlock(&e->mu);
        int nc = submit(e, s, cin, comp, cout, stream);
        pthread_mutex_lock(&e->mu);
        if (nc < 0) {
            e->dev_err = 1;
            note_err(e, fyg_error_string());
            s->state = S_FREE;
            rc = 1;
        } else if (nc == 0) {
            s->state = S_FREE;
        } else {
            s->state = S_SUBMITTED;
            s->writes_left = nc;
            push(e, (task_t){3, (int)(b % NSLOT), 0});
        }
        pthread_cond_broadcast(&e->cv_main);
        pthread_mutex_unlock(&e->mu);
    }
    if (nth) {
        /* drain: every slot read and written, then stop the threads */
        pthread_mutex_lock(&e->mu);
        for (int i = 0; i < NSLOT; i++) {
            slot_t *s = &e->sl[i];
            while (s->state == S_SUBMITTED || (s->state == S_READING && s->reads_left > 0))
                pthread_cond_wait(&e->cv_main, &e->mu);
        }
        for (int i = 0; i < nth; i++) push(e, (task_t){2, 0, 0});
        pthread_mutex_unlock(&e->mu);
        for (int i = 0; i < nth; i++) pthread_join(th[i], NULL);
        pthread_mutex_destroy(&e->mu);
        pthread_cond_destroy(&e->cv_task);
        pthread_cond_destroy(&e->cv_main);
    }
    if (comp && fyg_sync(comp)) rc = 1;
    fyg_stream_free(cin);
    fyg_stream_free(cout);
    for (int i = 0; i < NSLOT; i++) {
        slot_t *s = &e->sl[i];
        fyg_host_free(s->hin);
        fyg_host_free(s->hout);
        fyg_free(s->din);
        fyg_free(s->dout);
        fyg_event_destroy(s->e_in);
        fyg_event_destroy(s->e_comp);
        fyg_event_destroy(s->e_out);
        free(s->g);
        free(s->air);
        free(s->fail);
        free(s->cube_of);
        free(s->cubes);
        free(s->prm);
    }
    free(th);
    free(e->q);
    fy_gpu_zarr_cache_free(e->zc);
    if (e->io_fail) rc = 1;
    if (rc && e->err[0]) fyg_set_error(e->err);
    if (stats) {
        stats->tiles = t_end - t_begin;
        stats->written = e->written;
        stats->air = e->air;
        stats->io_fail = e->io_fail;
        stats->wall_s = now_s() - t_start;
        stats->read_s = e->read_s;
        stats->write_s = e->write_s;
    }
    return rc;
}
