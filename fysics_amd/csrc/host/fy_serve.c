/* fy_serve.c -- the served form of the tile chain: a request loop on one
 * stream with HBM-budget admission and continuous batching.
 *
 * Requests wait in a FIFO.  A poll first retires the waves whose completion
 * event has fired, then launches new waves while fewer than max_waves are in
 * flight: each wave takes the longest FIFO prefix (at most max_batch
 * requests) whose workspace fits the budget left by the waves still running,
 * and runs it as one fy_gpu_process_cubes call, so requests of one shape share
 * every launch.  At low load a request starts as soon as it arrives (a wave of
 * one); under load the queue that builds up while a wave runs becomes the next
 * wave.  Each cube owns its workspace slice and flag, so its output does not
 * depend on the requests it was batched with.
 */
#include <stdlib.h>
#include <string.h>

#include "fyh.h"

typedef struct {
    void *ev;                   /* completion event (reused) */
    long *tickets;
    int n;
    size_t bytes;
} wave;

struct fy_server {
    void *stream;
    fyg_stream *s;
    size_t budget, inflight;
    int max_batch, max_waves;
    /* pending FIFO (ring) */
    fy_cube *qc;
    long *qt;
    size_t *qb;
    int qh, qn, qcap;
    /* waves in flight (ring of max_waves) */
    wave *w;
    int wh, wn;
    fy_cube *batch;
};

fy_server *fy_serve_create(size_t hbm_budget, int max_batch, int max_waves, void *stream) {
    if (max_batch < 1 || max_waves < 1 || !hbm_budget) return NULL;
    fy_server *sv = (fy_server *)calloc(1, sizeof *sv);
    if (!sv) return NULL;
    sv->stream = stream;
    sv->s = fyg_stream_wrap(stream);
    sv->budget = hbm_budget;
    sv->max_batch = max_batch;
    sv->max_waves = max_waves;
    sv->qcap = 1024;
    sv->qc = (fy_cube *)malloc(sizeof(fy_cube) * (size_t)sv->qcap);
    sv->qt = (long *)malloc(sizeof(long) * (size_t)sv->qcap);
    sv->qb = (size_t *)malloc(sizeof(size_t) * (size_t)sv->qcap);
    sv->w = (wave *)calloc((size_t)max_waves, sizeof(wave));
    sv->batch = (fy_cube *)malloc(sizeof(fy_cube) * (size_t)max_batch);
    int ok = sv->s && sv->qc && sv->qt && sv->qb && sv->w && sv->batch;
    for (int i = 0; ok && i < max_waves; i++) {
        sv->w[i].tickets = (long *)malloc(sizeof(long) * (size_t)max_batch);
        ok = sv->w[i].tickets != NULL;
    }
    if (!ok) {
        fy_serve_destroy(sv);
        return NULL;
    }
    return sv;
}

void fy_serve_destroy(fy_server *sv) {
    if (!sv) return;
    if (sv->wn) (void)fyg_sync(sv->s);
    if (sv->w) {
        for (int i = 0; i < sv->max_waves; i++) {
            free(sv->w[i].tickets);
            fyg_event_destroy(sv->w[i].ev);
        }
    }
    free(sv->w); free(sv->qc); free(sv->qt); free(sv->qb); free(sv->batch);
    free(sv);
}

int fy_serve_submit(fy_server *sv, const fy_cube *cube, long ticket) {
    if (!sv || !cube) return 1;
    size_t b = fy_cube_workspace_bytes(cube);
    if (b > sv->budget) return 1;
    if (sv->qn == sv->qcap) {                     /* grow the ring, unrolled to start at 0 */
        int nc = sv->qcap * 2;
        fy_cube *qc = (fy_cube *)malloc(sizeof(fy_cube) * (size_t)nc);
        long *qt = (long *)malloc(sizeof(long) * (size_t)nc);
        size_t *qb = (size_t *)malloc(sizeof(size_t) * (size_t)nc);
        if (!qc || !qt || !qb) { free(qc); free(qt); free(qb); return 1; }
        for (int i = 0; i < sv->qn; i++) {
            int j = (sv->qh + i) % sv->qcap;
            qc[i] = sv->qc[j]; qt[i] = sv->qt[j]; qb[i] = sv->qb[j];
        }
        free(sv->qc); free(sv->qt); free(sv->qb);
        sv->qc = qc; sv->qt = qt; sv->qb = qb; sv->qh = 0; sv->qcap = nc;
    }
    int j = (sv->qh + sv->qn) % sv->qcap;
    sv->qc[j] = *cube;
    sv->qt[j] = ticket;
    sv->qb[j] = b;
    sv->qn++;
    return 0;
}

int fy_serve_outstanding(const fy_server *sv) {
    if (!sv) return 0;
    int n = sv->qn;
    for (int i = 0; i < sv->wn; i++) n += sv->w[(sv->wh + i) % sv->max_waves].n;
    return n;
}

int fy_serve_poll(fy_server *sv, long *done, int max_done) {
    if (!sv) return -1;
    int nd = 0;
    /* retire finished waves, oldest first (one stream: they finish in order) */
    while (sv->wn > 0) {
        wave *w = &sv->w[sv->wh];
        if (nd + w->n > max_done) break;
        int st = fyg_event_done(w->ev);
        if (st < 0) return -1;
        if (!st) break;
        memcpy(done + nd, w->tickets, sizeof(long) * (size_t)w->n);
        nd += w->n;
        sv->inflight -= w->bytes;
        w->n = 0;
        sv->wh = (sv->wh + 1) % sv->max_waves;
        sv->wn--;
    }
    /* admit and launch */
    while (sv->wn < sv->max_waves && sv->qn > 0) {
        int m = 0;
        size_t bytes = 0;
        while (m < sv->max_batch && m < sv->qn) {
            size_t b = sv->qb[(sv->qh + m) % sv->qcap];
            if (sv->inflight + bytes + b > sv->budget) break;
            bytes += b;
            m++;
        }
        if (m == 0) break;                        /* waits for running waves to free memory */
        wave *w = &sv->w[(sv->wh + sv->wn) % sv->max_waves];
        for (int i = 0; i < m; i++) {
            int j = (sv->qh + i) % sv->qcap;
            sv->batch[i] = sv->qc[j];
            w->tickets[i] = sv->qt[j];
        }
        if (fy_gpu_process_cubes(sv->batch, m, sv->stream)) return -1;
        if (fyg_event_record(sv->s, &w->ev)) return -1;
        w->n = m;
        w->bytes = bytes;
        sv->inflight += bytes;
        sv->wn++;
        sv->qh = (sv->qh + m) % sv->qcap;
        sv->qn -= m;
    }
    return nd;
}
