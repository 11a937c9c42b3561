/* fy_serve.c -- the served form of the tile chain: a request loop on one GPU
 * with HBM-budget admission and continuous batching over several lanes.
 *
 * Requests wait in a FIFO.  A lane is a stream of the server's own with its
 * own workspaces; one wave runs on a lane at a time, and waves on different
 * lanes run at the same time on the device (the kernels of one wave beside
 * the copies and kernels of another).  A poll first retires the waves whose
 * completion event has fired, in whatever order they finish, then, while a
 * lane is idle, launches the longest FIFO prefix (at most max_batch requests)
 * that fits the budget as one wave: one fy_gpu_process_cubes call, so requests
 * of one shape share every launch.  At low load a request starts as soon as it
 * arrives, alone on an idle lane; under load the queue that builds up while
 * the lanes are busy becomes the next waves.  Each cube owns its workspace
 * slice and all-air flag, so its output does not depend on the requests it
 * was batched with or on the lane it ran on.
 *
 * Host-resident requests (fy_serve_submit_host) bring host buffers: the wave
 * copies each input into the lane's device staging area on the lane's stream,
 * runs the chain and copies the inner tile back, so a request's PCIe transfers
 * overlap the kernels of the waves on other lanes.  With pinned (hipHostMalloc)
 * buffers the copies are asynchronous DMA.
 *
 * Admission counts what the lanes hold: workspaces are grow-only per lane, so
 * a wave of B bytes on a lane that holds C bytes makes the footprint grow by
 * max(B - C, 0); a wave is launched only if the footprint stays within the
 * budget.
 */
#include <stdlib.h>
#include <string.h>

#include "fyh.h"

#define FY_SERVE_MAX_LANES 8

typedef struct {
    fy_cube c;
    const unsigned char *h_in;  /* host-resident request: input and output on the host */
    unsigned char *h_out;
    long ticket;
    size_t ws;                  /* device bytes of the cube's workspace */
    size_t stage;               /* device staging bytes (host requests: region + inner tile) */
} request;

typedef struct {
    fyg_stream *s;              /* the lane's stream object (owned) */
    void *ev;                   /* completion event of the lane's wave (reused) */
    void *start;                /* the caller's stream at launch: the wave waits for it */
    long *tickets;
    int n;                      /* requests of the wave in flight (0: idle) */
    size_t cap;                 /* grow-only footprint: largest wave bytes run on this lane */
    fy_cube *batch;
} lane;

struct fy_server {
    fyg_stream *caller;         /* the stream the caller orders device inputs on */
    size_t budget;
    int max_batch, nlanes;
    request *q;                 /* pending FIFO (ring) */
    int qh, qn, qcap;
    lane lanes[FY_SERVE_MAX_LANES];
};

static size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

fy_server *fy_serve_create(size_t hbm_budget, int max_batch, int lanes, void *stream) {
    if (max_batch < 1 || lanes < 1 || !hbm_budget) return NULL;
    if (lanes > FY_SERVE_MAX_LANES) lanes = FY_SERVE_MAX_LANES;
    fy_server *sv = (fy_server *)calloc(1, sizeof *sv);
    if (!sv) return NULL;
    sv->caller = fyg_stream_wrap(stream);
    sv->budget = hbm_budget;
    sv->max_batch = max_batch;
    sv->nlanes = lanes;
    sv->qcap = 1024;
    sv->q = (request *)malloc(sizeof(request) * (size_t)sv->qcap);
    int ok = sv->caller && sv->q;
    for (int i = 0; ok && i < lanes; i++) {
        lane *l = &sv->lanes[i];
        l->s = fyg_stream_new();
        l->tickets = (long *)malloc(sizeof(long) * (size_t)max_batch);
        l->batch = (fy_cube *)malloc(sizeof(fy_cube) * (size_t)max_batch);
        ok = l->s && l->tickets && l->batch;
    }
    if (!ok) {
        fy_serve_destroy(sv);
        return NULL;
    }
    return sv;
}

void fy_serve_destroy(fy_server *sv) {
    if (!sv) return;
    for (int i = 0; i < FY_SERVE_MAX_LANES; i++) {
        lane *l = &sv->lanes[i];
        if (l->s) fyg_stream_free(l->s);          /* waits for the lane's work */
        fyg_event_destroy(l->ev);
        fyg_event_destroy(l->start);
        free(l->tickets);
        free(l->batch);
    }
    free(sv->q);
    free(sv);
}

static int enqueue(fy_server *sv, const request *r) {
    if (sv->qn == sv->qcap) {                     /* grow the ring, unrolled to start at 0 */
        int nc = sv->qcap * 2;
        request *q = (request *)malloc(sizeof(request) * (size_t)nc);
        if (!q) return 1;
        for (int i = 0; i < sv->qn; i++) q[i] = sv->q[(sv->qh + i) % sv->qcap];
        free(sv->q);
        sv->q = q;
        sv->qh = 0;
        sv->qcap = nc;
    }
    sv->q[(sv->qh + sv->qn) % sv->qcap] = *r;
    sv->qn++;
    return 0;
}

int fy_serve_submit(fy_server *sv, const fy_cube *cube, long ticket) {
    if (!sv || !cube) return 1;
    request r;
    memset(&r, 0, sizeof r);
    r.c = *cube;
    r.ticket = ticket;
    r.ws = fy_cube_workspace_bytes(cube);
    if (r.ws > sv->budget) return 1;
    return enqueue(sv, &r);
}

int fy_serve_submit_host(fy_server *sv, const fy_cube *cube, const unsigned char *h_in, unsigned char *h_out,
                         long ticket) {
    if (!sv || !cube || !h_in || !h_out || cube->hz <= 0 || cube->hy <= 0 || cube->hx <= 0) return 1;
    request r;
    memset(&r, 0, sizeof r);
    r.c = *cube;
    r.c.d_in = NULL;
    r.c.d_out = NULL;
    r.c.d_float = NULL;
    r.h_in = h_in;
    r.h_out = h_out;
    r.ticket = ticket;
    r.ws = fy_cube_workspace_bytes(cube);
    r.stage = al256((size_t)cube->hz * cube->hy * cube->hx) + al256((size_t)cube->tz * cube->ty * cube->tx);
    if (r.ws + r.stage > sv->budget) return 1;
    return enqueue(sv, &r);
}

int fy_serve_outstanding(const fy_server *sv) {
    if (!sv) return 0;
    int n = sv->qn;
    for (int i = 0; i < sv->nlanes; i++) n += sv->lanes[i].n;
    return n;
}

static size_t footprint(const fy_server *sv) {
    size_t f = 0;
    for (int i = 0; i < sv->nlanes; i++) f += sv->lanes[i].cap;
    return f;
}

/* queue one wave of m requests (the FIFO head) on lane l */
static int launch(fy_server *sv, lane *l, int m, size_t bytes) {
    size_t stage = 0;
    for (int i = 0; i < m; i++) stage += sv->q[(sv->qh + i) % sv->qcap].stage;
    unsigned char *dev = NULL;
    if (stage) {
        /* the lane's staging area for host requests (grow-only, slot WS_U8) */
        dev = (unsigned char *)fyg_ws(l->s, WS_U8, stage);
        if (!dev) return 1;
    }
    /* device inputs the caller queued before this poll are ready first */
    if (fyg_event_record(sv->caller, &l->start) || fyg_stream_wait(l->s, l->start)) return 1;
    size_t off = 0;
    for (int i = 0; i < m; i++) {
        request *r = &sv->q[(sv->qh + i) % sv->qcap];
        l->batch[i] = r->c;
        l->tickets[i] = r->ticket;
        if (r->h_in) {
            size_t hn = (size_t)r->c.hz * r->c.hy * r->c.hx;
            l->batch[i].d_in = dev + off;
            l->batch[i].d_out = dev + off + al256(hn);
            off += r->stage;
            if (fyg_h2d((void *)l->batch[i].d_in, r->h_in, hn, l->s)) return 1;
        }
    }
    if (fyh_process_cubes_serial(l->batch, m, l->s)) return 1;   /* the lanes are the concurrency */
    for (int i = 0; i < m; i++) {
        request *r = &sv->q[(sv->qh + i) % sv->qcap];
        if (r->h_out &&
            fyg_d2h(r->h_out, l->batch[i].d_out, (size_t)r->c.tz * r->c.ty * r->c.tx, l->s))
            return 1;
    }
    if (fyg_event_record(l->s, &l->ev)) return 1;
    l->n = m;
    if (bytes > l->cap) l->cap = bytes;
    sv->qh = (sv->qh + m) % sv->qcap;
    sv->qn -= m;
    return 0;
}

int fy_serve_poll(fy_server *sv, long *done, int max_done) {
    if (!sv) return -1;
    int nd = 0;
    /* retire finished waves in the order they finish */
    for (int i = 0; i < sv->nlanes; i++) {
        lane *l = &sv->lanes[i];
        if (!l->n || nd + l->n > max_done) continue;
        int st = fyg_event_done(l->ev);
        if (st < 0) return -1;
        if (!st) continue;
        memcpy(done + nd, l->tickets, sizeof(long) * (size_t)l->n);
        nd += l->n;
        l->n = 0;
    }
    /* admit and launch on idle lanes */
    for (int i = 0; i < sv->nlanes && sv->qn > 0; i++) {
        lane *l = &sv->lanes[i];
        if (l->n) continue;
        int m = 0;
        size_t bytes = 0;
        for (int pass = 0; pass < 2 && m == 0; pass++) {
            if (pass == 1) {
                /* the head request does not fit next to what the lanes hold:
                 * give back the workspaces of the idle lanes (their work has
                 * completed); busy lanes give theirs back once idle */
                for (int j = 0; j < sv->nlanes; j++) {
                    lane *o = &sv->lanes[j];
                    if (o->n || !o->cap) continue;
                    if (fyg_stream_trim(o->s)) return -1;
                    o->cap = 0;
                }
            }
            size_t others = footprint(sv) - l->cap;
            while (m < sv->max_batch && m < sv->qn) {
                const request *r = &sv->q[(sv->qh + m) % sv->qcap];
                size_t b = bytes + r->ws + r->stage;
                if (others + (b > l->cap ? b : l->cap) > sv->budget) break;
                bytes = b;
                m++;
            }
        }
        if (m == 0) break;                        /* waits for the busy lanes */
        if (launch(sv, l, m, bytes)) return -1;
    }
    return nd;
}
