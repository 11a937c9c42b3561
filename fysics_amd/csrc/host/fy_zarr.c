/* fy_zarr.c -- local zarr v2 raw-u8 volumes, the on-disk format the pass-2
 * tile loop reads and writes (zarr_io.c:1-9).
 *
 * Layout: <root>/0/.zarray (JSON) and one file per chunk at <root>/0/cz/cy/cx
 * holding the chunk's voxels, C order, no compression.  A chunk on the
 * volume's far faces holds only its true extent (min(chunk, shape - c*chunk)
 * per axis) and is addressed with those strides (zarr_io.c:280-304); a missing
 * chunk file, or the part of one past its end, reads as fill_value.
 *
 * Reads map each chunk file the region touches (a halo read faults in only the
 * pages of the sliver it copies) and copy rows straight into the caller's
 * buffer, which the pass-2 feeder allocates pinned so the bytes go to the GPU
 * without another copy.  S3 roots (zarr_io.c:20-36, built with FYSICS_S3) are
 * not supported: fy_gpu_zarr_open fails on them.  These functions keep the
 * reference's argument meaning under their own names (fy_gpu_zarr_*), so a
 * program that links both libraries keeps the reference's fy_zarr_* with S3.
 */
#define _GNU_SOURCE
#include <errno.h>
#include <fcntl.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include "fyh.h"

static int zerr(const char *what, const char *path) {
    char msg[256];
    snprintf(msg, sizeof msg, "zarr: %s %.200s", what, path);
    fyg_set_error(msg);
    return 1;
}

/* the JSON value of "key" (first occurrence), past the colon and blanks */
static const char *json_value(const char *js, const char *key) {
    char pat[64];
    snprintf(pat, sizeof pat, "\"%s\"", key);
    const char *p = strstr(js, pat);
    if (!p) return NULL;
    p += strlen(pat);
    while (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r') p++;
    if (*p != ':') return NULL;
    p++;
    while (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r') p++;
    return p;
}

/* "[a, b, c]" -> out[0..2] */
static int json_triple(const char *v, long out[3]) {
    if (!v || *v != '[') return 1;
    v++;
    for (int i = 0; i < 3; i++) {
        char *end;
        while (*v == ' ' || *v == '\t' || *v == '\n' || *v == '\r' || (i && *v == ',')) v++;
        long x = strtol(v, &end, 10);
        if (end == v) return 1;
        out[i] = x;
        v = end;
    }
    return 0;
}

static int json_is(const char *v, const char *lit) { return v && strncmp(v, lit, strlen(lit)) == 0; }

int fy_gpu_zarr_open(fy_zarr *z, const char *root) {
    memset(z, 0, sizeof *z);
    if (!root || strncmp(root, "s3://", 5) == 0) return zerr("S3 roots are not supported:", root ? root : "");
    if (strlen(root) >= sizeof z->root) return zerr("root path too long:", root);
    snprintf(z->root, sizeof z->root, "%s", root);
    char path[1100];
    snprintf(path, sizeof path, "%s/0/.zarray", root);
    FILE *f = fopen(path, "rb");
    if (!f) return zerr("no", path);
    char js[8192];
    size_t n = fread(js, 1, sizeof js - 1, f);
    fclose(f);
    js[n] = 0;
    if (json_triple(json_value(js, "shape"), z->shape) || json_triple(json_value(js, "chunks"), z->chunk))
        return zerr("bad shape/chunks in", path);
    for (int i = 0; i < 3; i++)
        if (z->shape[i] <= 0 || z->chunk[i] <= 0) return zerr("bad shape/chunks in", path);
    /* raw u8 only: a compressed or wider chunk read as raw bytes would be garbage */
    const char *c = json_value(js, "compressor");
    if (c && !json_is(c, "null")) return zerr("compressed chunks are not supported:", path);
    const char *d = json_value(js, "dtype");
    if (d && !(json_is(d, "\"|u1\"") || json_is(d, "\"<u1\"") || json_is(d, "\"u1\"") || json_is(d, "\"|b1\"")))
        return zerr("dtype is not u1 in", path);
    const char *fv = json_value(js, "fill_value");
    long fill = 0;
    if (fv && (*fv == '-' || (*fv >= '0' && *fv <= '9'))) fill = strtol(fv, NULL, 10);
    z->fill = (unsigned char)(fill > 0 ? (fill > 255 ? 255 : fill) : 0);
    return 0;
}

/* create every missing directory on the way to (not including) the last
 * component of path */
static int make_parents(const char *path) {
    char tmp[1100];
    snprintf(tmp, sizeof tmp, "%s", path);
    for (char *p = tmp + 1; *p; p++) {
        if (*p != '/') continue;
        *p = 0;
        if (mkdir(tmp, 0775) != 0 && errno != EEXIST) return 1;
        *p = '/';
    }
    return 0;
}

int fy_gpu_zarr_create(fy_zarr *z, const char *root, const long shape[3], const long chunk[3]) {
    memset(z, 0, sizeof *z);
    if (!root || strlen(root) >= sizeof z->root) return zerr("bad root", root ? root : "");
    snprintf(z->root, sizeof z->root, "%s", root);
    for (int i = 0; i < 3; i++) {
        if (shape[i] <= 0 || chunk[i] <= 0) return zerr("bad shape/chunks for", root);
        z->shape[i] = shape[i];
        z->chunk[i] = chunk[i];
    }
    char path[1100];
    snprintf(path, sizeof path, "%s/0/.zarray", root);
    if (make_parents(path)) return zerr("cannot create directories for", path);
    FILE *f = fopen(path, "wb");
    if (!f) return zerr("cannot create", path);
    int bad = fprintf(f,
                      "{\"chunks\":[%ld,%ld,%ld],\"compressor\":null,\"dtype\":\"|u1\",\"fill_value\":0,"
                      "\"order\":\"C\",\"shape\":[%ld,%ld,%ld],\"zarr_format\":2,\"dimension_separator\":\"/\"}",
                      chunk[0], chunk[1], chunk[2], shape[0], shape[1], shape[2]) < 0;
    bad |= fclose(f) != 0;
    return bad ? zerr("cannot write", path) : 0;
}

/* the bytes [off, off + len) of a chunk file's first `cap` bytes: mapped, read
 * into `scratch` when mapping fails, or absent (p == NULL: all fill) */
typedef struct {
    const unsigned char *p;     /* byte `off` of the chunk */
    size_t off, len;
    void *map;
    size_t maplen;
} chunk_bytes;

/* Only the pages holding bytes [lo, hi) of the chunk are mapped. */
static int chunk_open(const fy_zarr *z, long cz, long cy, long cx, size_t cap, size_t lo, size_t hi,
                      unsigned char **scratch, chunk_bytes *cb) {
    memset(cb, 0, sizeof *cb);
    char path[1100];
    snprintf(path, sizeof path, "%s/0/%ld/%ld/%ld", z->root, cz, cy, cx);
    int fd = open(path, O_RDONLY | O_CLOEXEC);
    if (fd < 0) {
        if (errno == ENOENT || errno == ENOTDIR) return 0;    /* sparse: fill */
        return zerr("cannot open", path);
    }
    struct stat st;
    if (fstat(fd, &st) != 0) {
        close(fd);
        return zerr("cannot stat", path);
    }
    size_t len = st.st_size > 0 ? (size_t)st.st_size : 0;
    if (len > cap) len = cap;
    if (hi > len) hi = len;
    if (lo >= hi) {                                          /* nothing stored there: fill */
        close(fd);
        return 0;
    }
    const size_t pg = (size_t)sysconf(_SC_PAGESIZE);
    const size_t off = lo / pg * pg;
    void *m = mmap(NULL, hi - off, PROT_READ, MAP_PRIVATE | (lo == 0 && hi == cap ? MAP_POPULATE : 0), fd, (off_t)off);
    if (m != MAP_FAILED) {
        close(fd);
        cb->map = m;
        cb->maplen = hi - off;
        cb->p = (const unsigned char *)m;
        cb->off = off;
        cb->len = hi - off;
        return 0;
    }
    /* no mapping (e.g. a filesystem without mmap): read the range into a
     * buffer of a whole chunk, which holds any chunk's extent */
    if (!*scratch && !(*scratch = (unsigned char *)malloc((size_t)z->chunk[0] * z->chunk[1] * z->chunk[2]))) {
        close(fd);
        return zerr("out of memory reading", path);
    }
    size_t got = 0;
    while (off + got < hi) {
        ssize_t r = pread(fd, *scratch + got, hi - off - got, (off_t)(off + got));
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) break;
        got += (size_t)r;
    }
    close(fd);
    cb->p = *scratch;
    cb->off = off;
    cb->len = got;
    return 0;
}

static void chunk_close(chunk_bytes *cb) {
    if (cb->map) munmap(cb->map, cb->maplen);
}

static long lmin(long a, long b) { return a < b ? a : b; }
static long lmax(long a, long b) { return a > b ? a : b; }

/* ---- shared chunk mappings for a tile loop.  Neighbouring tiles read the
 * same chunks (a tile-sized region with a halo touches up to 27), so the
 * exporter maps each chunk file once, whole, and lets every tile that needs it
 * copy from that mapping; unmapping (a TLB shootdown across the process's
 * threads) also happens once per chunk.  Entries no read holds are unmapped,
 * least recently used first, when the mapped bytes exceed the budget. */
typedef struct {
    long cz, cy, cx;
    int state;                  /* 0 empty, 1 loading, 2 ready, 3 failed */
    int refs;
    long stamp;
    chunk_bytes cb;             /* mapped, or (map == NULL) a malloc'd copy */
} zc_entry;

#define ZC_ENTRIES 256

struct fy_gpu_zarr_cache {
    pthread_mutex_t mu;
    pthread_cond_t cv;
    zc_entry e[ZC_ENTRIES];
    long clock;
    size_t bytes, budget;
};

static void zc_release(chunk_bytes *cb) {
    if (cb->map) munmap(cb->map, cb->maplen);
    else free((void *)cb->p);
    memset(cb, 0, sizeof *cb);
}

fy_gpu_zarr_cache *fy_gpu_zarr_cache_new(size_t budget) {
    fy_gpu_zarr_cache *c = (fy_gpu_zarr_cache *)calloc(1, sizeof *c);
    if (!c) return NULL;
    c->budget = budget;
    pthread_mutex_init(&c->mu, NULL);
    pthread_cond_init(&c->cv, NULL);
    return c;
}

void fy_gpu_zarr_cache_free(fy_gpu_zarr_cache *c) {
    if (!c) return;
    for (int i = 0; i < ZC_ENTRIES; i++) zc_release(&c->e[i].cb);
    pthread_mutex_destroy(&c->mu);
    pthread_cond_destroy(&c->cv);
    free(c);
}

/* the ready entry of chunk (cz,cy,cx), mapped whole and held until zc_put;
 * NULL on error.  A reader holds one entry at a time, so with fewer reading
 * threads than entries a free or unheld entry always turns up. */
static zc_entry *zc_get(fy_gpu_zarr_cache *c, const fy_zarr *z, long cz, long cy, long cx, size_t cap) {
    pthread_mutex_lock(&c->mu);
    zc_entry *slot;
    for (;;) {
        zc_entry *hit = NULL, *empty = NULL, *lru = NULL;
        for (int i = 0; i < ZC_ENTRIES; i++) {
            zc_entry *e = &c->e[i];
            if (!e->state) {
                if (!empty) empty = e;
            } else if (e->cz == cz && e->cy == cy && e->cx == cx) {
                hit = e;
            } else if (e->state >= 2 && e->refs == 0 && (!lru || e->stamp < lru->stamp)) {
                lru = e;
            }
        }
        if (hit && hit->state == 1) {               /* another thread is mapping it */
            pthread_cond_wait(&c->cv, &c->mu);
            continue;
        }
        if (hit) {
            if (hit->state == 3) {                  /* failed before: fails again */
                pthread_mutex_unlock(&c->mu);
                return NULL;
            }
            hit->refs++;
            hit->stamp = ++c->clock;
            pthread_mutex_unlock(&c->mu);
            return hit;
        }
        slot = (empty && (c->bytes <= c->budget || !lru)) ? empty : lru;
        if (slot) break;
        pthread_cond_wait(&c->cv, &c->mu);          /* every entry held or loading */
    }
    chunk_bytes old = slot->cb;
    c->bytes -= slot->state == 2 ? (old.map ? old.maplen : old.len) : 0;
    memset(slot, 0, sizeof *slot);
    slot->cz = cz;
    slot->cy = cy;
    slot->cx = cx;
    slot->state = 1;
    slot->refs = 1;
    slot->stamp = ++c->clock;
    pthread_mutex_unlock(&c->mu);
    zc_release(&old);
    chunk_bytes cb;
    unsigned char *scratch = NULL;
    int bad = chunk_open(z, cz, cy, cx, cap, 0, cap, &scratch, &cb);
    if (!bad && cb.p && !cb.map) {                  /* read, not mapped: keep a copy */
        unsigned char *own = (unsigned char *)malloc(cb.len ? cb.len : 1);
        if (own) memcpy(own, cb.p, cb.len);
        else bad = zerr("out of memory in the chunk cache", z->root);
        cb.p = own;
    }
    free(scratch);
    pthread_mutex_lock(&c->mu);
    slot->cb = bad ? (chunk_bytes){0} : cb;
    if (bad && cb.p && !cb.map) free((void *)cb.p);
    slot->state = bad ? 3 : 2;
    if (bad) slot->refs = 0;
    else c->bytes += cb.map ? cb.maplen : cb.len;
    pthread_cond_broadcast(&c->cv);
    pthread_mutex_unlock(&c->mu);
    return bad ? NULL : slot;
}

static void zc_put(fy_gpu_zarr_cache *c, zc_entry *e) {
    pthread_mutex_lock(&c->mu);
    e->refs--;
    pthread_cond_broadcast(&c->cv);
    pthread_mutex_unlock(&c->mu);
}

/* nr rows of w bytes; a halo read is mostly slivers a few bytes wide, so
 * narrow rows move as two overlapping words instead of a memcpy call each */
static void copy_rows(unsigned char *dst, size_t dpitch, const unsigned char *src, size_t spitch, size_t w,
                      long nr) {
    if (w == spitch && w == dpitch) {
        memcpy(dst, src, w * (size_t)nr);
    } else if (w >= 8 && w <= 16) {
        for (long j = 0; j < nr; j++, dst += dpitch, src += spitch) {
            unsigned long long a, b;
            memcpy(&a, src, 8);
            memcpy(&b, src + w - 8, 8);
            memcpy(dst, &a, 8);
            memcpy(dst + w - 8, &b, 8);
        }
    } else if (w >= 4 && w < 8) {
        for (long j = 0; j < nr; j++, dst += dpitch, src += spitch) {
            unsigned a, b;
            memcpy(&a, src, 4);
            memcpy(&b, src + w - 4, 4);
            memcpy(dst, &a, 4);
            memcpy(dst + w - 4, &b, 4);
        }
    } else if (w < 4) {
        for (long j = 0; j < nr; j++, dst += dpitch, src += spitch)
            for (size_t i = 0; i < w; i++) dst[i] = src[i];
    } else {
        for (long j = 0; j < nr; j++, dst += dpitch, src += spitch) memcpy(dst, src, w);
    }
}

int fy_gpu_zarr_read(const fy_zarr *z, long z0, long y0, long x0, long dz, long dy, long dx, unsigned char *out) {
    return fy_gpu_zarr_read_cached(z, NULL, z0, y0, x0, dz, dy, dx, out);
}

int fy_gpu_zarr_read_cached(const fy_zarr *z, fy_gpu_zarr_cache *zc, long z0, long y0, long x0, long dz, long dy, long dx,
                         unsigned char *out) {
    if (dz <= 0 || dy <= 0 || dx <= 0) return 0;
    const long C[3] = {z->chunk[0], z->chunk[1], z->chunk[2]};
    const long hi[3] = {z0 + dz, y0 + dy, x0 + dx};
    /* voxels outside the volume read as fill (the reference leaves them unwritten) */
    const long vlo[3] = {lmax(z0, 0), lmax(y0, 0), lmax(x0, 0)};
    const long vhi[3] = {lmin(hi[0], z->shape[0]), lmin(hi[1], z->shape[1]), lmin(hi[2], z->shape[2])};
    if (vlo[0] >= vhi[0] || vlo[1] >= vhi[1] || vlo[2] >= vhi[2] || z0 < 0 || y0 < 0 || x0 < 0 ||
        hi[0] > z->shape[0] || hi[1] > z->shape[1] || hi[2] > z->shape[2])
        memset(out, z->fill, (size_t)dz * dy * dx);
    if (vlo[0] >= vhi[0] || vlo[1] >= vhi[1] || vlo[2] >= vhi[2]) return 0;
    unsigned char *scratch = NULL;
    int rc = 0;
    for (long cz = vlo[0] / C[0]; cz * C[0] < vhi[0] && !rc; cz++)
        for (long cy = vlo[1] / C[1]; cy * C[1] < vhi[1] && !rc; cy++)
            for (long cx = vlo[2] / C[2]; cx * C[2] < vhi[2] && !rc; cx++) {
                /* the chunk's stored extent and the part of the region inside it */
                const long e[3] = {lmin(C[0], z->shape[0] - cz * C[0]), lmin(C[1], z->shape[1] - cy * C[1]),
                                   lmin(C[2], z->shape[2] - cx * C[2])};
                const long o[3] = {cz * C[0], cy * C[1], cx * C[2]};
                long a[3], b[3];
                for (int i = 0; i < 3; i++) {
                    a[i] = lmax(vlo[i], o[i]) - o[i];
                    b[i] = lmin(vhi[i], o[i] + e[i]) - o[i];
                }
                chunk_bytes cb;
                zc_entry *held = NULL;
                const size_t first = ((size_t)a[0] * e[1] + a[1]) * e[2] + a[2];
                const size_t last = ((size_t)(b[0] - 1) * e[1] + (b[1] - 1)) * e[2] + b[2];
                if (zc) {
                    if (!(held = zc_get(zc, z, cz, cy, cx, (size_t)e[0] * e[1] * e[2]))) {
                        rc = 1;
                        break;
                    }
                    cb = held->cb;
                } else if (chunk_open(z, cz, cy, cx, (size_t)e[0] * e[1] * e[2], first, last, &scratch, &cb)) {
                    rc = 1;
                    break;
                }
                const long w = b[2] - a[2], nr = b[1] - a[1];
                for (long k = a[0]; k < b[0]; k++) {
                    /* fast path: the plane's rows are all stored */
                    const size_t rel0 = ((size_t)k * e[1] + a[1]) * e[2] + a[2] - cb.off;
                    if (cb.p && rel0 + (size_t)(nr - 1) * e[2] + w <= cb.len) {
                        unsigned char *dst = out + ((size_t)(o[0] + k - z0) * dy + (size_t)(o[1] + a[1] - y0)) * dx +
                                             (size_t)(o[2] + a[2] - x0);
                        copy_rows(dst, (size_t)dx, cb.p + rel0, (size_t)e[2], (size_t)w, nr);
                        continue;
                    }
                    for (long j = a[1]; j < b[1]; j++) {
                        unsigned char *dst = out + ((size_t)(o[0] + k - z0) * dy + (size_t)(o[1] + j - y0)) * dx +
                                             (size_t)(o[2] + a[2] - x0);
                        const size_t src = ((size_t)k * e[1] + j) * e[2] + a[2];
                        const size_t rel = src - cb.off;           /* src >= first >= cb.off */
                        long have = cb.p && rel < cb.len ? (long)(cb.len - rel) : 0;
                        if (have > w) have = w;
                        if (have > 0) memcpy(dst, cb.p + rel, (size_t)have);
                        if (have < w) memset(dst + have, z->fill, (size_t)(w - have));
                    }
                }
                if (held) zc_put(zc, held);
                else chunk_close(&cb);
            }
    free(scratch);
    return rc;
}

int fy_gpu_zarr_write_chunk(const fy_zarr *z, long cz, long cy, long cx, const unsigned char *buf, long bz, long by,
                        long bx) {
    char path[1100];
    snprintf(path, sizeof path, "%s/0/%ld/%ld/%ld", z->root, cz, cy, cx);
    if (make_parents(path)) return zerr("cannot create directories for", path);
    /* write to a temporary name and rename, so a reader never sees half a chunk */
    char tmp[1200];
    snprintf(tmp, sizeof tmp, "%s.tmp%ld", path, (long)getpid());
    int fd = open(tmp, O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC, 0664);
    if (fd < 0) return zerr("cannot write", path);
    size_t n = (size_t)bz * by * bx, done = 0;
    while (done < n) {
        ssize_t w = write(fd, buf + done, n - done);
        if (w < 0 && errno == EINTR) continue;
        if (w <= 0) break;
        done += (size_t)w;
    }
    int bad = done != n;
    bad |= close(fd) != 0;
    if (!bad) bad = rename(tmp, path) != 0;
    if (bad) {
        unlink(tmp);
        return zerr("cannot write", path);
    }
    return 0;
}
