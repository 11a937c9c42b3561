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
 * not supported: fy_zarr_open fails on them.
 */
#define _GNU_SOURCE
#include <errno.h>
#include <fcntl.h>
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

int fy_zarr_open(fy_zarr *z, const char *root) {
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

int fy_zarr_create(fy_zarr *z, const char *root, const long shape[3], const long chunk[3]) {
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

/* a chunk file's bytes: mapped, read into `scratch` when mapping fails, or
 * absent (p == NULL: all fill) */
typedef struct {
    const unsigned char *p;
    size_t len;
    void *map;
    size_t maplen;
} chunk_bytes;

static int chunk_open(const fy_zarr *z, long cz, long cy, long cx, size_t cap, unsigned char **scratch,
                      chunk_bytes *cb) {
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
    if (st.st_size <= 0) {
        close(fd);
        return 0;
    }
    size_t len = (size_t)st.st_size;
    void *m = mmap(NULL, len, PROT_READ, MAP_PRIVATE, fd, 0);
    if (m != MAP_FAILED) {
        close(fd);
        cb->map = m;
        cb->maplen = len;
        cb->p = (const unsigned char *)m;
        cb->len = len < cap ? len : cap;
        return 0;
    }
    /* no mapping (e.g. a filesystem without mmap): read the chunk into a
     * buffer of a whole chunk, which holds any chunk's extent */
    if (!*scratch && !(*scratch = (unsigned char *)malloc((size_t)z->chunk[0] * z->chunk[1] * z->chunk[2]))) {
        close(fd);
        return zerr("out of memory reading", path);
    }
    size_t want = len < cap ? len : cap, got = 0;
    while (got < want) {
        ssize_t r = pread(fd, *scratch + got, want - got, (off_t)got);
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) break;
        got += (size_t)r;
    }
    close(fd);
    cb->p = *scratch;
    cb->len = got;
    return 0;
}

static void chunk_close(chunk_bytes *cb) {
    if (cb->map) munmap(cb->map, cb->maplen);
}

static long lmin(long a, long b) { return a < b ? a : b; }
static long lmax(long a, long b) { return a > b ? a : b; }

int fy_zarr_read(const fy_zarr *z, long z0, long y0, long x0, long dz, long dy, long dx, unsigned char *out) {
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
                if (chunk_open(z, cz, cy, cx, (size_t)e[0] * e[1] * e[2], &scratch, &cb)) {
                    rc = 1;
                    break;
                }
                const long w = b[2] - a[2];
                for (long k = a[0]; k < b[0]; k++)
                    for (long j = a[1]; j < b[1]; j++) {
                        unsigned char *dst = out + ((size_t)(o[0] + k - z0) * dy + (size_t)(o[1] + j - y0)) * dx +
                                             (size_t)(o[2] + a[2] - x0);
                        const size_t src = ((size_t)k * e[1] + j) * e[2] + a[2];
                        long have = cb.p && src < cb.len ? (long)(cb.len - src) : 0;
                        if (have > w) have = w;
                        if (have > 0) memcpy(dst, cb.p + src, (size_t)have);
                        if (have < w) memset(dst + have, z->fill, (size_t)(w - have));
                    }
                chunk_close(&cb);
            }
    free(scratch);
    return rc;
}

int fy_zarr_write_chunk(const fy_zarr *z, long cz, long cy, long cx, const unsigned char *buf, long bz, long by,
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
