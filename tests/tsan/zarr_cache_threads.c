/* zarr_cache_threads.c -- ThreadSanitizer driver for the shared chunk mappings
 * of fy_zarr.c (fy_gpu_zarr_cache), the one host structure the exporter's I/O
 * threads share.  Built with -fsanitize=thread together with fy_zarr.c alone
 * (no HIP): writes a 96x80x72 volume of 16x20x12 chunks (one chunk missing),
 * then eight threads read 400 overlapping regions through one cache whose
 * budget holds three chunks, so mappings are evicted and remapped while other
 * threads hold theirs, and every cached read is compared with an uncached one.
 * Exit status 0 and "ok" when every region matched; a data race makes
 * ThreadSanitizer report and exit with its own status. */
#define _GNU_SOURCE
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "fysics.h"
#include "fysics_gpu.h"

/* the library's error sink (the HIP runtime's in the product) */
void fyg_set_error(const char *msg) { (void)msg; }

enum { NZ = 96, NY = 80, NX = 72, CZ = 16, CY = 20, CX = 12, NREG = 400, NTHR = 8 };

static fy_zarr vol;
static fy_gpu_zarr_cache *cache;
static long reg[NREG][6];
static int bad;
static pthread_mutex_t bad_mu = PTHREAD_MUTEX_INITIALIZER;

static unsigned long long rng = 88172645463325252ULL;
static long rnd(long n) {                               /* xorshift64 */
    rng ^= rng << 13; rng ^= rng >> 7; rng ^= rng << 17;
    return (long)(rng % (unsigned long long)n);
}

static void *work(void *arg) {
    const long part = (long)arg;
    for (long k = part; k < NREG; k += NTHR) {
        const long *r = reg[k];
        const size_t n = (size_t)r[3] * r[4] * r[5];
        unsigned char *a = malloc(n), *b = malloc(n);
        int ok = a && b && fy_gpu_zarr_read_cached(&vol, cache, r[0], r[1], r[2], r[3], r[4], r[5], a) == 0 &&
                 fy_gpu_zarr_read(&vol, r[0], r[1], r[2], r[3], r[4], r[5], b) == 0 && memcmp(a, b, n) == 0;
        if (!ok) {
            pthread_mutex_lock(&bad_mu);
            bad++;
            pthread_mutex_unlock(&bad_mu);
        }
        free(a);
        free(b);
    }
    return NULL;
}

int main(int argc, char **argv) {
    char root[4096];
    snprintf(root, sizeof root, "%s/tsan_%d.zarr", argc > 1 ? argv[1] : "/tmp", (int)getpid());
    const long shape[3] = {NZ, NY, NX}, chunk[3] = {CZ, CY, CX};
    fy_zarr w;
    if (fy_gpu_zarr_create(&w, root, shape, chunk)) return 2;
    unsigned char *buf = malloc((size_t)CZ * CY * CX);
    for (long cz = 0; cz < NZ / CZ; cz++)
        for (long cy = 0; cy < NY / CY; cy++)
            for (long cx = 0; cx < NX / CX; cx++) {
                if (cz == 1 && cy == 1 && cx == 1) continue;        /* a missing chunk reads as fill */
                for (long i = 0; i < CZ * CY * CX; i++) buf[i] = (unsigned char)rnd(256);
                if (fy_gpu_zarr_write_chunk(&w, cz, cy, cx, buf, CZ, CY, CX)) return 3;
            }
    free(buf);
    if (fy_gpu_zarr_open(&vol, root)) return 4;
    for (int k = 0; k < NREG; k++) {
        long *r = reg[k];
        r[3] = 1 + rnd(40); r[4] = 1 + rnd(40); r[5] = 1 + rnd(40);
        r[0] = rnd(NZ + 8) - 4; r[1] = rnd(NY + 8) - 4; r[2] = rnd(NX + 8) - 4;     /* some reach outside */
    }
    cache = fy_gpu_zarr_cache_new((size_t)3 * CZ * CY * CX);
    if (!cache) return 5;
    pthread_t t[NTHR];
    for (long i = 0; i < NTHR; i++) pthread_create(&t[i], NULL, work, (void *)i);
    for (int i = 0; i < NTHR; i++) pthread_join(t[i], NULL);
    fy_gpu_zarr_cache_free(cache);
    printf("%s: %d of %d regions differ\n", bad ? "FAILED" : "ok", bad, NREG);
    return bad ? 1 : 0;
}
