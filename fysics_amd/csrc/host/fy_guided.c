/* fy_guided.c -- guided filter and box smooth entry points (guided.c:104-241)
 * plus the noise -> eps map (noise.c:197-202).  ws arguments are accepted for
 * ABI compatibility; device scratch belongs to the calling thread's stream.
 */
#include <string.h>

#include "fyh.h"

static int guided_dev(const float *d_in, float *d_out, int nz, int ny, int nx, int radius, double eps,
                      int sub, fyg_stream *s) {
    fyg_guided_args a;
    memset(&a, 0, sizeof a);
    a.nz = nz; a.ny = ny; a.nx = nx;
    a.r = radius < 1 ? 2 : radius;
    a.s = sub;
    a.eps = eps;
    a.in = d_in;
    a.em.quant = 0;
    a.em.out = d_out;
    a.ws = fyg_ws(s, WS_GUIDED, fyg_guided_ws_bytes(nz, ny, nx, sub));
    if (!a.ws) return 1;
    return fyg_guided_run(&a, s);
}

static int guided_host(const float *in, float *out, int nz, int ny, int nx, int radius, double eps, int sub) {
    if (nz <= 0 || ny <= 0 || nx <= 0) return 0;
    fyg_stream *s = fyg_thread_stream();
    if (!s) return 1;
    size_t n = (size_t)nz * ny * nx;
    float *d = (float *)fyg_ws(s, WS_IN, sizeof(float) * n);
    if (!d) return 1;
    if (fyg_h2d(d, in, sizeof(float) * n, s)) return 1;
    if (guided_dev(d, d, nz, ny, nx, radius, eps, sub, s)) return 1;
    if (fyg_d2h(out, d, sizeof(float) * n, s)) return 1;
    return fyg_sync(s);
}

int fy_guided_denoise_ws(const float *in, float *out, int nz, int ny, int nx, int radius, double eps, float *ws) {
    (void)ws;
    return guided_host(in, out, nz, ny, nx, radius, eps, 1);
}

int fy_guided_denoise(const float *in, float *out, int nz, int ny, int nx, int radius, double eps) {
    return guided_host(in, out, nz, ny, nx, radius, eps, 1);
}

size_t fy_guided_ws_floats(int nz, int ny, int nx) { return (size_t)4 * nz * ny * nx; }

int fy_guided_denoise_fast_ws(const float *in, float *out, int nz, int ny, int nx, int radius, double eps,
                              int s, float *ws) {
    (void)ws;
    return guided_host(in, out, nz, ny, nx, radius, eps, s <= 1 ? 1 : s);
}

void fy_box_smooth(const float *in, float *out, float *tmp, int nz, int ny, int nx, int r) {
    (void)tmp;
    if (nz <= 0 || ny <= 0 || nx <= 0) return;
    fyg_stream *s = fyg_thread_stream();
    if (!s) return;
    size_t n = (size_t)nz * ny * nx;
    float *d = (float *)fyg_ws(s, WS_IN, sizeof(float) * 2 * n);
    void *w = fyg_ws(s, WS_GUIDED, sizeof(float) * 2 * n);
    if (!d || !w) return;
    if (fyg_h2d(d, in, sizeof(float) * n, s)) return;
    if (fyg_box_mean_run(d, d + n, nz, ny, nx, r, w, s)) return;
    if (fyg_d2h(out, d + n, sizeof(float) * n, s)) return;
    fyg_sync(s);
}

/* eps ~ (3*noise_ref)^2 with a small positive floor */
double fy_guided_eps_for_noise(double noise_ref) {
    double e = (3.0 * noise_ref) * (3.0 * noise_ref);
    return e > 1e-8 ? e : 1e-8;
}

int fy_gpu_guided(const float *d_in, float *d_out, int nz, int ny, int nx, int radius, double eps, int sub,
                  void *stream) {
    if (nz <= 0 || ny <= 0 || nx <= 0) return 0;
    fyg_stream *s = fyg_stream_wrap(stream);
    if (!s) return 1;
    return guided_dev(d_in, d_out, nz, ny, nx, radius, eps, sub <= 1 ? 1 : sub, s);
}
