/* ref_harness.c -- checker-only access to static routines of the compiled
 * reference.  Built by oracle/Makefile into oracle/_ref/libfysics_ref.so together
 * with the reference's other sources, straight from where they lie (nothing is
 * copied into this repository).  Exposes:
 *   ref_deconv_radial  -- paganin.c:152 with either weight (wiener / matched),
 *                         so the anisotropic Paganin+unsharp case is reachable
 *   ref_quant_u8       -- pipeline.c:60
 *   ref_eps_for_tile   -- pipeline.c:447
 *   ref_air_cc_despeckle -- pipeline.c:227
 *   ref_process_tile   -- pipeline.c:252 (stage d, the air-zero, in isolation
 *                         when the cfg turns the other stages off)
 * TEST INFRASTRUCTURE ONLY.
 */
#include "paganin.c"
#include "pipeline.c"

int ref_deconv_radial(const float *in, float *out, int nz, int ny, int nx,
                      const fy_physics *p, double param, int kind, double cap,
                      double fz2_scale) {
    return deconv_radial(in, out, nz, ny, nx, p, param,
                         kind ? matched_transfer : wiener_weight, cap, fz2_scale);
}

unsigned char ref_quant_u8(float v, long z, long y, long x, int no_dither) {
    return quant_u8(v, z, y, x, no_dither);
}

double ref_eps_for_tile(const fy_pipeline_cfg *cfg, double ecy, double ecx,
                        long y0, long x0, long ty, long tx) {
    return eps_for_tile(cfg, ecy, ecx, y0, x0, ty, tx);
}

/* The per-tile chain of fy_process_buffer (pipeline.c:467-510 with process_tile
 * stages b-c, :252-297) with the deconvolution weight, regularisation, gain cap
 * and z frequency scale passed explicitly, so the Paganin+unsharp Wiener
 * weight can run with an anisotropic sigma_z (deconv_radial with wiener_weight
 * and fz2_scale).  Composed from the reference's own routines. */
int ref_chain_cube(const unsigned char *u8buf, long rz0, long ry0, long rx0, long hz, long hy, long hx,
                   long z0, long y0, long x0, long tz, long ty, long tx, unsigned char *out,
                   const fy_physics *p, int deconv, double param, double cap, double fz2_scale,
                   double wlo_d, double whi_d, double eps, int s, const fy_dering *ring, int no_dither) {
    size_t hn = (size_t)hz * hy * hx;
    if (!any_nonzero(u8buf, hn)) { memset(out, 0, (size_t)tz * ty * tx); return 0; }
    float *f = malloc(sizeof(float) * hn), *b1 = malloc(sizeof(float) * hn);
    float *ws = malloc(sizeof(float) * 4 * hn);
    if (!f || !b1 || !ws) { free(f); free(b1); free(ws); return 1; }
    u8_to_f01(u8buf, f, hn);
    if (ring && ring->have_rings) fy_dering_apply(ring, f, rz0, ry0, rx0, hz, hy, hx, 1.0 / 255.0);
    if (deconv) {
        int windowed = whi_d > wlo_d;
        float wlo = (float)wlo_d, wspan = (float)(whi_d - wlo_d);
        if (windowed) for (size_t i = 0; i < hn; i++) f[i] = wlo + f[i] * wspan;
        int rc = deconv_radial(f, b1, (int)hz, (int)hy, (int)hx, p, param,
                               deconv == 2 ? matched_transfer : wiener_weight, cap, fz2_scale);
        if (rc == 0) {
            if (windowed) { float inv = 1.0f / wspan; for (size_t i = 0; i < hn; i++) b1[i] = (b1[i] - wlo) * inv; }
            for (size_t i = 0; i < hn; i++) { float v = b1[i]; f[i] = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); }
        }
    }
    if (eps > 0 && fy_guided_denoise_fast_ws(f, b1, (int)hz, (int)hy, (int)hx, 2, eps, s, ws) == 0)
        memcpy(f, b1, sizeof(float) * hn);
    long iz0 = z0 - rz0, iy0 = y0 - ry0, ix0 = x0 - rx0, oi = 0;
    for (long zz = iz0; zz < iz0 + tz; zz++)
        for (long yy = iy0; yy < iy0 + ty; yy++)
            for (long xx = ix0; xx < ix0 + tx; xx++)
                out[oi++] = quant_u8(f[((size_t)zz * hy + yy) * hx + xx], rz0 + zz, ry0 + yy, rx0 + xx, no_dither);
    free(f); free(b1); free(ws);
    return 0;
}

/* n independent cubes of hz*hy*hx, one per OpenMP thread (the pass-2 tile loop's
 * parallel pattern, pipeline.c:1008-1020); inner tile = cube minus `halo` per
 * side.  Returns wall seconds. */
double ref_chain_many(int n, const unsigned char **u8, unsigned char **out, long hz, long hy, long hx, long halo,
                      const fy_physics *p, int deconv, double param, double cap, double fz2_scale,
                      double wlo, double whi, double eps, int s, const fy_dering *ring, int nthreads) {
    double t0 = omp_get_wtime();
    int fails = 0;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads) reduction(+:fails)
    for (int i = 0; i < n; i++)
        fails += ref_chain_cube(u8[i], 0, 0, 0, hz, hy, hx, halo, halo, halo, hz - 2 * halo, hy - 2 * halo,
                                hx - 2 * halo, out[i], p, deconv, param, cap, fz2_scale, wlo, whi, eps, s, ring, 0);
    double t = omp_get_wtime() - t0;
    return fails ? -t : t;
}

void ref_air_cc_despeckle(float *f, int nz, int ny, int nx, int min_size) {
    size_t n = (size_t)nz * ny * nx;
    int *labels = malloc(sizeof(int) * (n ? n : 1)), *queue = malloc(sizeof(int) * (n ? n : 1));
    if (labels && queue) air_cc_despeckle(f, nz, ny, nx, min_size, labels, queue);
    free(labels);
    free(queue);
}

int ref_process_tile(float *f, const float *orig, const unsigned char *u8orig, int nz, int ny, int nx,
                     const fy_pipeline_cfg *cfg) {
    size_t n = (size_t)nz * ny * nx;
    float *b1 = malloc(sizeof(float) * n), *b2 = malloc(sizeof(float) * n), *ws = malloc(sizeof(float) * 4 * n);
    int rc = !b1 || !b2 || !ws;
    if (!rc) process_tile(f, orig, u8orig, nz, ny, nx, cfg, cfg->have_dec_range, cfg->dec_lo, cfg->dec_hi,
                          b1, b2, ws, 0.0);
    free(b1); free(b2); free(ws);
    return rc;
}
