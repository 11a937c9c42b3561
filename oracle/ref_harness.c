/* ref_harness.c -- checker-only access to static routines of the compiled
 * reference.  Built by oracle/Makefile into oracle/_ref/libfysics_ref.so together
 * with the reference's other sources, straight from where they lie (nothing is
 * copied into this repository).  Exposes:
 *   ref_deconv_radial  -- paganin.c:152 with either weight (wiener / matched),
 *                         so the anisotropic Paganin+unsharp case is reachable
 *   ref_quant_u8       -- pipeline.c:60
 *   ref_eps_for_tile   -- pipeline.c:447
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
