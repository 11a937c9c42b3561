/* oracle.h -- CPU restatement of the fysics hot path, used ONLY as a checker.
 *
 * TEST INFRASTRUCTURE.  Nothing in the product (libfysics_amd.so) links, loads or
 * calls this code: it is imported by tests/, by __graft_entry__.smoke() and by
 * the cpu_baseline leg of bench.py, and there only to produce the expected
 * answer.  Every function restates one reference routine (cited file:line into
 * the fysics sources); the fp32 versions follow the reference's operation order,
 * the fp64 twins (suffix _f64) evaluate the same algorithm -- including the
 * 4097-entry LUT weight semantics of deconv_radial -- in double precision.
 *
 * Pinning (what stops this from being a retyped formula): tests/test_oracle.py
 * checks each function against golden vectors produced by the compiled
 * reference (oracle/_ref, tests/golden/) and against closed forms (naive DFT,
 * Parseval, brute-force box means, the nabu Paganin formula, integer hashes).
 * Functions without such a check are marked "parity unpinned" below.
 */
#ifndef FYSICS_ORACLE_H
#define FYSICS_ORACLE_H

#include <stddef.h>
#include "../include/fysics.h"       /* struct layouts only */
#include "../include/fysics_gpu.h"   /* fy_cube_params layout only */

#ifdef __cplusplus
extern "C" {
#endif

/* fft.c */
int  orc_next_fft_size(int n);
void orc_fft1d_f32(float *re, float *im, int n, int sign);
void orc_fft1d_f64(double *re, double *im, int n, int sign);
void orc_fft3d_f32(float *re, float *im, int nz, int ny, int nx, int sign);
void orc_fft3d_f64(double *re, double *im, int nz, int ny, int nx, int sign);
void orc_fft3d_normalize_f32(float *re, float *im, int nz, int ny, int nx);

/* paganin.c transfers */
double orc_paganin_transfer(double f, const fy_physics *p);
double orc_unsharp_transfer(double f, const fy_physics *p);
double orc_recon_transfer(double f, const fy_physics *p);
double orc_auto_reg(const fy_physics *p, double base);
double orc_auto_deltabeta_scale(const fy_physics *p);
int    orc_kernel_halo(const fy_physics *p);

/* paganin.c deconv_radial. kind 0 = Wiener of the recon transfer (fy_deconvolve),
 * kind 1 = matched PSF*unsharp (fy_deconvolve_matched). */
int orc_deconv_radial_f32(const float *in, float *out, int nz, int ny, int nx,
                          const fy_physics *p, double param, int kind,
                          double cap, double fz2_scale);
int orc_deconv_radial_f64(const float *in, double *out, int nz, int ny, int nx,
                          const fy_physics *p, double param, int kind,
                          double cap, double fz2_scale);
/* the LUT that deconv_radial builds (4097 floats) and its scale us */
void orc_deconv_lut(int pz, int py, int px, const fy_physics *p, double param, int kind,
                    double cap, double fz2_scale, float *lut, double *us);
int orc_deconvolve(const float *in, float *out, int nz, int ny, int nx,
                   const fy_physics *p, double reg);
int orc_deconvolve_matched(const float *in, float *out, int nz, int ny, int nx,
                           const fy_physics *p, double tikhonov);
/* parameters fy_deconvolve_matched passes to deconv_radial */
double orc_matched_fz2_scale(const fy_physics *p);

/* guided.c */
void orc_box_mean_f32(const float *in, float *out, int nz, int ny, int nx, int r);
void orc_box_mean_f64(const double *in, double *out, int nz, int ny, int nx, int r);
int  orc_guided_exact_f32(const float *in, float *out, int nz, int ny, int nx, int r, double eps);
int  orc_guided_fast_f32(const float *in, float *out, int nz, int ny, int nx, int r, double eps, int s);
int  orc_guided_exact_f64(const double *in, double *out, int nz, int ny, int nx, int r, double eps);
int  orc_guided_fast_f64(const double *in, double *out, int nz, int ny, int nx, int r, double eps, int s);

/* stream.c conversions */
void  orc_u8_to_float(const unsigned char *in, float *out, size_t n);
void  orc_float_to_u8(const float *in, unsigned char *out, size_t n);
void  orc_u8_to_phys(const unsigned char *in, float *out, size_t n, double lo, double hi);
void  orc_phys_to_u8(const float *in, unsigned char *out, size_t n, double lo, double hi);
float orc_phys_to_u8_level(double phys, double lo, double hi);

/* pipeline.c quantiser */
float         orc_dither01(long z, long y, long x);
unsigned char orc_quant_u8(float v, long gz, long gy, long gx, int no_dither);

/* dering.c */
int  orc_dering_init(fy_dering *d, long Z, long Y, long X, double cy, double cx, int slab_z, int ns);
void orc_dering_free(fy_dering *d);
int  orc_dering_tile_init(fy_dering *t, const fy_dering *d);
void orc_dering_accumulate_u8(fy_dering *t, const unsigned char *buf,
                              long y0, long x0, long dz, long dy, long dx, int ss);
void orc_dering_merge_tile(fy_dering *d, int slab, const fy_dering *t);
long orc_dering_finalize(fy_dering *d, int hp_win, unsigned int min_cnt, double min_amp, double max_amp);
void orc_dering_apply(const fy_dering *d, float *f, long z0, long y0, long x0,
                      long dz, long dy, long dx, double scale);

/* noise.c */
int    orc_estimate_noise(const float *in, int nz, int ny, int nx, int win,
                          double flat_pct, double ref_intensity, fy_noise_model *out);
double orc_guided_eps_for_noise(double noise_ref);
int    orc_noise_aniso(const float *v, int nz, int ny, int nx, double *rho_z, double *rho_xy);

/* stream.c histogram, normalisation, flat-block noise; zdrift.c accumulate and
 * finalize; the pass-1 u8 z-drift subsample sums of pipeline.c */
void   orc_hist_accumulate_u8(fy_hist_state *s, const unsigned char *chunk, size_t n);
int    orc_hist_percentile_u8(const fy_hist_state *s, double pct);
int    orc_hist_percentile_u8_inner(const fy_hist_state *s, double pct);
void   orc_norm_apply_u8(const unsigned char *in, float *out, size_t n, unsigned char lo, unsigned char hi);
void   orc_zdrift_accumulate(const float *chunk, int nz, int ny, int nx, int z0,
                             double *sums, long *counts, float thresh);
void   orc_zdrift_finalize(const double *sums, const long *counts, int nz, float *factor);
void   orc_zdrift_u8_sub(const unsigned char *pb, long dz, long dy, long dx, long z0,
                         double *sums, long *counts);
double orc_flat_noise(const float *v, int nz, int ny, int nx, int blk);

/* pipeline.c:735-740,851-900 pass-1 commits (radial eps fit, median floor,
 * guided eps; per-sample anisotropy ratio and its median commit).  Pinned by
 * closed forms in tests/test_oracle.py: exact linear data returns its own
 * line, each gate fails on its own, even/odd medians, the k = 4.2 default. */
void   orc_calib_commit_eps(fy_pipeline_cfg *cfg, const double *rads, const double *floors, int n);
double orc_aniso_ratio(double arz, double arxy);
void   orc_calib_commit_aniso(fy_pipeline_cfg *cfg, const double *ratios, int n);

/* stream.c valley depth; pipeline.c air-zero stage (d) and CC despeckle */
double orc_valley_depth(const long hist[256], int *dark_out, int *light_out, int *valley_out);
void   orc_air_cc_despeckle(float *f, int nz, int ny, int nx, int min_size);
void   orc_air_zero(float *f, const float *orig, int nz, int ny, int nx, const fy_cube_params *P);

/* downscale.c 2x LOD (method 0 box, 1 contrast-maintaining box) */
void   orc_downscale2x(const unsigned char *in, int nx, int ny, int nz, unsigned char *out,
                       int *ox, int *oy, int *oz, int method, float alpha);

/* phasecorr.c phase correlation (with the restated fft.c FFT) */
int    orc_phase_correlate(const float *fixed, const float *moving, int nz, int ny, int nx,
                           double *shift, double *peak, int window);

/* zdrift.c apply; pipeline.c per-tile eps and the tile chain */
void   orc_zdrift_apply(float *chunk, int nz, int ny, int nx, int z0, const float *factor);
double orc_eps_for_tile(const fy_pipeline_cfg *cfg, double ecy, double ecx,
                        long y0, long x0, long ty, long tx);
int    orc_process_buffer(const fy_pipeline_cfg *cfg, const unsigned char *u8buf,
                          long rz0, long ry0, long rx0, long hz, long hy, long hx,
                          long z0, long y0, long x0, long tz, long ty, long tx,
                          long vol_y, long vol_x, unsigned char *out, int *all_air);
/* the chain with explicit cube parameters; ftile receives the float tile after
 * the guided filter (before recenter and quantisation) */
int    orc_chain_cube(const fy_cube_params *P, const unsigned char *u8buf,
                      long rz0, long ry0, long rx0, long hz, long hy, long hx,
                      long z0, long y0, long x0, long tz, long ty, long tx,
                      unsigned char *out, int *all_air, float *ftile);
/* the same chain with deconvolution and guided filter in fp64: the double
 * float tile after the filter (ftile, may be NULL) and the u8 inner tile */
int    orc_chain_cube_f64(const fy_cube_params *P, const unsigned char *u8buf,
                          long rz0, long ry0, long rx0, long hz, long hy, long hx,
                          long z0, long y0, long x0, long tz, long ty, long tx,
                          unsigned char *out, double *ftile);
int    orc_cube_params_from_cfg(const fy_pipeline_cfg *cfg, long y0, long x0, long ty, long tx,
                                long vol_y, long vol_x, fy_cube_params *o);
/* same chain, also returning the float tile after the guided filter (hz*hy*hx) */
int    orc_process_buffer_f(const fy_pipeline_cfg *cfg, const unsigned char *u8buf,
                            long rz0, long ry0, long rx0, long hz, long hy, long hx,
                            long z0, long y0, long x0, long tz, long ty, long tx,
                            long vol_y, long vol_x, unsigned char *out, int *all_air,
                            float *ftile);

#ifdef __cplusplus
}
#endif
#endif
