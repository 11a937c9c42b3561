/* fysics.h -- MI355X-native drop-in for the volume-restoration hot path of fysics.
 *
 * Same C ABI as the reference header (fysics.h, extern "C", fysics.h:18-20): the
 * structs below are layout-identical and every function keeps the reference's
 * name, argument meaning, aliasing rules and 0/1 return convention, so viewers,
 * the pipeline tile loop and ctypes callers link against libfysics_amd.so
 * unchanged.  Each entry point is computed by hand-written HIP kernels on the
 * GPU (gfx950); host pointers in, host pointers out.  Device-resident, batched
 * and served forms live beside these in fysics_gpu.h.
 *
 * Only the hot-path subset is provided (FFT, transfers, Wiener deconvolution,
 * guided filter, u8/physical conversions, ring removal, noise fit, the pass-1
 * calibration statistics -- histogram, z-drift sums, flat-block noise floor --
 * the per-tile chain with its air-zero stage, the 2x LOD downscale and phase
 * correlation).  Registration, zarr I/O (local zarr I/O lives beside, under
 * fy_gpu_ names), fy_process_chunk, MUSICA and the whole-volume calibrate/run
 * drivers are not part of it.
 *
 * Layout: row-major, x fastest, idx = (z*ny + y)*nx + x (fysics.h:120).
 */
#ifndef FYSICS_H
#define FYSICS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Acquisition / reconstruction physics.  Layout of fysics.h:23-35 (8 doubles). */
typedef struct {
    double delta_beta;        /* Paganin delta/beta */
    double energy_kev;        /* beam energy, keV */
    double distance_mm;       /* propagation distance, mm */
    double pixel_um;          /* voxel size, micron (values < 0.1 are taken as mm) */
    double unsharp_sigma;     /* unsharp Gaussian sigma, voxels; 0 disables */
    double unsharp_coeff;     /* unsharp coefficient; 0 disables */
    double psf_sigma_vox;     /* system PSF sigma (matched deconvolution); <=0 -> 1 */
    double psf_sigma_z_vox;   /* z PSF sigma if anisotropic; 0 -> isotropic */
} fy_physics;

/* ---- FFT, sizes 2^k and 3*2^k (fysics.h:38-43; fft.c:25-38,170-228) ---- */
int  fy_is_pow2(int n);
int  fy_next_pow2(int n);
int  fy_next_fft_size(int n);
/* in place, split re/im, sign -1 forward / +1 inverse, unnormalised (fft.c:170) */
void fy_fft1d(float *re, float *im, int n, int sign);
/* in place over (nz,ny,nx), axis order X,Y,Z (fft.c:178) */
void fy_fft3d(float *re, float *im, int nz, int ny, int nx, int sign);
/* multiply by 1/(nz*ny*nx) (fft.c:223) */
void fy_fft3d_normalize(float *re, float *im, int nz, int ny, int nx);

/* ---- u8 <-> float and u8 <-> physical window (fysics.h:56-76; stream.c:27-63) ---- */
void  fy_u8_to_float(const unsigned char *in, float *out, size_t n);
void  fy_float_to_u8(const float *in, unsigned char *out, size_t n);
void  fy_u8_to_phys(const unsigned char *in, float *out, size_t n,
                    double f32_min, double f32_max);
void  fy_phys_to_u8(const float *in, unsigned char *out, size_t n,
                    double f32_min, double f32_max);
float fy_phys_to_u8_level(double phys, double f32_min, double f32_max);

/* ---- pass-1 calibration statistics (fysics.h:44-95; stream.c:65-200) ----
 * Global 256-bin histogram of u8 values: accumulated per chunk on the GPU,
 * merged and read on the host.  Percentiles: the first bin whose running count
 * reaches (long)(pct/100*total); the _inner form counts bins 1..254 only. */
typedef struct { long hist[256]; long total; } fy_hist_state;       /* fysics.h:54 */
void fy_hist_init(fy_hist_state *s);
void fy_hist_accumulate_u8(fy_hist_state *s, const unsigned char *chunk, size_t n);
void fy_hist_merge(fy_hist_state *dst, const fy_hist_state *src);
int  fy_hist_percentile_u8(const fy_hist_state *s, double pct);
int  fy_hist_percentile_u8_inner(const fy_hist_state *s, double pct);
void fy_norm_finalize(const fy_hist_state *s, double lo_pct, double hi_pct,
                      unsigned char *lo_out, unsigned char *hi_out);
/* out = clamp((u8 - lo) / (hi - lo), 0, 1) (1/(hi-lo) -> 1 when hi <= lo) */
void fy_norm_apply_u8(const unsigned char *in, float *out, size_t n,
                      unsigned char lo, unsigned char hi);
/* bimodal valley depth of a 256-bin histogram (rails 0/254/255 excluded):
 * 1 - h_valley/min(h_dark, h_light) of the 5-bin smoothed histogram, writing the
 * dark/light peak and valley bins (each may be NULL); -1 if not bimodal */
double fy_valley_depth(const long hist[256], int *dark_out, int *light_out, int *valley_out);
/* noise floor: 10th percentile of the stds of the non-overlapping blk^3 blocks
 * whose mean exceeds 0.1 (blk < 4 -> 8); 0 when there is no such block */
double fy_flat_noise(const float *v, int nz, int ny, int nx, int blk);

/* ---- transfer functions, f in cycles/voxel (fysics.h:98-117; paganin.c:17-91) ---- */
double fy_paganin_transfer(double f_cyc_per_voxel, const fy_physics *p);
double fy_unsharp_transfer(double f_cyc_per_voxel, const fy_physics *p);
double fy_recon_transfer(double f_cyc_per_voxel, const fy_physics *p);
double fy_auto_reg(const fy_physics *p, double base);
double fy_auto_deltabeta_scale(const fy_physics *p);

/* ---- Wiener deconvolution (fysics.h:131-146; paganin.c:93-118,152-342) ----
 * Reflect-pad to fy_next_fft_size per axis, real 3-D FFT, half-spectrum times the
 * 4097-entry LUT weight, inverse, crop.  out may equal in.  0 = success, 1 = failure
 * (device allocation / launch error). */
int fy_deconvolve(const float *in, float *out, int nz, int ny, int nx,
                  const fy_physics *p, double reg);
int fy_deconvolve_matched(const float *in, float *out, int nz, int ny, int nx,
                          const fy_physics *p, double tikhonov);
int fy_kernel_halo(const fy_physics *p);

/* ---- guided filter (fysics.h:152-167; guided.c:104-241, noise.c:197-202) ----
 * in/out may alias; ws is accepted for ABI compatibility and not touched
 * (device scratch is owned by the library). */
int    fy_guided_denoise(const float *in, float *out, int nz, int ny, int nx,
                         int radius, double eps);
int    fy_guided_denoise_ws(const float *in, float *out, int nz, int ny, int nx,
                            int radius, double eps, float *ws);
size_t fy_guided_ws_floats(int nz, int ny, int nx);
int    fy_guided_denoise_fast_ws(const float *in, float *out, int nz, int ny, int nx,
                                 int radius, double eps, int s, float *ws);
void   fy_box_smooth(const float *in, float *out, float *tmp, int nz, int ny, int nx, int r);
double fy_guided_eps_for_noise(double noise_ref);

/* ---- per-volume noise model (fysics.h:191-215; noise.c:75-269) ---- */
typedef struct {
    double g;             /* var = g*I + b */
    double b;
    double noise_ref;     /* noise std at ref_intensity */
    double ref_intensity;
    int    n_bins_used;
} fy_noise_model;

/* var = g*I + b from local (win^3) means/variances: 24 intensity bins, per-bin
 * flat_pct percentile floor, 4-iteration IRLS fit; noise_ref = sqrt(g*ref+b)*1.15
 * (noise.c:75-192).  0 on success. */
int fy_estimate_noise(const float *in, int nz, int ny, int nx,
                      int win, double flat_pct, double ref_intensity,
                      fy_noise_model *out);
/* lag-1 autocorrelation of the residual in flat 8^3 blocks (noise.c:210-269);
 * 1 if fewer than 64 blocks / 32 flat blocks */
int fy_noise_aniso(const float *v, int nz, int ny, int nx,
                   double *rho_z, double *rho_xy);

/* ---- z-drift (fysics.h:216-228; zdrift.c) ----
 * accumulate: per slice z0+z, the sum (double) and count of values above
 * papyrus_thresh; finalize: factor[z] = global mean / (mean over the slices within
 * nz/20+1 that have papyrus); apply: chunk *= factor[z0+z] (used inside the tile
 * chain); correct: the three over a whole volume.  Accumulate and apply run on the
 * GPU, finalize (one scalar per slice) on the host. */
void fy_zdrift_accumulate(const float *chunk, int nz_slab, int ny, int nx,
                          int z0, double *sums, long *counts, float papyrus_thresh);
void fy_zdrift_finalize(const double *sums, const long *counts, int nz, float *factor);
void fy_zdrift_apply(float *chunk, int nz_slab, int ny, int nx, int z0, const float *factor);
int  fy_correct_zdrift(float *vol, int nz, int ny, int nx, float papyrus_thresh);

/* ---- 2x LOD downscale (fysics.h:388-396; downscale.c) ----
 * Each output voxel is a function of its own 2x2x2 parent cell (edge cells
 * repeat the last plane/row/column); an all-zero cell stays zero.  BOX: the
 * cell mean; CBOX: the mean moved by alpha toward the cell's farthest-from-mean
 * voxel.  Rounded to u8 by +0.5 and truncation.  ox/oy/oz (may be NULL)
 * receive (n+1)/2 per axis.  Note the argument order nx, ny, nz. */
typedef enum { FY_DS_BOX = 0, FY_DS_CBOX } fy_ds_method;
void fy_downscale2x(const unsigned char *in, int nx, int ny, int nz,
                    unsigned char *out, int *ox, int *oy, int *oz,
                    fy_ds_method method, float alpha);

/* ---- sub-voxel translation by phase correlation (fysics.h:296-313; phasecorr.c:44-186) ----
 * shift (dz,dy,dx) such that moving(p) ~ fixed(p - shift): mean removal and an
 * optional separable Hann window, zero padding to the next power of two per
 * axis, the unit-magnitude cross-power spectrum conj(F)*M/|conj(F)*M|, its
 * inverse FFT, the first maximum of the surface unwrapped to a signed shift and
 * refined per axis by Foroosh's side-lobe estimator (clamped to +-0.5).
 * peak (may be NULL) = surface maximum / padded voxel count.  1 on bad
 * arguments (any axis < 2) or a flat input. */
int fy_phase_correlate(const float *fixed, const float *moving, int nz, int ny, int nx,
                       double *shift, double *peak, int window);

/* ---- ring removal (fysics.h:406-433; dering.c) ---- */
typedef struct {
    long Z, Y, X;
    double cy, cx;        /* rotation axis (voxels) */
    int slab_z, nslab;
    int ns, nr;           /* sectors, radial bins */
    float *sum;           /* [nslab][ns][nr] u8 sums */
    unsigned int *cnt;    /* [nslab][ns][nr] counts */
    float *ring;          /* [nslab][nr] detected profile (u8 units) */
    int have_rings;
} fy_dering;

int  fy_dering_init(fy_dering *d, long Z, long Y, long X,
                    double cy, double cx, int slab_z, int ns);
void fy_dering_free(fy_dering *d);
int  fy_dering_tile_init(fy_dering *t, const fy_dering *d);
void fy_dering_tile_reset(fy_dering *t);
void fy_dering_accumulate_u8(fy_dering *t, const unsigned char *buf,
                             long y0, long x0, long dz, long dy, long dx, int ss);
void fy_dering_merge_tile(fy_dering *d, int slab, const fy_dering *t);
long fy_dering_finalize(fy_dering *d, int hp_win, unsigned int min_cnt,
                        double min_amp, double max_amp);
void fy_dering_apply(const fy_dering *d, float *f, long z0, long y0, long x0,
                     long dz, long dy, long dx, double scale);

/* ---- pipeline configuration (layout of fysics.h:436-491) ---- */
typedef struct {
    double delta_beta, energy_kev, distance_mm, pixel_um, unsharp_sigma, unsharp_coeff;
    double machine_current_start, machine_current_stop;
    double window_lo, window_hi;
    int    do_deconv;
    int    use_matched_deconv;
    double psf_sigma_vox, deconv_tikhonov;
    double guided_eps;
    int    do_air_zero; int air_cut_u8, air_cut_band; double air_thresh;
    double air_cut_aggr;
    int    air_min_component;
    double denoise_k;
    int    scratch_passes;
    int    do_normalize; int norm_lo, norm_hi;
    int    do_zdrift; float *zdrift_factor;
    int    do_dering;
    double calib_budget_gb;
    double dering_cy, dering_cx;
    double psf_p5, psf_med;
    int    do_musica; double musica_p; int musica_levels; double musica_core;
    int    guided_subsample;
    int    no_dither;
    int    halo;
    int    have_norm, have_zdrift, have_dec_range;
    int    have_dering; fy_dering *dering;
    double psf_z_ratio;
    int    have_eps_r; double eps_fn_a, eps_fn_b, eps_fn_med;
    double dec_lo, dec_hi;
    long   vol_z;
    double noise_floor;
    const unsigned char *occ;
    long occ_ncz, occ_ncy, occ_ncx;
    long occ_cz, occ_cy, occ_cx;
} fy_pipeline_cfg;

/* Process one inner tile from an already-read halo'd u8 buffer (fysics.h:507-511;
 * pipeline.c:467-510): u8 -> f01 -> dering -> z-drift -> (deconv chain) -> guided
 * (per-tile eps) -> air-zero + despeckle (do_air_zero) -> recenter -> dithered
 * quantisation of the inner tile.  The MUSICA stage (do_musica) is not
 * provided: a cfg that enables it returns 1 without touching out. */
int fy_process_buffer(const fy_pipeline_cfg *cfg, const unsigned char *u8buf,
                      long rz0, long ry0, long rx0, long hz, long hy, long hx,
                      long z0, long y0, long x0, long tz, long ty, long tx,
                      long vol_y, long vol_x,
                      unsigned char *out, int *all_air);

/* The zarr volume handle (fysics.h:376-381, same layout).  This library reads
 * and writes local zarr volumes under its own names (fy_gpu_zarr_*,
 * fy_gpu_process_chunk in fysics_gpu.h); the reference's fy_zarr_* and
 * fy_process_chunk, which also reach S3, are not replaced. */
typedef struct {
    char root[1024];
    long shape[3];      /* z,y,x voxels */
    long chunk[3];      /* chunk dims */
    unsigned char fill; /* fill_value */
} fy_zarr;

#ifdef __cplusplus
}
#endif

#endif /* FYSICS_H */
