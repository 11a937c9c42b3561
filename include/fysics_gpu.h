/* fysics_gpu.h -- entry points that the CPU reference has no counterpart for:
 * device-resident forms of the hot path (inputs and outputs already in HBM,
 * work queued on a caller-supplied HIP stream) and the kernel-timing hooks.
 *
 * Pointers named d_* are device addresses (hipMalloc or a torch tensor's
 * data_ptr()).  `stream` is a hipStream_t; NULL selects the calling thread's
 * own stream.  Calls only enqueue work: results are ready after
 * fy_gpu_sync(stream) (or any later stream-ordered consumer).  Return 0 on
 * success, 1 on a device error or an unsupported size, like fysics.h.
 */
#ifndef FYSICS_GPU_H
#define FYSICS_GPU_H

#include <stddef.h>
#include "fysics.h"

#ifdef __cplusplus
extern "C" {
#endif

/* fy_deconvolve / fy_deconvolve_matched on device buffers (paganin.c:301-342).
 * d_out may equal d_in. */
int fy_gpu_deconvolve(const float *d_in, float *d_out, int nz, int ny, int nx,
                      const fy_physics *p, double reg, void *stream);
int fy_gpu_deconvolve_matched(const float *d_in, float *d_out, int nz, int ny, int nx,
                              const fy_physics *p, double tikhonov, void *stream);
/* the deconvolution core with an explicit weight (paganin.c:152-293):
 * kind 0 = Wiener of the recon transfer, kind 1 = matched PSF*unsharp;
 * cap <= 0 disables the gain clamp; fz2_scale = (sigma_z/sigma_xy)^2 */
int fy_gpu_deconv_radial(const float *d_in, float *d_out, int nz, int ny, int nx,
                         const fy_physics *p, double param, int kind, double cap,
                         double fz2_scale, void *stream);

/* fy_gpu_deconvolve that allocates nothing.  No reference counterpart: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The half
 * spectrum, the weight table, the frequency axes and the twiddles live in one
 * device array per GPU that the library holds until the process ends (80 MiB
 * of spectrum plus 256 KiB of tables); the tables are computed on the device
 * at each call.  Limit: the padded volume's half spectrum,
 * pz*py*roundup(px/2+1, 16)*8 bytes, must fit 80 MiB (256^3 fits, 384^3 does
 * not); larger sizes return 1 and fy_gpu_last_error() says why.  Calls that
 * overlap on one GPU share the array and race: serialise them (one stream, or
 * the caller's own ordering).  d_out may equal d_in. */
int fy_gpu_deconvolve_noalloc(const float *d_in, float *d_out, int nz, int ny, int nx,
                              const fy_physics *p, double reg, void *stream);

/* the all-FP64 variant of fy_gpu_deconv_radial: fp32 input, spectrum, weights
 * and output in double (the speed reference of the precision study; no
 * reference counterpart) */
int fy_gpu_deconv_radial_f64(const float *d_in, double *d_out, int nz, int ny, int nx,
                             const fy_physics *p, double param, int kind, double cap,
                             double fz2_scale, void *stream);

/* guided filter on device buffers; s <= 1 exact (guided.c:113), else fast (guided.c:152) */
int fy_gpu_guided(const float *d_in, float *d_out, int nz, int ny, int nx,
                  int radius, double eps, int s, void *stream);

/* fy_process_buffer on device buffers (pipeline.c:467-510): d_u8 is the halo'd
 * hz*hy*hx tile, d_out receives tz*ty*tx u8.  all_air, if non-NULL, makes the
 * call synchronous and reports whether the tile was all zero. */
int fy_gpu_process_buffer(const fy_pipeline_cfg *cfg, const unsigned char *d_u8,
                          long rz0, long ry0, long rx0, long hz, long hy, long hx,
                          long z0, long y0, long x0, long tz, long ty, long tx,
                          long vol_y, long vol_x, unsigned char *d_out, int *all_air,
                          void *stream);

/* ---- per-cube chain (the batched and served forms' unit of work) ----
 * One cube = a halo'd u8 region and the inner tile to emit, processed in
 * fy_process_buffer order: u8/255 -> ring subtraction -> z-drift -> dewindow ->
 * deconvolution -> rewindow / dec-range / clamp -> guided filter -> recenter ->
 * dithered quantisation (pipeline.c:252-297,413-421,467-510).  Unlike
 * fy_pipeline_cfg, the deconvolution weight, its regularisation and the z
 * frequency scale are explicit, so the Paganin+unsharp Wiener weight can be
 * combined with an anisotropic sigma_z (deconv_radial, paganin.c:152, with
 * wiener_weight and fz2_scale = (sigma_z/sigma_xy)^2).  The physics is used as
 * given (no delta/beta back-off is applied here). */
typedef struct {
    fy_physics phys;
    int    deconv;             /* 0 off, 1 Wiener of the recon transfer, 2 matched PSF*unsharp */
    double deconv_param;       /* reg (<=0 -> 0.015) or tikhonov (<=0 -> 0.05) */
    double deconv_cap;         /* gain clamp; <=0 none (fy_deconvolve uses 8) */
    double fz2_scale;          /* (sigma_z/sigma_xy)^2 applied to f_z^2; <=0 -> 1 */
    double window_lo, window_hi;          /* dewindow/rewindow around the deconvolution */
    int    have_dec_range; double dec_lo, dec_hi;
    double guided_eps;         /* <=0: no guided filter */
    int    guided_radius;      /* <1 -> 2 */
    int    guided_s;           /* 0 -> 2 (fast), 1 exact, else decimation */
    int    no_dither;
    int    recenter; int norm_lo, norm_hi;   /* pipeline.c:413-421 */
    const fy_dering *dering;   /* finalized ring state, or NULL */
    const float *zdrift_factor; long vol_z;  /* per-slice factors, or NULL */
    int    fp64;               /* all-FP64 variant of deconv + fast guided (needs guided_s >= 2) */
    /* (d) air-zero (pipeline.c:302-399): the pre-deconvolution tile is box-smoothed
     * on a 2x decimated grid, a cut is chosen from the smooth's histogram valley
     * (fy_valley_depth) and voxels whose upsampled smooth falls below it -- a
     * harder cut where the tile is locally uniform -- are zeroed; then 6-connected
     * material islands that do not touch the region's faces and are smaller than
     * air_min_component voxels are zeroed (pipeline.c:227-250) */
    int    air_zero;
    int    air_cut_u8, air_cut_band;   /* global cut and band; air_cut_u8 < 0: dark-peak rule */
    double air_thresh, air_cut_aggr;   /* physics floor (<= 0 -> u8 39), valley aggressiveness */
    int    air_min_component;          /* <= 0: no despeckle */
    int    scratch_passes;             /* scratch box radius (<= 0 -> 5) */
} fy_cube_params;

typedef struct {
    const unsigned char *d_in;            /* hz*hy*hx u8, device */
    long rz0, ry0, rx0, hz, hy, hx;       /* global origin and shape of the region */
    long z0, y0, x0, tz, ty, tx;          /* inner tile (global), inside the region */
    unsigned char *d_out;                 /* tz*ty*tx u8, device */
    float *d_float;                       /* optional: hz*hy*hx fp32 tile before quantisation */
    const fy_cube_params *params;
} fy_cube;

/* process n cubes on one stream.  Cubes of one region shape (and stage set)
 * share launches: one per stage per 16 cubes, in order of first appearance;
 * batches of several shapes run as one launch per stage (fy_gpu_batch_mix).  A
 * cube's output does not depend on what it is submitted with.  No reference
 * counterpart: the batched form of process_tile (pipeline.c:252-297). */
int fy_gpu_process_cubes(const fy_cube *cubes, int n, void *stream);
/* how fy_gpu_process_cubes (and its graph, and the served waves) launch a
 * batch of several shapes.  When every cube runs the u8 chain of
 * deconvolution and fast guided filter (fp32, s a power of two) at padded
 * lengths of 64, 96, 128, 192 or 256, each stage can be ONE launch over many
 * cubes of mixed shapes (at most 128 per launch): the grid is cut into one
 * segment per cube, largest first, and each segment runs the body its length
 * selects.  mode 2: one such launch per stage for the whole batch; mode 1
 * (default): the batch is packed into parts whose half spectra fit 96 MB and
 * each part is one launch per stage, the parts alternating between two
 * streams (the 256 MB last-level cache then keeps the spectra of the parts in
 * flight between their passes); mode 0: one launch per stage per shape group
 * and 16 cubes, the groups forked over helper streams.  Served waves use
 * mode 0.  All modes give the same bits.  Returns the previous mode.  No
 * reference counterpart. */
int fy_gpu_batch_mix(int mode);
/* the batched form captured: fy_gpu_graph_cubes runs the cubes once, then
 * records the same launches (one per stage per shape group and 16 cubes) into a
 * HIP graph; fy_graph_launch replays them on `stream` with whatever the cubes'
 * buffers hold at that time.  The cubes' buffers and parameters must stay valid
 * until fy_graph_destroy.  No reference counterpart. */
typedef struct fy_graph fy_graph;
fy_graph *fy_gpu_graph_cubes(const fy_cube *cubes, int n, void *stream);
int  fy_graph_launch(fy_graph *g);
void fy_graph_destroy(fy_graph *g);
/* device bytes one cube holds while it runs (workspace admission) */
size_t fy_cube_workspace_bytes(const fy_cube *cube);

/* ---- served form: a request loop with continuous batching over `lanes`
 * streams of its own (no reference counterpart).  fy_serve_submit queues a
 * request whose region and output are in device memory (its buffers and
 * params must stay valid until it is retired); fy_serve_submit_host queues
 * one whose input (hz*hy*hx u8) and output (tz*ty*tx u8) are host buffers --
 * the cube's d_in/d_out are ignored, the wave copies on its lane's stream
 * (pinned buffers make the copies asynchronous).  fy_serve_poll retires the
 * waves that have finished, in the order they finish, then, for each idle
 * lane, launches the pending requests that fit the HBM budget next to what the
 * other lanes hold (FIFO, at most max_batch) as one wave: waves on different
 * lanes overlap on the device.  A wave waits for the work queued on `stream`
 * (NULL: the calling thread's stream) before the poll that launched it.
 * Returns the number of tickets written to done, or -1 on error.
 * fy_serve_submit* return 1 if the request alone exceeds the budget. */
typedef struct fy_server fy_server;
fy_server *fy_serve_create(size_t hbm_budget, int max_batch, int lanes, void *stream);
int  fy_serve_submit(fy_server *sv, const fy_cube *cube, long ticket);
int  fy_serve_submit_host(fy_server *sv, const fy_cube *cube, const unsigned char *h_in,
                          unsigned char *h_out, long ticket);
int  fy_serve_poll(fy_server *sv, long *done, int max_done);
int  fy_serve_outstanding(const fy_server *sv);                          /* queued + in flight */
void fy_serve_destroy(fy_server *sv);
/* the cube parameters fy_process_buffer derives from a pipeline cfg
 * (delta/beta back-off, matched vs plain weight, per-tile eps) */
int fy_cube_params_from_cfg(const fy_pipeline_cfg *cfg, long y0, long x0, long ty, long tx,
                            long vol_y, long vol_x, fy_cube_params *out);

/* Pass-1 commits of fy_run_pipeline as functions (pipeline.c:735-740,851-900;
 * no reference counterpart: the reference runs them inline).  Host fp64.
 * fy_calib_commit_eps: from n flat-noise floors of sample sub-blocks and the
 * distances of their centres to the rotation axis, the gated least-squares fit
 * fn(r) = a + b*r (needs n >= 8, b > 0, correlation >= 0.5, a > 0 and
 * fn(rmax)/a >= 1.3; sets have_eps_r, eps_fn_a, eps_fn_b, or clears
 * have_eps_r), eps_fn_med = the median floor (0.015 with no samples),
 * noise_floor = 255 * median, and, when guided_eps <= 0 and denoise_k >= 0,
 * guided_eps = (k * median)^2 with k = denoise_k (4.2 when 0).  The inputs
 * are not modified. */
void   fy_calib_commit_eps(fy_pipeline_cfg *cfg, const double *rads, const double *floors, int n);
/* sigma_z/sigma_xy of one sample from fy_noise_aniso's lag-1 correlations:
 * sqrt(ln(1/arxy)/ln(1/arz)) when both lie in (0.05, 0.95) and the ratio in
 * (0.5, 3), else -1 */
double fy_aniso_ratio(double arz, double arxy);
/* psf_z_ratio = min(median of n ratios, 1.6) when n >= 5 and the median is
 * >= 1.05; otherwise cfg is left as it is */
void   fy_calib_commit_aniso(fy_pipeline_cfg *cfg, const double *ratios, int n);

/* ---- pass-1 calibration sweep on device tiles (pipeline.c:576-701).
 * A sweep holds device accumulators for the volume's u8 histogram and the
 * per-slice z-drift sums of Z slices.  fy_calib_accumulate_u8 queues the
 * statistics of one contiguous u8 ptile (dz*dy*dx, device memory, first plane
 * at global z0) on the sweep's stream:
 *   FY_CALIB_HIST    the tile's histogram is merged unless the tile is all zero
 *                    (pipeline.c:677-679, fy_hist_accumulate_u8 stream.c:76-86)
 *   FY_CALIB_ZDRIFT  per slice, the sum and count of v = u8/255 > 0.10 over the
 *                    4x4 (y, x) subsample (pipeline.c:686-701)
 * Sums are exact (integers, and multiples of 2^-27 below 2^26 per slice), so
 * they do not depend on the order of tiles or of the voxels within a tile.
 * fy_calib_read waits for the stream and adds the totals to hist / zsums /
 * zcnts (any may be NULL); fy_calib_reset clears the accumulators.
 * No reference counterpart: the reference accumulates on host threads. */
enum { FY_CALIB_HIST = 1, FY_CALIB_ZDRIFT = 2 };
typedef struct fy_calib fy_calib;
fy_calib *fy_calib_create(long Z, void *stream);
int  fy_calib_accumulate_u8(fy_calib *c, const unsigned char *d_tile, long z0,
                            long dz, long dy, long dx, int flags);
int  fy_calib_read(fy_calib *c, fy_hist_state *hist, double *zsums, long *zcnts);
int  fy_calib_reset(fy_calib *c);
void fy_calib_destroy(fy_calib *c);
/* fy_downscale2x on device buffers (downscale.c:27-45); d_out holds
 * ((nz+1)/2)*((ny+1)/2)*((nx+1)/2) bytes */
int  fy_gpu_downscale2x(const unsigned char *d_in, int nx, int ny, int nz, unsigned char *d_out,
                        fy_ds_method method, float alpha, void *stream);
/* fy_phase_correlate of two device volumes (phasecorr.c:44-186); waits for the stream */
int  fy_gpu_phase_correlate(const float *d_fixed, const float *d_moving, int nz, int ny, int nx,
                            double *shift, double *peak, int window, void *stream);
/* fy_flat_noise of a device volume (stream.c:174-200) */
int  fy_gpu_flat_noise(const float *d_v, int nz, int ny, int nx, int blk, double *out, void *stream);

/* ---- local zarr v2 raw-u8 volumes (the format of zarr_io.c:1-9; the
 * reference's fy_zarr_open / _create / _read / _write_chunk, fysics.h:375-389,
 * restricted to local roots and kept under names of their own so that a program
 * linking both libraries keeps the reference's S3-capable functions).
 * <root>/0/.zarray plus one uncompressed C-order file per chunk at
 * <root>/0/cz/cy/cx; chunks on the far faces hold their true extent; a missing
 * chunk (or the part past a short file's end) reads as fill_value.
 * fy_gpu_zarr_open fails on S3 roots, compressed chunks and dtypes other than
 * u1; fy_gpu_zarr_read fills voxels outside the volume with fill_value.
 * Volumes written by either library read back identically in the other. */
int fy_gpu_zarr_open(fy_zarr *z, const char *root);
int fy_gpu_zarr_create(fy_zarr *z, const char *root, const long shape[3], const long chunk[3]);
int fy_gpu_zarr_read(const fy_zarr *z, long z0, long y0, long x0,
                     long dz, long dy, long dx, unsigned char *out);
int fy_gpu_zarr_write_chunk(const fy_zarr *z, long cz, long cy, long cx,
                            const unsigned char *buf, long bz, long by, long bx);
/* chunk mappings shared by the reads of one tile loop: each chunk file a read
 * touches is mapped once and kept while the mapped bytes stay under
 * budget_bytes; fy_gpu_zarr_read_cached reads through it (zc NULL: uncached) */
typedef struct fy_gpu_zarr_cache fy_gpu_zarr_cache;
fy_gpu_zarr_cache *fy_gpu_zarr_cache_new(size_t budget_bytes);
void fy_gpu_zarr_cache_free(fy_gpu_zarr_cache *c);
int fy_gpu_zarr_read_cached(const fy_zarr *z, fy_gpu_zarr_cache *zc, long z0, long y0, long x0,
                            long dz, long dy, long dx, unsigned char *out);
/* fy_process_chunk (pipeline.c:518-538) on a local zarr: fy_process_buffer of
 * the inner tile^3 at (z0,y0,x0), clamped to the volume, with its halo
 * (cfg->halo, 8 if <= 0) read from zin.  Writes tz*ty*tx u8 to out and the
 * clamped sizes to out_t* (any may be NULL). */
int fy_gpu_process_chunk(const fy_zarr *zin, const fy_pipeline_cfg *cfg,
                         long z0, long y0, long x0, int tile,
                         unsigned char *out, int *out_tz, int *out_ty, int *out_tx, int *all_air);

/* ---- pass-2 export through the GPU: the tile loop of fy_run_pipeline
 * (pipeline.c:1000-1070) fed from zarr.  Tiles t in [t_begin, t_end) of the
 * pass-2 grid (t = (iz*nty + iy)*ntx + ix; t_end <= 0: all) are read with
 * their halo (cfg->halo, 8 if <= 0) from zin by `threads` host threads (<= 0:
 * one per core, at most 16) straight into pinned staging, copied to the device,
 * processed `batch` at a time (<= 0: 8) by fy_gpu_process_cubes on `stream`,
 * copied back and written as chunk (iz, iy, ix) of zout, which must have zin's
 * shape and tile^3 chunks.  Six batches are in flight, so reads, both PCIe
 * directions, kernels and writes overlap.  All-air tiles are skipped (their
 * chunks stay absent and read as 0), as in the reference; failed reads and
 * writes are counted and make the call return 1.  Output chunks are
 * bit-identical to fy_gpu_process_chunk's.  No reference counterpart (the
 * reference runs this loop on host threads). */
typedef struct {
    long tiles;                     /* tiles in the range */
    long written;                   /* chunks written */
    long air;                       /* all-air tiles skipped */
    long io_fail;                   /* failed reads or writes */
    double wall_s;                  /* wall time of the call */
    double read_s, write_s;         /* thread time spent reading / writing (summed) */
} fy_export_stats;
int fy_gpu_export(const fy_zarr *zin, const fy_zarr *zout, const fy_pipeline_cfg *cfg, int tile,
                  long t_begin, long t_end, int threads, int batch, fy_export_stats *stats,
                  void *stream);

/* the tile-shape autotuner of the fast guided filter (the coefficient march's z
 * chunk, the upsample's planes per workgroup; every candidate gives
 * bit-identical results).  mode 0: first candidates; 1: each new shape runs all
 * candidates once, timed, and keeps the fastest (default); 2 + k: candidate k.
 * Returns the previous mode.  The report lists each tuned shape with its
 * candidates' times; returns its length.  No reference counterpart. */
int fy_gpu_tuning(int mode);
int fy_gpu_tuning_report(char *buf, int len);
/* persist the tuned shapes: save writes every shape tuned or loaded so far to
 * a text file (one "key<TAB>coef<TAB>upsample" line each); load adds a file's
 * shapes, which later calls then use without timing; clear forgets them all.
 * Return 0 on success, 1 if the file cannot be written / read or has a bad
 * line (the good lines are still loaded). */
int fy_gpu_tuning_save(const char *path);
int fy_gpu_tuning_load(const char *path);
void fy_gpu_tuning_clear(void);

int fy_gpu_sync(void *stream);

/* forget the library's state for `stream` (NULL: the calling thread's own
 * stream, which is then destroyed): waits for the stream, then frees its
 * device workspaces, pinned staging, table slots and helper streams (each
 * helper holds a hardware queue of its own).  The next call on the stream
 * starts afresh.  No call on the stream may run meanwhile on another thread,
 * and a server or calibration sweep created on it must be destroyed first.
 * Returns 0.  No
 * reference counterpart: the reference's thread-local scratch is never freed
 * (fft.c:45,109-112; guided.c:49,83). */
int fy_gpu_stream_release(void *stream);

/* pass structure of the deconvolution for cubes whose padded X and Y axes
 * are one length of 64 or 96 (fp32): mode 1 (default) transforms whole z
 * planes along X and Y in one kernel each way, so the half spectrum makes one
 * HBM round trip fewer per direction; mode 0 runs the separate X and Y passes.
 * Both compute the same transforms; the fp32 rounding may differ in the last
 * bits (relative L2 ~1e-7 between them).  Returns the previous mode.  No
 * reference counterpart. */
int fy_gpu_fft_planes(int mode);

/* radix plans of the fp32 deconvolution passes.  Padded lengths 64..512 have
 * two plans: radix-16 stages behind one smaller radix (plan 0), or radix-8
 * stages behind one radix 2 or 4 (plan 1).  Each pass kind (X, Y, Z, the fused
 * XY planes) and length takes one, decided once per process: mode 0 plan 0
 * everywhere, 2 plan 1 everywhere, 1 (default) the tuner's choice -- the first
 * deconvolution that meets the length times both plans of the pass on a
 * synthetic problem of about 16 M voxels (a one-time ~200 MB allocation and a
 * few ms; never under graph capture or in fy_gpu_deconvolve_noalloc, which
 * take plan 0 for lengths not decided yet) and keeps plan 1 only when it is at
 * least 5 % faster.  The plans round differently (relative L2 ~1e-7 between
 * them), so a length keeps its plan for the process; the choices are saved and
 * loaded with the tile shapes (fy_gpu_tuning_save / _load: "fft<TAB>kind<TAB>
 * length<TAB>plan" lines) and listed by fy_gpu_tuning_report, and
 * fy_gpu_tuning_clear forgets them.  Returns the previous mode.  No reference
 * counterpart. */
int fy_gpu_fft_radix(int mode);

/* the process-wide cache of device tables (transfer tables of captured and
 * large deconvolutions, ring and z-drift blobs): stats (may be NULL) receives
 * 7 counts -- entries, evictable entries, entries held by queued work or a
 * live graph, evicted entries whose memory waits for their last readers,
 * evictions and frees so far, and the limit on evictable entries.  limit > 0
 * sets that limit (default 512).  Returns the limit before the call.  No
 * reference counterpart. */
long fy_gpu_table_cache(long *stats, long limit);

/* per-kernel timing: HIP events recorded around every launch on its own stream */
void fy_gpu_profile(int on);
/* names: '\0'-separated kernel names; ms/count per name; returns #entries */
int  fy_gpu_profile_read(char *names, int names_len, double *ms, long *count, int max);
void fy_gpu_profile_reset(void);
const char *fy_gpu_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
