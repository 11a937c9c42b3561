/* fyh.h -- internal helpers shared by the C host files of libfysics_amd. */
#ifndef FYH_H
#define FYH_H

#include <stddef.h>

#include "../../../include/fysics.h"
#include "../../../include/fysics_gpu.h"
#include "../fyg.h"

#define FYH_LUT_N 4096          /* 4097-entry weight table (paganin.c:138) */
#define FYH_MAX_GAIN 8.0        /* Wiener gain cap (paganin.c:50) */

/* workspace slots of a stream */
enum { WS_IN = 0, WS_SPEC = 1, WS_TILE = 2, WS_GUIDED = 3, WS_U8 = 4, WS_FLAG = 5, WS_OUT = 6, WS_MISC = 7,
       WS_AIR = 8 };

/* device-resident transfer tables of one (weight, physics, padded shape) */
typedef struct {
    int pz, py, px, hx, hxp;
    const void *lut;                /* fp32 (or fp64 for the all-FP64 variant) */
    double us;
    const double *fz2, *fy2, *fx2;
} fyh_tables;

/* kind 0: Wiener weight of the recon transfer (cap 8); kind 1: matched PSF*unsharp */
int fyh_tables_get(int nz, int ny, int nx, const fy_physics *p, double param, int kind,
                   double cap, double fz2_scale, int f64, fyg_stream *s, fyh_tables *t);
double fyh_matched_fz2_scale(const fy_physics *p);
/* the arguments of one fyh_tables_get call */
typedef struct {
    int nz, ny, nx, kind, f64;
    const fy_physics *p;
    double param, cap, fz2_scale;
} fyh_table_req;
/* LUT weights of the not yet cached tables among req, built ahead on host
 * threads; fyh_tables_get_m takes them from there instead of evaluating */
typedef struct fyh_lut_memo fyh_lut_memo;
fyh_lut_memo *fyh_tables_prefill(const fyh_table_req *req, int n, fyg_stream *s);
void fyh_lut_memo_free(fyh_lut_memo *m);
int fyh_tables_get_m(const fyh_table_req *r, const fyh_lut_memo *m, fyg_stream *s, fyh_tables *t);
/* the physics fy_process_buffer deconvolves with (pipeline.c:89-104) */
void fyh_scaled_physics(const fy_pipeline_cfg *c, fy_physics *p);

/* whether ring radii of a region may be evaluated in fp32 (see fy_dering.c) */
int fyh_ring_fast(double cy, double cx, long y0, long x0, long dy, long dx);

/* device copy of a small host array, cached by content */
const void *fyh_const_blob(char tag, const void *data, size_t bytes);

/* run the whole tile chain of fy_process_buffer on device buffers; gen and
 * nflag come from fyg_flags_reserve and are stepped by the chain itself */
int fyh_chain(const fy_pipeline_cfg *cfg, const unsigned char *d_u8,
              long rz0, long ry0, long rx0, long hz, long hy, long hx,
              long z0, long y0, long x0, long tz, long ty, long tx,
              long vol_y, long vol_x, unsigned char *d_out, int *d_flag, int *gen, int nflag,
              fyg_stream *s);

/* fy_gpu_process_cubes on a stream object */
int fyh_process_cubes(const fy_cube *cubes, int n, fyg_stream *s);
/* the same with every launch group on s itself (no helper streams) */
int fyh_process_cubes_serial(const fy_cube *cubes, int n, fyg_stream *s);

/* deconvolution on device buffers (fp32 in/out) */
int fyh_deconv_dev(const float *d_in, float *d_out, int nz, int ny, int nx, const fy_physics *p,
                   double param, int kind, double cap, double fz2_scale, fyg_stream *s);

#endif
