// fyg_elem.hip -- HBM-bound elementwise and reduction kernels of the tile chain:
// u8 <-> float / physical window (stream.c:27-57), z-drift apply (zdrift.c:64-71),
// ring subtraction (dering.c:200-248), ring (sector, radius) accumulation
// (dering.c:78-116) and the prologue-only path of fy_process_buffer when the
// deconvolution stage is off (pipeline.c:492-494).
#include "fyg_internal.hpp"

namespace {

constexpr int NT = 256;

unsigned grid1d(size_t n) {
    size_t b = (n + NT - 1) / NT;
    return (unsigned)(b < 16384 ? (b ? b : 1) : 16384);
}

__global__ void k_u8_to_f32(const unsigned char *in, float *out, size_t n, float lo, float scale) {
    #pragma clang fp contract(off)
    for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT)
        out[i] = lo + (float)in[i] * scale;   // two roundings: contraction is off here
}

// (in - lo)*k + 0.5 clamped to [0,255] and truncated (stream.c:31-36,48-57)
__global__ void k_f32_to_u8(const float *in, unsigned char *out, size_t n, float lo, float k) {
    #pragma clang fp contract(off)
    for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT) {
        float v = (in[i] - lo) * k + 0.5f;
        out[i] = v < 0.0f ? 0 : (v > 255.0f ? 255 : (unsigned char)v);
    }
}

__global__ void k_zdrift(float *f, size_t plane, size_t n, const float *factor) {
    for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT)
        f[i] *= factor[i / plane];
}

__device__ __forceinline__ float ring_sub(const fyg_prologue &P, float f, long gz, long gy, long gx) {
    #pragma clang fp contract(off)
    if (f <= 0.0f) return f;
    int ri;
    if (P.ring_fast) {
        float fy = (float)gy - (float)P.cy, fx = (float)gx - (float)P.cx;
        float fy2 = fy * fy;
        ri = (int)(sqrtf(fy2 + fx * fx) + 0.5f);
    } else {
        double dy = (double)gy - P.cy, dx = (double)gx - P.cx;
        double dy2 = dy * dy;
        ri = (int)(sqrt(dy2 + dx * dx) + 0.5);
    }
    if (ri >= P.nr) return f;
    long slab = gz / P.slab_z;
    if (slab < 0) slab = 0;
    if (slab >= P.nslab) slab = P.nslab - 1;
    float v = f - P.scale * P.ring[slab * P.nr + ri];
    return v > 0.0f ? v : 0.0f;
}

__global__ void k_dering_apply(float *f, long dz, long dy, long dx, fyg_prologue P) {
    size_t n = (size_t)dz * dy * dx;
    for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT) {
        long x = (long)(i % dx);
        size_t t = i / dx;
        long y = (long)(t % dy), z = (long)(t / dy);
        f[i] = ring_sub(P, f[i], P.gz0 + z, P.gy0 + y, P.gx0 + x);
    }
}

// the same two kernels on (row z*ny + y, x block) grids: a voxel's z and y (and
// its ring slab) are per-workgroup values instead of 64-bit divisions per voxel
__global__ __launch_bounds__(NT) void k_dering_apply_rows(float *f, int dy, long dx, fyg_prologue P) {
    const int row = blockIdx.x, z = row / dy, y = row - z * dy;
    const long x = (long)blockIdx.y * NT + threadIdx.x;
    if (x >= dx) return;
    const size_t i = (size_t)row * dx + x;
    f[i] = ring_sub(P, f[i], P.gz0 + z, P.gy0 + y, P.gx0 + x);
}

// ring_sub with the row's terms computed once: the y distance (squared, in the
// precision ring_sub uses) and the slab's ring row.  Coordinates are converted
// from int: the same values as from long for |c| < 2^31.
struct RingRow {
    const float *ring;                   // the slab's ring values (NULL: no ring removal)
    float fy2, cxf;                      // fast path (float radius)
    double dy2, cxd;                     // double radius
};
__device__ __forceinline__ RingRow ring_row(const fyg_prologue &P, int gz, int gy) {
    RingRow R;
    R.ring = nullptr;
    if (!P.dering) return R;
    int slab = gz / P.slab_z;
    slab = slab < 0 ? 0 : (slab >= P.nslab ? P.nslab - 1 : slab);
    R.ring = P.ring + (size_t)slab * P.nr;
    const float fy = (float)gy - (float)P.cy;
    R.fy2 = fy * fy;
    R.cxf = (float)P.cx;
    const double dy = (double)gy - P.cy;
    R.dy2 = dy * dy;
    R.cxd = P.cx;
    return R;
}
__device__ __forceinline__ float ring_sub_row(const fyg_prologue &P, const RingRow &R, float f, int gx) {
    #pragma clang fp contract(off)
    if (f <= 0.0f) return f;
    int ri;
    if (P.ring_fast) {
        const float fx = (float)gx - R.cxf;
        ri = (int)(sqrtf(R.fy2 + fx * fx) + 0.5f);
    } else {
        const double dx = (double)gx - R.cxd;
        ri = (int)(sqrt(R.dy2 + dx * dx) + 0.5);
    }
    if (ri >= P.nr) return f;
    const float v = f - P.scale * R.ring[ri];
    return v > 0.0f ? v : 0.0f;
}

// one row per wave (z*ny + y, NT/64 rows per workgroup), four consecutive
// voxels per lane and step (u8 word loads and float4 stores when the rows allow
// them: vec)
__global__ __launch_bounds__(NT) void k_prologue_rows(const unsigned char *in, float *out, int nrows, int ny,
                                                      int nx, fyg_prologue P, int *flag, const int *gen, int vec) {
    #pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
    const int z = row / ny, y = row - z * ny;
    int any = 0;
    for (int x0 = 4 * lane; row < nrows && x0 < nx; x0 += 256) {
        const size_t i0 = (size_t)row * nx + x0;
        const int gz = (int)(P.gz0 + z), gy = (int)(P.gy0 + y), gx0 = (int)(P.gx0 + x0);
        const RingRow R = ring_row(P, gz, gy);
        const float zf = P.zdrift ? P.zfactor[gz] : 1.0f;
        const bool full = vec && x0 + 3 < nx;
        unsigned char u[4];
        if (full) {
            const uchar4 w = *(const uchar4 *)(in + i0);
            u[0] = w.x; u[1] = w.y; u[2] = w.z; u[3] = w.w;
        } else {
            for (int e = 0; e < 4; e++) u[e] = x0 + e < nx ? in[i0 + e] : 0;
        }
        float o[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            any |= u[e];
            float f = (float)u[e] * (1.0f / 255.0f);
            if (P.dering) f = ring_sub_row(P, R, f, gx0 + e);
            if (P.zdrift) f *= zf;
            if (P.window) f = P.wlo + f * P.wspan;
            o[e] = f;
        }
        if (full) {
            *(float4 *)(out + i0) = make_float4(o[0], o[1], o[2], o[3]);
        } else {
            for (int e = 0; e < 4; e++)
                if (x0 + e < nx) out[i0 + e] = o[e];
        }
    }
    if (flag && __syncthreads_or(any) && threadIdx.x == 0) fyg_flag_set(flag, gen);
}

__global__ void k_prologue(const unsigned char *in, float *out, int nz, int ny, int nx, fyg_prologue P,
                           int *flag, const int *gen) {
    #pragma clang fp contract(off)
    size_t n = (size_t)nz * ny * nx;
    int any = 0;
    for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT) {
        long x = (long)(i % nx);
        size_t t = i / nx;
        long y = (long)(t % ny), z = (long)(t / ny);
        unsigned char u = in[i];
        any |= u;
        float f = (float)u * (1.0f / 255.0f);
        if (P.dering) f = ring_sub(P, f, P.gz0 + z, P.gy0 + y, P.gx0 + x);
        if (P.zdrift) f *= P.zfactor[P.gz0 + z];
        if (P.window) f = P.wlo + f * P.wspan;
        out[i] = f;
    }
    if (flag && __syncthreads_or(any) && threadIdx.x == 0) fyg_flag_set(flag, gen);
}

// Sector/radius histogram of nonzero u8 samples on the ss-subsampled (y,x) grid.
// Bins are privatised per workgroup in LDS (integer adds, so the totals are exact
// and independent of order), then merged with one global atomic per touched bin.
__global__ void k_dering_accum(const unsigned char *buf, long dz, long dy, long dx, int ss, long sy, long sx,
                               const int *map, int nbins, unsigned long long *gsum, unsigned int *gcnt) {
    extern __shared__ unsigned int lds[];
    unsigned int *lsum = lds, *lcnt = lds + nbins;
    for (int i = threadIdx.x; i < 2 * nbins; i += NT) lds[i] = 0u;
    __syncthreads();
    size_t n = (size_t)dz * sy * sx;
    for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT) {
        long xi = (long)(i % sx);
        size_t t = i / sx;
        long yi = (long)(t % sy), z = (long)(t / sy);
        unsigned char v = buf[((size_t)z * dy + (size_t)yi * ss) * dx + (size_t)xi * ss];
        if (!v) continue;
        int b = map[yi * sx + xi];
        if (b < 0) continue;
        atomicAdd(&lsum[b], (unsigned int)v);
        atomicAdd(&lcnt[b], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nbins; i += NT) {
        if (lcnt[i]) {
            atomicAdd(&gsum[i], (unsigned long long)lsum[i]);
            atomicAdd(&gcnt[i], lcnt[i]);
        }
    }
}

// wide bin ranges (large slices): same histogram with global atomics only
__global__ void k_dering_accum_global(const unsigned char *buf, long dz, long dy, long dx, int ss, long sy,
                                      long sx, const int *map, unsigned long long *gsum, unsigned int *gcnt) {
    size_t n = (size_t)dz * sy * sx;
    for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT) {
        long xi = (long)(i % sx);
        size_t t = i / sx;
        long yi = (long)(t % sy), z = (long)(t / sy);
        unsigned char v = buf[((size_t)z * dy + (size_t)yi * ss) * dx + (size_t)xi * ss];
        if (!v) continue;
        int b = map[yi * sx + xi];
        if (b < 0) continue;
        atomicAdd(&gsum[b], (unsigned long long)v);
        atomicAdd(&gcnt[b], 1u);
    }
}

// ---- 2x LOD downscale (downscale.c:15-45): one output voxel from its 2x2x2
// parent cell; an all-zero cell stays zero
__device__ __forceinline__ unsigned int ds_cell(const unsigned int c[8], int method, float alpha) {
    #pragma clang fp contract(off)
    unsigned int acc = c[0] | c[1] | c[2] | c[3] | c[4] | c[5] | c[6] | c[7];
    if (!acc) return 0;
    const float mean = (float)(int)(c[0] + c[1] + c[2] + c[3] + c[4] + c[5] + c[6] + c[7]) * 0.125f;
    float v = mean;
    if (method != 0) {
        float best = mean, bestdev = -1.0f;
        for (int i = 0; i < 8; i++) {
            const float d = (float)c[i] - mean, ad = d < 0 ? -d : d;
            if (ad > bestdev) {
                bestdev = ad;
                best = (float)c[i];
            }
        }
        v = mean + alpha * (best - mean);
    }
    return (unsigned int)(unsigned char)(v < 0 ? 0.0f : v > 255 ? 255.0f : (v + 0.5f));
}

// grid (x groups, output rows): a thread makes 4 consecutive outputs of one row
// from two planes x two rows x 8 bytes (one 8-byte load each where the rows are
// 8-byte aligned and the cells lie inside the volume)
__global__ void k_downscale2x(const unsigned char *in, int nx, int ny, int nz, unsigned char *out, int ox, int oy,
                              int method, float alpha, int vec) {
    const int row = blockIdx.y, z = row / oy, y = row - z * oy;
    const int x0 = (blockIdx.x * NT + threadIdx.x) * 4;
    if (x0 >= ox) return;
    const int za = 2 * z, zb = 2 * z + 1 < nz ? 2 * z + 1 : nz - 1;
    const int ya = 2 * y, yb = 2 * y + 1 < ny ? 2 * y + 1 : ny - 1;
    const unsigned char *r[4] = {in + ((size_t)za * ny + ya) * nx, in + ((size_t)za * ny + yb) * nx,
                                 in + ((size_t)zb * ny + ya) * nx, in + ((size_t)zb * ny + yb) * nx};
    unsigned int q[4];
    if (vec && 2 * x0 + 8 <= nx) {
        unsigned int w[4][2];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            uint2 t = *(const uint2 *)(r[k] + 2 * x0);
            w[k][0] = t.x;
            w[k][1] = t.y;
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int sh = (j & 1) * 16, h = j >> 1;
            // cell order: (dz, dy, dx) with x fastest
            unsigned int c[8] = {(w[0][h] >> sh) & 255, (w[0][h] >> (sh + 8)) & 255,
                                 (w[1][h] >> sh) & 255, (w[1][h] >> (sh + 8)) & 255,
                                 (w[2][h] >> sh) & 255, (w[2][h] >> (sh + 8)) & 255,
                                 (w[3][h] >> sh) & 255, (w[3][h] >> (sh + 8)) & 255};
            q[j] = ds_cell(c, method, alpha);
        }
    } else {
        for (int j = 0; j < 4; j++) {
            const int x = x0 + j;
            if (x >= ox) { q[j] = 0; continue; }
            const int xa = 2 * x, xb = 2 * x + 1 < nx ? 2 * x + 1 : nx - 1;
            unsigned int c[8] = {r[0][xa], r[0][xb], r[1][xa], r[1][xb], r[2][xa], r[2][xb], r[3][xa], r[3][xb]};
            q[j] = ds_cell(c, method, alpha);
        }
    }
    unsigned char *o = out + ((size_t)z * oy + y) * ox + x0;
    if (vec && x0 + 4 <= ox) {
        *(unsigned int *)o = q[0] | (q[1] << 8) | (q[2] << 16) | (q[3] << 24);
    } else {
        for (int j = 0; j < 4 && x0 + j < ox; j++) o[j] = (unsigned char)q[j];
    }
}

}  // namespace

extern "C" int fyg_downscale2x_run(const unsigned char *in, int nx, int ny, int nz, unsigned char *out, int method,
                                   float alpha, fyg_stream *s) {
    if (nx < 1 || ny < 1 || nz < 1) return 1;
    const int ox = (nx + 1) / 2, oy = (ny + 1) / 2, oz = (nz + 1) / 2;
    if ((long)oz * oy > 65535L * 32768L) return 1;
    // 8-byte row loads need 8-aligned rows; 4-byte stores 4-aligned output rows
    const int vec = nx % 8 == 0 && ((uintptr_t)in & 7) == 0 && ox % 4 == 0 && ((uintptr_t)out & 3) == 0;
    const unsigned gx = (unsigned)(((ox + 3) / 4 + NT - 1) / NT);
    if ((long)oz * oy > 65535) {
        // rows beyond the grid's y limit: one launch per slab of output planes
        const int per = 65535 / oy;
        for (int z0 = 0; z0 < oz; z0 += per) {
            const int nzs = oz - z0 < per ? oz - z0 : per;
            const int nin = 2 * nzs < nz - 2 * z0 ? 2 * nzs : nz - 2 * z0;
            fyg_launch(s, "downscale2x", k_downscale2x, dim3(gx, (unsigned)(nzs * oy)), dim3(NT), 0,
                       in + (size_t)2 * z0 * ny * nx, nx, ny, nin, out + (size_t)z0 * oy * ox, ox, oy, method, alpha,
                       vec);
        }
    } else {
        fyg_launch(s, "downscale2x", k_downscale2x, dim3(gx, (unsigned)(oz * oy)), dim3(NT), 0, in, nx, ny, nz, out,
                   ox, oy, method, alpha, vec);
    }
    return fyg_last_error();
}

extern "C" int fyg_u8_to_f32_run(const unsigned char *in, float *out, size_t n, float lo, float scale,
                                 fyg_stream *s) {
    if (!n) return 0;
    fyg_launch(s, "u8_to_f32", k_u8_to_f32, dim3(grid1d(n)), dim3(NT), 0, in, out, n, lo, scale);
    return fyg_last_error();
}

extern "C" int fyg_f32_to_u8_run(const float *in, unsigned char *out, size_t n, float lo, float k,
                                 fyg_stream *s) {
    if (!n) return 0;
    fyg_launch(s, "f32_to_u8", k_f32_to_u8, dim3(grid1d(n)), dim3(NT), 0, in, out, n, lo, k);
    return fyg_last_error();
}

extern "C" int fyg_zdrift_run(float *f, int nz, int ny, int nx, const float *factor_at_z0, fyg_stream *s) {
    size_t plane = (size_t)ny * nx, n = plane * nz;
    if (!n) return 0;
    fyg_launch(s, "zdrift", k_zdrift, dim3(grid1d(n)), dim3(NT), 0, f, plane, n, factor_at_z0);
    return fyg_last_error();
}

extern "C" int fyg_dering_apply_run(float *f, long dz, long dy, long dx, const fyg_prologue *pro, fyg_stream *s) {
    size_t n = (size_t)dz * dy * dx;
    if (!n) return 0;
    const long rows = dz * dy, xb = (dx + NT - 1) / NT;
    if (rows <= 0x7fffffffL && xb <= 65535)
        fyg_launch(s, "dering_apply", k_dering_apply_rows, dim3((unsigned)rows, (unsigned)xb), dim3(NT), 0, f,
                   (int)dy, dx, *pro);
    else
        fyg_launch(s, "dering_apply", k_dering_apply, dim3(grid1d(n)), dim3(NT), 0, f, dz, dy, dx, *pro);
    return fyg_last_error();
}

extern "C" int fyg_prologue_run(const unsigned char *in, float *out, int nz, int ny, int nx,
                                const fyg_prologue *pro, int *flag, const int *gen, fyg_stream *s) {
    size_t n = (size_t)nz * ny * nx;
    if (!n) return 0;
    const long rows = (long)nz * ny;
    // global coordinates are int in the row kernel
    const bool small = pro->gz0 + nz < 0x7fffffffL && pro->gy0 + ny < 0x7fffffffL && pro->gx0 + nx < 0x7fffffffL &&
                       pro->gz0 >= 0 && pro->gy0 >= 0 && pro->gx0 >= 0;
    const int vec = nx % 4 == 0 && (uintptr_t)in % 4 == 0 && (uintptr_t)out % 16 == 0;
    if (rows <= 0x7fffffffL - NT && small)
        fyg_launch(s, "prologue", k_prologue_rows, dim3((unsigned)((rows + NT / 64 - 1) / (NT / 64))), dim3(NT), 0,
                   in, out, (int)rows, ny, nx, *pro, flag, gen, vec);
    else
        fyg_launch(s, "prologue", k_prologue, dim3(grid1d(n)), dim3(NT), 0, in, out, nz, ny, nx, *pro, flag, gen);
    return fyg_last_error();
}

extern "C" int fyg_dering_accum_run(const unsigned char *buf, long dz, long dy, long dx, int ss,
                                    const int *map, int nbins, unsigned long long *sum, unsigned int *cnt,
                                    fyg_stream *s) {
    if (ss < 1) ss = 1;
    long sy = (dy + ss - 1) / ss, sx = (dx + ss - 1) / ss;
    size_t n = (size_t)dz * sy * sx;
    if (!n || nbins <= 0) return 0;
    size_t lds = sizeof(unsigned int) * 2 * (size_t)nbins;
    unsigned g = grid1d(n);
    if (lds > 48 * 1024) {
        fyg_launch(s, "dering_accum", k_dering_accum_global, dim3(g), dim3(NT), 0, buf, dz, dy, dx, ss, sy, sx,
                   map, sum, cnt);
        return fyg_last_error();
    }
    if (g > 1024) g = 1024;
    fyg_launch(s, "dering_accum", k_dering_accum, dim3(g), dim3(NT), lds, buf, dz, dy, dx, ss, sy, sx, map,
               nbins, sum, cnt);
    return fyg_last_error();
}
