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

__global__ void k_prologue(const unsigned char *in, float *out, int nz, int ny, int nx, fyg_prologue P,
                           int *flag, int gen) {
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
    if (flag && __syncthreads_or(any) && threadIdx.x == 0) *flag = gen;
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

}  // namespace

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
    fyg_launch(s, "dering_apply", k_dering_apply, dim3(grid1d(n)), dim3(NT), 0, f, dz, dy, dx, *pro);
    return fyg_last_error();
}

extern "C" int fyg_prologue_run(const unsigned char *in, float *out, int nz, int ny, int nx,
                                const fyg_prologue *pro, int *flag, int gen, fyg_stream *s) {
    size_t n = (size_t)nz * ny * nx;
    if (!n) return 0;
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
