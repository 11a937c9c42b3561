// fyg_calib.hip -- pass-1 calibration statistics on device tiles: the u8
// histogram (fy_hist_accumulate_u8, stream.c:76-86), the per-slice z-drift
// sums of the pipeline's 4x4 subsample (pipeline.c:686-701) and of
// fy_zdrift_accumulate (zdrift.c:23-35), the flat-block stds of fy_flat_noise
// (stream.c:174-200) and the global normalisation apply (stream.c:122-129).
//
// All of it is HBM-bound integer / byte work.  Counts are integers and the
// u8 z-drift sums are kept as integers in units of 2^-27 (every v = u8/255 above
// 0.1 is a multiple of 2^-27), so those results are exact and independent of
// the order in which tiles, workgroups and lanes add them.
#include "fyg_internal.hpp"

namespace {

constexpr int NT = 256;
// sub-histogram copies per workgroup: bin b of copy c lives at LDS word b*HC + c,
// i.e. in bank c, so the lanes of a half-wave (distinct c) never share a bank
constexpr int HC = 32;
constexpr float kZUnit = 134217728.0f;   // 2^27

__device__ __forceinline__ void hist_word(unsigned int *h, unsigned int w, int c) {
    atomicAdd(&h[(w & 255u) * HC + c], 1u);
    atomicAdd(&h[((w >> 8) & 255u) * HC + c], 1u);
    atomicAdd(&h[((w >> 16) & 255u) * HC + c], 1u);
    atomicAdd(&h[(w >> 24) * HC + c], 1u);
}

// grid-stride over the 16-byte-aligned body (uint4 loads); the unaligned head
// and the tail (< 32 bytes together) are taken by the first threads of the grid.
// Each workgroup writes its 256 counts to part[blockIdx.x]: no global atomics,
// which would serialise on the few bins that hold most voxels.
__global__ __launch_bounds__(NT) void k_hist_u8(const unsigned char *d, size_t n, size_t head,
                                                 unsigned int *part) {
    __shared__ unsigned int h[256 * HC];
    for (int i = threadIdx.x; i < 256 * HC; i += NT) h[i] = 0;
    __syncthreads();
    const int c = threadIdx.x % HC;
    const size_t nv = (n - head) / 16;
    const uint4 *v = (const uint4 *)(d + head);
    const size_t t0 = (size_t)blockIdx.x * NT + threadIdx.x, stride = (size_t)gridDim.x * NT;
    size_t i = t0;
    for (; i + 3 * stride < nv; i += 4 * stride) {       // four loads in flight
        uint4 w[4];
#pragma unroll
        for (int k = 0; k < 4; k++) w[k] = v[i + k * stride];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            hist_word(h, w[k].x, c);
            hist_word(h, w[k].y, c);
            hist_word(h, w[k].z, c);
            hist_word(h, w[k].w, c);
        }
    }
    for (; i < nv; i += stride) {
        uint4 w = v[i];
        hist_word(h, w.x, c);
        hist_word(h, w.y, c);
        hist_word(h, w.z, c);
        hist_word(h, w.w, c);
    }
    const size_t tail0 = head + nv * 16, ne = head + (n - tail0);
    if (t0 < ne) {
        size_t j = t0 < head ? t0 : tail0 + (t0 - head);
        atomicAdd(&h[d[j] * HC + c], 1u);
    }
    __syncthreads();
    // thread b sums the HC copies of bin b, starting at copy b % HC so that the
    // lanes of a wave read distinct banks
    const int b = threadIdx.x;
    unsigned int t = 0;
#pragma unroll 8
    for (int k = 0; k < HC; k++) t += h[b * HC + ((k + b) & (HC - 1))];
    part[(size_t)blockIdx.x * 256 + b] = t;
}

// hist[b] += sum over the rows of part: workgroup r takes rows r, r+R, ...
__global__ __launch_bounds__(NT) void k_hist_reduce(const unsigned int *part, int rows,
                                                     unsigned long long *hist) {
    const int b = threadIdx.x;
    unsigned long long t = 0;
    for (int r = blockIdx.x; r < rows; r += gridDim.x) t += part[(size_t)r * 256 + b];
    if (t) atomicAdd(&hist[b], t);
}

// acc[0..255] += tile[0..255], acc[256] += n, unless skip_zero and the tile is
// all zero (tile[0] == n, including n == 0: pipeline.c:677-679); tile is
// cleared for the next tile either way
__global__ void k_hist_merge(unsigned long long *tile, unsigned long long n, unsigned long long *acc,
                             int skip_zero) {
    const int b = threadIdx.x;
    const unsigned long long t = tile[b];
    const bool merge = !(skip_zero && tile[0] == n);
    __syncthreads();
    tile[b] = 0;
    if (merge) {
        acc[b] += t;
        if (b == 0) acc[256] += n;
    }
}

template <class T>
__device__ __forceinline__ T block_sum(T v, T *red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    T t = red[0];
    for (int k = 1; k < NT / 64; k++) t += red[k];
    return t;
}

// grid (row blocks, slices): sums of v = u8*(1/255) > 0.10 over (y, x) =
// multiples of ss, in units of 2^-27, and their count, added to zs[z], zc[z].
// Integer atomics: exact in any order.
constexpr int ZR = 8;                    // subsampled rows per workgroup
__global__ __launch_bounds__(NT) void k_zdrift_u8(const unsigned char *d, long dy, long dx, int ss,
                                                   unsigned long long *zs, unsigned long long *zc) {
    __shared__ unsigned long long red[NT / 64];
    const unsigned char *pl = d + (size_t)blockIdx.y * dy * dx;
    const int sy = (int)((dy + ss - 1) / ss), sx = (int)((dx + ss - 1) / ss);
    const int y0 = blockIdx.x * ZR, ny = min(ZR, sy - y0), np = ny * sx;
    unsigned long long s = 0, c = 0;
    for (int i = threadIdx.x; i < np; i += NT) {
        const int yi = y0 + i / sx, xi = i % sx;
        const float v = (float)pl[(size_t)yi * ss * dx + (size_t)xi * ss] * (1.0f / 255.0f);
        const bool ok = v > 0.10f;
        s += ok ? (unsigned long long)(v * kZUnit) : 0ull;
        c += ok;
    }
    const unsigned long long S = block_sum(s, red);
    __syncthreads();
    const unsigned long long Cn = block_sum(c, red);
    if (threadIdx.x == 0 && Cn) {
        atomicAdd(&zs[blockIdx.y], S);
        atomicAdd(&zc[blockIdx.y], Cn);
    }
}

// one workgroup per slice: double sum and count of the values above thresh
__global__ __launch_bounds__(NT) void k_zdrift_f32(const float *d, size_t plane, float thresh, double *sums,
                                                    long long *cnts) {
    __shared__ double redd[NT / 64];
    __shared__ long long redc[NT / 64];
    const float *pl = d + (size_t)blockIdx.x * plane;
    double s = 0.0;
    long long c = 0;
    for (size_t i = threadIdx.x; i < plane; i += NT) {
        const float v = pl[i];
        const bool ok = v > thresh;
        s += ok ? (double)v : 0.0;
        c += ok;
    }
    const double S = block_sum(s, redd);
    const long long Cn = block_sum(c, redc);
    if (threadIdx.x == 0) {
        sums[blockIdx.x] += S;
        cnts[blockIdx.x] += Cn;
    }
}

// one wave per blk^3 block (x fastest over the block grid): std of the block
// when its mean exceeds 0.1, else -1.  Squares are rounded to fp32 before they
// are summed in double, as the reference's `s2 += val*val` does.
__global__ __launch_bounds__(NT) void k_flat_blocks(const float *v, int ny, int nx, int blk, int nby, int nbx,
                                                     long nb, double *std_out) {
    #pragma clang fp contract(off)
    const long b = (long)blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (b >= nb) return;
    const int bx = (int)(b % nbx), by = (int)((b / nbx) % nby), bz = (int)(b / ((long)nbx * nby));
    const int n2 = blk * blk, n3 = n2 * blk;
    double s = 0.0, s2 = 0.0;
    for (int i = lane; i < n3; i += 64) {
        const int dz = i / n2, r = i - dz * n2, dy = r / blk, dx = r - dy * blk;
        const float val = v[((size_t)(bz * blk + dz) * ny + (by * blk + dy)) * nx + (bx * blk + dx)];
        const float sq = val * val;
        s += (double)val;
        s2 += (double)sq;
    }
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_xor(s, o);
        s2 += __shfl_xor(s2, o);
    }
    if (lane == 0) {
        const double mean = s / n3, var = s2 / n3 - mean * mean;
        std_out[b] = mean > 0.1 ? sqrt(var > 0.0 ? var : 0.0) : -1.0;
    }
}

__global__ void k_norm_u8(const unsigned char *in, float *out, size_t n, float flo, float inv) {
    #pragma clang fp contract(off)
    for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT) {
        const float v = ((float)in[i] - flo) * inv;
        out[i] = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
    }
}

unsigned grid_for(size_t work, unsigned cap) {
    size_t b = (work + NT - 1) / NT;
    return (unsigned)(b < 1 ? 1 : (b > cap ? cap : b));
}

}  // namespace

extern "C" size_t fyg_hist_part_bytes(size_t n) {
    (void)n;
    return sizeof(unsigned int) * 256 * 4096;
}

extern "C" int fyg_hist_u8_run(const unsigned char *d, size_t n, unsigned int *part, unsigned long long *hist,
                               fyg_stream *s) {
    if (!n) return 0;
    size_t head = (16 - ((uintptr_t)d & 15)) & 15;
    if (head > n) head = n;
    const size_t nv = (n - head) / 16, extra = n - nv * 16;
    // about 8 uint4 per thread (four 32 KB workgroups fit a CU), at most 4096
    // workgroups (the rows of part); enough threads for the head and tail bytes
    size_t want = nv / 8 > extra ? nv / 8 : extra;
    const unsigned g = grid_for(want, 4096);
    fyg_launch(s, "hist_u8", k_hist_u8, dim3(g), dim3(NT), 0, d, n, head, part);
    const unsigned r = g < 16 ? g : 16;
    fyg_launch(s, "hist_reduce", k_hist_reduce, dim3(r), dim3(NT), 0, (const unsigned int *)part, (int)g, hist);
    return fyg_last_error();
}

extern "C" int fyg_hist_reduce_run(const unsigned int *part, int rows, unsigned long long *hist, fyg_stream *s) {
    if (rows <= 0) return 0;
    const unsigned r = rows < 16 ? rows : 16;
    fyg_launch(s, "hist_reduce", k_hist_reduce, dim3(r), dim3(NT), 0, part, rows, hist);
    return fyg_last_error();
}

extern "C" int fyg_hist_merge_run(unsigned long long *tile, unsigned long long n, unsigned long long *acc,
                                  int skip_zero, fyg_stream *s) {
    fyg_launch(s, "hist_merge", k_hist_merge, dim3(1), dim3(256), 0, tile, n, acc, skip_zero);
    return fyg_last_error();
}

extern "C" int fyg_zdrift_u8_run(const unsigned char *d, long dz, long dy, long dx, int ss, unsigned long long *zs,
                                 unsigned long long *zc, fyg_stream *s) {
    if (dz <= 0 || dy <= 0 || dx <= 0) return 0;
    if (ss < 1) ss = 1;
    const long sy = (dy + ss - 1) / ss, sx = (dx + ss - 1) / ss;
    if (sx * ZR > 0x7fffffffL || dz > 65535) return 1;
    fyg_launch(s, "zdrift_u8", k_zdrift_u8, dim3((unsigned)((sy + ZR - 1) / ZR), (unsigned)dz), dim3(NT), 0, d, dy,
               dx, ss, zs, zc);
    return fyg_last_error();
}

extern "C" int fyg_zdrift_f32_run(const float *d, int nz, int ny, int nx, float thresh, double *sums,
                                  long long *cnts, fyg_stream *s) {
    if (nz <= 0 || ny <= 0 || nx <= 0) return 0;
    fyg_launch(s, "zdrift_f32", k_zdrift_f32, dim3((unsigned)nz), dim3(NT), 0, d, (size_t)ny * nx, thresh, sums,
               cnts);
    return fyg_last_error();
}

extern "C" int fyg_flat_blocks_run(const float *v, int nz, int ny, int nx, int blk, double *std_out,
                                   fyg_stream *s) {
    const long nbz = nz / blk, nby = ny / blk, nbx = nx / blk, nb = nbz * nby * nbx;
    if (nb <= 0) return 0;
    const long g = (nb + NT / 64 - 1) / (NT / 64);
    if (g > 0x7fffffffL) return 1;
    fyg_launch(s, "flat_blocks", k_flat_blocks, dim3((unsigned)g), dim3(NT), 0, v, ny, nx, blk, (int)nby,
               (int)nbx, nb, std_out);
    return fyg_last_error();
}

extern "C" int fyg_norm_u8_run(const unsigned char *in, float *out, size_t n, unsigned char lo, unsigned char hi,
                               fyg_stream *s) {
    if (!n) return 0;
    const float flo = lo, inv = hi > lo ? 1.0f / (float)(hi - lo) : 1.0f;
    fyg_launch(s, "norm_u8", k_norm_u8, dim3(grid_for(n, 16384)), dim3(NT), 0, in, out, n, flo, inv);
    return fyg_last_error();
}
