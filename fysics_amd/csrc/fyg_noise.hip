// fyg_noise.hip -- device side of the per-volume noise model (noise.c:20-192)
// and the noise-autocorrelation anisotropy (noise.c:210-269).
//
// Local statistics: the reference forms (2r+1)^3 box SUMS with double sliding
// windows along x, then y, then z.  Here one thread owns one line per pass and
// slides in the same order, so every sum -- and hence every local mean and
// variance -- is the same double the reference computes; no float
// reassociation enters.  The percentiles the fit needs (1%/99% of the local
// means, a low percentile of the local variance per intensity bin) are exact
// order statistics: one radix sort of (mean, variance) pairs by mean makes
// every intensity bin a contiguous segment, and a segmented radix sort orders
// the variances inside each segment.  The robust line fit itself (<= 24
// points, fp64) runs on the host.
#include <hipcub/hipcub.hpp>

#include "fyg_internal.hpp"

namespace {

constexpr int NT = 256;

unsigned grid1d(size_t n) {
    size_t b = (n + NT - 1) / NT;
    return (unsigned)(b ? b : 1);
}

// x pass: in (float) -> s1 = sum of in, s2 = sum of in*in (float product), per row
__global__ void k_box_x(const float *in, double *s1, double *s2, long nrows, int nx, int r) {
    #pragma clang fp contract(off)
    long row = (long)blockIdx.x * NT + threadIdx.x;
    if (row >= nrows) return;
    const float *a = in + (size_t)row * nx;
    double *o1 = s1 + (size_t)row * nx, *o2 = s2 + (size_t)row * nx;
    double u = 0, v = 0;
    for (int x = 0; x <= r && x < nx; x++) {
        u += a[x];
        v += (double)(a[x] * a[x]);
    }
    for (int x = 0; x < nx; x++) {
        o1[x] = u;
        o2[x] = v;
        int add = x + r + 1, sub = x - r;
        if (add < nx) { u += a[add]; v += (double)(a[add] * a[add]); }
        if (sub >= 0) { u -= a[sub]; v -= (double)(a[sub] * a[sub]); }
    }
}

// y or z pass over two double fields: lines of length len with element stride st;
// line index l maps to base offset (l / inner)*outer_stride + (l % inner)
__global__ void k_box_line(const double *i1, const double *i2, double *o1, double *o2, long nlines, long inner,
                           long outer_stride, int len, long st, int r) {
    #pragma clang fp contract(off)
    long l = (long)blockIdx.x * NT + threadIdx.x;
    if (l >= nlines) return;
    size_t b = (size_t)(l / inner) * outer_stride + (size_t)(l % inner);
    double u = 0, v = 0;
    for (int i = 0; i <= r && i < len; i++) {
        u += i1[b + (size_t)i * st];
        v += i2[b + (size_t)i * st];
    }
    for (int i = 0; i < len; i++) {
        o1[b + (size_t)i * st] = u;
        o2[b + (size_t)i * st] = v;
        int add = i + r + 1, sub = i - r;
        if (add < len) { u += i1[b + (size_t)add * st]; v += i2[b + (size_t)add * st]; }
        if (sub >= 0) { u -= i1[b + (size_t)sub * st]; v -= i2[b + (size_t)sub * st]; }
    }
}

// local mean m = s1/wc and variance max(s2/wc - m*m, 0) (noise.c:96-101)
__global__ void k_local(const double *s1, const double *s2, double *m, double *v, size_t n, double wc) {
    #pragma clang fp contract(off)
    for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT) {
        double mm = s1[i] / wc;
        double vv = s2[i] / wc - mm * mm;
        m[i] = mm;
        v[i] = vv > 0 ? vv : 0;
    }
}

// first index with sorted[idx] >= q (lower bound), for each query
__global__ void k_lower_bound(const double *sorted, long n, const double *q, int nq, long *idx) {
    int t = blockIdx.x * NT + threadIdx.x;
    if (t >= nq) return;
    double key = q[t];
    long lo = 0, hi = n;
    while (lo < hi) {
        long mid = lo + (hi - lo) / 2;
        if (sorted[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    idx[t] = lo;
}

__global__ void k_gather(const double *src, const long *idx, int n, double *out) {
    int t = blockIdx.x * NT + threadIdx.x;
    if (t < n) out[t] = src[idx[t]];
}

// per 8^3 block: sum, sum of squares (double, in the reference's dz,dy,dx order),
// positivity (noise.c:220-229)
__global__ void k_aniso_var(const float *v, int ny, int nx, int bz, int by, int bx, float *bvar, int *ok) {
    #pragma clang fp contract(off)
    long b = (long)blockIdx.x * NT + threadIdx.x;
    long nb = (long)bz * by * bx;
    if (b >= nb) return;
    int x = (int)(b % bx), y = (int)((b / bx) % by), z = (int)(b / ((long)bx * by));
    double s = 0, s2 = 0;
    int pos = 1;
    for (int dz = 0; dz < 8; dz++)
        for (int dy = 0; dy < 8; dy++) {
            const float *row = v + ((size_t)(z * 8 + dz) * ny + (y * 8 + dy)) * nx + (size_t)x * 8;
            for (int dx = 0; dx < 8; dx++) {
                float t = row[dx];
                if (t <= 0) pos = 0;
                s += t;
                s2 += (double)t * t;
            }
        }
    double m = s / 512;
    bvar[b] = (float)(s2 / 512 - m * m);
    ok[b] = pos;
}

// per flat block: lag-1 products of the mean-removed residual along z, y, x and
// the residual energy (noise.c:239-260); part[4*b + {0,1,2,3}] = den, z, y, x
__global__ void k_aniso_lag(const float *v, int ny, int nx, int bz, int by, int bx, float thresh, double *part,
                            int *used) {
    #pragma clang fp contract(off)
    long b = (long)blockIdx.x * NT + threadIdx.x;
    long nb = (long)bz * by * bx;
    if (b >= nb) return;
    int x = (int)(b % bx), y = (int)((b / bx) % by), z = (int)(b / ((long)bx * by));
    float blk[512];
    double s = 0, s2 = 0;
    int pos = 1;
    for (int dz = 0; dz < 8; dz++)
        for (int dy = 0; dy < 8; dy++) {
            const float *row = v + ((size_t)(z * 8 + dz) * ny + (y * 8 + dy)) * nx + (size_t)x * 8;
            for (int dx = 0; dx < 8; dx++) {
                float t = row[dx];
                if (t <= 0) pos = 0;
                blk[(dz * 8 + dy) * 8 + dx] = t;
                s += t;
                s2 += (double)t * t;
            }
        }
    used[b] = 0;
    part[4 * b] = part[4 * b + 1] = part[4 * b + 2] = part[4 * b + 3] = 0.0;
    if (!pos) return;
    double m = s / 512;
    if (s2 / 512 - m * m > thresh) return;
    float mf = (float)m;
    for (int i = 0; i < 512; i++) blk[i] -= mf;
    double den = 0, az = 0, ay = 0, ax = 0;
    for (int dz = 0; dz < 8; dz++)
        for (int dy = 0; dy < 8; dy++)
            for (int dx = 0; dx < 8; dx++) {
                float c = blk[(dz * 8 + dy) * 8 + dx];
                den += (double)c * c;
                if (dz + 1 < 8) az += (double)c * blk[((dz + 1) * 8 + dy) * 8 + dx];
                if (dy + 1 < 8) ay += (double)c * blk[(dz * 8 + dy + 1) * 8 + dx];
                if (dx + 1 < 8) ax += (double)c * blk[(dz * 8 + dy) * 8 + dx + 1];
            }
    part[4 * b] = den;
    part[4 * b + 1] = az;
    part[4 * b + 2] = ay;
    part[4 * b + 3] = ax;
    used[b] = 1;
}

}  // namespace

extern "C" int fyg_noise_local(const float *d_in, int nz, int ny, int nx, int r, double wc, double *m, double *v,
                               double *t0, double *t1, double *t2, double *t3, fyg_stream *s) {
    size_t n = (size_t)nz * ny * nx;
    long rows = (long)nz * ny;
    fyg_launch(s, "noise_box_x", k_box_x, dim3(grid1d(rows)), dim3(NT), 0, d_in, t0, t1, rows, nx, r);
    // y: lines (z, x), inner = nx, outer stride = ny*nx, element stride nx
    long ly = (long)nz * nx;
    fyg_launch(s, "noise_box_y", k_box_line, dim3(grid1d(ly)), dim3(NT), 0, (const double *)t0,
               (const double *)t1, t2, t3, ly, (long)nx, (long)ny * nx, ny, (long)nx, r);
    // z: lines (y, x), inner = ny*nx, element stride ny*nx
    long lz = (long)ny * nx;
    fyg_launch(s, "noise_box_z", k_box_line, dim3(grid1d(lz)), dim3(NT), 0, (const double *)t2,
               (const double *)t3, t0, t1, lz, lz, 0L, nz, lz, r);
    unsigned g = grid1d(n);
    if (g > 16384) g = 16384;
    fyg_launch(s, "noise_local", k_local, dim3(g), dim3(NT), 0, (const double *)t0, (const double *)t1, m, v, n,
               wc);
    return fyg_last_error();
}

extern "C" size_t fyg_sort_pairs_temp_bytes(size_t n) {
    size_t t = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, t, (const double *)nullptr, (double *)nullptr,
                                             (const double *)nullptr, (double *)nullptr, (int)n);
    return t;
}

extern "C" int fyg_sort_pairs_f64(const double *kin, double *kout, const double *vin, double *vout, size_t n,
                                  void *temp, size_t temp_bytes, fyg_stream *s) {
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, kin, kout, vin, vout, (int)n, 0, 64,
                                                      s->s);
    return e != hipSuccess || fyg_last_error();
}

extern "C" size_t fyg_segsort_temp_bytes(size_t n, int nseg) {
    size_t t = 0;
    (void)hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, t, (const double *)nullptr, (double *)nullptr,
                                                     (int)n, nseg, (const int *)nullptr, (const int *)nullptr);
    return t;
}

extern "C" int fyg_segsort_f64(const double *in, double *out, size_t n, int nseg, const int *d_begin,
                               const int *d_end, void *temp, size_t temp_bytes, fyg_stream *s) {
    hipError_t e = hipcub::DeviceSegmentedRadixSort::SortKeys(temp, temp_bytes, in, out, (int)n, nseg, d_begin,
                                                              d_end, 0, 64, s->s);
    return e != hipSuccess || fyg_last_error();
}

extern "C" int fyg_lower_bounds_f64(const double *sorted, size_t n, const double *d_q, int nq, long *d_idx,
                                    fyg_stream *s) {
    fyg_launch(s, "lower_bound", k_lower_bound, dim3(grid1d(nq)), dim3(NT), 0, sorted, (long)n, d_q, nq, d_idx);
    return fyg_last_error();
}

extern "C" int fyg_gather_f64(const double *src, const long *d_idx, int n, double *d_out, fyg_stream *s) {
    fyg_launch(s, "gather", k_gather, dim3(grid1d(n)), dim3(NT), 0, src, d_idx, n, d_out);
    return fyg_last_error();
}

extern "C" int fyg_aniso_blocks(const float *v, int nz, int ny, int nx, float *bvar, int *ok, fyg_stream *s) {
    int bz = nz / 8, by = ny / 8, bx = nx / 8;
    long nb = (long)bz * by * bx;
    if (nb <= 0) return 0;
    fyg_launch(s, "aniso_var", k_aniso_var, dim3(grid1d(nb)), dim3(NT), 0, v, ny, nx, bz, by, bx, bvar, ok);
    return fyg_last_error();
}

extern "C" int fyg_aniso_lag(const float *v, int nz, int ny, int nx, float thresh, double *part, int *used,
                             fyg_stream *s) {
    int bz = nz / 8, by = ny / 8, bx = nx / 8;
    long nb = (long)bz * by * bx;
    if (nb <= 0) return 0;
    fyg_launch(s, "aniso_lag", k_aniso_lag, dim3(grid1d(nb)), dim3(NT), 0, v, ny, nx, bz, by, bx, thresh, part,
               used);
    return fyg_last_error();
}
