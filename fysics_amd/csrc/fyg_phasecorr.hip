// fyg_phasecorr.hip -- phase correlation on the device (phasecorr.c:44-145):
// mean removal and windowing into the zero-padded power-of-two cube, two
// forward complex FFTs, the unit-magnitude cross-power spectrum, the inverse FFT
// and the first maximum of the correlation surface with its six circular
// neighbours.  The sub-voxel refinement (a few scalar divisions) is left to the
// caller.  Reductions run in a fixed order (per-block partials, then one block),
// so results do not depend on scheduling.
#include "fyg_internal.hpp"

namespace {

constexpr int NT = 256;
constexpr int NB = 1024;                 // partial blocks of the reductions

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// per-block double sums of fixed and moving -> part[2*blockIdx.x + {0,1}]
__global__ __launch_bounds__(NT) void k_pc_sums(const float *f, const float *m, size_t n, double *part) {
    __shared__ double red[2][NT / 64];
    double a = 0.0, b = 0.0;
    for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT) {
        a += (double)f[i];
        b += (double)m[i];
    }
    a = wave_sum(a);
    b = wave_sum(b);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = a;
        red[1][threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0.0, sb = 0.0;
        for (int k = 0; k < NT / 64; k++) { sa += red[0][k]; sb += red[1][k]; }
        part[2 * blockIdx.x] = sa;
        part[2 * blockIdx.x + 1] = sb;
    }
}

// means (fp32, as the reference rounds them) from the partial sums
__global__ void k_pc_means(const double *part, int nb, double n, float *means) {
    if (threadIdx.x != 0) return;
    double sa = 0.0, sb = 0.0;
    for (int k = 0; k < nb; k++) { sa += part[2 * k]; sb += part[2 * k + 1]; }
    means[0] = (float)(sa / n);
    means[1] = (float)(sb / n);
}

// padded (val - mean) * wz*wy*wx as complex (re, 0); zero outside the volume
__global__ void k_pc_fill(const float *f, const float *m, int nz, int ny, int nx, int pz, int py, int px,
                          const float *wz, const float *wy, const float *wx, const float *means, float2 *F,
                          float2 *M) {
    #pragma clang fp contract(off)
    const size_t pn = (size_t)pz * py * px;
    const float mf = means[0], mm = means[1];
    for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < pn; i += (size_t)gridDim.x * NT) {
        const int x = (int)(i % px), t = (int)(i / px), y = t % py, z = t / py;
        float a = 0.0f, b = 0.0f;
        if (z < nz && y < ny && x < nx) {
            const float w = wz[z] * wy[y] * wx[x];
            const size_t j = ((size_t)z * ny + y) * nx + x;
            a = (f[j] - mf) * w;
            b = (m[j] - mm) * w;
        }
        F[i] = make_float2(a, 0.0f);
        M[i] = make_float2(b, 0.0f);
    }
}

// F <- conj(F)*M / |conj(F)*M| (0 below 1e-12); per-block maxima of |.|
__global__ __launch_bounds__(NT) void k_pc_cross(float2 *F, const float2 *M, size_t pn, float *pmax) {
    #pragma clang fp contract(off)
    __shared__ float red[NT / 64];
    float mx = 0.0f;
    for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < pn; i += (size_t)gridDim.x * NT) {
        const float2 a = F[i], b = M[i];
        const float cr = a.x * b.x + a.y * b.y, ci = a.x * b.y - a.y * b.x;
        const float mag = sqrtf(cr * cr + ci * ci);
        mx = mag > mx ? mag : mx;
        F[i] = mag > 1e-12f ? make_float2(cr / mag, ci / mag) : make_float2(0.0f, 0.0f);
    }
    for (int o = 32; o > 0; o >>= 1) {
        const float t = __shfl_xor(mx, o);
        mx = t > mx ? t : mx;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        float b = red[0];
        for (int k = 1; k < NT / 64; k++) b = red[k] > b ? red[k] : b;
        pmax[blockIdx.x] = b;
    }
}

// (value, index) of the larger; on equal values the smaller index
__device__ __forceinline__ void pc_better(float &v, long long &i, float v2, long long i2) {
    if (v2 > v || (v2 == v && i2 < i)) {
        v = v2;
        i = i2;
    }
}

// per-block first maximum of the real part
__global__ __launch_bounds__(NT) void k_pc_argmax(const float2 *F, size_t pn, float *pv, long long *pi) {
    __shared__ float rv[NT / 64];
    __shared__ long long ri[NT / 64];
    float v = -INFINITY;
    long long idx = 0x7fffffffffffffffLL;
    for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < pn; i += (size_t)gridDim.x * NT)
        pc_better(v, idx, F[i].x, (long long)i);
    for (int o = 32; o > 0; o >>= 1) {
        const float v2 = __shfl_xor(v, o);
        const long long i2 = __shfl_xor(idx, o);
        pc_better(v, idx, v2, i2);
    }
    if ((threadIdx.x & 63) == 0) {
        rv[threadIdx.x >> 6] = v;
        ri[threadIdx.x >> 6] = idx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < NT / 64; k++) pc_better(v, idx, rv[k], ri[k]);
        pv[blockIdx.x] = v;
        pi[blockIdx.x] = idx;
    }
}

// the global maximum, the largest magnitude and the peak's neighbourhood
__global__ void k_pc_final(const float2 *F, int pz, int py, int px, const float *pv, const long long *pi,
                           const float *pmax, int nb, fyg_pc_result *out) {
    if (threadIdx.x != 0) return;
    float v = pv[0], mx = pmax[0];
    long long idx = pi[0];
    for (int k = 1; k < nb; k++) {
        pc_better(v, idx, pv[k], pi[k]);
        mx = pmax[k] > mx ? pmax[k] : mx;
    }
    if (idx < 0 || idx >= (long long)pz * py * px) idx = 0;   // only if every value was NaN
    out->best = v;
    out->maxmag = mx;
    out->index = idx;
    const int kz = (int)(idx / ((long long)py * px)), ky = (int)((idx / px) % py), kx = (int)(idx % px);
    auto at = [&](int z, int y, int x) { return F[((size_t)z * py + y) * px + x].x; };
    out->c[0] = at(kz, ky, kx);
    out->c[1] = at((kz - 1 + pz) % pz, ky, kx);
    out->c[2] = at((kz + 1) % pz, ky, kx);
    out->c[3] = at(kz, (ky - 1 + py) % py, kx);
    out->c[4] = at(kz, (ky + 1) % py, kx);
    out->c[5] = at(kz, ky, (kx - 1 + px) % px);
    out->c[6] = at(kz, ky, (kx + 1) % px);
}

size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

extern "C" size_t fyg_pc_ws_bytes(int pz, int py, int px) {
    const size_t pn = (size_t)pz * py * px;
    return 2 * al256(sizeof(float2) * pn) + al256(sizeof(double) * 2 * NB) + al256(sizeof(float) * 2) +
           al256(sizeof(float) * NB) * 2 + al256(sizeof(long long) * NB) + al256(sizeof(fyg_pc_result));
}

extern "C" int fyg_phase_correlate_run(const float *fixed, const float *moving, int nz, int ny, int nx, int pz,
                                       int py, int px, const float *wz, const float *wy, const float *wx,
                                       void *ws, fyg_pc_result *out, fyg_stream *s) {
    const size_t n = (size_t)nz * ny * nx, pn = (size_t)pz * py * px;
    char *w = (char *)ws;
    float2 *F = (float2 *)w;
    w += al256(sizeof(float2) * pn);
    float2 *M = (float2 *)w;
    w += al256(sizeof(float2) * pn);
    double *part = (double *)w;
    w += al256(sizeof(double) * 2 * NB);
    float *means = (float *)w;
    w += al256(sizeof(float) * 2);
    float *pmax = (float *)w;
    w += al256(sizeof(float) * NB);
    float *pv = (float *)w;
    w += al256(sizeof(float) * NB);
    long long *pi = (long long *)w;
    w += al256(sizeof(long long) * NB);
    fyg_pc_result *dres = (fyg_pc_result *)w;
    auto grid = [](size_t work) {
        size_t b = (work + NT - 1) / NT;
        return dim3((unsigned)(b < 1 ? 1 : (b > NB ? NB : b)));
    };
    const dim3 gs = grid(n), gp = grid(pn);
    fyg_launch(s, "pc_sums", k_pc_sums, gs, dim3(NT), 0, fixed, moving, n, part);
    fyg_launch(s, "pc_means", k_pc_means, dim3(1), dim3(64), 0, (const double *)part, (int)gs.x, (double)n, means);
    fyg_launch(s, "pc_fill", k_pc_fill, gp, dim3(NT), 0, fixed, moving, nz, ny, nx, pz, py, px, wz, wy, wx,
               (const float *)means, F, M);
    if (fyg_cfft3d_run(F, pz, py, px, -1, s) || fyg_cfft3d_run(M, pz, py, px, -1, s)) return 1;
    fyg_launch(s, "pc_cross", k_pc_cross, gp, dim3(NT), 0, F, (const float2 *)M, pn, pmax);
    if (fyg_cfft3d_run(F, pz, py, px, 1, s)) return 1;
    fyg_launch(s, "pc_argmax", k_pc_argmax, gp, dim3(NT), 0, (const float2 *)F, pn, pv, pi);
    fyg_launch(s, "pc_final", k_pc_final, dim3(1), dim3(64), 0, (const float2 *)F, pz, py, px, (const float *)pv,
               (const long long *)pi, (const float *)pmax, (int)gp.x, dres);
    if (fyg_last_error()) return 1;
    return fyg_d2h(out, dres, sizeof(fyg_pc_result), s) || fyg_sync(s);
}
