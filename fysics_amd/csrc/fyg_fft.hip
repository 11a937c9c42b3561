// fyg_fft.hip -- the fp32 instantiation of the deconvolution passes
// (fyg_fft_engine.hpp; deconv_radial, paganin.c:152-293) and the complex FFTs of
// fft.c (fy_fft1d / fy_fft3d, fft.c:170-228) behind the C ABI of fyg.h.
#include <atomic>

#include "fyg_fft_engine.hpp"

namespace {

struct Rows {
    float2 *data;
    long nlines;
    float sg;
    const float2 *tw;
};

template <int L>
__global__ __launch_bounds__(256) void k_cfft_rows(Rows a) {
    constexpr int NL = Shape<L>::NL_RM;
    __shared__ float2 buf[NL * Shape<L>::LD];
    __shared__ float2 tw_s[L];
    const TwPrefetch<L, float2> twp(a.tw);
    const long g0 = (long)blockIdx.x * NL;
    auto first = [&](int l, int p0, int off) -> float2 {
        long g = g0 + l;
        return g < a.nlines ? a.data[(size_t)g * L + p0 + off] : make_float2(0.f, 0.f);
    };
    auto last = [&](int l, int p0, int off, float2 v) {
        long g = g0 + l;
        if (g < a.nlines) a.data[(size_t)g * L + p0 + off] = v;
    };
    auto hook = [&] { twp.store(tw_s); };
    if (a.sg < 0)
        run_from<L, NL, false, false, 0>(-1.0f, tw_s, buf, first, last, hook);
    else
        run_from<L, NL, false, false, 0>(1.0f, tw_s, buf, first, last, hook);
}

__global__ void k_interleave(const float *re, const float *im, float2 *c, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        c[i] = make_float2(re[i], im[i]);
}
__global__ void k_deinterleave(const float2 *c, float *re, float *im, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float2 v = c[i];
        re[i] = v.x;
        im[i] = v.y;
    }
}
__global__ void k_scale(float *d, size_t n, float k) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        d[i] *= k;
}

}  // namespace

static std::atomic<int> g_plane_mode{1};
extern "C" int fyg_fft_plane_mode(void) { return g_plane_mode.load(std::memory_order_relaxed); }
extern "C" int fyg_fft_plane_set(int mode) { return g_plane_mode.exchange(mode ? 1 : 0); }

extern "C" int fyg_fft_supported(int L) {
    if (L < 1 || L > FYG_FFT_MAX) return 0;
    int m = L % 3 == 0 ? L / 3 : L;
    return (m & (m - 1)) == 0;
}

extern "C" int fyg_deconv_hxp(int px) {
    int hx = px / 2 + 1;
    return (hx + 15) / 16 * 16;
}

extern "C" size_t fyg_deconv_spec_bytes(int pz, int py, int px) {
    return sizeof(float2) * (size_t)pz * py * fyg_deconv_hxp(px);
}

extern "C" size_t fyg_deconv_spec_bytes64(int pz, int py, int px) {
    return sizeof(double2) * (size_t)pz * py * fyg_deconv_hxp(px);
}

extern "C" int fyg_deconv_run_batch(const fyg_deconv_args *a, int n, fyg_stream *s) {
    if (n < 1) return 0;
    const fyg_deconv_args &h = a[0];
    if (!fyg_fft_supported(h.px) || !fyg_fft_supported(h.py) || !fyg_fft_supported(h.pz)) return 1;
    if (h.f64) return fyg_deconv_run_batch_f64(a, n, s);
    return deconv_batch<float>(a, n, s);
}

extern "C" int fyg_deconv_run(const fyg_deconv_args *a, fyg_stream *s) { return fyg_deconv_run_batch(a, 1, s); }

extern "C" size_t fyg_ring_image_bytes(const fyg_prologue *p, int nz, int ny, int nx) {
    if (!p->dering) return 0;
    int s0 = ring_slab(*p, p->gz0), s1 = ring_slab(*p, p->gz0 + nz - 1);
    return sizeof(float) * (size_t)(s1 - s0 + 1) * ny * nx;
}

// in-place complex 3-D FFT of interleaved data, axis order X, Y, Z (fft.c:178-213)
extern "C" int fyg_cfft3d_run(void *data, int nz, int ny, int nx, int sign, fyg_stream *s) {
    if (!fyg_fft_supported(nx) || !fyg_fft_supported(ny) || !fyg_fft_supported(nz)) return 1;
    float2 *c = (float2 *)data;
    float sg = sign < 0 ? -1.0f : 1.0f;
    Rows rx{c, (long)nz * ny, sg, (const float2 *)fyg_twiddles(nx)};
    dispatch(nx, [&](auto cl) {
        constexpr int L = decltype(cl)::value;
        unsigned g = (unsigned)((rx.nlines + Shape<L>::NL_RM - 1) / Shape<L>::NL_RM);
        fyg_launch(s, "cfft_rows", k_cfft_rows<L>, dim3(g), dim3(256), 0, rx);
    });
    if (ny > 1) {
        Cols<float> cy{c, nx, nx, (long)ny * nx, sg, (const float2 *)fyg_twiddles(ny)};
        dispatch(ny, [&](auto cl) { Launch<decltype(cl)::value, float>::cols(s, &cy, 1, nz, "cfft_cols"); });
    }
    if (nz > 1) {
        Cols<float> cz{c, (long)ny * nx, (long)ny * nx, 0, sg, (const float2 *)fyg_twiddles(nz)};
        dispatch(nz, [&](auto cl) { Launch<decltype(cl)::value, float>::cols(s, &cz, 1, 1, "cfft_cols"); });
    }
    return fyg_last_error();
}

extern "C" int fyg_fft3d_run(float *d_re, float *d_im, int nz, int ny, int nx, int sign, void *work,
                             fyg_stream *s) {
    if (!fyg_fft_supported(nx) || !fyg_fft_supported(ny) || !fyg_fft_supported(nz)) return 1;
    size_t n = (size_t)nz * ny * nx;
    float2 *c = (float2 *)work;
    fyg_launch(s, "interleave", k_interleave, dim3(grid1d(n)), dim3(256), 0, (const float *)d_re,
               (const float *)d_im, c, n);
    if (fyg_cfft3d_run(c, nz, ny, nx, sign, s)) return 1;
    fyg_launch(s, "deinterleave", k_deinterleave, dim3(grid1d(n)), dim3(256), 0, (const float2 *)c, d_re,
               d_im, n);
    return fyg_last_error();
}

extern "C" int fyg_scale_run(float *d, size_t n, float k, fyg_stream *s) {
    fyg_launch(s, "scale", k_scale, dim3(grid1d(n)), dim3(256), 0, d, n, k);
    return fyg_last_error();
}
