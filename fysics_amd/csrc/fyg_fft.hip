// fyg_fft.hip -- the fp32 instantiation of the deconvolution passes
// (fyg_fft_engine.hpp; deconv_radial, paganin.c:152-293) and the complex FFTs of
// fft.c (fy_fft1d / fy_fft3d, fft.c:170-228) behind the C ABI of fyg.h.
#include <atomic>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>

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

// ---- radix plans of the fp32 deconvolution passes.  A (kind, length) with two
// plans is decided once per process: by the mode, by a tuning file, or by
// timing both plans of that pass on a synthetic problem of about 16 M voxels
// at that length (zero data: the time of a pass does not depend on the values).
// The choice changes rounding (both plans are exact factorizations), so it is
// made before the first result of the length and kept.
namespace {

enum RadixKind { RK_X = FYG_RK_X, RK_Y = FYG_RK_Y, RK_Z = FYG_RK_Z, RK_XY = FYG_RK_XY, RK_N };
const char *const kRadixKind[RK_N] = {"x", "y", "z", "xy"};

struct RadixRec {
    int pv = 0;
    float us[2] = {0.0f, 0.0f};
    bool loaded = false;
};

std::mutex g_rad_mu;
std::map<std::pair<int, int>, RadixRec> g_rad;
std::atomic<int> g_rad_mode{1};

// the supported FFT length nearest (in ratio) to x
int nearest_length(double x) {
    int best = 16;
    double bd = 1e30;
    for (int L = 16; L <= 1024; L++) {
        if (!fyg_fft_supported(L)) continue;
        double d = std::fabs(std::log((double)L / x));
        if (d < bd) {
            bd = d;
            best = L;
        }
    }
    return best;
}

float elapsed_us(hipEvent_t a, hipEvent_t b) {
    float ms = 1e30f;
    if (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess) return 1e30f;
    return 1000.0f * ms;
}

// both plans of one pass kind at length L: us[pv] = best of three timed runs
// after one untimed run (which also loads the kernels' code).  Returns 0 when
// timed, 1 when the synthetic problem could not be set up.
int time_plans(int kind, int L, fyg_stream *s, float us[2]) {
    int nz, ny, nx;
    if (kind == RK_XY) {
        nz = (1 << 24) / (L * L);
        ny = nx = L;
    } else {
        const int M = nearest_length(std::sqrt((double)(1 << 24) / L));
        nz = ny = nx = M;
        if (kind == RK_X) nx = L;
        if (kind == RK_Y) ny = L;
        if (kind == RK_Z) nz = L;
    }
    const int hx = nx / 2 + 1, hxp = fyg_deconv_hxp(nx);
    const size_t nvox = (size_t)nz * ny * nx;
    const size_t b_in = sizeof(float) * nvox, b_spec = fyg_deconv_spec_bytes(nz, ny, nx);
    const int nf = nz > ny ? (nz > hx ? nz : hx) : (ny > hx ? ny : hx);
    const size_t b_lut = sizeof(float) * 4100, b_f2 = sizeof(double) * (size_t)nf;
    const size_t total = 2 * b_in + b_spec + b_lut + b_f2 + 4 * 256;
    char *base = nullptr;
    if (hipMalloc((void **)&base, total) != hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    char *p = base;
    float *in = (float *)p;
    p += up(b_in);
    float *out = (float *)p;
    p += up(b_in);
    void *spec = p;
    p += up(b_spec);
    float *lut = (float *)p;
    p += up(b_lut);
    double *f2 = (double *)p;
    int rc = hipMemsetAsync(base, 0, total, s->s) != hipSuccess;
    fyg_deconv_args a;
    memset(&a, 0, sizeof a);
    a.nz = nz; a.ny = ny; a.nx = nx;
    a.pz = nz; a.py = ny; a.px = nx; a.hx = hx; a.hxp = hxp;
    a.lut = lut; a.us = 1.0; a.fz2 = f2; a.fy2 = f2; a.fx2 = f2;
    a.in = in; a.out = out; a.spec = spec;
    Passes<float> ps;
    rc |= ps.build(&a, 1);
    hipEvent_t e[4] = {};
    for (int i = 0; i < 4 && !rc; i++) rc |= hipEventCreate(&e[i]) != hipSuccess;
    for (int pv = 0; pv < 2 && !rc; pv++) {
        auto run = [&] {
            switch (kind) {
            case RK_X: ps.x_fwd(s, pv); ps.x_inv(s, pv); break;
            case RK_Y: ps.y_fwd(s, pv); ps.y_inv(s, pv); break;
            case RK_Z: ps.z(s, pv); break;
            default: ps.xy_fwd(s, pv); ps.yx_inv(s, pv); break;
            }
        };
        run();
        float best = 1e30f;
        for (int r = 0; r < 3; r++) {
            (void)hipEventRecord(e[0], s->s);
            run();
            (void)hipEventRecord(e[1], s->s);
            float t = elapsed_us(e[0], e[1]);
            if (t < best) best = t;
        }
        us[pv] = best;
    }
    rc |= fyg_last_error();
    (void)hipStreamSynchronize(s->s);
    for (int i = 0; i < 4; i++)
        if (e[i]) (void)hipEventDestroy(e[i]);
    (void)hipFree(base);
    (void)hipGetLastError();
    return rc;
}

// plan variant of (kind, L) for a deconvolution on s
int radix_plan(int kind, int L, fyg_stream *s, bool may_tune) {
    if (!plan_alt(L)) return 0;
    const int mode = g_rad_mode.load(std::memory_order_relaxed);
    if (mode == 0) return 0;
    if (mode == 2) return 1;
    std::lock_guard<std::mutex> lk(g_rad_mu);
    auto it = g_rad.find({kind, L});
    if (it != g_rad.end()) return it->second.pv;
    if (!may_tune) return 0;
    RadixRec r;
    if (time_plans(kind, L, s, r.us) == 0)
        // plan 1 only when clearly faster: plans within a few percent of each other
        // would otherwise flip with the timing noise from one process to the next
        r.pv = r.us[1] < 0.95f * r.us[0] ? 1 : 0;
    else
        r.us[0] = r.us[1] = -1.0f;                      // not timed: plan 0
    g_rad[{kind, L}] = r;
    return r.pv;
}

}  // namespace

extern "C" int fyg_radix_plan(int kind, int L, fyg_stream *s, int may_tune) {
    return radix_plan(kind, L, s, may_tune != 0);
}

static std::atomic<int> g_mix_mode{1};
extern "C" int fyg_mix_mode(void) { return g_mix_mode.load(std::memory_order_relaxed); }
extern "C" int fyg_mix_set(int mode) { return g_mix_mode.exchange(mode < 0 ? 0 : mode > 2 ? 2 : mode); }

extern "C" int fyg_fft_radix_mode(int mode) { return g_rad_mode.exchange(mode < 0 ? 1 : mode > 2 ? 2 : mode); }

extern "C" void fyg_fft_radix_clear(void) {
    std::lock_guard<std::mutex> lk(g_rad_mu);
    g_rad.clear();
}

extern "C" int fyg_fft_radix_lines(char *buf, int len) {
    std::lock_guard<std::mutex> lk(g_rad_mu);
    std::string out;
    char line[96];
    for (auto &kv : g_rad) {
        snprintf(line, sizeof line, "fft\t%s\t%d\t%d\n", kRadixKind[kv.first.first], kv.first.second,
                 kv.second.pv);
        out += line;
    }
    if (buf && len > 0) snprintf(buf, (size_t)len, "%s", out.c_str());
    return (int)out.size();
}

extern "C" int fyg_fft_radix_load_line(const char *line) {
    char kind[8];
    int L = 0, pv = -1;
    if (sscanf(line, "%7s %d %d", kind, &L, &pv) != 3) return 1;
    int k = 0;
    while (k < RK_N && strcmp(kind, kRadixKind[k])) k++;
    if (k == RK_N || !plan_alt(L) || pv < 0 || pv > 1) return 1;
    RadixRec r;
    r.pv = pv;
    r.loaded = true;
    std::lock_guard<std::mutex> lk(g_rad_mu);
    g_rad[{k, L}] = r;
    return 0;
}

extern "C" int fyg_fft_radix_report(char *buf, int len) {
    std::lock_guard<std::mutex> lk(g_rad_mu);
    std::string out;
    char line[160];
    for (auto &kv : g_rad) {
        const RadixRec &r = kv.second;
        const char *k = kRadixKind[kv.first.first];
        const int L = kv.first.second;
        if (r.loaded)
            snprintf(line, sizeof line, "fft %s %d: plan %d (from a tuning file)\n", k, L, r.pv);
        else if (r.us[0] < 0)
            snprintf(line, sizeof line, "fft %s %d: plan 0 (not timed: no memory for the problem)\n", k, L);
        else
            snprintf(line, sizeof line, "fft %s %d: plan %d (radix-16 %.1f us, radix-8 %.1f us)\n", k, L, r.pv,
                     r.us[0], r.us[1]);
        out += line;
    }
    if (buf && len > 0) snprintf(buf, (size_t)len, "%s", out.c_str());
    return (int)out.size();
}

extern "C" int fyg_deconv_run_batch(const fyg_deconv_args *a, int n, fyg_stream *s) {
    if (n < 1) return 0;
    const fyg_deconv_args &h = a[0];
    if (!fyg_fft_supported(h.px) || !fyg_fft_supported(h.py) || !fyg_fft_supported(h.pz)) return 1;
    if (h.f64) return fyg_deconv_run_batch_f64(a, n, s);
    const bool may_tune = s && !s->capturing && !h.no_tune;
    PassPlans pp;
    if (use_planes<float>(h.px, h.py))
        pp.xy = radix_plan(RK_XY, h.px, s, may_tune);
    else {
        pp.x = radix_plan(RK_X, h.px, s, may_tune);
        pp.y = radix_plan(RK_Y, h.py, s, may_tune);
    }
    pp.z = radix_plan(RK_Z, h.pz, s, may_tune);
    return deconv_batch<float>(a, n, s, pp);
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
