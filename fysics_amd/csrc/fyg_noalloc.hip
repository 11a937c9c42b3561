// fyg_noalloc.hip -- the Wiener deconvolution for callers that cannot allocate
// device memory (code under graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  Everything the passes need beyond the caller's input and
// output -- the half spectrum, the 4097-entry weight table, the squared
// frequency axes and the three twiddle tables -- lives in `scratch`, one
// file-scope device array per GPU that the library holds until the process
// ends.  The tables are computed on the device at the start of each call from
// the physics passed by value, so a call queues two kinds of work (one table
// launch, then the five passes of deconv_radial) and copies nothing from the
// host.
//
// Every call on a GPU uses that GPU's array: calls that overlap on one GPU
// (two streams, two threads) race.  Nothing in the array has to start at zero:
// the table launch writes every table entry the passes read, and the X pass
// writes every spectrum element (all pz*py rows, columns < hx) before a later
// pass reads it; the pitch columns hx..hxp-1 are never read.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "fyg_internal.hpp"

namespace {
constexpr size_t kTabBytes = (size_t)256 << 10;        // weights, axes and twiddles (<= 197 KB used)
constexpr size_t kLutOff = 0;                           // float[4100]
constexpr size_t kAxesOff = 16640;                      // double fz2[pz], fy2[py], fx2[hx]
constexpr size_t kTwOff = kAxesOff + 8 * (2 * FYG_FFT_MAX + FYG_FFT_MAX / 2 + 1 + 31);   // float2 tables
constexpr int kMaxDev = 64;
static_assert(kTwOff % 8 == 0, "twiddle tables are float2 aligned");
static_assert(kTwOff + 8 * 3 * (size_t)FYG_FFT_MAX <= kTabBytes, "tables fit the table area");
}  // namespace

static __device__ __align__(256) unsigned char scratch[FYG_NOALLOC_SPEC_MAX + kTabBytes];

// device address of `scratch` on each GPU, taken once with that GPU current
static unsigned char *g_scratch[kMaxDev];
static std::mutex g_scratch_mu;

static unsigned char *scratch_ptr() {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= kMaxDev) {
        fyg_set_error("fy_gpu_deconvolve_noalloc: no current HIP device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            char msg[200];
            snprintf(msg, sizeof msg, "fy_gpu_deconvolve_noalloc: hipGetSymbolAddress(scratch): %s",
                     hipGetErrorString(e));
            fyg_set_error(msg);
            return nullptr;
        }
        g_scratch[dev] = (unsigned char *)p;
    }
    return g_scratch[dev];
}

namespace {
struct NoallocTables {
    double db, e_kev, d_mm, px_um, u_sigma, u_coeff;    // fy_physics fields the weight reads
    double reg, cap, us;
    int pz, py, px, hx;
    float *lut;
    double *fz2, *fy2, *fx2;
    float2 *twx, *twy, *twz;
};

// the recon transfer max(T*U, 1e-4) of paganin.c:17-44 and the Wiener weight
// H/(H^2+reg) of paganin.c:296-299, in double with the operation order of the
// host tables (no contraction), for one frequency in cycles per voxel
__device__ double wiener_weight_dev(double f, const NoallocTables &t) {
#pragma clang fp contract(off)
    double lam = 1.23984199e-3 / t.e_kev;
    double dist = t.d_mm * 1000.0;
    double px = t.px_um;
    if (px > 0.0 && px < 0.1) px *= 1000.0;
    double fu = f / px;
    double T = 1.0 / (1.0 + t.db * lam * dist * M_PI * (fu * fu));
    double U = 1.0;
    if (!(t.u_coeff == 0.0 || t.u_sigma == 0.0)) {
        double g = exp(-2.0 * M_PI * M_PI * t.u_sigma * t.u_sigma * f * f);
        U = 1.0 + t.u_coeff * (1.0 - g);
    }
    double h = T * U;
    if (h < 1e-4) h = 1e-4;
    return h / (h * h + t.reg);
}

__device__ void twiddle(float2 *tw, int L, int m) {
#pragma clang fp contract(off)
    double a = 2.0 * M_PI * (double)m / (double)L;
    tw[m] = make_float2((float)cos(a), (float)sin(a));
}

__device__ double axis_f(int i, int n) { return (i <= n / 2) ? (double)i / n : (double)(i - n) / n; }

__global__ void k_noalloc_tables(NoallocTables t) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= 4096) {
        double w = wiener_weight_dev(sqrt((double)i / t.us), t);
        if (t.cap > 0 && w > t.cap) w = t.cap;
        t.lut[i] = (float)w;
    } else if (i < 4100) {
        t.lut[i] = 0.0f;
    }
    if (i < t.pz) {
        double f = axis_f(i, t.pz);
        t.fz2[i] = f * f * 1.0;
    }
    if (i < t.py) {
        double f = axis_f(i, t.py);
        t.fy2[i] = f * f;
    }
    if (i < t.hx) {
        double f = (double)i / t.px;
        t.fx2[i] = f * f;
    }
    if (i < t.px) twiddle(t.twx, t.px, i);
    if (i < t.py) twiddle(t.twy, t.py, i);
    if (i < t.pz) twiddle(t.twz, t.pz, i);
}
}  // namespace

extern "C" int fyg_noalloc_deconv(const float *in, float *out, int nz, int ny, int nx, int pz, int py, int px,
                                  const double *phys8, double reg, double cap, double us, fyg_stream *s) {
    if (!s || !fyg_fft_supported(pz) || !fyg_fft_supported(py) || !fyg_fft_supported(px)) return 1;
    if (fyg_deconv_spec_bytes(pz, py, px) > FYG_NOALLOC_SPEC_MAX) return 1;
    unsigned char *base = scratch_ptr();
    if (!base) return 1;
    unsigned char *tab = base + FYG_NOALLOC_SPEC_MAX;
    const int hx = px / 2 + 1;
    NoallocTables t;
    t.db = phys8[0]; t.e_kev = phys8[1]; t.d_mm = phys8[2]; t.px_um = phys8[3];
    t.u_sigma = phys8[4]; t.u_coeff = phys8[5];
    t.reg = reg; t.cap = cap; t.us = us;
    t.pz = pz; t.py = py; t.px = px; t.hx = hx;
    t.lut = (float *)(tab + kLutOff);
    t.fz2 = (double *)(tab + kAxesOff);
    t.fy2 = t.fz2 + pz;
    t.fx2 = t.fy2 + py;
    t.twx = (float2 *)(tab + kTwOff);
    t.twy = t.twx + px;
    t.twz = t.twy + py;
    int count = 4100;
    if (px > count) count = px;
    if (py > count) count = py;
    if (pz > count) count = pz;
    fyg_launch(s, "noalloc_tables", k_noalloc_tables, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, t);
    if (fyg_last_error()) return 1;

    fyg_deconv_args a;
    memset(&a, 0, sizeof a);
    a.nz = nz; a.ny = ny; a.nx = nx;
    a.pz = pz; a.py = py; a.px = px; a.hx = hx; a.hxp = fyg_deconv_hxp(px);
    a.lut = t.lut; a.us = us; a.fz2 = t.fz2; a.fy2 = t.fy2; a.fx2 = t.fx2;
    a.in = in; a.in_u8 = 0;
    a.out = out;
    a.spec = base;
    a.twx = t.twx; a.twy = t.twy; a.twz = t.twz;
    a.no_tune = 1;                  /* timing the radix plans would allocate */
    return fyg_deconv_run(&a, s);
}
