// fyg_air.hip -- stage (d) of process_tile on the device: the air-zero value
// cut (pipeline.c:302-399) and the connected-component despeckle
// (pipeline.c:227-250).
//
//   k_air_decimate   orig -> every second voxel of every second row and plane
//   k_box_axis  x3   box mean of the scratch, one sliding line per thread, the
//                    window sum updated add-then-subtract in fp32 exactly as
//                    guided.c:57-91 (so the scratch equals the reference's)
//   k_air_hist       256 bins of (int)(smooth*255 + 0.5)
//   k_air_cut        smoothed histogram in LDS, then one thread: valley depth
//                    (stream.c:138-170) -> cut
//   k_air_apply      trilinear upsample of the scratch, texture-split threshold
//                    (without the despeckle; with it, k_cc_init applies the rule)
//   k_cc_*           6-connected components of f > 0: x runs labelled by a wave
//                    scan, runs joined along y and z by union-find (parents only
//                    ever decrease: L[x] <= x, a root is its component's smallest
//                    index), labels flattened, face contact (from the face voxels
//                    only) and sizes (up to min_size) accumulated per root with
//                    wave-aggregated atomics, then the small interior islands are
//                    zeroed.  Per-voxel kernels run on (row, x block) grids, so a
//                    voxel's (z, y) is a per-workgroup 32-bit division.
// The arithmetic of the scratch, the blend and the thresholds is written in the
// reference's order with contraction off, so the decisions match the oracle.
#include "fyg_internal.hpp"

namespace {

constexpr int NT = 256;
constexpr int HC = 32;                   // LDS sub-histogram copies (bank-spread)
constexpr int kEdge = 1 << 30;           // face-contact bit of a component count

struct AirCut {
    float uni, tex;                      // cuts for uniform / textured voxels
};

}  // namespace

// valley depth of a 256-bin histogram (stream.c:138-170) in two parts: the
// 5-bin smoothing of one bin, and the search over the smoothed histogram hf (rails
// 0, 254, 255 cleared here): the two tallest peaks above 5% of the maximum
// (peaks within 10 bins merged), the lowest bin between them.  Returns -1 when
// there are fewer than two peaks.  peaks: scratch of 256 ints.
__host__ __device__ double fyg_valley_smooth(const long long *hist, int i) {
    double sm = 0.0;
    int cnt = 0;
    for (int k = -2; k <= 2; k++) {
        const int j = i + k;
        if (j >= 0 && j < 256) {
            sm += (double)hist[j];
            cnt++;
        }
    }
    return sm / cnt;
}

__host__ __device__ double fyg_valley_search(double *hf, int *peaks, int *dark, int *light, int *valley) {
    hf[0] = hf[254] = hf[255] = 0.0;
    double mx = 0.0;
    for (int i = 0; i < 256; i++) mx = hf[i] > mx ? hf[i] : mx;
    if (mx <= 0.0) return -1.0;
    int np = 0;
    for (int u = 3; u < 253; u++) {
        if (!(hf[u] >= hf[u - 1] && hf[u] >= hf[u + 1] && hf[u] > 0.05 * mx)) continue;
        if (np && u - peaks[np - 1] <= 10) {
            if (hf[u] > hf[peaks[np - 1]]) peaks[np - 1] = u;
        } else {
            peaks[np++] = u;
        }
    }
    if (np < 2) return -1.0;
    int p1 = peaks[0], p2 = peaks[1];
    for (int i = 0; i < np; i++) {
        const int p = peaks[i];
        if (hf[p] > hf[p1]) {
            p2 = p1;
            p1 = p;
        } else if (hf[p] > hf[p2] && p != p1) {
            p2 = p;
        }
    }
    const int a = p1 < p2 ? p1 : p2, b = p1 < p2 ? p2 : p1;
    int v = a;
    double vmin = hf[a];
    for (int u = a; u <= b; u++)
        if (hf[u] < vmin) {
            vmin = hf[u];
            v = u;
        }
    if (dark) *dark = a;
    if (light) *light = b;
    if (valley) *valley = v;
    const double mn = hf[a] < hf[b] ? hf[a] : hf[b];
    return 1.0 - vmin / (mn + 1e-12);
}

namespace {

// grid (low-res rows z*ly + y, x blocks)
__global__ __launch_bounds__(NT) void k_air_decimate(const float *orig, float *low, int ny, int nx, int ly,
                                                      int lx) {
    const int row = blockIdx.x, z = row / ly, y = row - z * ly;
    const int x = blockIdx.y * NT + threadIdx.x;
    if (x >= lx) return;
    low[(size_t)row * lx + x] = orig[((size_t)(2 * z) * ny + 2 * y) * nx + 2 * x];
}

// one line of the box mean: the window sum slides add-then-subtract in fp32
// (guided.c:57-91).  ld(i) reads input i; st(i, v) receives output i, called
// after the step's updates, by which time input i - r is no longer needed.
template <class Ld, class St>
__device__ __forceinline__ void box_line(Ld ld, St st, int len, int r) {
    #pragma clang fp contract(off)
    auto rc = [&](int i) {
        const int lo = i - r < 0 ? 0 : i - r, hi = i + r > len - 1 ? len - 1 : i + r;
        return 1.0f / (float)(hi - lo + 1);
    };
    float acc = 0.0f;
    for (int i = 0; i <= r && i < len; i++) acc += ld(i);
    // the same updates as one loop with both conditions, split where they change
    // so the middle loop has no branches and its loads can be issued ahead
    const int m0 = r < len ? r : len;                         // i < r: add only
    const int m1 = len - r - 1 > m0 ? len - r - 1 : m0;       // r <= i < len-r-1: add and subtract
    int i = 0;
    for (; i < m0; i++) {
        const float v = acc * rc(i);
        if (i + r + 1 < len) acc += ld(i + r + 1);
        st(i, v);
    }
    const float rmid = 1.0f / (float)(2 * r + 1);
#pragma unroll 8
    for (; i < m1; i++) {
        const float v = acc * rmid;
        acc += ld(i + r + 1);
        acc -= ld(i - r);
        st(i, v);
    }
    for (; i < len; i++) {
        const float v = acc * rc(i);
        if (i + r + 1 < len) acc += ld(i + r + 1);
        if (i - r >= 0) acc -= ld(i - r);
        st(i, v);
    }
}

// axis 0: x (lines = rows), 1: y, 2: z; one line per thread
__global__ void k_box_axis(const float *src, float *dst, int nz, int ny, int nx, int axis, int r) {
    const size_t plane = (size_t)ny * nx;
    const int len = axis == 0 ? nx : (axis == 1 ? ny : nz);
    const size_t st = axis == 0 ? 1 : (axis == 1 ? (size_t)nx : plane);
    const size_t nlines = plane * nz / len;
    const size_t line = (size_t)blockIdx.x * NT + threadIdx.x;
    if (line >= nlines) return;
    const size_t base = axis == 0 ? line * nx : (axis == 1 ? (line / nx) * plane + line % nx : line);
    const float *in = src + base;
    float *out = dst + base;
    box_line([&](int i) { return in[(size_t)i * st]; }, [&](int i, float v) { out[(size_t)i * st] = v; }, len, r);
}

// The x axis through LDS: one wave stages BXR rows with coalesced loads, each
// lane slides along its own row (row pitch nx + 1 words: the lanes' reads fall
// in different banks), and the results leave with coalesced stores.  (A thread
// per row reading global memory directly touches 64 cache lines per load.)
// Output i goes over input i - r, which the slide no longer needs; outputs
// i < r wait in a side area of BOX_RMAX words per row.
constexpr int BXR = 64;
constexpr int BOX_RMAX = 8;
constexpr int BOX_X_MAX = 240;           // 64 * (241 + 8) floats: within 64 KB of LDS
// (all four waves stage the rows, eight loads in flight per lane; wave 0 slides)
__global__ __launch_bounds__(NT) void k_box_x(const float *src, float *dst, long nrows, int nx, int r) {
    extern __shared__ float sh[];
    const int P = nx + 1;
    float *buf = sh, *head = sh + BXR * P;
    const long row0 = (long)blockIdx.x * BXR;
    const int nr = nrows - row0 < BXR ? (int)(nrows - row0) : BXR;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int cnt = nr * nx;
    const float *s0 = src + (size_t)row0 * nx;
    for (int k0 = threadIdx.x; k0 < cnt; k0 += 8 * NT) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int k = k0 + u * NT;
            v[u] = k < cnt ? s0[k] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int k = k0 + u * NT;
            if (k < cnt) {
                const int rr = k / nx;
                buf[rr * P + (k - rr * nx)] = v[u];
            }
        }
    }
    __syncthreads();
    if (wv == 0 && lane < nr) {
        float *b = buf + lane * P, *h = head + lane * BOX_RMAX;
        box_line([&](int i) { return b[i]; },
                 [&](int i, float v) {
                     if (i >= r) b[i - r] = v;
                     else h[i] = v;
                 },
                 nx, r);
    }
    __syncthreads();
    float *d0 = dst + (size_t)row0 * nx;
    for (int k = threadIdx.x; k < cnt; k += NT) {
        const int rr = k / nx, x = k - rr * nx;
        d0[k] = x >= r ? buf[rr * P + x - r] : head[rr * BOX_RMAX + x];
    }
}

// The y and z axes the same way: a workgroup takes BXR lines that neighbour in
// x (line l: element i at base + l + i*st), stages them with loads of BXR
// consecutive floats per element index, slides them in LDS (wave 0, a lane per
// line; line pitch len + 1 words) and stores them back.
__global__ __launch_bounds__(NT) void k_box_yz(const float *src, float *dst, int nx, int len, long st, long ost,
                                               int r) {
    extern __shared__ float sh[];
    const int P = len + 1;
    float *buf = sh, *head = sh + BXR * P;
    const int x0 = blockIdx.x * BXR;
    const int nl = nx - x0 < BXR ? nx - x0 : BXR;
    const size_t base = (size_t)blockIdx.y * ost + x0;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int cnt = len * BXR;
    for (int k0 = threadIdx.x; k0 < cnt; k0 += 8 * NT) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int k = k0 + u * NT, i = k / BXR, l = k - i * BXR;
            v[u] = (k < cnt && l < nl) ? src[base + (size_t)i * st + l] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int k = k0 + u * NT, i = k / BXR, l = k - i * BXR;
            if (k < cnt) buf[l * P + i] = v[u];
        }
    }
    __syncthreads();
    if (wv == 0 && lane < nl) {
        float *b = buf + lane * P, *h = head + lane * BOX_RMAX;
        box_line([&](int i) { return b[i]; },
                 [&](int i, float v) {
                     if (i >= r) b[i - r] = v;
                     else h[i] = v;
                 },
                 len, r);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < cnt; k += NT) {
        const int i = k / BXR, l = k - i * BXR;
        if (l < nl) dst[base + (size_t)i * st + l] = i >= r ? buf[l * P + i - r] : head[l * BOX_RMAX + i];
    }
}

// per-workgroup bins of (int)(v*255 + 0.5) clamped to [0, 255] -> part[blockIdx.x]
__global__ __launch_bounds__(NT) void k_air_hist(const float *v, size_t n, unsigned int *part) {
    #pragma clang fp contract(off)
    __shared__ unsigned int h[256 * HC];
    for (int i = threadIdx.x; i < 256 * HC; i += NT) h[i] = 0;
    __syncthreads();
    const int c = threadIdx.x % HC;
    for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT) {
        int b = (int)(v[i] * 255.0f + 0.5f);
        b = b < 0 ? 0 : (b > 255 ? 255 : b);
        atomicAdd(&h[b * HC + c], 1u);
    }
    __syncthreads();
    const int b = threadIdx.x;
    unsigned int t = 0;
    for (int k = 0; k < HC; k++) t += h[b * HC + ((k + b) & (HC - 1))];
    part[(size_t)blockIdx.x * 256 + b] = t;
}

struct CutParams {
    int cut_u8, band, min_component;
    double thresh, aggr;
};

// the cut of pipeline.c:329-363 from the scratch histogram.  The valley search
// of fyg_valley_search runs across the workgroup: 256 threads smooth one bin
// each, the maximum and the peak candidates (local maxima above 5 % of it) come
// from wave reductions and ballots, one thread merges the candidates in bin
// order (peaks within 10 bins, the same sequential rule), and the valley -- the
// first minimum between the two tallest peaks -- is a reduction again.  The
// results equal the sequential search's.
__global__ __launch_bounds__(256) void k_air_cut(const unsigned long long *hist, CutParams p, AirCut *out) {
    #pragma clang fp contract(off)
    __shared__ long long h[256];
    __shared__ double hf[256];
    __shared__ int peaks[256];
    __shared__ double wred[4];
    __shared__ int wpos[4];
    __shared__ unsigned long long cand[4];
    __shared__ int s_ab[3];                  // a, b, found (two peaks)
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    h[t] = (long long)hist[t];
    __syncthreads();
    const double v = (t == 0 || t >= 254) ? 0.0 : fyg_valley_smooth(h, t);     // rails 0, 254, 255
    hf[t] = v;
    double m = v;
    for (int o = 32; o > 0; o >>= 1) {
        const double q = __shfl_xor(m, o);
        m = q > m ? q : m;
    }
    if (lane == 0) wred[w] = m;
    __syncthreads();
    double mx = 0.0;
    for (int k = 0; k < 4; k++) mx = wred[k] > mx ? wred[k] : mx;
    const bool c = t >= 3 && t < 253 && v >= hf[t - 1] && v >= hf[t + 1] && v > 0.05 * mx;
    const unsigned long long bal = __ballot(c);
    if (lane == 0) cand[w] = bal;
    __syncthreads();
    if (t == 0) {
        int np = 0;
        if (mx > 0.0) {
            for (int k = 0; k < 4; k++) {
                unsigned long long mm = cand[k];
                while (mm) {
                    const int u = 64 * k + __ffsll((long long)mm) - 1;
                    mm &= mm - 1;
                    if (np && u - peaks[np - 1] <= 10) {
                        if (hf[u] > hf[peaks[np - 1]]) peaks[np - 1] = u;
                    } else {
                        peaks[np++] = u;
                    }
                }
            }
        }
        s_ab[2] = np >= 2;
        if (np >= 2) {
            int p1 = peaks[0], p2 = peaks[1];
            for (int i = 0; i < np; i++) {
                const int q = peaks[i];
                if (hf[q] > hf[p1]) {
                    p2 = p1;
                    p1 = q;
                } else if (hf[q] > hf[p2] && q != p1) {
                    p2 = q;
                }
            }
            s_ab[0] = p1 < p2 ? p1 : p2;
            s_ab[1] = p1 < p2 ? p2 : p1;
        }
    }
    __syncthreads();
    const bool found = s_ab[2] != 0;
    const int a = s_ab[0], b = s_ab[1];
    // the first minimum of hf over [a, b]: smallest value, then smallest bin
    double vm = (found && t >= a && t <= b) ? v : 1e300;
    int pm = (found && t >= a && t <= b) ? t : 1 << 30;
    for (int o = 32; o > 0; o >>= 1) {
        const double qv = __shfl_xor(vm, o);
        const int qp = __shfl_xor(pm, o);
        if (qv < vm || (qv == vm && qp < pm)) {
            vm = qv;
            pm = qp;
        }
    }
    if (lane == 0) {
        wred[w] = vm;
        wpos[w] = pm;
    }
    __syncthreads();
    if (t != 0) return;
    int dark = 0, light = 0, valley = 0;
    double d = -1.0;
    if (found) {
        double vmin = wred[0];
        int vi = wpos[0];
        for (int k = 1; k < 4; k++)
            if (wred[k] < vmin || (wred[k] == vmin && wpos[k] < vi)) {
                vmin = wred[k];
                vi = wpos[k];
            }
        dark = a;
        light = b;
        valley = vi;
        const double mn = hf[a] < hf[b] ? hf[a] : hf[b];
        d = 1.0 - vmin / (mn + 1e-12);
    }
    int cut;
    if (p.cut_u8 >= 0) {
        const int gg = p.cut_u8, band = p.band;
        if (d >= 0) {
            const int pf = p.thresh > 0 ? (int)(p.thresh * 255 + 0.5) : 39;
            double a = p.aggr;
            a = a < 0 ? 0 : (a > 2 ? 2 : a);
            int local;
            if (valley <= pf) local = pf;
            else if (a <= 1.0) local = (int)(pf + a * (valley - pf) + 0.5);
            else local = (int)(valley + (a - 1.0) * (light > valley ? light - valley : 0) + 0.5);
            cut = local < gg - band ? gg - band : (local > gg + band ? gg + band : local);
        } else {
            cut = gg;
        }
    } else if (d >= 0) {
        const int mid = (dark + valley) / 2;
        cut = dark + 8 < mid ? dark + 8 : mid;
    } else {
        cut = (int)((p.thresh > 0 ? p.thresh : 0.05) * 255);
    }
    const int pf2 = p.thresh > 0 ? (int)(p.thresh * 255 + 0.5) : 39;
    const int spread = valley > pf2 ? (valley - pf2) / 2 : 8;
    int cu = cut + spread;
    if (valley > cut && cu > valley) cu = valley;
    int ct = cut - spread;
    if (ct < pf2) ct = pf2;
    out->uni = cu / 255.0f;
    out->tex = ct / 255.0f;
}

// The air rule: the scratch at the voxel (a trilinear upsample of the
// decimated grid when lowres) below the cut of the voxel's texture class.
struct AirGrid {
    const float *orig, *sm;
    int ny, nx, lz, ly, lx, lowres;
};

// the scratch at voxels x = xc + lane of row (z, y), one wave (xc a multiple of
// 64; every lane must call it).  At half resolution each lane forms the blend
// of the four low-res rows at one column and the voxels take the two columns
// around them by shuffles: 4 loads per lane instead of 8 per voxel, the same
// expressions in the same order as a voxel's own blend.
__device__ __forceinline__ float air_scratch_row(const AirGrid &g, int z, int y, int xc, int lane) {
    #pragma clang fp contract(off)
    const int x = xc + lane;
    if (!g.lowres) return x < g.nx ? g.sm[((size_t)z * g.ny + y) * g.nx + x] : 0.0f;
    int z0 = z >> 1, y0 = y >> 1;
    z0 = z0 > g.lz - 2 ? g.lz - 2 : z0;
    y0 = y0 > g.ly - 2 ? g.ly - 2 : y0;
    float zf = 0.5f * (float)z - (float)z0, yf = 0.5f * (float)y - (float)y0;
    zf = zf > 1.0f ? 1.0f : zf;
    yf = yf > 1.0f ? 1.0f : yf;
    const size_t r00 = ((size_t)z0 * g.ly + y0) * g.lx, r01 = r00 + g.lx, r10 = r00 + (size_t)g.ly * g.lx,
                 r11 = r10 + g.lx;
    const float w00 = (1 - zf) * (1 - yf), w01 = (1 - zf) * yf, w10 = zf * (1 - yf), w11 = zf * yf;
    // columns base .. base+32 cover the knots of the wave's voxels
    const int base = (xc >> 1) < g.lx - 2 ? (xc >> 1) : g.lx - 2;
    const int q = base + lane < g.lx - 1 ? base + lane : g.lx - 1;
    const float *sm = g.sm;
    const float bl = w00 * sm[r00 + q] + w01 * sm[r01 + q] + w10 * sm[r10 + q] + w11 * sm[r11 + q];
    int x0 = x >> 1;
    x0 = x0 > g.lx - 2 ? g.lx - 2 : x0;
    float fx = 0.5f * (float)x - (float)x0;
    fx = fx > 1.0f ? 1.0f : fx;
    const float a = __shfl(bl, x0 - base), b = __shfl(bl, x0 + 1 - base);
    return a + (b - a) * fx;
}

__device__ __forceinline__ bool air_zeroes(const AirGrid &g, const AirCut &cut, float v, size_t i) {
    #pragma clang fp contract(off)
    float tex = g.orig[i] - v;
    tex = tex < 0 ? -tex : tex;
    return v < (tex < 0.03f ? cut.uni : cut.tex);
}

// zero f where the air rule holds; grid (rows, x blocks of NT)
__global__ __launch_bounds__(NT) void k_air_apply(float *f, AirGrid g, const AirCut *cutp) {
    const AirCut cut = *cutp;
    const int row = blockIdx.x, z = row / g.ny, y = row - z * g.ny;
    const int lane = threadIdx.x & 63, xc = blockIdx.y * NT + (threadIdx.x & ~63);
    if (xc >= g.nx) return;                             // whole waves
    const float v = air_scratch_row(g, z, y, xc, lane);
    const int x = xc + lane;
    if (x >= g.nx) return;
    const size_t i = (size_t)row * g.nx + x;
    if (air_zeroes(g, cut, v, i)) f[i] = 0.0f;
}

// ---- connected components (6-connectivity) of f > 0
__device__ __forceinline__ int cc_load(int *L, int i) {
    return __hip_atomic_load(&L[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// root of i with path halving: each visited voxel is re-pointed to its
// grandparent (atomicMin keeps L[x] <= x and x's root unchanged).  Parents
// strictly decrease along a chain, so this ends within i steps.
__device__ __forceinline__ int cc_find(int *L, int i) {
    int cur = cc_load(L, i);
    if (cur == i) return i;
    int prev = i, next;
    while (cur > (next = cc_load(L, cur))) {
        atomicMin(&L[prev], next);
        prev = cur;
        cur = next;
    }
    return cur;
}

// link the roots of a and b, the larger under the smaller; a root that was
// linked meanwhile is followed to its new parent, so every retry moves to a
// strictly smaller index and the loop ends (the cap only reports a broken
// invariant)
__device__ void cc_unite(int *L, int a, int b, int *err) {
    for (int it = 0; it < (1 << 26); it++) {
        a = cc_find(L, a);
        b = cc_find(L, b);
        if (a == b) return;
        if (a > b) {
            const int t = a;
            a = b;
            b = t;
        }
        const int old = atomicMin(&L[b], a);
        if (old == b) return;
        b = old;
    }
    atomicOr(err, 1);
}

// x runs first: L[i] = the first voxel of i's run of material along x (a run
// start is its own parent), so the unions below only join runs.  One wave per
// row chunk of 64 voxels; the last gap before each voxel comes from a max-scan
// across the wave, carried along the row.  Also clears the counts of the run
// starts: only run starts are ever roots, and counts are kept per root.
// The air rule is applied first, in the same pass (k_air_apply's zeros,
// written where they change f's bits), so f is read once for both.
__global__ __launch_bounds__(NT) void k_cc_init(float *f, int *L, int *cnt, long nrows, int nx, AirGrid ag,
                                                const AirCut *cutp) {
    const int lane = threadIdx.x & 63;
    const long row = (long)blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
    if (row >= nrows) return;                           // whole waves leave together
    const size_t base = (size_t)row * nx;
    const AirCut cut = *cutp;
    const int z = (int)(row / ag.ny), y = (int)(row - (long)z * ag.ny);
    int carry = -1;                                     // last gap (x) before this chunk
    // four chunks at a time: their loads and air decisions are independent,
    // so all are issued before the scans that carry the gap along the row
    for (int xg = 0; xg < nx; xg += 4 * 64) {
        bool mm[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int xc = xg + 64 * c, x = xc + lane;
            mm[c] = false;
            if (xc >= nx) continue;                     // whole waves
            const float v = air_scratch_row(ag, z, y, xc, lane);
            if (x < nx) {
                const float fv = f[base + x];
                mm[c] = fv > 0.0f;
                if (air_zeroes(ag, cut, v, base + x)) {
                    if (__float_as_uint(fv) != 0u) f[base + x] = 0.0f;
                    mm[c] = false;
                }
            }
        }
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int x = xg + 64 * c + lane;
            const bool in = x < nx, m = mm[c];
            int g = m ? carry : x;                      // gap positions; carry for material
            for (int o = 1; o < 64; o <<= 1) {
                const int t = __shfl_up(g, o);
                if (lane >= o) g = g > t ? g : t;
            }
            if (in) {
                L[base + x] = m ? (int)(base + g + 1) : -1;
                if (m && g + 1 == x) cnt[base + x] = 0;
            }
            carry = __shfl(g, 63);
        }
    }
}

__device__ __forceinline__ bool cc_mat(const int *L, size_t j) { return L[j] >= 0; }

// grid (rows z*ny + y, x blocks of NT): the voxel's coordinates and index (n < 2^30)
struct RowPos {
    int x, y, z, i;
    bool in;
};
__device__ __forceinline__ RowPos row_pos(int ny, int nx) {
    RowPos p;
    const int row = blockIdx.x;
    p.z = row / ny;
    p.y = row - p.z * ny;
    p.x = blockIdx.y * NT + threadIdx.x;
    p.in = p.x < nx;
    p.i = row * nx + p.x;
    return p;
}

// join runs that touch along y and z; where the voxel before i and its
// neighbour are both material, that pair already joined the same two runs.
// The six labels a voxel's decision needs are loaded at once (indices clamped
// into the volume, results masked), one memory round trip instead of a chain.
__global__ __launch_bounds__(NT) void k_cc_union(int *L, int ny, int nx, int *err) {
    const RowPos q = row_pos(ny, nx);
    if (!q.in) return;
    const int i = q.i, plane = ny * nx;
    const bool hx = q.x > 0, hy = q.y > 0, hz = q.z > 0;
    const int c = L[i], l = L[hx ? i - 1 : i], u = L[hy ? i - nx : i], ul = L[hx && hy ? i - 1 - nx : i],
              b = L[hz ? i - plane : i], bl = L[hx && hz ? i - 1 - plane : i];
    if (c < 0) return;
    const bool left = hx && l >= 0;
    if (hy && u >= 0 && !(left && ul >= 0)) cc_unite(L, i, i - nx, err);
    if (hz && b >= 0 && !(left && bl >= 0)) cc_unite(L, i, i - plane, err);
}

// Labels to roots in two steps.  Every ancestor of a voxel is a run start (only
// run starts are ever roots), so once the run starts point at their roots, any
// voxel's root is L[L[i]] (a run start's L is already its root, and a root's L
// is itself); k_cc_count takes that second step and stores it.
__global__ __launch_bounds__(NT) void k_cc_flatten_starts(int *L, int ny, int nx) {
    const RowPos q = row_pos(ny, nx);
    if (!q.in) return;
    const int c = cc_load(L, q.i), l = cc_load(L, q.x != 0 ? q.i - 1 : q.i);
    if (c < 0 || (q.x != 0 && l >= 0)) return;                         // air, or not a run start
    L[q.i] = cc_find(L, q.i);
}


// wave-aggregated count of per-root sizes: lanes that share a root add once
// through the first of them.  The loop bound is wave-uniform.
__device__ __forceinline__ void cc_wave_count(int *cnt, bool act, int r, int min_size) {
    const int lane = threadIdx.x & 63;
    unsigned long long pending = __ballot(act);
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const int r0 = __shfl(r, leader);
        const unsigned long long grp = __ballot(act && r == r0);
        if (lane == leader) {
            const int c = cnt[r0];                      // kept components and met sizes need no count
            if (!(c & kEdge) && c < min_size) atomicAdd(&cnt[r0], __popcll(grp));
        }
        pending &= ~grp;
    }
}

// roots of components that touch a face of the region, from the face voxels
// only, one thread each: the two z faces, then the two y faces without their
// z-face rows, then the two x faces without either.  A workgroup collects its
// distinct roots in LDS first (most face voxels belong to a few components)
// and sets each one's face bit once.
constexpr int EROOTS = 64;               // distinct roots a workgroup collects before it spills
struct Faces {
    long zf, yf, xf;                     // voxels on the z faces, the rest of the y faces, the rest of the x faces
};
__host__ __device__ inline Faces faces_of(int nz, int ny, int nx) {
    const long iz = nz > 2 ? nz - 2 : 0, iy = ny > 2 ? ny - 2 : 0;
    return Faces{2L * ny * nx, 2L * iz * nx, 2L * iz * iy};
}
__global__ __launch_bounds__(NT) void k_cc_edge(const int *L, int *cnt, int nz, int ny, int nx, int per) {
    __shared__ int roots[EROOTS];
    __shared__ int nroots;
    const int lane = threadIdx.x & 63;
    if (threadIdx.x == 0) nroots = 0;
    // -1, never a root: a slot read between its reservation and its store
    // matches nothing, so the reader adds its root itself (a duplicate at worst)
    for (int j = threadIdx.x; j < EROOTS; j += NT) roots[j] = -1;
    __syncthreads();
    const Faces F = faces_of(nz, ny, nx);
    const long nf = F.zf + F.yf + F.xf;
    // each workgroup takes `per` consecutive chunks of NT face voxels
    for (int it = 0; it < per; it++) {
    long f = ((long)blockIdx.x * per + it) * NT + threadIdx.x;
    if ((long)blockIdx.x * per * NT + (long)it * NT >= nf) break;          // workgroup-uniform
    int r = -1;
    if (f < nf) {
        int z, y, x;
        if (f < F.zf) {                                  // z = 0 and z = nz - 1
            const long pl = (long)ny * nx, q = f % pl;
            z = f < pl ? 0 : nz - 1;
            y = (int)(q / nx);
            x = (int)(q - (long)y * nx);
        } else if ((f -= F.zf) < F.yf) {                 // y = 0 and y = ny - 1, 0 < z < nz - 1
            const long half = F.yf / 2, q = f % half;
            y = f < half ? 0 : ny - 1;
            z = 1 + (int)(q / nx);
            x = (int)(q - (long)(z - 1) * nx);
        } else {                                         // x = 0 and x = nx - 1, 0 < z < nz - 1, 0 < y < ny - 1
            f -= F.yf;
            const long half = F.xf / 2, q = f % half;
            const int iy = ny - 2;
            x = f < half ? 0 : nx - 1;
            z = 1 + (int)(q / iy);
            y = 1 + (int)(q - (long)(z - 1) * iy);
        }
        r = L[((size_t)z * ny + y) * nx + x];
        if (r >= 0) r = L[r];                           // the run start's root
    }
    unsigned long long pending = __ballot(r >= 0);
    while (pending) {                                    // one leader per distinct root of the wave
        const int leader = __ffsll((long long)pending) - 1;
        const int r0 = __shfl(r, leader);
        pending &= ~__ballot(r == r0);
        if (lane == leader) {
            const int k = __hip_atomic_load(&nroots, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            bool seen = false;
            for (int j = 0; j < k && j < EROOTS && !seen; j++)
                seen = __hip_atomic_load(&roots[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == r0;
            if (!seen) {
                const int slot = atomicAdd(&nroots, 1);
                if (slot < EROOTS)
                    __hip_atomic_store(&roots[slot], r0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                else
                    atomicOr(&cnt[r0], kEdge);
            }
        }
    }
    }
    __syncthreads();
    const int n = nroots < EROOTS ? nroots : EROOTS;
    for (int j = threadIdx.x; j < n; j += NT) {
        const int q = roots[j];
        // duplicates (two waves adding one root at once) only repeat an idempotent OR
        if (!(__hip_atomic_load(&cnt[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & kEdge)) atomicOr(&cnt[q], kEdge);
    }
}

// sizes of the components that do not touch a face, counted up to min_size;
// every material voxel's label becomes its root here (the second flatten step)
__global__ void k_cc_count(int *L, int *cnt, size_t n, int min_size) {
    const size_t stride = (size_t)gridDim.x * NT;
    for (size_t i0 = (size_t)blockIdx.x * NT + (threadIdx.x & ~63u); i0 < n; i0 += stride) {
        const size_t i = i0 + (threadIdx.x & 63);
        const int p = i < n ? L[i] : -1;
        int r = p;
        if (p >= 0) {
            r = L[p];
            if (r != p) L[i] = r;
        }
        cc_wave_count(cnt, r >= 0, r, min_size);
    }
}

__global__ void k_cc_zero(float *f, const int *L, const int *cnt, size_t n, int min_size) {
    for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT) {
        const int r = L[i];
        if (r < 0) continue;
        const int c = cnt[r];
        if (!(c & kEdge) && c < min_size) f[i] = 0.0f;
    }
}

unsigned grid_for(size_t work, unsigned cap) {
    size_t b = (work + NT - 1) / NT;
    return (unsigned)(b < 1 ? 1 : (b > cap ? cap : b));
}

size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

extern "C" double fyg_valley_depth(const long *hist, int *dark, int *light, int *valley) {
    long long h[256];
    double hf[256];
    int peaks[256];
    for (int b = 0; b < 256; b++) h[b] = hist[b];
    for (int b = 0; b < 256; b++) hf[b] = fyg_valley_smooth(h, b);
    return fyg_valley_search(hf, peaks, dark, light, valley);
}

// workspace: decimated copy, scratch, box temporaries, histogram rows, the cut,
// component labels and counts
extern "C" size_t fyg_air_ws_bytes(int nz, int ny, int nx) {
    const size_t n = (size_t)nz * ny * nx;
    const size_t sc = n;                                 // the scratch grid never exceeds n
    return 3 * al256(sizeof(float) * sc) + al256(sizeof(unsigned int) * 256 * 1024) +
           al256(sizeof(unsigned long long) * 256) + al256(sizeof(AirCut)) + al256(sizeof(int)) +
           2 * al256(sizeof(int) * n);
}

extern "C" int fyg_air_zero_run(float *f, const float *orig, int nz, int ny, int nx, const fyg_air_args *a,
                                void *ws, fyg_stream *s) {
    const size_t n = (size_t)nz * ny * nx;
    if (!n) return 0;
    if (n >= (1ULL << 30)) return 1;                     // int labels; counts stay below the edge bit
    char *w = (char *)ws;
    float *low = (float *)w;
    w += al256(sizeof(float) * n);
    float *sm = (float *)w;
    w += al256(sizeof(float) * n);
    float *tmp = (float *)w;
    w += al256(sizeof(float) * n);
    unsigned int *part = (unsigned int *)w;
    w += al256(sizeof(unsigned int) * 256 * 1024);
    unsigned long long *hist = (unsigned long long *)w;
    w += al256(sizeof(unsigned long long) * 256);
    AirCut *cut = (AirCut *)w;
    w += al256(sizeof(AirCut));
    int *err = (int *)w;
    w += al256(sizeof(int));
    int *L = (int *)w;
    w += al256(sizeof(int) * n);
    int *cnt = (int *)w;

    const int sr = a->scratch_passes > 0 ? a->scratch_passes : 5;
    const int lz = (nz + 1) / 2, ly = (ny + 1) / 2, lx = (nx + 1) / 2;
    const int lowres = lz >= 2 && ly >= 2 && lx >= 2;
    int gz = nz, gy = ny, gx = nx, r = sr;
    const float *src = orig;
    if (lowres) {
        fyg_launch(s, "air_decimate", k_air_decimate, dim3((unsigned)((long)lz * ly), (unsigned)((lx + NT - 1) / NT)),
                   dim3(NT), 0, orig, low, ny, nx, ly, lx);
        gz = lz; gy = ly; gx = lx;
        r = (sr + 1) / 2 < 1 ? 1 : (sr + 1) / 2;
        src = low;
    }
    // x into tmp, y into the decimated copy's buffer (or sm), z into sm
    const size_t ns = (size_t)gz * gy * gx;
    float *mid = lowres ? low : sm;
    auto lines = [&](int len) { return dim3((unsigned)((ns / len + NT - 1) / NT)); };   // one line per thread
    if (gx <= BOX_X_MAX && r <= BOX_RMAX) {
        const long nrows = (long)gz * gy;
        fyg_launch(s, "air_box", k_box_x, dim3((unsigned)((nrows + BXR - 1) / BXR)), dim3(NT),
                   sizeof(float) * BXR * (gx + 1 + BOX_RMAX), src, tmp, nrows, gx, r);
    } else {
        fyg_launch(s, "air_box", k_box_axis, lines(gx), dim3(NT), 0, src, tmp, gz, gy, gx, 0, r);
    }
    const unsigned xc = (unsigned)((gx + BXR - 1) / BXR);
    const size_t gplane = (size_t)gy * gx;
    if (gy <= BOX_X_MAX && r <= BOX_RMAX)
        fyg_launch(s, "air_box", k_box_yz, dim3(xc, (unsigned)gz), dim3(NT), sizeof(float) * BXR * (gy + 1 + BOX_RMAX),
                   (const float *)tmp, mid, gx, gy, (long)gx, (long)gplane, r);
    else
        fyg_launch(s, "air_box", k_box_axis, lines(gy), dim3(NT), 0, (const float *)tmp, mid, gz, gy, gx, 1, r);
    float *fin = lowres ? sm : tmp;
    if (gz <= BOX_X_MAX && r <= BOX_RMAX)
        fyg_launch(s, "air_box", k_box_yz, dim3(xc, (unsigned)gy), dim3(NT), sizeof(float) * BXR * (gz + 1 + BOX_RMAX),
                   (const float *)mid, fin, gx, gz, (long)gplane, (long)gx, r);
    else
        fyg_launch(s, "air_box", k_box_axis, lines(gz), dim3(NT), 0, (const float *)mid, fin, gz, gy, gx, 2, r);
    const unsigned hg = grid_for(ns / 16, 1024);
    fyg_launch(s, "air_hist", k_air_hist, dim3(hg), dim3(NT), 0, (const float *)fin, ns, part);
    if (fyg_zero(hist, sizeof(unsigned long long) * 256, s) || fyg_hist_reduce_run(part, (int)hg, hist, s)) return 1;
    CutParams cp{a->cut_u8, a->band, a->min_component, a->thresh, a->aggr};
    fyg_launch(s, "air_cut", k_air_cut, dim3(1), dim3(256), 0, (const unsigned long long *)hist, cp, cut);
    if ((long)nz * ny > 0x7fffffffL || (nx + NT - 1) / NT > 65535) return 1;
    dim3 ag((unsigned)((long)nz * ny), (unsigned)((nx + NT - 1) / NT));
    const AirGrid agr{orig, (const float *)fin, ny, nx, lz, ly, lx, lowres};
    if (a->min_component <= 0) {
        fyg_launch(s, "air_apply", k_air_apply, ag, dim3(NT), 0, f, agr, (const AirCut *)cut);
    } else {                                              // pipeline.c:396-398
        const unsigned g = grid_for(n, 16384);
        if (fyg_zero(err, sizeof(int), s)) return 1;
        const long nrows = (long)nz * ny;
        const dim3 wrows((unsigned)((nrows + NT / 64 - 1) / (NT / 64)));     // one wave per row
        // the air rule and the x runs of the components in one pass over f
        fyg_launch(s, "air_cc_init", k_cc_init, wrows, dim3(NT), 0, f, L, cnt, nrows, nx, agr,
                   (const AirCut *)cut);
        fyg_launch(s, "cc_union", k_cc_union, ag, dim3(NT), 0, L, ny, nx, err);
        fyg_launch(s, "cc_flatten", k_cc_flatten_starts, ag, dim3(NT), 0, L, ny, nx);
        const Faces fc = faces_of(nz, ny, nx);
        const long chunks = (fc.zf + fc.yf + fc.xf + NT - 1) / NT, per = chunks > 1024 ? (chunks + 1023) / 1024 : 1;
        fyg_launch(s, "cc_edge", k_cc_edge, dim3((unsigned)((chunks + per - 1) / per)), dim3(NT), 0, (const int *)L,
                   cnt, nz, ny, nx, (int)per);
        fyg_launch(s, "cc_count", k_cc_count, dim3(g), dim3(NT), 0, L, cnt, n, a->min_component);
        fyg_launch(s, "cc_zero", k_cc_zero, dim3(g), dim3(NT), 0, f, (const int *)L, (const int *)cnt, n,
                   a->min_component);
    }
    return fyg_last_error();
}

extern "C" int fyg_air_err_read(void *ws, int nz, int ny, int nx, int *out, fyg_stream *s) {
    const size_t n = (size_t)nz * ny * nx;
    char *w = (char *)ws + 3 * al256(sizeof(float) * n) + al256(sizeof(unsigned int) * 256 * 1024) +
              al256(sizeof(unsigned long long) * 256) + al256(sizeof(AirCut));
    return fyg_d2h(out, w, sizeof(int), s) || fyg_sync(s);
}
