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
//   k_cc_*           6-connected components of f > 0: x runs labelled by a wave
//                    scan, runs joined along y and z by union-find (parents only
//                    ever decrease: L[x] <= x, a root is its component's smallest
//                    index), labels flattened, face contact and sizes (up to
//                    min_size) accumulated per root with wave-aggregated atomics,
//                    then the small interior islands are zeroed
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

__global__ void k_air_decimate(const float *orig, float *low, int ny, int nx, int lz, int ly, int lx) {
    const size_t nl = (size_t)lz * ly * lx;
    for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < nl; i += (size_t)gridDim.x * NT) {
        const int x = (int)(i % lx), t = (int)(i / lx), y = t % ly, z = t / ly;
        low[i] = orig[((size_t)(2 * z) * ny + 2 * y) * nx + 2 * x];
    }
}

// axis 0: x (lines = rows), 1: y, 2: z
__global__ void k_box_axis(const float *src, float *dst, int nz, int ny, int nx, int axis, int r) {
    #pragma clang fp contract(off)
    const size_t plane = (size_t)ny * nx;
    const int len = axis == 0 ? nx : (axis == 1 ? ny : nz);
    const size_t st = axis == 0 ? 1 : (axis == 1 ? (size_t)nx : plane);
    const size_t nlines = plane * nz / len;
    const size_t line = (size_t)blockIdx.x * NT + threadIdx.x;
    if (line >= nlines) return;
    const size_t base = axis == 0 ? line * nx : (axis == 1 ? (line / nx) * plane + line % nx : line);
    const float *in = src + base;
    float *out = dst + base;
    auto rc = [&](int i) {
        const int lo = i - r < 0 ? 0 : i - r, hi = i + r > len - 1 ? len - 1 : i + r;
        return 1.0f / (float)(hi - lo + 1);
    };
    float acc = 0.0f;
    for (int i = 0; i <= r && i < len; i++) acc += in[(size_t)i * st];
    // the same updates as one loop with both conditions, split where they change
    // so the middle loop has no branches and its loads can be issued ahead
    const int m0 = r < len ? r : len;                         // i < r: add only
    const int m1 = len - r - 1 > m0 ? len - r - 1 : m0;       // r <= i < len-r-1: add and subtract
    int i = 0;
    for (; i < m0; i++) {
        out[(size_t)i * st] = acc * rc(i);
        if (i + r + 1 < len) acc += in[(size_t)(i + r + 1) * st];
    }
    const float rmid = 1.0f / (float)(2 * r + 1);
#pragma unroll 8
    for (; i < m1; i++) {
        out[(size_t)i * st] = acc * rmid;
        acc += in[(size_t)(i + r + 1) * st];
        acc -= in[(size_t)(i - r) * st];
    }
    for (; i < len; i++) {
        out[(size_t)i * st] = acc * rc(i);
        if (i + r + 1 < len) acc += in[(size_t)(i + r + 1) * st];
        if (i - r >= 0) acc -= in[(size_t)(i - r) * st];
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

// the cut of pipeline.c:329-363 from the scratch histogram: the histogram is
// staged in LDS by 256 threads, then one thread runs the valley search and the
// cut rules
__global__ __launch_bounds__(256) void k_air_cut(const unsigned long long *hist, CutParams p, AirCut *out) {
    #pragma clang fp contract(off)
    __shared__ long long h[256];
    __shared__ double hf[256];
    __shared__ int peaks[256];
    h[threadIdx.x] = (long long)hist[threadIdx.x];
    __syncthreads();
    hf[threadIdx.x] = fyg_valley_smooth(h, threadIdx.x);
    __syncthreads();
    if (threadIdx.x != 0) return;
    int dark = 0, light = 0, valley = 0;
    const double d = fyg_valley_search(hf, peaks, &dark, &light, &valley);
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

// zero f where the (upsampled) scratch is below the cut of its texture class;
// grid (rows, x blocks): the row's (z, y) blend weights are uniform
__global__ __launch_bounds__(NT) void k_air_apply(float *f, const float *orig, const float *sm, int nz, int ny,
                                                   int nx, int lz, int ly, int lx, int lowres, const AirCut *cutp) {
    #pragma clang fp contract(off)
    const AirCut cut = *cutp;
    const int row = blockIdx.x, z = row / ny, y = row - z * ny;
    const int x = blockIdx.y * NT + threadIdx.x;
    if (x >= nx) return;
    const size_t i = (size_t)row * nx + x;
    float v;
    if (lowres) {
        int z0 = z >> 1, y0 = y >> 1, x0 = x >> 1;
        z0 = z0 > lz - 2 ? lz - 2 : z0;
        y0 = y0 > ly - 2 ? ly - 2 : y0;
        x0 = x0 > lx - 2 ? lx - 2 : x0;
        float zf = 0.5f * (float)z - (float)z0, yf = 0.5f * (float)y - (float)y0, fx = 0.5f * (float)x - (float)x0;
        zf = zf > 1.0f ? 1.0f : zf;
        yf = yf > 1.0f ? 1.0f : yf;
        fx = fx > 1.0f ? 1.0f : fx;
        const size_t r00 = ((size_t)z0 * ly + y0) * lx, r01 = r00 + lx, r10 = r00 + (size_t)ly * lx, r11 = r10 + lx;
        const float w00 = (1 - zf) * (1 - yf), w01 = (1 - zf) * yf, w10 = zf * (1 - yf), w11 = zf * yf;
        const float a = w00 * sm[r00 + x0] + w01 * sm[r01 + x0] + w10 * sm[r10 + x0] + w11 * sm[r11 + x0];
        const float b = w00 * sm[r00 + x0 + 1] + w01 * sm[r01 + x0 + 1] + w10 * sm[r10 + x0 + 1] +
                        w11 * sm[r11 + x0 + 1];
        v = a + (b - a) * fx;
    } else {
        v = sm[i];
    }
    float tex = orig[i] - v;
    tex = tex < 0 ? -tex : tex;
    if (v < (tex < 0.03f ? cut.uni : cut.tex)) f[i] = 0.0f;
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
// across the wave, carried along the row.  Also clears the counts.
__global__ __launch_bounds__(NT) void k_cc_init(const float *f, int *L, int *cnt, long nrows, int nx) {
    const int lane = threadIdx.x & 63;
    const long row = (long)blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
    if (row >= nrows) return;                           // whole waves leave together
    const size_t base = (size_t)row * nx;
    int carry = -1;                                     // last gap (x) before this chunk
    for (int x0 = 0; x0 < nx; x0 += 64) {
        const int x = x0 + lane;
        const bool in = x < nx;
        const bool m = in && f[base + x] > 0.0f;
        int g = m ? carry : x;                          // gap positions; carry for material
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(g, o);
            if (lane >= o) g = g > t ? g : t;
        }
        if (in) {
            L[base + x] = m ? (int)(base + g + 1) : -1;
            cnt[base + x] = 0;
        }
        carry = __shfl(g, 63);
    }
}

__device__ __forceinline__ bool cc_mat(const int *L, size_t j) { return L[j] >= 0; }

// join runs that touch along y and z; where the voxel before i and its
// neighbour are both material, that pair already joined the same two runs
__global__ void k_cc_union(int *L, int nz, int ny, int nx, int *err) {
    const size_t plane = (size_t)ny * nx, n = plane * nz;
    for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT) {
        if (!cc_mat(L, i)) continue;
        const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / plane);
        const bool left = x > 0 && cc_mat(L, i - 1);
        if (y > 0 && cc_mat(L, i - nx) && !(left && cc_mat(L, i - 1 - nx))) cc_unite(L, (int)i, (int)(i - nx), err);
        if (z > 0 && cc_mat(L, i - plane) && !(left && cc_mat(L, i - 1 - plane)))
            cc_unite(L, (int)i, (int)(i - plane), err);
    }
}

// Labels to roots in two steps.  Every ancestor of a voxel is a run start (only
// run starts are ever roots), so once the run starts point at their roots, any
// other voxel reaches its root in one hop.
__global__ void k_cc_flatten_starts(int *L, int nx, size_t n) {
    for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT) {
        if (cc_load(L, (int)i) < 0) continue;
        if (i % nx != 0 && cc_load(L, (int)(i - 1)) >= 0) continue;     // not a run start
        L[i] = cc_find(L, (int)i);
    }
}

__global__ void k_cc_flatten_rest(int *L, int nx, size_t n) {
    for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT) {
        const int p = L[i];
        if (p < 0 || i % nx == 0 || L[i - 1] < 0) continue;
        L[i] = L[p];
    }
}

// wave-aggregated update of per-root words: lanes that share a root add (or
// set) once through the first of them.  The loop bound is wave-uniform.
template <bool EDGE>
__device__ __forceinline__ void cc_wave_add(int *cnt, bool act, int r, int min_size) {
    const int lane = threadIdx.x & 63;
    unsigned long long pending = __ballot(act);
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const int r0 = __shfl(r, leader);
        const unsigned long long grp = __ballot(act && r == r0);
        if (lane == leader) {
            if (EDGE) {
                // most face voxels belong to a few components: set the bit once
                if (!(__hip_atomic_load(&cnt[r0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & kEdge))
                    atomicOr(&cnt[r0], kEdge);
            } else {
                const int c = cnt[r0];                  // kept components and met sizes need no count
                if (!(c & kEdge) && c < min_size) atomicAdd(&cnt[r0], __popcll(grp));
            }
        }
        pending &= ~grp;
    }
}

// roots of components that touch a face of the region
__global__ void k_cc_edge(const int *L, int *cnt, int nz, int ny, int nx) {
    const size_t plane = (size_t)ny * nx, n = plane * nz;
    const size_t stride = (size_t)gridDim.x * NT;
    for (size_t i0 = (size_t)blockIdx.x * NT + (threadIdx.x & ~63u); i0 < n; i0 += stride) {
        const size_t i = i0 + (threadIdx.x & 63);
        bool act = false;
        int r = -1;
        if (i < n && (r = L[i]) >= 0) {
            const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / plane);
            act = x == 0 || x == nx - 1 || y == 0 || y == ny - 1 || z == 0 || z == nz - 1;
        }
        cc_wave_add<true>(cnt, act, r, 0);
    }
}

// sizes of the components that do not touch a face, counted up to min_size
__global__ void k_cc_count(const int *L, int *cnt, size_t n, int min_size) {
    const size_t stride = (size_t)gridDim.x * NT;
    for (size_t i0 = (size_t)blockIdx.x * NT + (threadIdx.x & ~63u); i0 < n; i0 += stride) {
        const size_t i = i0 + (threadIdx.x & 63);
        const int r = i < n ? L[i] : -1;
        cc_wave_add<false>(cnt, r >= 0, r, min_size);
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
        const size_t nl = (size_t)lz * ly * lx;
        fyg_launch(s, "air_decimate", k_air_decimate, dim3(grid_for(nl, 4096)), dim3(NT), 0, orig, low, ny, nx, lz,
                   ly, lx);
        gz = lz; gy = ly; gx = lx;
        r = (sr + 1) / 2 < 1 ? 1 : (sr + 1) / 2;
        src = low;
    }
    // x into tmp, y into the decimated copy's buffer (or sm), z into sm
    const size_t ns = (size_t)gz * gy * gx;
    float *mid = lowres ? low : sm;
    auto lines = [&](int len) { return dim3((unsigned)((ns / len + NT - 1) / NT)); };   // one line per thread
    fyg_launch(s, "air_box", k_box_axis, lines(gx), dim3(NT), 0, src, tmp, gz, gy, gx, 0, r);
    fyg_launch(s, "air_box", k_box_axis, lines(gy), dim3(NT), 0, (const float *)tmp, mid, gz, gy, gx, 1, r);
    float *fin = lowres ? sm : tmp;
    fyg_launch(s, "air_box", k_box_axis, lines(gz), dim3(NT), 0, (const float *)mid, fin, gz, gy, gx, 2, r);
    const unsigned hg = grid_for(ns / 16, 1024);
    fyg_launch(s, "air_hist", k_air_hist, dim3(hg), dim3(NT), 0, (const float *)fin, ns, part);
    if (fyg_zero(hist, sizeof(unsigned long long) * 256, s) || fyg_hist_reduce_run(part, (int)hg, hist, s)) return 1;
    CutParams cp{a->cut_u8, a->band, a->min_component, a->thresh, a->aggr};
    fyg_launch(s, "air_cut", k_air_cut, dim3(1), dim3(256), 0, (const unsigned long long *)hist, cp, cut);
    if ((long)nz * ny > 0x7fffffffL || (nx + NT - 1) / NT > 65535) return 1;
    dim3 ag((unsigned)((long)nz * ny), (unsigned)((nx + NT - 1) / NT));
    fyg_launch(s, "air_apply", k_air_apply, ag, dim3(NT), 0, f, orig, (const float *)fin, nz, ny, nx, lz, ly, lx,
               lowres, (const AirCut *)cut);
    if (a->min_component > 0) {                           // pipeline.c:396-398
        const unsigned g = grid_for(n, 16384);
        if (fyg_zero(err, sizeof(int), s)) return 1;
        const long nrows = (long)nz * ny;
        fyg_launch(s, "cc_init", k_cc_init, dim3((unsigned)((nrows + NT / 64 - 1) / (NT / 64))), dim3(NT), 0,
                   (const float *)f, L, cnt, nrows, nx);
        fyg_launch(s, "cc_union", k_cc_union, dim3(g), dim3(NT), 0, L, nz, ny, nx, err);
        fyg_launch(s, "cc_flatten", k_cc_flatten_starts, dim3(g), dim3(NT), 0, L, nx, n);
        fyg_launch(s, "cc_flatten", k_cc_flatten_rest, dim3(g), dim3(NT), 0, L, nx, n);
        fyg_launch(s, "cc_edge", k_cc_edge, dim3(g), dim3(NT), 0, (const int *)L, cnt, nz, ny, nx);
        fyg_launch(s, "cc_count", k_cc_count, dim3(g), dim3(NT), 0, (const int *)L, cnt, n, a->min_component);
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
