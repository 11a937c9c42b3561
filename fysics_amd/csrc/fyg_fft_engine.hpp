// fyg_fft_engine.hpp -- the FFT engine and the deconvolution passes, templated
// on the working precision P (float for the product path, double for the
// all-FP64 variant).  Included by fyg_fft.hip (P = float, plus the complex
// FFTs of fft.c) and fyg_fft64.hip (P = double), so the two precisions compile
// in parallel.
//
// Engine: batched Stockham FFT of length L (2^k or 3*2^k, L <= 4096).  Each
// workgroup owns NL lines.  Stage s (radix R, span NS) reads x[j + r*L/R],
// multiplies by w^(r*(j%NS)) with w = exp(sign*2*pi*i/(NS*R)), runs the R-point
// DFT in registers and writes y[(j/NS)*NS*R + j%NS + r*NS]; the first stage
// reads HBM, the last writes HBM, the ones between go through LDS (in place:
// read, barrier, write, barrier).  Radices are 16 where possible (2 stages for
// 256), so a 256-line makes a single LDS round trip.  Twiddles are copies of
// double-evaluated cos/sin (the fft.c:47-61 scheme, fp32 for P = float),
// staged in LDS.
//
// Passes of one deconvolution over the half spectrum S[z][y][k] (pitch hxp):
//   k_rfft_x    rows (y, y+1) of a reflect-padded z plane packed re/im, one
//               complex FFT, untangled to two half-spectrum rows; the u8 tile
//               prologue (f = u8/255, rings, z-drift, dewindow) is fused into
//               the gather                                   (paganin.c:166-194)
//   k_cfft_cols Y forward over all z                         (paganin.c:197-203)
//   k_zfused    Z forward, LUT weight, Z inverse in one kernel: the weight is
//               pointwise, so the spectrum never returns to HBM between the
//               forward and inverse Z transforms             (paganin.c:204-253)
//   k_cfft_cols Y inverse for z < nz                         (paganin.c:254-260)
//   k_irfft_x   Hermitian repack of rows (y, y+1), complex inverse, 1/(pz*py*px),
//               crop, fused rewindow / dec-range / clamp     (paganin.c:263-289)
#pragma once
#include <cstdio>
#include <type_traits>
#include <vector>

#include "fyg_internal.hpp"

namespace {

// ------------------------------------------------------------------ precision
template <class P>
struct Cx;
template <>
struct Cx<float> {
    using V = float2;
    static __host__ __device__ __forceinline__ float2 make(float x, float y) { return make_float2(x, y); }
};
template <>
struct Cx<double> {
    using V = double2;
    static __host__ __device__ __forceinline__ double2 make(double x, double y) { return make_double2(x, y); }
};
template <class V>
struct ScalarOf;
template <>
struct ScalarOf<float2> {
    using T = float;
};
template <>
struct ScalarOf<double2> {
    using T = double;
};
template <class V>
using scal_t = typename ScalarOf<V>::T;

// ------------------------------------------------------------------ plans
struct Radices {
    int n;
    int r[10];
};

constexpr int ilog2c(int x) {
    int l = 0;
    while ((1 << l) < x) l++;
    return l;
}

// The radix plan of a length: a factor 3 first when there is one, then
//   pv 0: radix-16 stages behind one smaller radix (2, 4, 8, or 4 and 8 where
//         a lone 2 would be left);
//   pv 1: radix-8 stages behind one radix 2 or 4.
// Both are exact factorizations; they round differently, so one length's plan
// is fixed before its first use in a process.
constexpr Radices make_plan(int L, int pv = 0) {
    Radices p{};
    p.n = 0;
    int m = L;
    if (m % 3 == 0) {
        p.r[p.n++] = 3;
        m /= 3;
    }
    int lg = ilog2c(m);
    if (pv == 1) {
        if (lg % 3) p.r[p.n++] = 1 << (lg % 3);
        for (int i = 0; i < lg / 3; i++) p.r[p.n++] = 8;
    } else {
        int q = lg / 4, rem = lg % 4;
        if (rem == 1 && q >= 1) {
            p.r[p.n++] = 4;
            p.r[p.n++] = 8;
            q -= 1;
        } else if (rem) {
            p.r[p.n++] = 1 << rem;
        }
        for (int i = 0; i < q; i++) p.r[p.n++] = 16;
    }
    if (p.n == 0) p.r[p.n++] = 1;
    return p;
}

constexpr bool same_plan(const Radices &a, const Radices &b) {
    if (a.n != b.n) return false;
    for (int i = 0; i < a.n; i++)
        if (a.r[i] != b.r[i]) return false;
    return true;
}

// lengths with a second plan to choose from (fp32 only): the cube sizes of the
// batched and served forms; elsewhere pv 1 means pv 0
constexpr bool plan_alt(int L) { return L >= 64 && L <= 512 && !same_plan(make_plan(L, 0), make_plan(L, 1)); }

// stage s of the forward plan (REV=false) or of the reversed plan (REV=true)
template <int L, bool REV, int PV = 0>
struct Plan {
    static constexpr Radices P = make_plan(L, PV);
    static constexpr int n = P.n;
    static constexpr int R(int s) { return REV ? P.r[P.n - 1 - s] : P.r[s]; }
    static constexpr int NS(int s) {
        int v = 1;
        for (int i = 0; i < s; i++) v *= R(i);
        return v;
    }
};

// lines per workgroup: column passes keep 16 contiguous fp32 columns (128-byte
// rows; 8 in fp64), row passes fill ~4096 fp32 elements (2048 in fp64), so a
// workgroup holds the same LDS bytes in either precision
template <int L, class P = float>
struct Shape {
    static constexpr int T = 256;
    static constexpr int Q = sizeof(P) / sizeof(float);
    static constexpr int NL_CM = (L <= 256 ? 16 : (4096 / L)) / Q < 1 ? 1 : (L <= 256 ? 16 : (4096 / L)) / Q;
    // the fused Z pass keeps full 128-byte rows at L = 512 too (2 workgroups
    // of 74 KB per CU): it measured faster than half rows at 4 per CU
    static constexpr int NL_ZF = L == 512 ? (16 / Q < 1 ? 1 : 16 / Q) : NL_CM;
    static constexpr int RM0 = (4096 / Q) / L;
    static constexpr int NL_RM = RM0 > 256 ? 256 : RM0 < 1 ? 1 : RM0;
    static constexpr int LD = L + (L >> 4) + 1;   // padded line pitch in LDS (complex elements)
};

__device__ __forceinline__ int lds_pos(int pos) { return pos + (pos >> 4); }

// ------------------------------------------------------------------ small DFTs
template <class V>
__device__ __forceinline__ V f2add(V a, V b) { return V(a.x + b.x, a.y + b.y); }
template <class V>
__device__ __forceinline__ V f2sub(V a, V b) { return V(a.x - b.x, a.y - b.y); }
// a * (sg * i)
template <class V>
__device__ __forceinline__ V f2rot(V a, scal_t<V> sg) { return V(-sg * a.y, sg * a.x); }
// a * (c + i*sg*s)
template <class V>
__device__ __forceinline__ V f2tw(V a, scal_t<V> c, scal_t<V> s, scal_t<V> sg) {
    scal_t<V> ss = sg * s;
    return V(a.x * c - a.y * ss, a.x * ss + a.y * c);
}

template <int R>
struct DFT;

template <>
struct DFT<1> {
    template <class V>
    static __device__ __forceinline__ void run(V *, scal_t<V>) {}
};

template <>
struct DFT<2> {
    template <class V>
    static __device__ __forceinline__ void run(V *v, scal_t<V>) {
        V a = v[0], b = v[1];
        v[0] = f2add(a, b);
        v[1] = f2sub(a, b);
    }
};

template <>
struct DFT<3> {
    template <class V>
    static __device__ __forceinline__ void run(V *v, scal_t<V> sg) {
        using S = scal_t<V>;
        const S c3 = (S)0.86602540378443864676;
        const S h = (S)0.5;
        V t = f2add(v[1], v[2]);
        V m = V(v[0].x - h * t.x, v[0].y - h * t.y);
        V d = f2sub(v[1], v[2]);
        d = V(-sg * c3 * d.y, sg * c3 * d.x);
        v[0] = f2add(v[0], t);
        v[1] = f2add(m, d);
        v[2] = f2sub(m, d);
    }
};

template <>
struct DFT<4> {
    template <class V>
    static __device__ __forceinline__ void run(V *v, scal_t<V> sg) {
        V t0 = f2add(v[0], v[2]), t1 = f2sub(v[0], v[2]);
        V t2 = f2add(v[1], v[3]), t3 = f2rot(f2sub(v[1], v[3]), sg);
        v[0] = f2add(t0, t2);
        v[2] = f2sub(t0, t2);
        v[1] = f2add(t1, t3);
        v[3] = f2sub(t1, t3);
    }
};

template <>
struct DFT<8> {
    template <class V>
    static __device__ __forceinline__ void run(V *v, scal_t<V> sg) {
        using S = scal_t<V>;
        const S c = (S)0.70710678118654752440;
        V e[4] = {v[0], v[2], v[4], v[6]};
        V o[4] = {v[1], v[3], v[5], v[7]};
        DFT<4>::run(e, sg);
        DFT<4>::run(o, sg);
        o[1] = f2tw(o[1], c, c, sg);
        o[2] = f2rot(o[2], sg);
        o[3] = f2tw(o[3], -c, c, sg);
        for (int k = 0; k < 4; k++) {
            v[k] = f2add(e[k], o[k]);
            v[k + 4] = f2sub(e[k], o[k]);
        }
    }
};

// W16 constants: C[m] = cos(2*pi*m/16)
template <>
struct DFT<16> {
    template <class V>
    static __device__ __forceinline__ void run(V *v, scal_t<V> sg) {
        using S = scal_t<V>;
        constexpr S C[16] = {
            (S)1.0, (S)0.92387953251128675613, (S)0.70710678118654752440, (S)0.38268343236508977173,
            (S)0.0, (S)-0.38268343236508977173, (S)-0.70710678118654752440, (S)-0.92387953251128675613,
            (S)-1.0, (S)-0.92387953251128675613, (S)-0.70710678118654752440, (S)-0.38268343236508977173,
            (S)0.0, (S)0.38268343236508977173, (S)0.70710678118654752440, (S)0.92387953251128675613};
        V y[4][4];
#pragma unroll
        for (int n2 = 0; n2 < 4; n2++) {
            y[n2][0] = v[n2];
            y[n2][1] = v[n2 + 4];
            y[n2][2] = v[n2 + 8];
            y[n2][3] = v[n2 + 12];
            DFT<4>::run(y[n2], sg);
        }
#pragma unroll
        for (int n2 = 1; n2 < 4; n2++)
#pragma unroll
            for (int k1 = 1; k1 < 4; k1++) {
                const int m = n2 * k1;
                // W16^m = cos(2 pi m/16) + i*sg*sin(2 pi m/16); sin(x) = cos(x - pi/2)
                y[n2][k1] = f2tw(y[n2][k1], C[m & 15], C[(m + 12) & 15], sg);
            }
#pragma unroll
        for (int k1 = 0; k1 < 4; k1++) {
            V z[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
            DFT<4>::run(z, sg);
#pragma unroll
            for (int k2 = 0; k2 < 4; k2++) v[k1 + 4 * k2] = z[k2];
        }
    }
};

template <int R, class V>
__device__ __forceinline__ void dft(V *v, scal_t<V> sg) {
    DFT<R>::run(v, sg);
}

// ------------------------------------------------------------------ stage engine
// item i -> (line, butterfly j); CM: consecutive threads walk lines (columns),
// RM: consecutive threads walk j within a line (contiguous rows)
template <int NL, int NB, bool CM>
__device__ __forceinline__ void item_of(int i, int &l, int &j) {
    if (CM) {
        l = i % NL;
        j = i / NL;
    } else {
        l = i / NB;
        j = i % NB;
    }
}

// one Stockham stage.  The loader ld(l, p0, off) returns element p0 + off of line
// l and the storer st(l, p0, off, v) writes it: p0 is the item's base position
// and off = r*L/R (loads) or r*NS (stores) a compile-time constant, so address
// work that depends only on (l, p0) is done once per item.  SYNC: barrier
// between the read and the write phase (needed when both touch the same LDS
// buffer).
template <int L, int NL, int T, bool CM, int R, int NS, bool SYNC, class V, class LdF, class StF>
__device__ __forceinline__ void stage(scal_t<V> sg, const V *__restrict__ tw, LdF ld, StF st) {
    constexpr int NB = L / R;
    constexpr int TOT = NL * NB;
    constexpr int BPT = (TOT + T - 1) / T;
    constexpr int TWS = L / (NS * R);
    V v[BPT][R];
#pragma unroll
    for (int b = 0; b < BPT; b++) {
        int i = threadIdx.x + b * T;
        if (TOT % T == 0 || i < TOT) {
            int l, j;
            item_of<NL, NB, CM>(i, l, j);
#pragma unroll
            for (int r = 0; r < R; r++) v[b][r] = ld(l, j, r * NB);
        }
    }
    if (SYNC) __syncthreads();
#pragma unroll
    for (int b = 0; b < BPT; b++) {
        int i = threadIdx.x + b * T;
        if (TOT % T == 0 || i < TOT) {
            int l, j;
            item_of<NL, NB, CM>(i, l, j);
            const int k = j % NS;
            if (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    V w = tw[r * k * TWS];
                    v[b][r] = f2tw(v[b][r], w.x, w.y, sg);
                }
            }
            dft<R>(v[b], sg);
            const int ob = (j / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) st(l, ob, r * NS, v[b][r]);
        }
    }
}

template <int L, int NL, class V>
struct LdsRd {
    const V *buf;
    __device__ V operator()(int l, int p0, int off) const { return buf[l * Shape<L>::LD + lds_pos(p0 + off)]; }
};
template <int L, int NL, class V>
struct LdsWr {
    V *buf;
    __device__ void operator()(int l, int p0, int off, V v) const { buf[l * Shape<L>::LD + lds_pos(p0 + off)] = v; }
};

struct NoHook {
    __device__ void operator()() const {}
};

// stages [S, END) of a plan (END < 0: all); stage 0 reads `first`, the final
// stage writes `last`, stages between go through LDS.  `hook` runs after the
// first stage's stores and before its barrier: stage 0 applies no twiddles
// (span 1), so the kernels store their twiddle table there and its load does
// not delay the first data loads.
template <int L, int NL, bool CM, bool REV, int S, int END = -1, int PV = 0, class V, class First, class Last,
          class Hook = NoHook>
__device__ __forceinline__ void run_from(scal_t<V> sg, const V *tw, V *buf, First first, Last last,
                                         Hook hook = Hook{}) {
    using PL = Plan<L, REV, PV>;
    constexpr int E = END < 0 ? PL::n : END;
    constexpr int T = Shape<L>::T;
    if constexpr (S < E) {
        constexpr int R = PL::R(S), NS = PL::NS(S);
        constexpr bool FIRST = S == 0, FINAL = S == E - 1;
        static_assert(!FIRST || NS == 1, "the first stage applies no twiddles");
        if constexpr (FIRST && FINAL) {
            stage<L, NL, T, CM, R, NS, false>(sg, tw, first, last);
            hook();
        } else if constexpr (FIRST) {
            stage<L, NL, T, CM, R, NS, false>(sg, tw, first, LdsWr<L, NL, V>{buf});
            hook();
            __syncthreads();
            run_from<L, NL, CM, REV, S + 1, END, PV>(sg, tw, buf, first, last);
        } else if constexpr (FINAL) {
            // a final stage that stores back into the same LDS lines (a partial
            // plan, END < n) needs the barrier between its read and write phases
            constexpr bool IN_PLACE = std::is_same<Last, LdsWr<L, NL, V>>::value;
            stage<L, NL, T, CM, R, NS, IN_PLACE>(sg, tw, LdsRd<L, NL, V>{buf}, last);
        } else {
            stage<L, NL, T, CM, R, NS, true>(sg, tw, LdsRd<L, NL, V>{buf}, LdsWr<L, NL, V>{buf});
            __syncthreads();
            run_from<L, NL, CM, REV, S + 1, END, PV>(sg, tw, buf, first, last);
        }
    }
}

template <int L, class V>
__device__ __forceinline__ void load_twiddles(V *tw_s, const V *__restrict__ tw) {
    for (int i = threadIdx.x; i < L; i += Shape<L>::T) tw_s[i] = tw[i];
}

// a twiddle table loaded into registers at kernel start (L <= 1024: at most 4
// per thread) and stored to LDS later, as run_from's hook
template <int L, class V>
struct TwPrefetch {
    static constexpr int T = Shape<L>::T, K = L <= 1024 ? (L + T - 1) / T : 0;
    V r[K > 0 ? K : 1];
    const V *src;
    __device__ explicit TwPrefetch(const V *__restrict__ tw) : src(tw) {
        if constexpr (K > 0) {
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int i = threadIdx.x + k * T;
                if (i < L) r[k] = tw[i];
            }
        }
    }
    __device__ void store(V *tw_s) const {
        if constexpr (K > 0) {
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int i = threadIdx.x + k * T;
                if (i < L) tw_s[i] = r[k];
            }
        } else {
            load_twiddles<L>(tw_s, src);
        }
    }
};

__device__ __forceinline__ int reflect_idx(int i, int n) { return i < n ? i : 2 * n - 1 - i; }

// u8 tile prologue: f = u8/255 -> ring subtraction -> z-drift -> dewindow
// (pipeline.c:492-494,266-268).  c = scale*ring[slab][radius] comes from the
// ring correction image (0 without rings); the subtraction applies to f > 0
// only and clamps at 0 (dering.c:226-230).  Select form, so the unrolled gather
// keeps all of its loads in flight.  Evaluated in fp32 in both precisions: it
// is the reference's input conversion.
__device__ __forceinline__ float prologue(const fyg_prologue &P, unsigned char u, float c, float zf) {
    #pragma clang fp contract(off)
    float f = (float)u * (1.0f / 255.0f);
    float v = f - c;
    v = v > 0.0f ? v : 0.0f;
    f = f > 0.0f ? v : f;
    f *= zf;
    float w = P.wlo + f * P.wspan;
    return P.window ? w : f;
}

__host__ __device__ __forceinline__ int ring_slab(const fyg_prologue &P, long gz) {
    long slab = gz / P.slab_z;
    if (slab < 0) slab = 0;
    if (slab >= P.nslab) slab = P.nslab - 1;
    return (int)slab;
}

// ring correction image of a buffer region: img[s][y][x] = scale*ring[s0+s][ri]
// for the slabs s0.. the region's planes fall in, ri = (int)(sqrt(dy^2+dx^2)+0.5)
// in double as dering.c:207-213 evaluates it, 0 where ri >= nr.  The X pass then
// subtracts one coalesced, cache-resident float per voxel instead of
// evaluating a radius and gathering from the ring table.
struct RingImg {
    fyg_prologue P;
    int ny, nx, s0, nimg;
};

// voxel x of row y of image si
__device__ __forceinline__ void ring_image_at(const RingImg &r, int x, int y, int si) {
    #pragma clang fp contract(off)
    const fyg_prologue &P = r.P;
    if (x >= r.nx || y >= r.ny || si >= r.nimg) return;
    double dy = (double)(P.gy0 + y) - P.cy, dx = (double)(P.gx0 + x) - P.cx;
    double dy2 = dy * dy;
    int ri = (int)(sqrt(dy2 + dx * dx) + 0.5);
    float c = 0.0f;
    if (ri < P.nr) c = P.scale * P.ring[(long)(r.s0 + si) * P.nr + ri];
    P.cimg[((long)si * r.ny + y) * r.nx + x] = c;
}

// grid: (x / 256, y, cube * mslab + slab).  step: the call's generation word
// to step (workgroup 0, before any thread returns), or null
__global__ __launch_bounds__(256) void k_ring_image(Batch<RingImg> b, int mslab, int *step, int nflag) {
    if (step && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0) fyg_gen_step_block(step, nflag);
    ring_image_at(b.a[blockIdx.z / mslab], blockIdx.x * 256 + threadIdx.x, blockIdx.y, blockIdx.z % mslab);
}

// rewindow, dec-range rescale, clamp (pipeline.c:273-283), in the working precision
template <class P>
__device__ __forceinline__ P epilogue(const fyg_epilogue &E, P v) {
    #pragma clang fp contract(off)
    if (E.window) v = (v - (P)E.wlo) * (P)E.winv;
    if (E.dec) v = (v - (P)E.dlo) * (P)E.dinv;
    if (E.clamp01) v = v < (P)0 ? (P)0 : (v > (P)1 ? (P)1 : v);
    return v;
}

// ------------------------------------------------------------------ kernels
template <class P>
struct XFwd {
    using V = typename Cx<P>::V;
    const void *in;
    int nz, ny, nx, pz, py, hx, hxp;
    long nlines, npairs;
    V *spec;
    const V *tw;
    fyg_prologue pro;
    int *flag;
    const int *gen;
};

// per-line constants of a row-pair line, computed once per workgroup
struct XLine {
    long a, b;          // input offsets of rows y0 and y0+1 (b == a: no second row)
    long ca, cb;        // ring correction image offsets of the two rows
    long so;            // half-spectrum offset of row y0
    float zf;           // z-drift factor of the plane
    int pair, valid;
};

// LDS of one row-pass workgroup (X forward and X inverse)
template <int L, class P, class Line>
struct RowLds {
    using V = typename Cx<P>::V;
    static constexpr int NL = Shape<L, P>::NL_RM;
    V buf[NL * Shape<L>::LD];
    V tw_s[L];
    Line meta[NL];
};

// MODE 0: fp32 input; 1: u8 input with the prologue; 2: as 1 plus rings.
// Workgroup bx of one cube's X forward pass.
template <int L, int MODE, class P, int PV = 0>
__device__ __forceinline__ void rfft_x_body(const XFwd<P> &a, long bx, RowLds<L, P, XLine> &sm) {
    using V = typename Cx<P>::V;
    constexpr int NL = Shape<L, P>::NL_RM;
    constexpr bool U8 = MODE != 0, RING = MODE == 2;
    V *buf = sm.buf, *tw_s = sm.tw_s;
    XLine *meta = sm.meta;
    const TwPrefetch<L, V> twp(a.tw);
    const long g0 = bx * NL;
    int s0 = 0;
    if (RING) s0 = ring_slab(a.pro, a.pro.gz0);
    for (int l = threadIdx.x; l < NL; l += Shape<L>::T) {
        XLine m;
        long g = g0 + l;
        m.valid = g < a.nlines;
        if (!m.valid) g = 0;
        int z = (int)(g / a.npairs);
        int y0 = 2 * (int)(g - (long)z * a.npairs);
        int rz = reflect_idx(z, a.nz), ry0 = reflect_idx(y0, a.ny);
        m.pair = y0 + 1 < a.py;
        int ry1 = m.pair ? reflect_idx(y0 + 1, a.ny) : ry0;
        m.a = ((long)rz * a.ny + ry0) * a.nx;
        m.b = ((long)rz * a.ny + ry1) * a.nx;
        m.so = ((long)z * a.py + y0) * a.hxp;
        long gz = a.pro.gz0 + rz;
        m.zf = (U8 && a.pro.zdrift) ? a.pro.zfactor[gz] : 1.0f;
        long cs = RING ? (long)(ring_slab(a.pro, gz) - s0) * a.ny : 0;
        m.ca = (cs + ry0) * a.nx;
        m.cb = (cs + ry1) * a.nx;
        meta[l] = m;
    }
    __syncthreads();
    int any = 0;
    auto first = [&](int l, int p0, int off) -> V {
        const XLine &m = meta[l];
        int rx = reflect_idx(p0 + off, a.nx);
        float va, vb;
        if (U8) {
            const unsigned char *u = (const unsigned char *)a.in;
            unsigned char ua = u[m.a + rx], ub = u[m.b + rx];
            float ca = 0.0f, cb = 0.0f;
            if (RING) {
                ca = a.pro.cimg[m.ca + rx];
                cb = a.pro.cimg[m.cb + rx];
            }
            any |= m.valid ? (ua | ub) : 0;
            va = prologue(a.pro, ua, ca, m.zf);
            vb = prologue(a.pro, ub, cb, m.zf);
        } else {
            const float *f = (const float *)a.in;
            va = f[m.a + rx];
            vb = f[m.b + rx];
        }
        // No selects on the loaded values: a line past the end transforms line 0
        // again and is not stored; without a second row the imaginary part
        // repeats row y0, which the untangle separates from the real part exactly
        // as it would separate zeros, and B is not stored.
        return Cx<P>::make((P)va, (P)vb);
    };
    run_from<L, NL, false, false, 0, -1, PV>((P)-1, tw_s, buf, first, LdsWr<L, NL, V>{buf}, [&] { twp.store(tw_s); });
    __syncthreads();
    if (U8 && a.flag) {
        if (__syncthreads_or(any) && threadIdx.x == 0) fyg_flag_set(a.flag, a.gen);
    }
    // untangle A(k) = (X(k) + conj X(-k))/2, B(k) = (X(k) - conj X(-k))/(2i)
    const int dl = Shape<L>::T / a.hxp, dk = Shape<L>::T - dl * a.hxp;
    int l = threadIdx.x / a.hxp, k = threadIdx.x - l * a.hxp;
    const P h = (P)0.5;
    for (; l < NL; l += dl, k += dk) {
        if (k >= a.hxp) { k -= a.hxp; l++; if (l >= NL) break; }
        const XLine &m = meta[l];
        // pitch padding k >= hx is never read by the later passes: not written
        if (!m.valid || k >= a.hx) continue;
        int nk = k ? L - k : 0;
        V p = buf[l * Shape<L>::LD + lds_pos(k)];
        V q = buf[l * Shape<L>::LD + lds_pos(nk)];
        V A = Cx<P>::make(h * (p.x + q.x), h * (p.y - q.y));
        V B = Cx<P>::make(h * (p.y + q.y), h * (q.x - p.x));
        a.spec[m.so + k] = A;
        if (m.pair) a.spec[m.so + a.hxp + k] = B;
    }
}

// grid: (line groups, cube)
template <int L, int MODE, class P, int PV = 0>
__global__ __launch_bounds__(256) void k_rfft_x(Batch<XFwd<P>> bt) {
    __shared__ RowLds<L, P, XLine> sm;
    const XFwd<P> a = bt.a[blockIdx.y];             // into registers once
    rfft_x_body<L, MODE, P, PV>(a, blockIdx.x, sm);
}

template <class P>
struct Cols {
    using V = typename Cx<P>::V;
    V *data;
    long ncols, estride, bstride;
    P sg;
    const V *tw;
};

// Lines longer than 256 leave a workgroup 8 fp32 columns (64 bytes of each
// 128-byte row line; 4 fp64 columns past 256 in double), so two neighbouring
// workgroups share every line.
// Workgroups are dispatched to the 8 XCDs round-robin, each XCD with its own
// L2; pair_block renumbers them so that the workgroups of logical groups 2k
// and 2k+1 are dispatched 8 apart, i.e. together on one XCD, and the second
// one's half of each line is an L2 hit.  Identity on a grid's last partial 16.
__device__ __forceinline__ long pair_block(long h, long T) {
    const long c = h >> 4, i = h & 15;
    if ((c + 1) << 4 > T) return h;
    return (c << 4) + 2 * (i & 7) + (i >> 3);
}

// LDS of one column-pass workgroup
template <int L, class P>
struct ColLds {
    using V = typename Cx<P>::V;
    static constexpr int NL = Shape<L, P>::NL_CM;
    V buf[NL * Shape<L>::LD];
    V tw_s[L];
};

// workgroup (bx, by) of one cube's column pass: column group bx of batch by
template <int L, class P, int PV = 0>
__device__ __forceinline__ void cols_body(const Cols<P> &a, long bx, long by, ColLds<L, P> &sm) {
    using V = typename Cx<P>::V;
    constexpr int NL = Shape<L, P>::NL_CM;
    V *buf = sm.buf, *tw_s = sm.tw_s;
    const TwPrefetch<L, V> twp(a.tw);
    const long c0 = bx * NL;
    V *base = a.data + (size_t)by * a.bstride + c0;
    const long nc = a.ncols - c0;
    const size_t es = (size_t)a.estride;
    // Columns past ncols (the last group of a row) load column nc-1 instead --
    // in bounds, no branch around the load -- and are never stored; the lines
    // of a workgroup are independent transforms.
    auto first = [&](int l, int p0, int off) -> V {
        int lc = l < nc ? l : (int)nc - 1;
        return base[(size_t)p0 * es + lc + (size_t)off * es];
    };
    auto last = [&](int l, int p0, int off, V v) {
        if (l < nc) base[(size_t)p0 * es + l + (size_t)off * es] = v;
    };
    auto hook = [&] { twp.store(tw_s); };
    if (a.sg < 0)
        run_from<L, NL, true, false, 0, -1, PV>((P)-1, tw_s, buf, first, last, hook);
    else
        run_from<L, NL, true, false, 0, -1, PV>((P)1, tw_s, buf, first, last, hook);
}

// grid: (column groups, batches, cube)
template <int L, class P, int PV = 0>
__global__ __launch_bounds__(256) void k_cfft_cols(Batch<Cols<P>> bt) {
    constexpr bool PAIR = Shape<L, P>::NL_CM * 2 * sizeof(P) < 128;
    long bx = blockIdx.x, by = blockIdx.y;
    if (PAIR) {
        const long T = (long)gridDim.x * gridDim.y;
        const long h = pair_block(blockIdx.x + (long)gridDim.x * blockIdx.y, T);
        bx = h % gridDim.x;
        by = h / gridDim.x;
    }
    __shared__ ColLds<L, P> sm;
    cols_body<L, P, PV>(bt.a[blockIdx.z], bx, by, sm);
}

// the LUT holds fp32 weights for P = float (the reference table) and the same
// weights evaluated in double for P = double
template <class P>
struct ZFused {
    using V = typename Cx<P>::V;
    V *data;
    int py, hx, hxp;
    long estride;               // py*hxp
    const P *lut;
    double us;
    const double *fz2, *fy2, *fx2;
    const V *tw;
};

// LDS of one fused Z workgroup
template <int L, class P>
struct ZLds {
    using V = typename Cx<P>::V;
    static constexpr int NL = Shape<L, P>::NL_ZF;
    V buf[NL * Shape<L>::LD];
    V tw_s[L];
    double s_fy[NL], s_fx[NL];
    int s_ok[NL];
};

// forward Z (plan order), weight, inverse Z (reversed plan), all in one
// workgroup: column group bx of one cube
template <int L, class P, int PV = 0>
__device__ __forceinline__ void zfused_body(const ZFused<P> &a, long bx, ZLds<L, P> &sm) {
    using V = typename Cx<P>::V;
    constexpr int NL = Shape<L, P>::NL_ZF;
    constexpr int T = Shape<L>::T;
    using PF = Plan<L, false, PV>;
    constexpr int n = PF::n;
    constexpr int RL = PF::R(n - 1);          // last forward radix == first inverse radix
    constexpr int NSL = PF::NS(n - 1);
    V *buf = sm.buf, *tw_s = sm.tw_s;
    double *s_fy = sm.s_fy, *s_fx = sm.s_fx;
    int *s_ok = sm.s_ok;
    // the table's loads overlap the per-column metadata loads that the barrier
    // below waits for anyway
    load_twiddles<L>(tw_s, a.tw);
    const long c0 = bx * NL;
    if (threadIdx.x < NL) {
        long c = c0 + threadIdx.x;
        int y = (int)(c / a.hxp), k = (int)(c - (long)y * a.hxp);
        int ok = y < a.py && k < a.hx;     // padding columns k >= hx are neither read nor written
        s_ok[threadIdx.x] = ok;
        s_fy[threadIdx.x] = ok ? a.fy2[y] : 0.0;
        s_fx[threadIdx.x] = ok ? a.fx2[k] : 0.0;
    }
    __syncthreads();
    V *base = a.data + c0;
    const size_t es = (size_t)a.estride;
    auto valid = [&](int l) { return s_ok[l] != 0; };
    // Pitch-padding columns (k >= hx) of a group are transformed and not
    // stored.  When NL does not divide py*hxp (lines of 384 and 768), the last
    // group reaches past the plane's columns: those lines load its last column.
    const long left = (long)a.py * a.hxp - c0;
    const int nv = left < NL ? (int)left : NL;
    auto first = [&](int l, int p0, int off) -> V {
        const int lc = l < nv ? l : nv - 1;
        return base[(size_t)p0 * es + lc + (size_t)off * es];
    };
    auto last = [&](int l, int p0, int off, V v) {
        if (valid(l)) base[(size_t)p0 * es + l + (size_t)off * es] = v;
    };
    // LUT weight of deconv_radial (paganin.c:231-242) at z = pos:
    // t = ((fz2[z] + fy2[y]) + fx2[k]) * us, linear interpolation between knots.
    // The knots an item needs are loaded before the forward stages, so their
    // latency hides behind the spectrum's loads instead of following them.
    // t and its knot are formed in double, as the reference does: the knot
    // index is part of the weight's semantics.
    constexpr int NB = L / RL;
    constexpr int TOT = NL * NB;
    constexpr int BPT = (TOT + T - 1) / T;
    P k0[BPT][RL], k1[BPT][RL], kf[BPT][RL];
#pragma unroll
    for (int b = 0; b < BPT; b++) {
        int i = threadIdx.x + b * T;
        if (TOT % T == 0 || i < TOT) {
            int l, j;
            item_of<NL, NB, true>(i, l, j);
            const double fy = s_fy[l], fx = s_fx[l];
#pragma unroll
            for (int r = 0; r < RL; r++) {
                double fzy = a.fz2[j + r * NB] + fy;
                double t = (fzy + fx) * a.us;
                int ti = (int)t;
                if (ti >= 4096) ti = 4095;
                k0[b][r] = a.lut[ti];
                k1[b][r] = a.lut[ti + 1];
                kf[b][r] = (P)(t - (double)ti);
            }
        }
    }
    // Forward stages 0 .. n-2 (HBM -> LDS), then one fused step: the last forward
    // stage (radix RL, span NSL) leaves item j's outputs at positions j + r*L/RL --
    // exactly the inputs of the first inverse stage (radix RL, span 1) for the same
    // item -- so the weight multiply and that inverse stage run in registers and
    // the spectrum makes a single LDS round trip around them.
    V v[BPT][RL];
    if constexpr (n > 1) {
        run_from<L, NL, true, false, 0, n - 1, PV>((P)-1, tw_s, buf, first, LdsWr<L, NL, V>{buf});
        __syncthreads();
    }
#pragma unroll
    for (int b = 0; b < BPT; b++) {
        int i = threadIdx.x + b * T;
        if (TOT % T == 0 || i < TOT) {
            int l, j;
            item_of<NL, NB, true>(i, l, j);
#pragma unroll
            for (int r = 0; r < RL; r++) {
                if constexpr (n > 1) v[b][r] = buf[l * Shape<L>::LD + lds_pos(j + r * NB)];
                else v[b][r] = first(l, j, r * NB);
            }
            if constexpr (NSL > 1) {
                constexpr int TWS = L / (NSL * RL);
                const int k = j % NSL;
#pragma unroll
                for (int r = 1; r < RL; r++) {
                    V w = tw_s[r * k * TWS];
                    v[b][r] = f2tw(v[b][r], w.x, w.y, (P)-1);
                }
            }
            dft<RL>(v[b], (P)-1);
#pragma unroll
            for (int r = 0; r < RL; r++) {
                P w = k0[b][r] + (k1[b][r] - k0[b][r]) * kf[b][r];
                v[b][r].x *= w;
                v[b][r].y *= w;
            }
            dft<RL>(v[b], (P)1);
        }
    }
    if constexpr (n > 1) {
        __syncthreads();
#pragma unroll
        for (int b = 0; b < BPT; b++) {
            int i = threadIdx.x + b * T;
            if (TOT % T == 0 || i < TOT) {
                int l, j;
                item_of<NL, NB, true>(i, l, j);
#pragma unroll
                for (int r = 0; r < RL; r++) buf[l * Shape<L>::LD + lds_pos(j * RL + r)] = v[b][r];
            }
        }
        __syncthreads();
        run_from<L, NL, true, true, 1, -1, PV>((P)1, tw_s, buf, first, last);
    } else {
#pragma unroll
        for (int b = 0; b < BPT; b++) {
            int i = threadIdx.x + b * T;
            if (TOT % T == 0 || i < TOT) {
                int l, j;
                item_of<NL, NB, true>(i, l, j);
#pragma unroll
                for (int r = 0; r < RL; r++) last(l, j * RL, r, v[b][r]);
            }
        }
    }
}

// grid: (column groups, cube)
template <int L, class P, int PV = 0>
__global__ __launch_bounds__(256) void k_zfused(Batch<ZFused<P>> bt) {
    const long bx = Shape<L, P>::NL_ZF * 2 * sizeof(P) < 128 ? pair_block(blockIdx.x, gridDim.x) : (long)blockIdx.x;
    __shared__ ZLds<L, P> sm;
    zfused_body<L, P, PV>(bt.a[blockIdx.y], bx, sm);
}

template <class P>
struct XInv {
    using V = typename Cx<P>::V;
    const V *spec;
    P *out;
    int nz, ny, nx, py, hx, hxp;
    long nlines, npairs;
    P invn;
    fyg_epilogue epi;
    const V *tw;
};

struct XInvLine {
    long so;        // half-spectrum offset of row y0
    long sb;        // half-spectrum offset of row y0+1 (so when it does not exist)
    long oo;        // output offset of row y0
    int pair;       // row y0+1 exists
    int valid;
};

// workgroup bx of one cube's X inverse pass
template <int L, class P, int PV = 0>
__device__ __forceinline__ void irfft_x_body(const XInv<P> &a, long bx, RowLds<L, P, XInvLine> &sm) {
    using V = typename Cx<P>::V;
    constexpr int NL = Shape<L, P>::NL_RM;
    V *buf = sm.buf, *tw_s = sm.tw_s;
    XInvLine *meta = sm.meta;
    const TwPrefetch<L, V> twp(a.tw);
    const long g0 = bx * NL;
    for (int l = threadIdx.x; l < NL; l += Shape<L>::T) {
        XInvLine m;
        long g = g0 + l;
        m.valid = g < a.nlines;
        if (!m.valid) g = 0;
        int z = (int)(g / a.npairs);
        int y0 = 2 * (int)(g - (long)z * a.npairs);
        m.pair = y0 + 1 < a.ny;
        m.so = ((long)z * a.py + y0) * a.hxp;
        m.sb = y0 + 1 < a.py ? m.so + a.hxp : m.so;
        m.oo = ((long)z * a.ny + y0) * a.nx;
        meta[l] = m;
    }
    __syncthreads();
    // L(x) = A(x) + i B(x) for x <= px/2, Hermitian extension above (paganin.c:271-280)
    auto first = [&](int l, int p0, int off) -> V {
        const XInvLine &m = meta[l];
        const int x = p0 + off;
        int k = x < a.hx ? x : L - x;
        // Loads are unconditional so the unrolled gather keeps them all in flight:
        // a line past the end recomputes line 0, and a row y0+1 >= ny still packs
        // spectrum row y0+1 (or row y0 again) as the imaginary part -- a Hermitian
        // half spectrum, so the real part stays row y0's -- and neither is stored.
        V A = a.spec[m.so + k];
        V B = a.spec[m.sb + k];
        return x < a.hx ? Cx<P>::make(A.x - B.y, A.y + B.x) : Cx<P>::make(A.x + B.y, B.x - A.y);
    };
    auto last = [&](int l, int p0, int off, V v) {
        const XInvLine &m = meta[l];
        const int x = p0 + off;
        if (x >= a.nx || !m.valid) return;
        a.out[m.oo + x] = epilogue<P>(a.epi, v.x * a.invn);
        if (m.pair) a.out[m.oo + a.nx + x] = epilogue<P>(a.epi, v.y * a.invn);
    };
    run_from<L, NL, false, false, 0, -1, PV>((P)1, tw_s, buf, first, last, [&] { twp.store(tw_s); });
}

// grid: (line groups, cube)
template <int L, class P, int PV = 0>
__global__ __launch_bounds__(256) void k_irfft_x(Batch<XInv<P>> bt) {
    __shared__ RowLds<L, P, XInvLine> sm;
    const XInv<P> a = bt.a[blockIdx.y];             // into registers once
    irfft_x_body<L, P, PV>(a, blockIdx.x, sm);
}

unsigned grid1d(size_t n) {
    size_t b = (n + 255) / 256;
    return (unsigned)(b < 8192 ? (b ? b : 1) : 8192);
}

// ------------------------------------------------------------------ fused planes (short axes)
// For cubes whose padded X and Y axes are one length L of 64 or 96 (fp32), a z
// plane's half spectrum (L rows x L/2+1 columns, at most 37.6 KB) fits in LDS,
// so the X and Y transforms of a plane run in one workgroup and the spectrum
// visits HBM once between them instead of twice:
//   k_rfft_xy : u8 gather + prologue, R2C along x in chunks of NLX row pairs,
//               untangled into the plane held column by column, then the
//               column FFTs along y in place; the last stage writes the plane.
//   k_irfft_yx: the inverse: column FFTs of a plane (first stage read from HBM),
//               then the Hermitian repack and C2R along x of rows y < ny with
//               the epilogue, stored as the tile.
// Every line's arithmetic (radix order, twiddles, untangle, repack) is that
// of k_rfft_x + k_cfft_cols and k_cfft_cols + k_irfft_x, so the results are
// bit-identical to the separate passes.
template <int L>
struct PlaneShape {
    static constexpr int HX = L / 2 + 1;                    // half-spectrum columns
    static constexpr int NLX = L >= 96 ? 8 : 16;            // X row pairs per chunk (LDS: 2-3 groups per CU)
    // 128 fits too (81 KB, two groups per CU) but measured slower than the
    // separate passes: eight chunk barriers per plane at 8 waves per CU
    static constexpr bool OK = L == 64 || L == 96;
};

template <int L>
__device__ __forceinline__ float2 *plane_col(float2 *col, int k, int y) {
    return col + k * Shape<L>::LD + lds_pos(y);
}

// LDS of one plane workgroup
template <int L, class Line>
struct PlaneLds {
    float2 col[PlaneShape<L>::HX * Shape<L>::LD];
    float2 xb[PlaneShape<L>::NLX * Shape<L>::LD];
    float2 tw_s[L];
    Line meta[PlaneShape<L>::NLX];
};

// plane z of one cube's fused X+Y forward pass
template <int L, int MODE, int PV = 0>
__device__ __forceinline__ void rfft_xy_body(const XFwd<float> &a, int z, PlaneLds<L, XLine> &sm) {
    using V = float2;
    using PS = PlaneShape<L>;
    constexpr int HX = PS::HX, NLX = PS::NLX, LD = Shape<L>::LD;
    constexpr bool U8 = MODE != 0, RING = MODE == 2;
    V *col = sm.col, *xb = sm.xb, *tw_s = sm.tw_s;
    XLine *meta = sm.meta;
    const TwPrefetch<L, V> twp(a.tw);                       // px == py: one table
    const int rz = reflect_idx(z, a.nz);
    const long gz = a.pro.gz0 + rz;
    const float zf = (U8 && a.pro.zdrift) ? a.pro.zfactor[gz] : 1.0f;
    long cs = 0;
    if (RING) cs = (long)(ring_slab(a.pro, gz) - ring_slab(a.pro, a.pro.gz0)) * a.ny;
    int any = 0;
    for (int c0 = 0; c0 < L / 2; c0 += NLX) {
        for (int l = threadIdx.x; l < NLX; l += Shape<L>::T) {
            XLine m;
            const int y0 = 2 * (c0 + l);
            const int ry0 = reflect_idx(y0, a.ny), ry1 = reflect_idx(y0 + 1, a.ny);
            m.valid = 1;
            m.pair = 1;
            m.a = ((long)rz * a.ny + ry0) * a.nx;
            m.b = ((long)rz * a.ny + ry1) * a.nx;
            m.so = 0;
            m.zf = zf;
            m.ca = (cs + ry0) * a.nx;
            m.cb = (cs + ry1) * a.nx;
            meta[l] = m;
        }
        __syncthreads();
        auto first = [&](int l, int p0, int off) -> V {
            const XLine &m = meta[l];
            int rx = reflect_idx(p0 + off, a.nx);
            float va, vb;
            if (U8) {
                const unsigned char *u = (const unsigned char *)a.in;
                unsigned char ua = u[m.a + rx], ub = u[m.b + rx];
                float ca = 0.0f, cb = 0.0f;
                if (RING) {
                    ca = a.pro.cimg[m.ca + rx];
                    cb = a.pro.cimg[m.cb + rx];
                }
                any |= ua | ub;
                va = prologue(a.pro, ua, ca, m.zf);
                vb = prologue(a.pro, ub, cb, m.zf);
            } else {
                const float *f = (const float *)a.in;
                va = f[m.a + rx];
                vb = f[m.b + rx];
            }
            return make_float2(va, vb);
        };
        const bool c_first = c0 == 0;
        run_from<L, NLX, false, false, 0, -1, PV>(-1.0f, tw_s, xb, first, LdsWr<L, NLX, V>{xb}, [&] {
            if (c_first) twp.store(tw_s);
        });
        __syncthreads();
        // untangle A(k) = (X(k) + conj X(-k))/2 -> row y0, B(k) = (X(k) - conj X(-k))/(2i) -> row y0+1
        for (int i = threadIdx.x; i < NLX * HX; i += Shape<L>::T) {
            const int l = i / HX, k = i - l * HX;
            const int nk = k ? L - k : 0;
            const V p = xb[l * LD + lds_pos(k)];
            const V q = xb[l * LD + lds_pos(nk)];
            const int y0 = 2 * (c0 + l);
            *plane_col<L>(col, k, y0) = make_float2(0.5f * (p.x + q.x), 0.5f * (p.y - q.y));
            *plane_col<L>(col, k, y0 + 1) = make_float2(0.5f * (p.y + q.y), 0.5f * (q.x - p.x));
        }
        __syncthreads();
    }
    if (U8 && a.flag) {
        if (__syncthreads_or(any) && threadIdx.x == 0) fyg_flag_set(a.flag, a.gen);
    }
    // column FFTs along y, in place: the first stage with a barrier between its
    // reads and writes, then the remaining stages; the last one writes the plane
    using PL = Plan<L, false, PV>;
    stage<L, HX, Shape<L>::T, true, PL::R(0), 1, true>(-1.0f, tw_s, LdsRd<L, HX, V>{col}, LdsWr<L, HX, V>{col});
    __syncthreads();
    V *plane = a.spec + (size_t)z * a.py * a.hxp;
    const long hxp = a.hxp;
    auto last = [&](int l, int p0, int off, V v) { plane[(size_t)(p0 + off) * hxp + l] = v; };
    if constexpr (PL::n > 1) run_from<L, HX, true, false, 1, -1, PV>(-1.0f, tw_s, col, LdsRd<L, HX, V>{col}, last);
    else
        for (int i = threadIdx.x; i < HX * L; i += Shape<L>::T) last(i % HX, i / HX, 0, col[(i % HX) * LD + lds_pos(i / HX)]);
}

// grid: (pz planes, cube)
template <int L, int MODE, int PV = 0>
__global__ __launch_bounds__(256) void k_rfft_xy(Batch<XFwd<float>> bt) {
    __shared__ PlaneLds<L, XLine> sm;
    const XFwd<float> a = bt.a[blockIdx.y];
    rfft_xy_body<L, MODE, PV>(a, blockIdx.x, sm);
}

// plane z of one cube's fused Y+X inverse pass
template <int L, int PV = 0>
__device__ __forceinline__ void irfft_yx_body(const XInv<float> &a, int z, PlaneLds<L, XInvLine> &sm) {
    using V = float2;
    using PS = PlaneShape<L>;
    constexpr int HX = PS::HX, NLX = PS::NLX, LD = Shape<L>::LD;
    V *col = sm.col, *xb = sm.xb, *tw_s = sm.tw_s;
    XInvLine *meta = sm.meta;
    const TwPrefetch<L, V> twp(a.tw);
    const V *plane = a.spec + (size_t)z * a.py * a.hxp;
    const long hxp = a.hxp;
    // column FFTs of the plane (k < hx), inverse sign, result kept in LDS
    auto firstg = [&](int l, int p0, int off) -> V { return plane[(size_t)(p0 + off) * hxp + l]; };
    run_from<L, HX, true, false, 0, -1, PV>(1.0f, tw_s, col, firstg, LdsWr<L, HX, V>{col}, [&] { twp.store(tw_s); });
    __syncthreads();
    const int npairs = (a.ny + 1) / 2;
    for (int c0 = 0; c0 < npairs; c0 += NLX) {
        for (int l = threadIdx.x; l < NLX; l += Shape<L>::T) {
            XInvLine m;
            const int g = c0 + l;
            m.valid = g < npairs;
            const int y0 = 2 * (m.valid ? g : 0);
            m.pair = y0 + 1 < a.ny;
            m.so = y0;                          // plane rows of y0 and y0+1 (y0+1 < L: L is even)
            m.sb = y0 + 1;
            m.oo = ((long)z * a.ny + y0) * a.nx;
            meta[l] = m;
        }
        __syncthreads();
        auto first = [&](int l, int p0, int off) -> V {
            const XInvLine &m = meta[l];
            const int x = p0 + off;
            const int k = x < HX ? x : L - x;
            const V A = *plane_col<L>(col, k, (int)m.so);
            const V B = *plane_col<L>(col, k, (int)m.sb);
            return x < HX ? make_float2(A.x - B.y, A.y + B.x) : make_float2(A.x + B.y, B.x - A.y);
        };
        auto last = [&](int l, int p0, int off, V v) {
            const XInvLine &m = meta[l];
            const int x = p0 + off;
            if (x >= a.nx || !m.valid) return;
            a.out[m.oo + x] = epilogue<float>(a.epi, v.x * a.invn);
            if (m.pair) a.out[m.oo + a.nx + x] = epilogue<float>(a.epi, v.y * a.invn);
        };
        run_from<L, NLX, false, false, 0, -1, PV>(1.0f, tw_s, xb, first, last);
        __syncthreads();
    }
}

// grid: (nz planes, cube)
template <int L, int PV = 0>
__global__ __launch_bounds__(256) void k_irfft_yx(Batch<XInv<float>> bt) {
    __shared__ PlaneLds<L, XInvLine> sm;
    const XInv<float> a = bt.a[blockIdx.y];
    irfft_yx_body<L, PV>(a, blockIdx.x, sm);
}

// ------------------------------------------------------------------ dispatch over L
template <class F>
bool dispatch(int L, F f) {
    switch (L) {
#define FYG_CASE(n) \
    case n: f(std::integral_constant<int, n>{}); return true;
        FYG_CASE(1) FYG_CASE(2) FYG_CASE(3) FYG_CASE(4) FYG_CASE(6) FYG_CASE(8) FYG_CASE(12)
        FYG_CASE(16) FYG_CASE(24) FYG_CASE(32) FYG_CASE(48) FYG_CASE(64) FYG_CASE(96)
        FYG_CASE(128) FYG_CASE(192) FYG_CASE(256) FYG_CASE(384) FYG_CASE(512) FYG_CASE(768)
        FYG_CASE(1024) FYG_CASE(1536) FYG_CASE(2048) FYG_CASE(3072) FYG_CASE(4096)
#undef FYG_CASE
        default: return false;
    }
}

// f(std::integral_constant<int, PV>) for the plan variant pv of length L:
// lengths with one plan, and fp64, always take plan 0
template <int L, class P, class F>
void with_plan(int pv, F f) {
    if constexpr (plan_alt(L) && std::is_same<P, float>::value) {
        if (pv == 1) {
            f(std::integral_constant<int, 1>{});
            return;
        }
    }
    f(std::integral_constant<int, 0>{});
}

template <int L, class P>
struct Launch {
    static void rfft_x(fyg_stream *s, const XFwd<P> *a, int n, bool u8, int pv = 0) {
        unsigned g = (unsigned)((a[0].nlines + Shape<L, P>::NL_RM - 1) / Shape<L, P>::NL_RM);
        with_plan<L, P>(pv, [&](auto c) {
            constexpr int PV = decltype(c)::value;
            for (int i = 0; i < n; i += FYG_BATCH) {
                int m = n - i < FYG_BATCH ? n - i : FYG_BATCH;
                Batch<XFwd<P>> b{};
                for (int j = 0; j < m; j++) b.a[j] = a[i + j];
                if (u8 && a[0].pro.dering)
                    fyg_launch(s, "rfft_x_u8", k_rfft_x<L, 2, P, PV>, dim3(g, m), dim3(256), 0, b);
                else if (u8)
                    fyg_launch(s, "rfft_x_u8", k_rfft_x<L, 1, P, PV>, dim3(g, m), dim3(256), 0, b);
                else
                    fyg_launch(s, "rfft_x", k_rfft_x<L, 0, P, PV>, dim3(g, m), dim3(256), 0, b);
            }
        });
    }
    static void cols(fyg_stream *s, const Cols<P> *a, int n, int nbatch, const char *name, int pv = 0) {
        unsigned g = (unsigned)((a[0].ncols + Shape<L, P>::NL_CM - 1) / Shape<L, P>::NL_CM);
        with_plan<L, P>(pv, [&](auto c) {
            constexpr int PV = decltype(c)::value;
            for (int i = 0; i < n; i += FYG_BATCH) {
                int m = n - i < FYG_BATCH ? n - i : FYG_BATCH;
                Batch<Cols<P>> b{};
                for (int j = 0; j < m; j++) b.a[j] = a[i + j];
                fyg_launch(s, name, k_cfft_cols<L, P, PV>, dim3(g, nbatch, m), dim3(256), 0, b);
            }
        });
    }
    static void zfused(fyg_stream *s, const ZFused<P> *a, int n, int pv = 0) {
        long nc = (long)a[0].py * a[0].hxp;
        unsigned g = (unsigned)((nc + Shape<L, P>::NL_ZF - 1) / Shape<L, P>::NL_ZF);
        with_plan<L, P>(pv, [&](auto c) {
            constexpr int PV = decltype(c)::value;
            for (int i = 0; i < n; i += FYG_BATCH) {
                int m = n - i < FYG_BATCH ? n - i : FYG_BATCH;
                Batch<ZFused<P>> b{};
                for (int j = 0; j < m; j++) b.a[j] = a[i + j];
                fyg_launch(s, "zfused", k_zfused<L, P, PV>, dim3(g, m), dim3(256), 0, b);
            }
        });
    }
    // fused X+Y passes of whole planes (fp32, px == py == L in {64, 96})
    static void rfft_xy(fyg_stream *s, const XFwd<P> *a, int n, bool u8, int pv = 0) {
        if constexpr (PlaneShape<L>::OK && std::is_same<P, float>::value) {
            with_plan<L, P>(pv, [&](auto c) {
                constexpr int PV = decltype(c)::value;
                for (int i = 0; i < n; i += FYG_BATCH) {
                    int m = n - i < FYG_BATCH ? n - i : FYG_BATCH;
                    Batch<XFwd<P>> b{};
                    for (int j = 0; j < m; j++) b.a[j] = a[i + j];
                    dim3 g((unsigned)a[0].pz, (unsigned)m);
                    if (u8 && a[0].pro.dering)
                        fyg_launch(s, "rfft_xy_u8", k_rfft_xy<L, 2, PV>, g, dim3(256), 0, b);
                    else if (u8)
                        fyg_launch(s, "rfft_xy_u8", k_rfft_xy<L, 1, PV>, g, dim3(256), 0, b);
                    else
                        fyg_launch(s, "rfft_xy", k_rfft_xy<L, 0, PV>, g, dim3(256), 0, b);
                }
            });
        }
    }
    static void irfft_yx(fyg_stream *s, const XInv<P> *a, int n, int pv = 0) {
        if constexpr (PlaneShape<L>::OK && std::is_same<P, float>::value) {
            with_plan<L, P>(pv, [&](auto c) {
                constexpr int PV = decltype(c)::value;
                for (int i = 0; i < n; i += FYG_BATCH) {
                    int m = n - i < FYG_BATCH ? n - i : FYG_BATCH;
                    Batch<XInv<P>> b{};
                    for (int j = 0; j < m; j++) b.a[j] = a[i + j];
                    fyg_launch(s, "irfft_yx", k_irfft_yx<L, PV>, dim3((unsigned)a[0].nz, (unsigned)m), dim3(256),
                               0, b);
                }
            });
        }
    }
    static void irfft_x(fyg_stream *s, const XInv<P> *a, int n, int pv = 0) {
        unsigned g = (unsigned)((a[0].nlines + Shape<L, P>::NL_RM - 1) / Shape<L, P>::NL_RM);
        with_plan<L, P>(pv, [&](auto c) {
            constexpr int PV = decltype(c)::value;
            for (int i = 0; i < n; i += FYG_BATCH) {
                int m = n - i < FYG_BATCH ? n - i : FYG_BATCH;
                Batch<XInv<P>> b{};
                for (int j = 0; j < m; j++) b.a[j] = a[i + j];
                fyg_launch(s, "irfft_x", k_irfft_x<L, P, PV>, dim3(g, m), dim3(256), 0, b);
            }
        });
    }
};

// squared-frequency tables and twiddles of precision P
template <class P>
const typename Cx<P>::V *twiddles(int L);
template <>
inline const float2 *twiddles<float>(int L) { return (const float2 *)fyg_twiddles(L); }
template <>
inline const double2 *twiddles<double>(int L) { return (const double2 *)fyg_twiddles64(L); }

// plan variant per pass kind of a deconvolution: X (rfft_x / irfft_x), Y (the
// column passes), Z (zfused), XY (the fused plane kernels)
struct PassPlans {
    int x = 0, y = 0, z = 0, xy = 0;
};

// the kernel descriptors of n deconvolutions of one shape in precision P
template <class P>
struct Passes {
    using V = typename Cx<P>::V;
    std::vector<XFwd<P>> xf;
    std::vector<Cols<P>> yf, yi;
    std::vector<ZFused<P>> zf;
    std::vector<XInv<P>> xi;
    int n = 0;
    int px = 0, py = 0, pz = 0;
    bool u8 = false;

    // 0, or 1 when the cubes do not share one shape and kernel variant
    int build(const fyg_deconv_args *a, int na) {
        const fyg_deconv_args &h = a[0];
        for (int i = 1; i < na; i++) {              // one shape and one kernel variant per batch
            const fyg_deconv_args &b = a[i];
            if (b.nz != h.nz || b.ny != h.ny || b.nx != h.nx || b.pz != h.pz || b.py != h.py || b.px != h.px ||
                b.hx != h.hx || b.hxp != h.hxp || (b.in_u8 != 0) != (h.in_u8 != 0) ||
                (b.in_u8 && (b.pro.dering != 0) != (h.pro.dering != 0)) || (b.f64 != 0) != (h.f64 != 0))
                return 1;
        }
        const V *twx = h.twx ? (const V *)h.twx : twiddles<P>(h.px);
        const V *twy = h.twy ? (const V *)h.twy : twiddles<P>(h.py);
        const V *twz = h.twz ? (const V *)h.twz : twiddles<P>(h.pz);
        if (!twx || !twy || !twz) return 1;
        n = na;
        px = h.px;
        py = h.py;
        pz = h.pz;
        u8 = h.in_u8 != 0;
        const long plane = (long)h.py * h.hxp;
        xf.resize(n);
        yf.resize(n);
        yi.resize(n);
        zf.resize(n);
        xi.resize(n);
        for (int i = 0; i < n; i++) {
            const fyg_deconv_args *d = &a[i];
            V *spec = (V *)d->spec;
            xf[i] = XFwd<P>{d->in, d->nz, d->ny, d->nx, d->pz, d->py, d->hx, d->hxp,
                            (long)d->pz * ((d->py + 1) / 2), (long)(d->py + 1) / 2, spec, twx, d->pro,
                            d->nonzero_flag, d->nonzero_gen};
            // columns k < hx: the pitch padding stays idle
            yf[i] = Cols<P>{spec, d->hx, d->hxp, plane, (P)-1, twy};
            zf[i] = ZFused<P>{spec, d->py, d->hx, d->hxp, plane, (const P *)d->lut, d->us, d->fz2, d->fy2, d->fx2,
                              twz};
            yi[i] = Cols<P>{spec, d->hx, d->hxp, plane, (P)1, twy};
            // 1/(pz*py*px) as paganin.c:285 forms it in fp32; exact in fp64
            P invn = sizeof(P) == sizeof(float) ? (P)(1.0f / ((float)d->pz * (float)d->py * (float)d->px))
                                                : (P)(1.0 / ((double)d->pz * d->py * d->px));
            xi[i] = XInv<P>{spec, (P *)d->out, d->nz, d->ny, d->nx, d->py, d->hx, d->hxp,
                            (long)d->nz * ((d->ny + 1) / 2), (long)(d->ny + 1) / 2, invn, d->epi, twx};
        }
        return 0;
    }
    void x_fwd(fyg_stream *s, int pv) {
        dispatch(px, [&](auto c) { Launch<decltype(c)::value, P>::rfft_x(s, xf.data(), n, u8, pv); });
    }
    void y_fwd(fyg_stream *s, int pv) {
        dispatch(py, [&](auto c) { Launch<decltype(c)::value, P>::cols(s, yf.data(), n, pz, "cfft_y_fwd", pv); });
    }
    void z(fyg_stream *s, int pv) {
        dispatch(pz, [&](auto c) { Launch<decltype(c)::value, P>::zfused(s, zf.data(), n, pv); });
    }
    void y_inv(fyg_stream *s, int pv) {
        const int nz = xi[0].nz;
        dispatch(py, [&](auto c) { Launch<decltype(c)::value, P>::cols(s, yi.data(), n, nz, "cfft_y_inv", pv); });
    }
    void x_inv(fyg_stream *s, int pv) {
        dispatch(px, [&](auto c) { Launch<decltype(c)::value, P>::irfft_x(s, xi.data(), n, pv); });
    }
    void xy_fwd(fyg_stream *s, int pv) {
        dispatch(px, [&](auto c) { Launch<decltype(c)::value, P>::rfft_xy(s, xf.data(), n, u8, pv); });
    }
    void yx_inv(fyg_stream *s, int pv) {
        dispatch(px, [&](auto c) { Launch<decltype(c)::value, P>::irfft_yx(s, xi.data(), n, pv); });
    }
};

// the fused plane kernels take cubes whose X and Y axes are one length of 64
// or 96 (fp32, unless switched off)
template <class P>
bool use_planes(int px, int py) {
    return std::is_same<P, float>::value && px == py && (px == 64 || px == 96) && fyg_fft_plane_mode() != 0;
}

// n deconvolutions of one shape in precision P: ring images, then the five
// passes, each one launch per FYG_BATCH cubes
template <class P>
int deconv_batch(const fyg_deconv_args *a, int n, fyg_stream *s, PassPlans pp = PassPlans{}) {
    if (n < 1) return 0;
    const fyg_deconv_args &h = a[0];
    Passes<P> ps;
    if (ps.build(a, n)) return 1;
    if (h.step_gen && !(h.in_u8 && h.pro.dering) && fyg_gen_step(h.step_gen, h.step_nflag, s)) return 1;
    if (h.in_u8 && h.pro.dering) {
        std::vector<RingImg> ri(n);
        int mslab = 1;
        for (int i = 0; i < n; i++) {
            const fyg_prologue &PR = a[i].pro;
            if (!PR.cimg) return 1;
            int s0 = ring_slab(PR, PR.gz0), s1 = ring_slab(PR, PR.gz0 + a[i].nz - 1);
            ri[i] = RingImg{PR, a[i].ny, a[i].nx, s0, s1 - s0 + 1};
            if (s1 - s0 + 1 > mslab) mslab = s1 - s0 + 1;
        }
        for (int i = 0; i < n; i += FYG_BATCH) {
            int m = n - i < FYG_BATCH ? n - i : FYG_BATCH;
            Batch<RingImg> b{};
            for (int j = 0; j < m; j++) b.a[j] = ri[i + j];
            int *step = i == 0 ? h.step_gen : nullptr;   // the call's first launch steps the generation
            fyg_launch(s, "ring_image", k_ring_image, dim3((unsigned)((h.nx + 255) / 256), (unsigned)h.ny,
                       (unsigned)(m * mslab)), dim3(256), 0, b, mslab, step, h.step_nflag);
        }
    }
    const bool planes = use_planes<P>(h.px, h.py);
    if (planes)
        ps.xy_fwd(s, pp.xy);
    else {
        ps.x_fwd(s, pp.x);
        ps.y_fwd(s, pp.y);
    }
    ps.z(s, pp.z);
    if (planes)
        ps.yx_inv(s, pp.xy);
    else {
        ps.y_inv(s, pp.y);
        ps.x_inv(s, pp.x);
    }
    return fyg_last_error();
}

}  // namespace
