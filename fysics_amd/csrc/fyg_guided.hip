// fyg_guided.hip -- the self-guided filter of guided.c (He/Sun/Tang), exact and
// fast (He & Sun, s-decimated coefficients) variants, the plain box mean, and
// the per-tile emit stage (recenter + dithered u8 quantisation, pipeline.c:53-65,
// 413-421, 502-507).
//
// Box means (guided.c:22-98) are (2r+1)^3 windows that shrink at the volume
// faces: mean = (sum over in-bounds voxels) * 1/cnt_x * 1/cnt_y * 1/cnt_z.  The
// GPU computes them on LDS tiles: a workgroup loads its 32x8x4 output tile plus
// an r-voxel halo once, then sums along x, y and z inside LDS, so each input
// voxel is read from HBM about once per kernel instead of once per axis pass.
// The moment kernels centre the second moment on a per-tile constant c0:
// var = E[(I-c0)^2] - (E[I]-c0)^2 is the same quantity with far less
// cancellation than E[I^2] - E[I]^2 in fp32.  E[I] itself is summed from the
// samples, as the reference sums it, and forms b = E[I](1-a).
//
// Exact path (guided.c:113-137):   k_moments(I) -> a,b ; k_box2(a,b) -> out = ma*I + mb
// Fast path  (guided.c:152-231):   k_moments(I decimated by s) -> a,b (low-res);
//                                  k_box2 -> ma, mb (low-res); k_upsample -> out
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>

#include <vector>

#include "fyg_internal.hpp"

namespace {

constexpr int TX = 32, TY = 8, TZ = 4, NT = 256;

template <int RHO>
struct Tile {
    static constexpr int EX = TX + 2 * RHO, EY = TY + 2 * RHO, EZ = TZ + 2 * RHO;
    static constexpr int AREG = EZ * EY * EX;     // loaded region per field
    static constexpr int BREG = EZ * EY * TX;     // after the x pass
};

struct Grid {
    int nz, ny, nx;     // box grid (full or decimated)
    int s;              // decimation of the source (1 = none)
    int fny, fnx;       // full-resolution source dims (row pitch)
};

template <class P = float>
__device__ __forceinline__ P count_recip(int i, int n, int rho) {
    int lo = i - rho < 0 ? 0 : i - rho;
    int hi = i + rho > n - 1 ? n - 1 : i + rho;
    return (P)1 / (P)(hi - lo + 1);
}

// the same reciprocals from a table (window lengths up to 7): for a
// wave-uniform position this is one scalar load instead of a division
__constant__ float kRecipF[8] = {0.0f, 1.0f, 1.0f / 2, 1.0f / 3, 1.0f / 4, 1.0f / 5, 1.0f / 6, 1.0f / 7};
__constant__ double kRecipD[8] = {0.0, 1.0, 1.0 / 2, 1.0 / 3, 1.0 / 4, 1.0 / 5, 1.0 / 6, 1.0 / 7};
template <class P>
__device__ __forceinline__ P count_recip_tab(int i, int n, int rho) {
    int lo = i - rho < 0 ? 0 : i - rho;
    int hi = i + rho > n - 1 ? n - 1 : i + rho;
    if constexpr (sizeof(P) == sizeof(float)) return kRecipF[hi - lo + 1];
    else return kRecipD[hi - lo + 1];
}

// a / (a + e) for the coefficient a = var/(var+eps): the hardware reciprocal
// (1 ulp) in fp32, a correctly rounded division in fp64
__device__ __forceinline__ float ratio(float v, float e) { return v * __builtin_amdgcn_rcpf(v + e); }
__device__ __forceinline__ double ratio(double v, double e) { return v / (v + e); }

constexpr int gcd_c(int a, int b) { return b == 0 ? a : gcd_c(b, a % b); }

// Generic tiled box over NF fields.  load(gz,gy,gx,v[NF]) fills in-grid values;
// cells outside the grid contribute 0.  out(gz,gy,gx,sum[NF],recip) is called for
// every in-grid output voxel of the tile.
template <int RHO, int NF, class Load, class Out>
__device__ __forceinline__ void box_tile(const Grid &g, float *sm, Load load, Out out) {
    using Tl = Tile<RHO>;
    float *A = sm;                     // [NF][AREG]
    float *B = sm + NF * Tl::AREG;     // [NF][BREG]
    const int bx = blockIdx.x * TX, by = blockIdx.y * TY, bz = blockIdx.z * TZ;
    {
        // every load of a thread is issued before the first is used: cells
        // outside the grid load the tile origin and are zeroed after
        constexpr int K = (Tl::AREG + NT - 1) / NT;
        float v[K][NF];
#pragma unroll
        for (int k = 0; k < K; k++) {
            int i = threadIdx.x + k * NT;
            int ex = i % Tl::EX, t = i / Tl::EX, ey = t % Tl::EY, ez = t / Tl::EY;
            int gx = bx + ex - RHO, gy = by + ey - RHO, gz = bz + ez - RHO;
            bool in = i < Tl::AREG && gx >= 0 && gx < g.nx && gy >= 0 && gy < g.ny && gz >= 0 && gz < g.nz;
            load(in ? gz : bz, in ? gy : by, in ? gx : bx, v[k]);
#pragma unroll
            for (int f = 0; f < NF; f++) v[k][f] = in ? v[k][f] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < K; k++) {
            int i = threadIdx.x + k * NT;
            if (i < Tl::AREG) {
#pragma unroll
                for (int f = 0; f < NF; f++) A[f * Tl::AREG + i] = v[k][f];
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < Tl::BREG; i += NT) {           // x
        int tx = i % TX, row = i / TX;                            // row = ez*EY + ey
#pragma unroll
        for (int f = 0; f < NF; f++) {
            const float *a = A + f * Tl::AREG + row * Tl::EX + tx;
            float s = 0.0f;
#pragma unroll
            for (int d = 0; d <= 2 * RHO; d++) s += a[d];
            B[f * Tl::BREG + i] = s;
        }
    }
    __syncthreads();
    constexpr int CREG = Tl::EZ * TY * TX;                        // y pass -> A
    for (int i = threadIdx.x; i < CREG; i += NT) {
        int tx = i % TX, t = i / TX, ty = t % TY, ez = t / TY;
#pragma unroll
        for (int f = 0; f < NF; f++) {
            const float *b = B + f * Tl::BREG + (ez * Tl::EY + ty) * TX + tx;
            float s = 0.0f;
#pragma unroll
            for (int d = 0; d <= 2 * RHO; d++) s += b[d * TX];
            A[f * CREG + i] = s;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < TZ * TY * TX; i += NT) {       // z + output
        int tx = i % TX, t = i / TX, ty = t % TY, tz = t / TY;
        int gx = bx + tx, gy = by + ty, gz = bz + tz;
        if (gx >= g.nx || gy >= g.ny || gz >= g.nz) continue;
        float s[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) {
            const float *c = A + f * CREG + (tz * TY + ty) * TX + tx;
            float acc = 0.0f;
#pragma unroll
            for (int d = 0; d <= 2 * RHO; d++) acc += c[d * TY * TX];
            s[f] = acc;
        }
        float rc = count_recip(gx, g.nx, RHO) * count_recip(gy, g.ny, RHO) * count_recip(gz, g.nz, RHO);
        out(gz, gy, gx, s, rc);
    }
}

// ------------------------------------------------------------------ emit
// Dither hash of pipeline.c:53-59.  The low 32 bits of the 64-bit products
// (long)c * k are the 32-bit products of the low words, so it runs in 32-bit
// arithmetic; the coordinate terms are split so that a row computes its (z,y)
// part once and steps the x part by its constant.
constexpr unsigned kHz = 73856093u, kHy = 19349663u, kHx = 83492791u;

__device__ __forceinline__ float dither_mix(unsigned int h) {
    h ^= h >> 16;
    h *= 0x7feb352dU;
    h ^= h >> 15;
    h *= 0x846ca68bU;
    h ^= h >> 16;
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}

__device__ __forceinline__ float dither01(long z, long y, long x) {
    return dither_mix(((unsigned int)z * kHz) ^ ((unsigned int)y * kHy) ^ ((unsigned int)x * kHx));
}

// v*255 + d rounded as two separate operations, as written in pipeline.c:61
__device__ __forceinline__ unsigned char quant_d(float v, float d) {
    #pragma clang fp contract(off)
    float t = v * 255.0f + d;
    if (t <= 0.0f) return 0;
    int q = (int)t;
    return q > 255 ? 255 : (unsigned char)q;
}

__device__ __forceinline__ unsigned char quant_u8(float v, long gz, long gy, long gx, int no_dither) {
    return quant_d(v, no_dither ? 0.5f : dither01(gz, gy, gx));
}

// recenter of pipeline.c:413-421
__device__ __forceinline__ float recenter(const fyg_emit &E, float v) {
    #pragma clang fp contract(off)
    if (v <= 0.0f) return 0.0f;
    float w = 0.5f + (v - E.rmid) * E.rk;
    return w < 0.0f ? 0.0f : (w > 1.0f ? 1.0f : w);
}

// quantised output of one voxel at local (z,y,x): recenter, dither, all-air
__device__ __forceinline__ unsigned char emit_q(const fyg_emit &E, bool keep, long z, long y, long x, float v) {
    if (!keep) return 0;
    if (E.recenter) v = recenter(E, v);
    return quant_u8(v, E.gz0 + z, E.gy0 + y, E.gx0 + x, E.no_dither);
}

__device__ __forceinline__ bool emit_keep(const fyg_emit &E) {
    return !E.nonzero_flag || *E.nonzero_flag == *E.nonzero_gen;
}

// write one filtered voxel (local coords z,y,x of an nz*ny*nx buffer)
__device__ __forceinline__ void emit(const fyg_emit &E, int ny, int nx, long z, long y, long x, float v) {
    if (!E.quant) {
        E.out[((size_t)z * ny + y) * nx + x] = v;
        return;
    }
    long lz = z - E.iz0, ly = y - E.iy0, lx = x - E.ix0;
    if (lz < 0 || lz >= E.tz || ly < 0 || ly >= E.ty || lx < 0 || lx >= E.tx) return;
    E.q[((size_t)lz * E.ty + ly) * E.tx + lx] = emit_q(E, emit_keep(E), z, y, x, v);
}

// write 4 consecutive voxels x..x+3 of one row (all inside the region)
__device__ __forceinline__ void emit4(const fyg_emit &E, int ny, int nx, long z, long y, long x, const float *v,
                                      bool vec, bool keep) {
    if (!E.quant) {
        float *o = E.out + ((size_t)z * ny + y) * nx + x;
        if (vec) *(float4 *)o = make_float4(v[0], v[1], v[2], v[3]);
        else for (int j = 0; j < 4; j++) o[j] = v[j];
        return;
    }
    long lz = z - E.iz0, ly = y - E.iy0, lx = x - E.ix0;
    unsigned char *o = E.q + ((size_t)lz * E.ty + ly) * E.tx + lx;
    const unsigned hzy = ((unsigned)(E.gz0 + z) * kHz) ^ ((unsigned)(E.gy0 + y) * kHy);
    unsigned hx = (unsigned)(E.gx0 + x) * kHx;
    unsigned char q[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        float w = E.recenter ? recenter(E, v[j]) : v[j];
        q[j] = keep ? quant_d(w, E.no_dither ? 0.5f : dither_mix(hzy ^ hx)) : 0;
        hx += kHx;
    }
    if (vec) *(uchar4 *)o = make_uchar4(q[0], q[1], q[2], q[3]);
    else for (int j = 0; j < 4; j++) o[j] = q[j];
}

// ------------------------------------------------------------------ kernels
struct MomArgs {
    Grid g;
    const float *in;        // full-resolution source
    double eps;
    float *a, *b;           // grid-sized outputs
};

// box(I), box((I-c0)^2) -> var = box((I-c0)^2) - (box(I)-c0)^2, a = var/(var+eps),
// b = box(I)*(1-a)   (guided.c:122-132).  The mean that forms b is summed from
// the samples themselves, as the reference sums it: of samples >= 0 it is > 0
// unless every one is 0, which the recenter's "air stays 0" test
// (pipeline.c:417) reads; only the second moment is centred.
template <int RHO>
__global__ __launch_bounds__(NT) void k_moments(MomArgs p) {
    extern __shared__ float sm[];
    const Grid &g = p.g;
    // tile-local centring constant: the source voxel at the tile origin
    const int bx = blockIdx.x * TX, by = blockIdx.y * TY, bz = blockIdx.z * TZ;
    const float c0 = p.in[((size_t)bz * g.s * g.fny + (size_t)by * g.s) * g.fnx + (size_t)bx * g.s];
    auto load = [&](int z, int y, int x, float *v) {
        float u = p.in[((size_t)z * g.s * g.fny + (size_t)y * g.s) * g.fnx + (size_t)x * g.s];
        float t = u - c0;
        v[0] = u;
        v[1] = t * t;
    };
    const float epsf = (float)p.eps;
    auto out = [&](int z, int y, int x, const float *s, float rc) {
        float mean = s[0] * rc, q = s[1] * rc;
        float m = mean - c0;
        float var = q - m * m;
        if (var < 0.0f) var = 0.0f;
        float ai = var / (var + epsf);
        size_t o = ((size_t)z * g.ny + y) * g.nx + x;
        p.a[o] = ai;
        p.b[o] = mean * (1.0f - ai);
    };
    box_tile<RHO, 2>(g, sm, load, out);
}

struct Box2Args {
    Grid g;
    const float *a, *b;
    int apply;              // 0: store ma, mb;  1: out = ma*I + mb (exact path)
    float *ma, *mb;
    const float *in;        // apply: the guide/input at full resolution (== grid)
    fyg_emit em;
};

template <int RHO>
__global__ __launch_bounds__(NT) void k_box2(Box2Args p) {
    extern __shared__ float sm[];
    const Grid &g = p.g;
    auto load = [&](int z, int y, int x, float *v) {
        size_t o = ((size_t)z * g.ny + y) * g.nx + x;
        v[0] = p.a[o];
        v[1] = p.b[o];
    };
    auto out = [&](int z, int y, int x, const float *s, float rc) {
        size_t o = ((size_t)z * g.ny + y) * g.nx + x;
        float ma = s[0] * rc, mb = s[1] * rc;
        if (p.apply) {
            emit(p.em, g.ny, g.nx, z, y, x, ma * p.in[o] + mb);
        } else {
            p.ma[o] = ma;
            p.mb[o] = mb;
        }
    };
    box_tile<RHO, 2>(g, sm, load, out);
}

// ---- fused low-res coefficient field of the fast path (guided.c:163-182).
// A workgroup owns a 16x16 column block of the decimated grid and marches
// along z through ZC output planes (plus a 2*rho halo on each side).  Per
// sample plane it
//   1. stages the samples I and the centred samples v = I - c0 of an
//      (16+2rho)^2 patch in LDS,
//   2. forms each column's x-then-y window sums of I and v^2 and keeps the last
//      2*rho+1 of them in registers, so their sum is box(I), box(v^2) centred
//      rho planes back -> var = box(v^2) - (box(I) - c0)^2, a = var/(var+eps),
//      b = box(I)(1-a) for that plane,
//   3. stages a, b in LDS and sums them the same way for the (16-2*rho)^2 inner
//      columns -> mean_a, mean_b one more rho plane back.
// The halo costs 1.3x in samples instead of the 3.4x of a 3-D tile and a and b
// never reach HBM.  Cells outside the grid contribute 0 and windows shrink at
// the faces (guided.c:22-98).
template <int RHO, int ZC_>
struct March {
    static constexpr int AW = 16;                   // a, b columns per side (NT = AW*AW)
    static constexpr int OW = AW - 2 * RHO;         // output columns per side
    static constexpr int SW = AW + 2 * RHO;         // sample columns per side
    static constexpr int NS = SW * SW;
    static constexpr int KS = (NS + NT - 1) / NT;   // sample slots per thread
    static constexpr int ZC = ZC_;                  // output planes per workgroup
    static constexpr int Q = 2 * RHO + 1;           // window length
    static constexpr int STEPS = ZC + 4 * RHO;      // sample planes of a full chunk
    // sample planes in flight + 1; divides STEPS so the buffer ring rotates by
    // full unrolling (a short last chunk runs idle steps)
    static constexpr int PD = STEPS % 5 == 0 ? 5 : STEPS % 6 == 0 ? 6 : STEPS % 4 == 0 ? 4 : 1;
    // steps per unrolled loop body: the sample ring (PD) and the window queues
    // (Q) both rotate by renaming, so no value moves between registers
    static constexpr int U = PD / gcd_c(PD, Q) * Q;
    // LDS row pitches: the two rows of 16 columns a half-wave reads land in
    // disjoint banks (pitch 48 = 16 mod 32 for samples, 16 for a and b)
    static constexpr int SP = 48, AP = 16;
};

template <class P>
struct CoefArgs {
    Grid g;                 // decimated grid; source rows are s*y, planes s*z, columns s*x
    const P *in;
    double eps;
    P *ma, *mb;
};

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup h runs on XCD
// h % 8) and each XCD has its own L2.  xcd_tile renumbers them so that XCD k
// runs the k-th contiguous eighth of the logical tiles: neighbouring tiles,
// which read each other's halo rows, then meet in one L2.
__device__ __forceinline__ int xcd_tile(int h, int T) {
    const int c = T / 8, r = T % 8, x = h % 8, i = h / 8;
    return (x < r ? x * (c + 1) : r * (c + 1) + (x - r) * c) + i;
}

// LDS of one coefficient-march workgroup
template <int RHO, class P, int ZC>
struct CoefLds {
    using M = March<RHO, ZC>;
    P sv[M::SW][M::SP];                 // centred samples I - c0
    P sr[M::SW][M::SP];                 // the samples I
    P sa[M::AW][M::AP], sb[M::AW][M::AP];
};

// column block (tx, ty), z chunk zchunk of one cube's coefficient march
template <int RHO, class P, int ZC>
__device__ __forceinline__ void coef_body(const CoefArgs<P> &p, int tx, int ty, int zchunk, CoefLds<RHO, P, ZC> &sm) {
    using M = March<RHO, ZC>;
    static_assert(M::AW * M::AW == NT, "one thread per a, b column");
    static_assert(M::SW <= M::SP, "sample row fits its pitch");
    auto &sv = sm.sv;
    auto &sr = sm.sr;
    auto &sa = sm.sa;
    auto &sb = sm.sb;
    const Grid &g = p.g;
    const int bx = tx * M::OW, by = ty * M::OW, zs = zchunk * M::ZC;
    const int ze = min(zs + M::ZC, g.nz);
    const size_t sz = (size_t)g.s * g.fny * g.fnx, sy = (size_t)g.s * g.fnx;
    // samples are centred on one sample of the column block (plane 0, so the
    // result does not depend on the z chunking)
    const P c0 = p.in[(size_t)by * sy + (size_t)bx * g.s];
    const P epsf = (P)p.eps;

    // sample slots: in-plane source offset (0 outside the grid) and LDS cell
    size_t soff[M::KS];
    bool sok[M::KS];
    int scell[M::KS];
#pragma unroll
    for (int k = 0; k < M::KS; k++) {
        int i = threadIdx.x + k * NT;
        int ex = i % M::SW, ey = i / M::SW;
        int gx = bx - 2 * RHO + ex, gy = by - 2 * RHO + ey;
        sok[k] = i < M::NS && gx >= 0 && gx < g.nx && gy >= 0 && gy < g.ny;
        soff[k] = sok[k] ? (size_t)gy * sy + (size_t)gx * g.s : 0;
        scell[k] = i < M::NS ? ey * M::SP + ex : -1;
    }
    // this thread's a, b column (and output column when inner)
    const int ax = threadIdx.x % M::AW, ay = threadIdx.x / M::AW;
    const int gx = bx - RHO + ax, gy = by - RHO + ay;
    const bool colin = gx >= 0 && gx < g.nx && gy >= 0 && gy < g.ny;
    const P rcxy = colin ? count_recip<P>(gx, g.nx, RHO) * count_recip<P>(gy, g.ny, RHO) : 0.0f;
    const bool inner = ax >= RHO && ax < M::AW - RHO && ay >= RHO && ay < M::AW - RHO && colin;

    P q1[M::Q], q2[M::Q], qa[M::Q], qb[M::Q];
#pragma unroll
    for (int j = 0; j < M::Q; j++) q1[j] = q2[j] = qa[j] = qb[j] = 0.0f;

    auto fetch = [&](int zp, P *dst) {
        size_t zo = (size_t)(zp >= 0 && zp < g.nz ? zp : zs) * sz;
#pragma unroll
        for (int k = 0; k < M::KS; k++) dst[k] = p.in[zo + soff[k]];
    };
    // u: the step's position in the unrolled loop body; the window queues are
    // rings whose newest entry is slot u % Q (steps before zs push zeros).  The
    // window sums run oldest to newest plane, so a column's arithmetic does not
    // depend on where its z chunk starts (nor on the chunk length).
    auto step = [&](int zp, const P *cur, int u) {
        const int slot = u % M::Q;
        const bool zin = zp >= 0 && zp < g.nz;
#pragma unroll
        for (int k = 0; k < M::KS; k++)
            if (scell[k] >= 0) {
                (&sv[0][0])[scell[k]] = (zin && sok[k]) ? cur[k] - c0 : 0.0f;
                (&sr[0][0])[scell[k]] = (zin && sok[k]) ? cur[k] : 0.0f;
            }
        __syncthreads();
        // x-then-y window sums of v and v^2 at plane zp
        P s1 = 0.0f, s2 = 0.0f;
#pragma unroll
        for (int dy = 0; dy < M::Q; dy++) {
            P r1 = 0.0f, r2 = 0.0f;
#pragma unroll
            for (int dx = 0; dx < M::Q; dx++) {
                P v = sv[ay + dy][ax + dx];
                r1 += sr[ay + dy][ax + dx];
                r2 += v * v;
            }
            s1 += r1;
            s2 += r2;
        }
        q1[slot] = s1;
        q2[slot] = s2;
        if (zp >= zs) {                             // a, b at plane za = zp - rho
            const int za = zp - RHO;
            P av = 0.0f, bv = 0.0f;
            if (colin && za >= 0 && za < g.nz) {
                P t1 = 0.0f, t2 = 0.0f;
#pragma unroll
                for (int j = 1; j <= M::Q; j++) { t1 += q1[(slot + j) % M::Q]; t2 += q2[(slot + j) % M::Q]; }
                P rc = rcxy * count_recip_tab<P>(za, g.nz, RHO);
                P mean = t1 * rc, q = t2 * rc;
                P m = mean - c0;
                P var = q - m * m;
                if (var < 0.0f) var = 0.0f;
                av = ratio(var, epsf);
                bv = mean * (1.0f - av);
            }
            sa[ay][ax] = av;
            sb[ay][ax] = bv;
        }
        __syncthreads();
        {                                           // window sums of a, b; output plane zp - 2 rho
            P u1 = 0.0f, u2 = 0.0f;
            if (inner && zp >= zs) {
#pragma unroll
                for (int dy = -RHO; dy <= RHO; dy++) {
                    P r1 = 0.0f, r2 = 0.0f;
#pragma unroll
                    for (int dx = -RHO; dx <= RHO; dx++) {
                        r1 += sa[ay + dy][ax + dx];
                        r2 += sb[ay + dy][ax + dx];
                    }
                    u1 += r1;
                    u2 += r2;
                }
            }
            qa[slot] = u1;
            qb[slot] = u2;
            const int zo = zp - 2 * RHO;
            if (inner && zo >= zs && zo < ze) {
                P t1 = 0.0f, t2 = 0.0f;
#pragma unroll
                for (int j = 1; j <= M::Q; j++) { t1 += qa[(slot + j) % M::Q]; t2 += qb[(slot + j) % M::Q]; }
                P rc = rcxy * count_recip_tab<P>(zo, g.nz, RHO);
                size_t o = ((size_t)zo * g.ny + gy) * g.nx + gx;
                p.ma[o] = t1 * rc;
                p.mb[o] = t2 * rc;
            }
        }
    };
    // sample planes PD-1 ahead are in flight while a plane is processed; the
    // buffer ring rotates by full unrolling, so no register copy waits on a load
    constexpr int PD = M::PD, U = M::U;
    const int zfirst = zs - 2 * RHO, zlast = ze - 1 + 2 * RHO;
    P buf[PD][M::KS];
#pragma unroll
    for (int d = 0; d < PD - 1; d++) fetch(zfirst + d, buf[d]);
    for (int z0 = zfirst; z0 <= zlast; z0 += U) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            fetch(z0 + u + PD - 1, buf[(u + PD - 1) % PD]);
            step(z0 + u, buf[u % PD], u);
        }
    }
}

// grid: (x blocks, y blocks, cube * nzc + z chunk), renumbered by xcd_tile
template <int RHO, class P, int ZC>
__global__ __launch_bounds__(NT) void k_coef(Batch<CoefArgs<P>> bt, int nzc) {
    const int T = gridDim.x * gridDim.y * gridDim.z;
    const int L = xcd_tile(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z), T);
    const int tx = L % gridDim.x, ty = (L / gridDim.x) % gridDim.y, tz = L / (gridDim.x * gridDim.y);
    const CoefArgs<P> p = bt.a[tz / nzc];          // into registers once
    __shared__ CoefLds<RHO, P, ZC> sm;
    coef_body<RHO, P, ZC>(p, tx, ty, tz % nzc, sm);
}

struct Box1Args {
    Grid g;
    const float *in;
    float *out;
};

template <int RHO>
__global__ __launch_bounds__(NT) void k_box1(Box1Args p) {
    extern __shared__ float sm[];
    const Grid &g = p.g;
    auto load = [&](int z, int y, int x, float *v) { v[0] = p.in[((size_t)z * g.ny + y) * g.nx + x]; };
    auto out = [&](int z, int y, int x, const float *s, float rc) {
        p.out[((size_t)z * g.ny + y) * g.nx + x] = s[0] * rc;
    };
    box_tile<RHO, 1>(g, sm, load, out);
}

template <class P>
struct UpArgs {
    int nz, ny, nx, lz, ly, lx, s;
    const P *ma, *mb, *in;
    int z0, y0, x0, dz, dy, dx;      // region of output voxels to produce
    int vin, vout;                   // 16-byte input rows / aligned 4-voxel output groups
    fyg_emit em;
};

constexpr int UPX = 4;               // consecutive x outputs per thread
constexpr int URY = 8;               // output rows of one workgroup at the full chunk width
// Shape of one upsample workgroup: LPR lanes per row chunk (64, 32 or 16), so a
// chunk is XW = LPR*UPX voxels wide and a wave covers 64/LPR rows; narrow
// volumes (64-128 voxels in x) then keep every lane busy instead of leaving a
// 256-wide chunk three quarters empty.  The workgroup covers URZ planes x RY
// rows x XW voxels; each thread owns RPT rows of UPX voxels.  The low-res
// patch these rows touch (s >= 2): UPZ planes x UPY rows x UPK columns.
template <int URZ, int LPR>
struct UpShape {
    static constexpr int SW = 64 / LPR, XW = LPR * UPX, RY = URY * SW;
    static constexpr int RPT = URZ * URY / 4;
    static constexpr int UPZ = URZ / 2 + 2, UPY = RY / 2 + 2, UPK = XW / 2 + 4, UPH = UPK / 2;
};
// patch rows store even columns first, then odd ones: a wave's reads of
// columns 2q+e (s = 2) then hit consecutive words instead of every other one
template <int UPH>
__device__ __forceinline__ int up_col(int c) { return (c & 1) * UPH + (c >> 1); }

// trilinear upsample of (mean_a, mean_b) from the decimated grid and apply
// out = a*I + b (guided.c:188-228): u = c/s, i0 = min((int)u, l-2), f = min(u-i0, 1).
// A workgroup produces URZ planes x RY rows x XW voxels.  Those rows share a
// patch of at most UPZ x UPY low-res rows, which is loaded into LDS once.  As in
// the reference, each output row blends its four (z,y) low-res rows first and
// then lerps along x; a thread blends only the three low-res columns its four
// consecutive voxels need.  (Four voxels span at most two knots when s > 2, or
// when s = 2 and the quad starts on an even x: up_args starts an odd region
// one voxel early at s = 2, and that voxel is not emitted.)  Thread (q, g) --
// q the lane within its row chunk, g its sub-wave -- owns x quad q of rows g,
// g+4*SW, g+8*SW, ...
// P2: s = 2^sh, so c/s is exactly c*(1/s) and needs no division; the knot
// is then c >> sh and the fraction (c - (knot << sh)) / s, both exact, so the
// row arithmetic (uniform across a wave when LPR = 64) stays on the scalar unit.
template <bool P2, class P = float>
struct UpCoord {
    P fs, inv;
    int sh;
    __device__ P u(int c) const { return P2 ? (P)c * inv : (P)c / fs; }
    __device__ int i0(int c, int l) const {
        int i = P2 ? c >> sh : (int)u(c);
        return i > l - 2 ? l - 2 : i;
    }
    // min(u(c) - i, 1) for the knot i = i0(c, l)
    __device__ P frac(int c, int i) const {
        P f = P2 ? (P)(c - (i << sh)) * inv : u(c) - (P)i;
        return f > (P)1 ? (P)1 : f;
    }
};

// LDS of one upsample workgroup
template <class P, int URZ, int LPR>
struct UpLds {
    using S = UpShape<URZ, LPR>;
    P pa[S::UPZ][S::UPY][S::UPK], pb[S::UPZ][S::UPY][S::UPK];
};

// x chunk bx, y block by, z block bz of one cube's upsample
template <bool P2, class P, int URZ, int LPR>
__device__ __forceinline__ void upsample_body(const UpArgs<P> &p, int bx, int by, int bz, UpLds<P, URZ, LPR> &sm) {
    using S = UpShape<URZ, LPR>;
    constexpr int RPT = S::RPT, UPZ = S::UPZ, UPY = S::UPY, UPK = S::UPK, SW = S::SW, RY = S::RY;
    auto &pa = sm.pa;
    auto &pb = sm.pb;
    const UpCoord<P2, P> U{(P)p.s, (P)1 / (P)p.s, P2 ? __builtin_ctz((unsigned)p.s) : 0};
    const int Z0 = p.z0 + bz * URZ, Y0 = p.y0 + by * RY;
    const int zl = min(Z0 + URZ, p.z0 + p.dz) - 1, yl = min(Y0 + RY, p.y0 + p.dy) - 1;
    const int xc0 = p.x0 + bx * S::XW;
    const int xc1 = min(xc0 + S::XW, p.x0 + p.dx) - 1;
    const int kb = U.i0(xc0, p.lx), nk = U.i0(xc1, p.lx) - kb + 2;
    const int zlo = U.i0(Z0, p.lz), nzp = U.i0(zl, p.lz) - zlo + 2;
    const int ylo = U.i0(Y0, p.ly), nyp = U.i0(yl, p.ly) - ylo + 2;
    // with whole-wave row chunks the row index is uniform across a wave: row and
    // patch-row arithmetic below then runs on the scalar unit
    const int q = threadIdx.x % LPR;
    int g = threadIdx.x / LPR;
    if constexpr (SW == 1) g = __builtin_amdgcn_readfirstlane(g);
    const int xb = xc0 + q * UPX;
    // the first x emitted: a quantised region may start one voxel after x0
    const int xe = p.em.quant ? (int)p.em.ix0 : p.x0;
    const bool xin = xb <= xc1, full = xb >= xe && xb + UPX - 1 <= xc1;

    // input voxels of the thread's rows: issued first, used last
    P iv[RPT][UPX];
#pragma unroll
    for (int j = 0; j < RPT; j++) {
        int r = g + 4 * SW * j, z = Z0 + r / RY, y = Y0 + r % RY;
        const bool rok = z <= zl && y <= yl;
        bool ok = xin && rok;
        const P *irow = p.in + ((size_t)(rok ? z : Z0) * p.ny + (rok ? y : Y0)) * p.nx;
        if (ok && full && p.vin) {
            if constexpr (sizeof(P) == sizeof(float)) {
                float4 t = *(const float4 *)(irow + xb);
                iv[j][0] = t.x; iv[j][1] = t.y; iv[j][2] = t.z; iv[j][3] = t.w;
            } else {
                double2 t0 = *(const double2 *)(irow + xb), t1 = *(const double2 *)(irow + xb + 2);
                iv[j][0] = t0.x; iv[j][1] = t0.y; iv[j][2] = t1.x; iv[j][3] = t1.y;
            }
        } else {
#pragma unroll
            for (int e = 0; e < UPX; e++) iv[j][e] = (ok && xb + e <= xc1) ? irow[xb + e] : 0.0f;
        }
    }
    // low-res patch: wave w loads patch rows w, w+4, ..; its lanes split the
    // columns.  The row index is wave-uniform, so row arithmetic is scalar.
    {
        constexpr int RW = (UPZ * UPY + 3) / 4, CW = (UPK + 63) / 64;
        const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
        const int nrow = nzp * nyp;
        P va[RW][CW], vb[RW][CW];
#pragma unroll
        for (int i = 0; i < RW; i++) {
            const int rr = wv + 4 * i;
            int pz = 0;
#pragma unroll
            for (int k = 1; k < UPZ; k++) pz += rr >= k * nyp;
            const int py = rr - pz * nyp;
            const bool rok = rr < nrow;
            const size_t ro = rok ? ((size_t)(zlo + pz) * p.ly + (ylo + py)) * p.lx + kb : (size_t)zlo * p.ly * p.lx;
            const P *ra = p.ma + ro, *rb = p.mb + ro;
#pragma unroll
            for (int c = 0; c < CW; c++) {
                int col = lane + 64 * c;
                int cc = (rok && col < nk) ? col : 0;
                va[i][c] = ra[cc];
                vb[i][c] = rb[cc];
            }
        }
#pragma unroll
        for (int i = 0; i < RW; i++) {
            const int rr = wv + 4 * i;
            int pz = 0;
#pragma unroll
            for (int k = 1; k < UPZ; k++) pz += rr >= k * nyp;
            const int py = rr - pz * nyp;
#pragma unroll
            for (int c = 0; c < CW; c++) {
                int col = lane + 64 * c;
                if (rr < nrow && col < nk && col < UPK) {
                    pa[pz][py][up_col<S::UPH>(col)] = va[i][c];
                    pb[pz][py][up_col<S::UPH>(col)] = vb[i][c];
                }
            }
        }
    }
    __syncthreads();
    if (!xin) return;
    // x-dependent weights, shared by the thread's rows
    const int k0 = U.i0(xb, p.lx) - kb;
    P fxs[UPX];
    bool dsel[UPX];
#pragma unroll
    for (int e = 0; e < UPX; e++) {
        int x = xb + e, x0 = U.i0(x, p.lx);
        P fx = U.frac(x, x0);
        fxs[e] = fx > 1.0f ? 1.0f : fx;
        dsel[e] = x0 - kb - k0 != 0;            // 0 or 1
    }
    const int c0 = up_col<S::UPH>(min(k0, nk - 1)), c1 = up_col<S::UPH>(min(k0 + 1, nk - 1)),
              c2 = up_col<S::UPH>(min(k0 + 2, nk - 1));
    const bool keep = emit_keep(p.em);
#pragma unroll
    for (int j = 0; j < RPT; j++) {
        int r = g + 4 * SW * j, z = Z0 + r / RY, y = Y0 + r % RY;
        if (z > zl || y > yl) continue;
        int z0 = U.i0(z, p.lz), y0 = U.i0(y, p.ly);
        P zf = U.frac(z, z0);
        P yf = U.frac(y, y0);
        const P w00 = (1 - zf) * (1 - yf), w01 = (1 - zf) * yf, w10 = zf * (1 - yf), w11 = zf * yf;
        const int pz = z0 - zlo, py = y0 - ylo;
        auto blend = [&](const P (*Q)[UPY][UPK], int c) {
            return w00 * Q[pz][py][c] + w01 * Q[pz][py + 1][c] + w10 * Q[pz + 1][py][c] + w11 * Q[pz + 1][py + 1][c];
        };
        const P A0 = blend(pa, c0), A1 = blend(pa, c1), A2 = blend(pa, c2);
        const P B0 = blend(pb, c0), B1 = blend(pb, c1), B2 = blend(pb, c2);
        P ov[UPX];
#pragma unroll
        for (int e = 0; e < UPX; e++) {
            P fx = fxs[e], gx = 1 - fx;
            P a0 = dsel[e] ? A1 : A0, a1 = dsel[e] ? A2 : A1;
            P b0 = dsel[e] ? B1 : B0, b1 = dsel[e] ? B2 : B1;
            P av = gx * a0 + fx * a1, bv = gx * b0 + fx * b1;
            ov[e] = av * iv[j][e] + bv;
        }
        float of[UPX];                          // emitted (and quantised) in fp32
#pragma unroll
        for (int e = 0; e < UPX; e++) of[e] = (float)ov[e];
        if (full) {
            emit4(p.em, p.ny, p.nx, z, y, xb, of, p.vout != 0, keep);
        } else {
            for (int e = 0; e < UPX && xb + e <= xc1; e++) emit(p.em, p.ny, p.nx, z, y, xb + e, of[e]);
        }
    }
}

// grid: (x chunks, y blocks, cube * nzb + z block)
template <bool P2, class P, int URZ, int LPR>
__global__ __launch_bounds__(NT) void k_upsample(Batch<UpArgs<P>> bt, int nzb) {
    __shared__ UpLds<P, URZ, LPR> sm;
    const UpArgs<P> p = bt.a[blockIdx.z / nzb];     // into registers once, not re-read per use
    upsample_body<P2, P, URZ, LPR>(p, blockIdx.x, blockIdx.y, blockIdx.z % nzb, sm);
}

struct EmitArgs {
    const float *f;
    int nz, ny, nx;
    int z0, y0, x0, dz, dy, dx;
    fyg_emit em;
};

__global__ __launch_bounds__(NT) void k_emit(EmitArgs p) {
    const int y = p.y0 + blockIdx.y * 4 + threadIdx.y;
    const int z = p.z0 + blockIdx.z;
    if (y >= p.y0 + p.dy) return;
    const float *row = p.f + ((size_t)z * p.ny + y) * p.nx;
    const int xb = p.x0 + (blockIdx.x * 64 + threadIdx.x) * UPX;
#pragma unroll
    for (int j = 0; j < UPX; j++) {
        int x = xb + j;
        if (x >= p.x0 + p.dx) break;
        emit(p.em, p.ny, p.nx, z, y, x, row[x]);
    }
}

template <int RHO, int ZC>
dim3 march_grid(const Grid &g) {
    using M = March<RHO, ZC>;
    return dim3((unsigned)((g.nx + M::OW - 1) / M::OW), (unsigned)((g.ny + M::OW - 1) / M::OW),
                (unsigned)((g.nz + M::ZC - 1) / M::ZC));
}

dim3 row_grid(int dx, int dy, int dz) {
    return dim3((unsigned)((dx + 64 * UPX - 1) / (64 * UPX)), (unsigned)((dy + 3) / 4), (unsigned)dz);
}

// ---- global-memory separable fallback for radii the LDS tile does not cover
struct PassArgs {
    int nz, ny, nx, rho, axis, nf;
    const float *in[2];
    float *out[2];
};

__global__ void k_box_pass(PassArgs p) {
    size_t n = (size_t)p.nz * p.ny * p.nx;
    for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT) {
        int x = (int)(i % p.nx);
        size_t t = i / p.nx;
        int y = (int)(t % p.ny), z = (int)(t / p.ny);
        int c = p.axis == 0 ? x : (p.axis == 1 ? y : z);
        int len = p.axis == 0 ? p.nx : (p.axis == 1 ? p.ny : p.nz);
        size_t st = p.axis == 0 ? 1 : (p.axis == 1 ? (size_t)p.nx : (size_t)p.nx * p.ny);
        int lo = c - p.rho < 0 ? 0 : c - p.rho, hi = c + p.rho > len - 1 ? len - 1 : c + p.rho;
        for (int f = 0; f < p.nf; f++) {
            float s = 0.0f;
            for (int d = lo; d <= hi; d++) s += p.in[f][i + (size_t)(d - c) * st];
            p.out[f][i] = s / (float)(hi - lo + 1);
        }
    }
}

struct PointArgs {
    int mode;               // 0: I,(I-c0)^2 from decimated source; 1: moments -> a,b; 2: apply
    Grid g;
    const float *in;        // the guide; c0 = in[0], the centring constant, is read here on the device
    double eps;
    const float *m0, *m1;
    float *o0, *o1;
    fyg_emit em;
};

__global__ void k_point(PointArgs p) {
    const Grid &g = p.g;
    size_t n = (size_t)g.nz * g.ny * g.nx;
    // modes 0 and 1 write only workspace, so the guide's first voxel is intact
    const float c0 = p.mode < 2 ? p.in[0] : 0.0f;
    for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT) {
        int x = (int)(i % g.nx);
        size_t t = i / g.nx;
        int y = (int)(t % g.ny), z = (int)(t / g.ny);
        if (p.mode == 0) {
            float u = p.in[((size_t)z * g.s * g.fny + (size_t)y * g.s) * g.fnx + (size_t)x * g.s];
            float v = u - c0;
            p.o0[i] = u;
            p.o1[i] = v * v;
        } else if (p.mode == 1) {
            float mean = p.m0[i], q = p.m1[i];
            float m = mean - c0;
            float var = q - m * m;
            if (var < 0.0f) var = 0.0f;
            float ai = var / (var + (float)p.eps);
            p.o0[i] = ai;
            p.o1[i] = mean * (1.0f - ai);
        } else {
            emit(p.em, g.ny, g.nx, z, y, x, p.m0[i] * p.in[i] + p.m1[i]);
        }
    }
}

unsigned grid1d(size_t n) {
    size_t b = (n + NT - 1) / NT;
    return (unsigned)(b < 16384 ? (b ? b : 1) : 16384);
}

dim3 tile_grid(const Grid &g) {
    return dim3((unsigned)((g.nx + TX - 1) / TX), (unsigned)((g.ny + TY - 1) / TY), (unsigned)((g.nz + TZ - 1) / TZ));
}

template <int RHO>
size_t tile_lds(int nf) {
    return sizeof(float) * (size_t)nf * (Tile<RHO>::AREG + Tile<RHO>::BREG);
}

constexpr int kMaxTileRho = 2;   // 3+ exceeds 64 KB of LDS per tile; wider radii use the global passes

template <int RHO>
void launch_moments(fyg_stream *s, const MomArgs &a) {
    fyg_launch(s, "guided_moments", k_moments<RHO>, tile_grid(a.g), dim3(NT), tile_lds<RHO>(2), a);
}
template <int RHO>
void launch_box2(fyg_stream *s, const Box2Args &a) {
    fyg_launch(s, a.apply ? "guided_apply" : "guided_box_ab", k_box2<RHO>, tile_grid(a.g), dim3(NT),
               tile_lds<RHO>(2), a);
}
template <int RHO>
void launch_box1(fyg_stream *s, const Box1Args &a) {
    fyg_launch(s, "box_mean", k_box1<RHO>, tile_grid(a.g), dim3(NT), tile_lds<RHO>(1), a);
}

void moments(fyg_stream *s, const MomArgs &a, int rho) {
    switch (rho) {
        case 1: launch_moments<1>(s, a); break;
        case 2: launch_moments<2>(s, a); break;
    }
}
void box2(fyg_stream *s, const Box2Args &a, int rho) {
    switch (rho) {
        case 1: launch_box2<1>(s, a); break;
        case 2: launch_box2<2>(s, a); break;
    }
}

// separable global-memory passes: in[nf] -> out[nf] box means over grid g
void box_global(fyg_stream *s, const Grid &g, int rho, int nf, const float *in0, const float *in1,
                float *t0, float *t1, float *o0, float *o1) {
    size_t n = (size_t)g.nz * g.ny * g.nx;
    PassArgs px{g.nz, g.ny, g.nx, rho, 0, nf, {in0, in1}, {t0, t1}};
    fyg_launch(s, "box_pass", k_box_pass, dim3(grid1d(n)), dim3(NT), 0, px);
    PassArgs py{g.nz, g.ny, g.nx, rho, 1, nf, {t0, t1}, {o0, o1}};
    fyg_launch(s, "box_pass", k_box_pass, dim3(grid1d(n)), dim3(NT), 0, py);
    PassArgs pz{g.nz, g.ny, g.nx, rho, 2, nf, {o0, o1}, {t0, t1}};
    fyg_launch(s, "box_pass", k_box_pass, dim3(grid1d(n)), dim3(NT), 0, pz);
}

}  // namespace

extern "C" size_t fyg_guided_ws_bytes64(int nz, int ny, int nx, int s) {
    size_t lz = (nz + s - 1) / s, ly = (ny + s - 1) / s, lx = (nx + s - 1) / s;
    return sizeof(double) * 4 * lz * ly * lx;
}

extern "C" size_t fyg_guided_ws_bytes(int nz, int ny, int nx, int s) {
    size_t n = (size_t)nz * ny * nx;
    if (s > 1) {
        size_t lz = (nz + s - 1) / s, ly = (ny + s - 1) / s, lx = (nx + s - 1) / s;
        size_t nl = lz * ly * lx;
        if (lz >= 2 && ly >= 2 && lx >= 2) return sizeof(float) * 8 * nl;
    }
    return sizeof(float) * 6 * n;       // a, b (+ 4 temporaries for the wide-radius fallback)
}

// upsample + emit arguments of one fast-path filter (region = inner tile when quantising)
template <class P>
static const P *guide(const fyg_guided_args *a) {
    if constexpr (sizeof(P) == sizeof(float)) return a->in;
    else return a->in64;
}

template <class P>
static UpArgs<P> up_args(const fyg_guided_args *a, const P *MA, const P *MB) {
    const int sd = a->s, nz = a->nz, ny = a->ny, nx = a->nx;
    const int lz = (nz + sd - 1) / sd, ly = (ny + sd - 1) / sd, lx = (nx + sd - 1) / sd;
    const P *in = guide<P>(a);
    UpArgs<P> u{nz, ny, nx, lz, ly, lx, sd, MA, MB, in, 0, 0, 0, nz, ny, nx, 0, 0, a->em};
    if (a->em.quant) {
        u.z0 = (int)a->em.iz0; u.y0 = (int)a->em.iy0; u.x0 = (int)a->em.ix0;
        u.dz = (int)a->em.tz; u.dy = (int)a->em.ty; u.dx = (int)a->em.tx;
    }
    // At s = 2 four consecutive voxels from an odd x span three knots, one more
    // than the upsample blends: start such a region one voxel early (its first
    // voxel is computed and not emitted) so every quad starts on an even x.
    const bool shifted = sd == 2 && (u.x0 & 1);
    if (shifted) {
        u.x0 -= 1;
        u.dx += 1;
    }
    u.vin = (nx % 4 == 0) && (u.x0 % 4 == 0) && ((uintptr_t)in % 16 == 0);
    u.vout = a->em.quant ? (!shifted && a->em.tx % 4 == 0 && (uintptr_t)a->em.q % 4 == 0)
                         : (u.vin && (uintptr_t)a->em.out % 16 == 0);
    return u;
}

// ---- tile-shape autotuning.  The candidates change only how the work is cut
// into workgroups -- the coefficient march's z chunk, the upsample's planes per
// workgroup -- never the arithmetic of an output voxel, so every choice gives
// bit-identical results and the tuner may pick by speed alone.  The first time a
// (precision, radius, decimation, grid, region) shape is filtered, each
// candidate runs on the real data -- once to warm up, then twice timed with
// events on the launch stream -- and the fastest is kept for the process.
template <int RHO>
struct CoefCand;
template <>
struct CoefCand<1> {
    static constexpr int z[3] = {8, 14, 20};
};
template <>
struct CoefCand<2> {
    static constexpr int z[3] = {12, 7, 22};
};
constexpr int kNCoef = 3, kNUp = 2;
constexpr int kUpRZ[kNUp] = {2, 4};

struct TuneRec {
    int coef = 0, up = 0;
    bool coef_done = false, up_done = false;
    bool loaded = false;                     // from a tuning file, not timed here
    float coef_us[kNCoef] = {}, up_us[kNUp] = {};
};

// [p, p + n) and [q, q + m) share a byte
static bool overlaps(const void *p, size_t n, const void *q, size_t m) {
    const uintptr_t a = (uintptr_t)p, b = (uintptr_t)q;
    return n && m && a < b + m && b < a + n;
}

// the upsample candidates of a group may only be raced against each other when
// no filter's output overlaps any filter's guide: each candidate rereads the
// guides and rewrites the outputs
template <class P>
static bool up_rerunnable(const UpArgs<P> *u, int m) {
    for (int j = 0; j < m; j++) {
        const fyg_emit &e = u[j].em;
        const void *o = e.quant ? (const void *)e.q : (const void *)e.out;
        const size_t on = e.quant ? (size_t)e.tz * e.ty * e.tx : sizeof(float) * u[j].nz * u[j].ny * u[j].nx;
        for (int k = 0; k < m; k++)
            if (overlaps(o, on, u[k].in, sizeof(P) * u[k].nz * u[k].ny * u[k].nx)) return false;
    }
    return true;
}
static std::mutex g_tune_mu;
static std::map<std::string, TuneRec> g_tune;
static std::atomic<int> g_tune_mode{1};   // 0: first candidates, 1: autotune, 2 + k: candidate k

// a candidate's time: one untimed launch first (the first launch of a kernel
// also loads its code object), then the best of two timed launches
template <class F>
static float timed1(fyg_stream *s, F launch);
template <class F>
static float timed(fyg_stream *s, F launch) {
    launch();
    const float a = timed1(s, launch), b = timed1(s, launch);
    return a < b ? a : b;
}
template <class F>
static float timed1(fyg_stream *s, F launch) {
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return 1e30f;
    if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return 1e30f; }
    (void)hipEventRecord(e0, s->s);
    launch();
    (void)hipEventRecord(e1, s->s);
    float ms = 1e30f;
    if (hipEventSynchronize(e1) == hipSuccess) (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 1000.0f * ms;
}

template <int RHO, class P, int ZC>
static void coef_go(fyg_stream *s, const Batch<CoefArgs<P>> &b, int m, const Grid &gl) {
    dim3 g = march_grid<RHO, ZC>(gl);
    int nzc = (int)g.z;
    g.z *= m;
    fyg_launch(s, "guided_coef", k_coef<RHO, P, ZC>, g, dim3(NT), 0, b, nzc);
}

template <int RHO, class P>
static void launch_coef(fyg_stream *s, int v, const Batch<CoefArgs<P>> &b, int m, const Grid &gl) {
    switch (v) {
    case 1: coef_go<RHO, P, CoefCand<RHO>::z[1]>(s, b, m, gl); break;
    case 2: coef_go<RHO, P, CoefCand<RHO>::z[2]>(s, b, m, gl); break;
    default: coef_go<RHO, P, CoefCand<RHO>::z[0]>(s, b, m, gl); break;
    }
}

template <class P, int URZ, int LPR>
static void up_go_w(fyg_stream *s, const Batch<UpArgs<P>> &b, int m, int sd) {
    using S = UpShape<URZ, LPR>;
    const UpArgs<P> &u = b.a[0];
    const int nzb = (u.dz + URZ - 1) / URZ;
    dim3 grid((unsigned)((u.dx + S::XW - 1) / S::XW), (unsigned)((u.dy + S::RY - 1) / S::RY), (unsigned)(nzb * m));
    if ((sd & (sd - 1)) == 0)
        fyg_launch(s, "guided_upsample", k_upsample<true, P, URZ, LPR>, grid, dim3(NT), 0, b, nzb);
    else
        fyg_launch(s, "guided_upsample", k_upsample<false, P, URZ, LPR>, grid, dim3(NT), 0, b, nzb);
}

// the row chunk width that wastes the fewest lanes on a region dx wide (ties:
// the wider chunk)
static int up_lanes(int dx) {
    int best = 64;
    long waste = -1;
    for (int lpr = 64; lpr >= 16; lpr /= 2) {
        long xw = (long)lpr * UPX, w = (dx + xw - 1) / xw * xw - dx;
        if (waste < 0 || w < waste) {
            waste = w;
            best = lpr;
        }
    }
    return best;
}

template <class P, int URZ>
static void up_go(fyg_stream *s, const Batch<UpArgs<P>> &b, int m, int sd) {
    switch (up_lanes(b.a[0].dx)) {
    case 16: up_go_w<P, URZ, 16>(s, b, m, sd); break;
    case 32: up_go_w<P, URZ, 32>(s, b, m, sd); break;
    default: up_go_w<P, URZ, 64>(s, b, m, sd); break;
    }
}

template <class P>
static void launch_up_group(fyg_stream *s, int v, const Batch<UpArgs<P>> &b, int m, int sd) {
    if (v == 1) up_go<P, kUpRZ[1]>(s, b, m, sd);
    else up_go<P, kUpRZ[0]>(s, b, m, sd);
}

// upsample launches for n filters whose output regions have one shape
template <class P>
static void launch_upsample(fyg_stream *s, const UpArgs<P> *u, int n, int sd, int v = 0, int from = 0) {
    if (n < 1 || (size_t)u[0].dz * u[0].dy * u[0].dx == 0) return;
    for (int i = from; i < n; i += FYG_BATCH) {
        int m = n - i < FYG_BATCH ? n - i : FYG_BATCH;
        Batch<UpArgs<P>> b{};
        for (int j = 0; j < m; j++) b.a[j] = u[i + j];
        launch_up_group(s, v, b, m, sd);
    }
}

template <class P>
static int fast_batch(const fyg_guided_args *a, int n, fyg_stream *s);

extern "C" int fyg_guided_run(const fyg_guided_args *a, fyg_stream *s) {
    if (a->f64) return fyg_guided_run_batch(a, 1, s);
    int r = a->r < 1 ? 2 : a->r;
    int sd = a->s;
    int nz = a->nz, ny = a->ny, nx = a->nx;
    size_t n = (size_t)nz * ny * nx;
    if (n == 0) return 0;
    int lz = 0, ly = 0, lx = 0;
    bool fast = false;
    if (sd > 1) {
        lz = (nz + sd - 1) / sd;
        ly = (ny + sd - 1) / sd;
        lx = (nx + sd - 1) / sd;
        fast = lz >= 2 && ly >= 2 && lx >= 2;
    }
    float *ws = (float *)a->ws;
    if (!fast) {
        Grid g{nz, ny, nx, 1, ny, nx};
        float *A = ws, *B = ws + n;
        if (r <= kMaxTileRho) {
            MomArgs m{g, a->in, a->eps, A, B};
            moments(s, m, r);
            Box2Args b{g, A, B, 1, nullptr, nullptr, a->in, a->em};
            box2(s, b, r);
        } else {
            float *T0 = ws + 2 * n, *T1 = ws + 3 * n, *O0 = ws + 4 * n, *O1 = ws + 5 * n;
            PointArgs p0{0, g, a->in, a->eps, nullptr, nullptr, O0, O1, a->em};
            fyg_launch(s, "box_point", k_point, dim3(grid1d(n)), dim3(NT), 0, p0);
            box_global(s, g, r, 2, O0, O1, T0, T1, A, B);   // means in T0, T1
            PointArgs p1{1, g, a->in, a->eps, T0, T1, A, B, a->em};
            fyg_launch(s, "box_point", k_point, dim3(grid1d(n)), dim3(NT), 0, p1);
            box_global(s, g, r, 2, A, B, O0, O1, T0, T1);   // means in O0, O1
            PointArgs p2{2, g, a->in, a->eps, O0, O1, nullptr, nullptr, a->em};
            fyg_launch(s, "box_point", k_point, dim3(grid1d(n)), dim3(NT), 0, p2);
        }
        return fyg_last_error();
    }
    size_t nl = (size_t)lz * ly * lx;
    int rl = r / sd;
    if (rl < 1) rl = 1;
    if (rl <= 2) return fast_batch<float>(a, 1, s);
    float *A = ws, *B = ws + nl, *MA = ws + 2 * nl, *MB = ws + 3 * nl;
    Grid gl{lz, ly, lx, sd, ny, nx};
    {
        float *T0 = ws + 4 * nl, *T1 = ws + 5 * nl, *O0 = ws + 6 * nl, *O1 = ws + 7 * nl;
        PointArgs p0{0, gl, a->in, a->eps, nullptr, nullptr, O0, O1, a->em};
        fyg_launch(s, "box_point", k_point, dim3(grid1d(nl)), dim3(NT), 0, p0);
        box_global(s, gl, rl, 2, O0, O1, T0, T1, A, B);
        PointArgs p1{1, gl, a->in, a->eps, T0, T1, A, B, a->em};
        fyg_launch(s, "box_point", k_point, dim3(grid1d(nl)), dim3(NT), 0, p1);
        box_global(s, gl, rl, 2, A, B, O0, O1, T0, T1);
        if (fyg_d2d(MA, O0, sizeof(float) * nl, s) || fyg_d2d(MB, O1, sizeof(float) * nl, s)) return 1;
    }
    UpArgs<float> u = up_args<float>(a, MA, MB);
    launch_upsample(s, &u, 1, sd);
    return fyg_last_error();
}

// fast path with rl = r/s <= 2 for n filters of one shape: per stage one launch
// per FYG_BATCH cubes (coefficients, then upsample + emit), in precision P
template <class P>
static int fast_batch(const fyg_guided_args *a, int n, fyg_stream *s) {
    const fyg_guided_args &h = a[0];
    const int sd = h.s, r = h.r < 1 ? 2 : h.r;
    const int lz = (h.nz + sd - 1) / sd, ly = (h.ny + sd - 1) / sd, lx = (h.nx + sd - 1) / sd;
    const size_t nl = (size_t)lz * ly * lx;
    int rl = r / sd;
    if (rl < 1) rl = 1;
    const Grid gl{lz, ly, lx, sd, h.ny, h.nx};
    std::vector<CoefArgs<P>> c(n);
    std::vector<UpArgs<P>> u(n);
    for (int i = 0; i < n; i++) {
        P *ws = (P *)a[i].ws;
        P *MA = ws + 2 * nl, *MB = ws + 3 * nl;
        c[i] = CoefArgs<P>{gl, guide<P>(&a[i]), a[i].eps, MA, MB};
        u[i] = up_args<P>(&a[i], MA, MB);
    }
    // tile-shape choice: cached, tuned on first use, or forced
    char kb[160];
    snprintf(kb, sizeof kb, "%s r%d s%d grid %dx%dx%d region %dx%dx%d %s", sizeof(P) == 8 ? "fp64" : "fp32", rl,
             sd, lz, ly, lx, u[0].dz, u[0].dy, u[0].dx, u[0].em.quant ? "u8" : "fp32");
    const std::string key(kb);
    const int mode = g_tune_mode.load();
    const int m0 = n < FYG_BATCH ? n : FYG_BATCH;
    TuneRec rec;
    bool tune_coef = false, tune_up = false;
    if (mode == 1) {
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            auto it = g_tune.find(key);
            if (it != g_tune.end()) rec = it->second;
        }
        tune_coef = !rec.coef_done && !s->capturing;
        tune_up = !rec.up_done && !s->capturing && up_rerunnable(u.data(), m0);
    } else if (mode >= 2) {
        rec.coef = mode - 2 < kNCoef ? mode - 2 : kNCoef - 1;
        rec.up = mode - 2 < kNUp ? mode - 2 : kNUp - 1;
    }
    for (int i = 0; i < n; i += FYG_BATCH) {
        int m = n - i < FYG_BATCH ? n - i : FYG_BATCH;
        Batch<CoefArgs<P>> b{};
        for (int j = 0; j < m; j++) b.a[j] = c[i + j];
        auto go = [&](int v) {
            if (rl == 1) launch_coef<1, P>(s, v, b, m, gl);
            else launch_coef<2, P>(s, v, b, m, gl);
        };
        if (tune_coef && i == 0) {             // every candidate writes the same coefficients
            int best = 0;
            for (int v = 0; v < kNCoef; v++) {
                rec.coef_us[v] = timed(s, [&] { go(v); });
                if (rec.coef_us[v] < rec.coef_us[best]) best = v;
            }
            rec.coef = best;
            rec.coef_done = true;
        } else {
            go(rec.coef);
        }
    }
    int from = 0;
    if (tune_up) {                              // the same for the upsample's first group
        Batch<UpArgs<P>> b{};
        for (int j = 0; j < m0; j++) b.a[j] = u[j];
        int best = 0;
        for (int v = 0; v < kNUp; v++) {
            rec.up_us[v] = timed(s, [&] { launch_up_group(s, v, b, m0, sd); });
            if (rec.up_us[v] < rec.up_us[best]) best = v;
        }
        rec.up = best;
        rec.up_done = true;
        from = FYG_BATCH;
    }
    if (tune_coef || tune_up) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune[key] = rec;
    }
    launch_upsample(s, u.data(), n, sd, rec.up, from);
    return fyg_last_error();
}

// ---- the fast filter of cubes of mixed shapes: one launch for every cube's
// coefficient march and one for every cube's upsample (segments as in
// fyg_mix.hip).  The tile shapes are the first candidates (every candidate
// gives the same bits); the upsample's row chunk follows the region's width
// as in the grouped launches.
namespace {
// kind: the window radius on the decimated grid (1 or 2); nbx, nby: the
// cube's x and y column blocks, nbt its workgroups (a segment starts on a
// multiple of 8, so xcd_tile's XCD assignment holds within the cube)
template <int C>
__global__ __launch_bounds__(NT) void k_mix_coef(MixBatch<CoefArgs<float>, C> mb) {
    const MixSegs<C> &t = mb.t;
    const int sg = mix_seg(t, blockIdx.x);
    const int l = mix_uniform((int)blockIdx.x - t.start[sg]), T = mix_uniform(t.nbt[sg]);
    if (l >= T) return;
    const int gx = mix_uniform(t.nbx[sg]), gy = mix_uniform(t.nby[sg]);
    const int L = mix_uniform(xcd_tile(l, T));
    const int tx = mix_uniform(L % gx), ty = mix_uniform((L / gx) % gy), tz = mix_uniform(L / (gx * gy));
    const CoefArgs<float> p = mb.a[t.idx[sg]];      // into registers once
    if (t.kind[sg] == 2) {
        constexpr int ZC = CoefCand<2>::z[0];
        coef_body<2, float, ZC>(p, tx, ty, tz, mix_lds_as<CoefLds<2, float, ZC>>());
    } else {
        constexpr int ZC = CoefCand<1>::z[0];
        coef_body<1, float, ZC>(p, tx, ty, tz, mix_lds_as<CoefLds<1, float, ZC>>());
    }
}

// kind: lanes per row chunk (64, 32 or 16); nbx, nby: x chunks and y blocks
template <int C>
__global__ __launch_bounds__(NT) void k_mix_up(MixBatch<UpArgs<float>, C> mb) {
    constexpr int URZ = kUpRZ[0];
    const MixSegs<C> &t = mb.t;
    const int sg = mix_seg(t, blockIdx.x);
    const int l = mix_uniform((int)blockIdx.x - t.start[sg]);
    const int gx = mix_uniform(t.nbx[sg]), gy = mix_uniform(t.nby[sg]);
    const int bx = mix_uniform(l % gx), by = mix_uniform((l / gx) % gy), bz = mix_uniform(l / (gx * gy));
    // read from the arguments where used: a register copy of every field, live
    // across the three bodies, exceeds the scalar registers and spills
    const UpArgs<float> &p = mb.a[t.idx[sg]];
    switch (t.kind[sg]) {
    case 16: upsample_body<true, float, URZ, 16>(p, bx, by, bz, mix_lds_as<UpLds<float, URZ, 16>>()); break;
    case 32: upsample_body<true, float, URZ, 32>(p, bx, by, bz, mix_lds_as<UpLds<float, URZ, 32>>()); break;
    default: upsample_body<true, float, URZ, 64>(p, bx, by, bz, mix_lds_as<UpLds<float, URZ, 64>>()); break;
    }
}

}  // namespace

// the cube's fast filter can run in a mixed launch: fp32, s a power of two,
// a decimated window radius of 1 or 2, a decimated grid of at least 2 per axis
static_assert(sizeof(MixBatch<CoefArgs<float>, FYG_MIX_MAX>) <= kMixArgMax, "mixed-launch arguments too large");
static_assert(sizeof(MixBatch<UpArgs<float>, FYG_MIX_MAX>) <= kMixArgMax, "mixed-launch arguments too large");

extern "C" int fyg_mix_guided_ok(const fyg_guided_args *a) {
    const int sd = a->s, r = a->r < 1 ? 2 : a->r;
    if (a->f64 || sd < 2 || (sd & (sd - 1)) || r / sd > 2) return 0;
    return (a->nz + sd - 1) / sd >= 2 && (a->ny + sd - 1) / sd >= 2 && (a->nx + sd - 1) / sd >= 2;
}

extern "C" int fyg_mix_guided_run(const fyg_guided_args *a, int n, fyg_stream *s) {
    if (n < 1) return 0;
    if (n > FYG_MIX_MAX) return 1;
    for (int i = 0; i < n; i++)
        if (!fyg_mix_guided_ok(&a[i])) return 1;
    std::vector<CoefArgs<float>> c(n);
    std::vector<UpArgs<float>> u(n);
    std::vector<int> rl(n), ord(n);
    for (int i = 0; i < n; i++) {
        const int sd = a[i].s, r = a[i].r < 1 ? 2 : a[i].r;
        const int lz = (a[i].nz + sd - 1) / sd, ly = (a[i].ny + sd - 1) / sd, lx = (a[i].nx + sd - 1) / sd;
        const size_t nl = (size_t)lz * ly * lx;
        rl[i] = r / sd < 1 ? 1 : r / sd;
        float *ws = (float *)a[i].ws;
        float *MA = ws + 2 * nl, *MB = ws + 3 * nl;
        c[i] = CoefArgs<float>{Grid{lz, ly, lx, sd, a[i].ny, a[i].nx}, guide<float>(&a[i]), a[i].eps, MA, MB};
        u[i] = up_args<float>(&a[i], MA, MB);
        ord[i] = i;
    }
    // largest volume first
    std::stable_sort(ord.begin(), ord.end(), [&](int p, int q) {
        return (size_t)a[p].nz * a[p].ny * a[p].nx > (size_t)a[q].nz * a[q].ny * a[q].nx;
    });
    MixStage sc, su;
    for (int i : ord) {
        const dim3 g = rl[i] == 2 ? march_grid<2, CoefCand<2>::z[0]>(c[i].g) : march_grid<1, CoefCand<1>::z[0]>(c[i].g);
        const size_t lds = rl[i] == 2 ? sizeof(CoefLds<2, float, CoefCand<2>::z[0]>)
                                      : sizeof(CoefLds<1, float, CoefCand<1>::z[0]>);
        sc.add(i, rl[i], (long)g.x * g.y * g.z, lds, (int)g.x, (int)g.y, 8);
        const UpArgs<float> &p = u[i];
        if ((size_t)p.dz * p.dy * p.dx == 0) continue;
        const int lpr = up_lanes(p.dx), urz = kUpRZ[0];
        const int xw = lpr * UPX, ry = URY * (64 / lpr);
        const int gx = (p.dx + xw - 1) / xw, gy = (p.dy + ry - 1) / ry, gz = (p.dz + urz - 1) / urz;
        const size_t ul = lpr == 16 ? sizeof(UpLds<float, kUpRZ[0], 16>)
                          : lpr == 32 ? sizeof(UpLds<float, kUpRZ[0], 32>) : sizeof(UpLds<float, kUpRZ[0], 64>);
        su.add(i, lpr, (long)gx * gy * gz, ul, gx, gy);
    }
    if (sc.total > 0x7fffffff || su.total > 0x7fffffff) return 1;
    // launches of capacity C (the kernel-argument bytes follow the capacity)
    auto run = [&](auto cap) {
        constexpr int C = decltype(cap)::value;
        if (sc.total) {
            auto mb = std::make_unique<MixBatch<CoefArgs<float>, C>>();
            sc.fill(*mb, c.data(), n);
            fyg_launch(s, "mix_coef", k_mix_coef<C>, dim3((unsigned)sc.total), dim3(NT), sc.lds, *mb);
        }
        if (su.total) {
            auto mb = std::make_unique<MixBatch<UpArgs<float>, C>>();
            su.fill(*mb, u.data(), n);
            fyg_launch(s, "mix_up", k_mix_up<C>, dim3((unsigned)su.total), dim3(NT), su.lds, *mb);
        }
    };
    if (n <= FYG_MIX_SMALL) run(std::integral_constant<int, FYG_MIX_SMALL>{});
    else run(std::integral_constant<int, FYG_MIX_MAX>{});
    return fyg_last_error();
}

extern "C" int fyg_tune_mode(int mode) { return g_tune_mode.exchange(mode); }

// Tuning file: one line per shape, "key<TAB>coef<TAB>up" (up -1: not tuned),
// then one line per decided FFT radix plan, "fft<TAB>kind<TAB>length<TAB>plan".
// Saving writes every shape tuned or loaded so far; loading adds the file's
// shapes, which are then used without timing (a later process starts tuned).
extern "C" int fyg_tune_save(const char *path) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    std::string tmp = std::string(path) + ".tmp";
    FILE *f = fopen(tmp.c_str(), "w");
    if (!f) return 1;
    int bad = fprintf(f, "# fysics_amd guided-filter tile shapes: key\tcoef candidate\tupsample candidate\n") < 0;
    for (auto &kv : g_tune)
        if (kv.second.coef_done)
            bad |= fprintf(f, "%s\t%d\t%d\n", kv.first.c_str(), kv.second.coef,
                           kv.second.up_done ? kv.second.up : -1) < 0;
    const int nf = fyg_fft_radix_lines(nullptr, 0);
    std::string fl((size_t)nf + 1, '\0');
    fyg_fft_radix_lines(&fl[0], nf + 1);
    bad |= fputs(fl.c_str(), f) < 0;
    bad |= fclose(f) != 0;
    if (!bad) bad = rename(tmp.c_str(), path) != 0;
    if (bad) remove(tmp.c_str());
    return bad;
}

extern "C" int fyg_tune_load(const char *path) {
    FILE *f = fopen(path, "r");
    if (!f) return 1;
    char line[512];
    int bad = 0;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    while (fgets(line, sizeof line, f)) {
        if (line[0] == '#' || line[0] == '\n') continue;
        if (!strncmp(line, "fft\t", 4)) {
            bad |= fyg_fft_radix_load_line(line + 4);
            continue;
        }
        char *t1 = strchr(line, '\t');
        char *t2 = t1 ? strchr(t1 + 1, '\t') : nullptr;
        if (!t1 || !t2) { bad = 1; continue; }
        *t1 = 0;
        const int coef = atoi(t1 + 1), up = atoi(t2 + 1);
        if (coef < 0 || coef >= kNCoef || up < -1 || up >= kNUp) { bad = 1; continue; }
        TuneRec r;
        r.coef = coef;
        r.coef_done = true;
        r.up = up < 0 ? 0 : up;
        r.up_done = up >= 0;
        r.loaded = true;
        g_tune[std::string(line)] = r;
    }
    fclose(f);
    return bad;
}

extern "C" void fyg_tune_clear(void) {
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune.clear();
    }
    fyg_fft_radix_clear();
}

// one line per tuned shape: key, chosen candidates and their times (us)
extern "C" int fyg_tune_report(char *buf, int len) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    std::string out;
    char line[320];
    for (auto &kv : g_tune) {
        const TuneRec &r = kv.second;
        if (r.loaded)
            snprintf(line, sizeof line, "%s: coef zchunk #%d, upsample planes #%d%s (from a tuning file)\n",
                     kv.first.c_str(), r.coef, r.up, r.up_done ? "" : " (untuned)");
        else if (r.up_done)
            snprintf(line, sizeof line, "%s: coef zchunk #%d (%.1f %.1f %.1f us), upsample planes #%d (%.1f %.1f us)\n",
                     kv.first.c_str(), r.coef, r.coef_us[0], r.coef_us[1], r.coef_us[2], r.up, r.up_us[0],
                     r.up_us[1]);
        else
            snprintf(line, sizeof line, "%s: coef zchunk #%d (%.1f %.1f %.1f us), upsample planes #0 (untuned: "
                     "outputs overlap guides)\n", kv.first.c_str(), r.coef, r.coef_us[0], r.coef_us[1], r.coef_us[2]);
        out += line;
    }
    const int nf = fyg_fft_radix_report(nullptr, 0);
    std::string fl((size_t)nf + 1, '\0');
    fyg_fft_radix_report(&fl[0], nf + 1);
    out += fl.c_str();
    if (buf && len > 0) {
        snprintf(buf, (size_t)len, "%s", out.c_str());
    }
    return (int)out.size();
}

extern "C" int fyg_guided_run_batch(const fyg_guided_args *a, int n, fyg_stream *s) {
    if (n < 1) return 0;
    const fyg_guided_args &h = a[0];
    const int r = h.r < 1 ? 2 : h.r;
    bool batch = h.s > 1 && (h.nz + h.s - 1) / h.s >= 2 && (h.ny + h.s - 1) / h.s >= 2 &&
                 (h.nx + h.s - 1) / h.s >= 2 && r / h.s <= 2;
    for (int i = 1; i < n && batch; i++) {
        const fyg_guided_args &b = a[i];
        batch = b.nz == h.nz && b.ny == h.ny && b.nx == h.nx && b.s == h.s && (b.r < 1 ? 2 : b.r) == r &&
                (b.em.quant != 0) == (h.em.quant != 0) &&
                (!h.em.quant || (b.em.tz == h.em.tz && b.em.ty == h.em.ty && b.em.tx == h.em.tx));
    }
    if (h.f64) {                     // all-FP64 variant: fast path only
        for (int i = 0; i < n && batch; i++) batch = a[i].f64 && a[i].in64;
        return batch ? fast_batch<double>(a, n, s) : 1;
    }
    if (batch) return fast_batch<float>(a, n, s);
    for (int i = 0; i < n; i++)
        if (fyg_guided_run(&a[i], s)) return 1;
    return 0;
}

extern "C" int fyg_emit_run(const float *f, int nz, int ny, int nx, const fyg_emit *em, fyg_stream *s) {
    EmitArgs e{f, nz, ny, nx, 0, 0, 0, nz, ny, nx, *em};
    if (em->quant) {
        e.z0 = (int)em->iz0; e.y0 = (int)em->iy0; e.x0 = (int)em->ix0;
        e.dz = (int)em->tz; e.dy = (int)em->ty; e.dx = (int)em->tx;
    }
    if ((size_t)e.dz * e.dy * e.dx) fyg_launch(s, "emit", k_emit, row_grid(e.dx, e.dy, e.dz), dim3(64, 4), 0, e);
    return fyg_last_error();
}

extern "C" int fyg_box_mean_run(const float *in, float *out, int nz, int ny, int nx, int r, void *ws,
                                fyg_stream *s) {
    Grid g{nz, ny, nx, 1, ny, nx};
    size_t n = (size_t)nz * ny * nx;
    if (!n) return 0;
    if (r < 0) r = 0;
    if (r >= 1 && r <= kMaxTileRho) {
        Box1Args b{g, in, out};
        switch (r) {
            case 1: launch_box1<1>(s, b); break;
            case 2: launch_box1<2>(s, b); break;
        }
    } else {
        float *t0 = (float *)ws, *t1 = t0 + n;
        box_global(s, g, r, 1, in, nullptr, t0, nullptr, t1, nullptr);
        if (fyg_d2d(out, t0, sizeof(float) * n, s)) return 1;
    }
    return fyg_last_error();
}
