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
// The moment kernel subtracts a per-tile constant c0 from I before forming
// box(I-c0) and box((I-c0)^2): var = E[(I-c0)^2] - E[I-c0]^2 is the same
// quantity with far less cancellation than E[I^2] - E[I]^2 in fp32.
//
// Exact path (guided.c:113-137):   k_moments(I) -> a,b ; k_box2(a,b) -> out = ma*I + mb
// Fast path  (guided.c:152-231):   k_moments(I decimated by s) -> a,b (low-res);
//                                  k_box2 -> ma, mb (low-res); k_upsample -> out
#include <cstdio>

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

__device__ __forceinline__ float count_recip(int i, int n, int rho) {
    int lo = i - rho < 0 ? 0 : i - rho;
    int hi = i + rho > n - 1 ? n - 1 : i + rho;
    return 1.0f / (float)(hi - lo + 1);
}

// Generic tiled box over NF fields.  load(gz,gy,gx,v[NF]) fills in-grid values;
// cells outside the grid contribute 0.  out(gz,gy,gx,sum[NF],recip) is called for
// every in-grid output voxel of the tile.
template <int RHO, int NF, class Load, class Out>
__device__ __forceinline__ void box_tile(const Grid &g, float *sm, Load load, Out out) {
    using Tl = Tile<RHO>;
    float *A = sm;                     // [NF][AREG]
    float *B = sm + NF * Tl::AREG;     // [NF][BREG]
    const int bx = blockIdx.x * TX, by = blockIdx.y * TY, bz = blockIdx.z * TZ;
    for (int i = threadIdx.x; i < Tl::AREG; i += NT) {
        int ex = i % Tl::EX, t = i / Tl::EX, ey = t % Tl::EY, ez = t / Tl::EY;
        int gx = bx + ex - RHO, gy = by + ey - RHO, gz = bz + ez - RHO;
        float v[NF];
        if (gx >= 0 && gx < g.nx && gy >= 0 && gy < g.ny && gz >= 0 && gz < g.nz) {
            load(gz, gy, gx, v);
        } else {
#pragma unroll
            for (int f = 0; f < NF; f++) v[f] = 0.0f;
        }
#pragma unroll
        for (int f = 0; f < NF; f++) A[f * Tl::AREG + i] = v[f];
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
__device__ __forceinline__ float dither01(long z, long y, long x) {
    unsigned int h = (unsigned int)(z * 73856093L) ^ (unsigned int)(y * 19349663L) ^ (unsigned int)(x * 83492791L);
    h ^= h >> 16;
    h *= 0x7feb352dU;
    h ^= h >> 15;
    h *= 0x846ca68bU;
    h ^= h >> 16;
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}

__device__ __forceinline__ unsigned char quant_u8(float v, long gz, long gy, long gx, int no_dither) {
    #pragma clang fp contract(off)
    // v*255 + d rounded as two separate operations, as written in pipeline.c:61
    float t = v * 255.0f + (no_dither ? 0.5f : dither01(gz, gy, gx));
    if (t <= 0.0f) return 0;
    int q = (int)t;
    return q > 255 ? 255 : (unsigned char)q;
}

// quantised output of one voxel at local (z,y,x): recenter, dither, all-air
__device__ __forceinline__ unsigned char emit_q(const fyg_emit &E, bool keep, long z, long y, long x, float v) {
    #pragma clang fp contract(off)
    if (!keep) return 0;
    if (E.recenter) {
        if (v <= 0.0f) v = 0.0f;
        else {
            float w = 0.5f + (v - E.rmid) * E.rk;
            v = w < 0.0f ? 0.0f : (w > 1.0f ? 1.0f : w);
        }
    }
    return quant_u8(v, E.gz0 + z, E.gy0 + y, E.gx0 + x, E.no_dither);
}

__device__ __forceinline__ bool emit_keep(const fyg_emit &E) {
    return !E.nonzero_flag || *E.nonzero_flag == E.nonzero_gen;
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
    unsigned char q[4];
    for (int j = 0; j < 4; j++) q[j] = emit_q(E, keep, z, y, x + j, v[j]);
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

// box(I-c0), box((I-c0)^2) -> a = var/(var+eps), b = mean*(1-a)   (guided.c:122-132)
template <int RHO>
__global__ __launch_bounds__(NT) void k_moments(MomArgs p) {
    extern __shared__ float sm[];
    const Grid &g = p.g;
    // tile-local centring constant: the source voxel at the tile origin
    const int bx = blockIdx.x * TX, by = blockIdx.y * TY, bz = blockIdx.z * TZ;
    const float c0 = p.in[((size_t)bz * g.s * g.fny + (size_t)by * g.s) * g.fnx + (size_t)bx * g.s];
    auto load = [&](int z, int y, int x, float *v) {
        float t = p.in[((size_t)z * g.s * g.fny + (size_t)y * g.s) * g.fnx + (size_t)x * g.s] - c0;
        v[0] = t;
        v[1] = t * t;
    };
    const float epsf = (float)p.eps;
    auto out = [&](int z, int y, int x, const float *s, float rc) {
        float m = s[0] * rc, q = s[1] * rc;
        float var = q - m * m;
        if (var < 0.0f) var = 0.0f;
        float ai = var / (var + epsf);
        size_t o = ((size_t)z * g.ny + y) * g.nx + x;
        p.a[o] = ai;
        p.b[o] = (m + c0) * (1.0f - ai);
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

// ---- fused low-res coefficient field of the fast path (guided.c:163-182):
// one workgroup produces mean_a, mean_b for a 32x8x4 tile of the decimated grid.
// It loads the decimated, tile-centred samples with a 2*rho halo, forms
// box(I-c0), box((I-c0)^2) -> a, b on the tile plus a rho halo, then box(a),
// box(b) on the tile -- all in LDS, so a and b never reach HBM.
template <int RHO>
struct CoefTile {
    static constexpr int X2 = TX + 4 * RHO, Y2 = TY + 4 * RHO, Z2 = TZ + 4 * RHO;   // samples
    static constexpr int X1 = TX + 2 * RHO, Y1 = TY + 2 * RHO, Z1 = TZ + 2 * RHO;   // a, b
    static constexpr int P0 = 2 * Z2 * Y2 * X1;     // largest intermediate
    static constexpr int P1 = (2 * Z2 * Y1 * X1 > Z2 * Y2 * X2) ? 2 * Z2 * Y1 * X1 : Z2 * Y2 * X2;
    static constexpr size_t bytes = sizeof(float) * (P0 + P1);
};

struct CoefArgs {
    Grid g;                 // decimated grid; source rows are s*y, planes s*z, columns s*x
    const float *in;
    double eps;
    float *ma, *mb;
};

template <int RHO>
__global__ __launch_bounds__(NT) void k_coef(CoefArgs p) {
    using CT = CoefTile<RHO>;
    extern __shared__ float sm[];
    float *P0 = sm, *P1 = sm + CT::P0;
    const Grid &g = p.g;
    const int bx = blockIdx.x * TX, by = blockIdx.y * TY, bz = blockIdx.z * TZ;
    const size_t sz = (size_t)g.s * g.fny * g.fnx, sy = (size_t)g.s * g.fnx;
    const float c0 = p.in[(size_t)bz * sz + (size_t)by * sy + (size_t)bx * g.s];
    // 1. centred samples over the tile + 2*rho halo (zero outside the grid)
    for (int i = threadIdx.x; i < CT::Z2 * CT::Y2 * CT::X2; i += NT) {
        int ex = i % CT::X2, t = i / CT::X2, ey = t % CT::Y2, ez = t / CT::Y2;
        int gx = bx + ex - 2 * RHO, gy = by + ey - 2 * RHO, gz = bz + ez - 2 * RHO;
        float v = 0.0f;
        if (gx >= 0 && gx < g.nx && gy >= 0 && gy < g.ny && gz >= 0 && gz < g.nz)
            v = p.in[(size_t)gz * sz + (size_t)gy * sy + (size_t)gx * g.s] - c0;
        P1[i] = v;
    }
    __syncthreads();
    // 2. x sums of v and v^2: [Z2][Y2][X1] x 2
    constexpr int NX = CT::Z2 * CT::Y2 * CT::X1;
    for (int i = threadIdx.x; i < NX; i += NT) {
        int x = i % CT::X1, row = i / CT::X1;
        const float *a = P1 + row * CT::X2 + x;
        float s1 = 0.0f, s2 = 0.0f;
#pragma unroll
        for (int d = 0; d <= 2 * RHO; d++) { s1 += a[d]; s2 += a[d] * a[d]; }
        P0[i] = s1;
        P0[NX + i] = s2;
    }
    __syncthreads();
    // 3. y sums: [Z2][Y1][X1] x 2
    constexpr int NY = CT::Z2 * CT::Y1 * CT::X1;
    for (int i = threadIdx.x; i < NY; i += NT) {
        int x = i % CT::X1, t = i / CT::X1, y = t % CT::Y1, z = t / CT::Y1;
        const float *a = P0 + (z * CT::Y2 + y) * CT::X1 + x;
        float s1 = 0.0f, s2 = 0.0f;
#pragma unroll
        for (int d = 0; d <= 2 * RHO; d++) { s1 += a[d * CT::X1]; s2 += a[NX + d * CT::X1]; }
        P1[i] = s1;
        P1[NY + i] = s2;
    }
    __syncthreads();
    // 4. z sums -> a, b on the tile + rho halo (zero outside the grid): [Z1][Y1][X1] x 2
    constexpr int NA = CT::Z1 * CT::Y1 * CT::X1;
    const float epsf = (float)p.eps;
    for (int i = threadIdx.x; i < NA; i += NT) {
        int x = i % CT::X1, t = i / CT::X1, y = t % CT::Y1, z = t / CT::Y1;
        int gx = bx + x - RHO, gy = by + y - RHO, gz = bz + z - RHO;
        float av = 0.0f, bv = 0.0f;
        if (gx >= 0 && gx < g.nx && gy >= 0 && gy < g.ny && gz >= 0 && gz < g.nz) {
            const float *a = P1 + (z * CT::Y1 + y) * CT::X1 + x;
            float s1 = 0.0f, s2 = 0.0f;
#pragma unroll
            for (int d = 0; d <= 2 * RHO; d++) {
                s1 += a[d * CT::Y1 * CT::X1];
                s2 += a[NY + d * CT::Y1 * CT::X1];
            }
            float rc = count_recip(gx, g.nx, RHO) * count_recip(gy, g.ny, RHO) * count_recip(gz, g.nz, RHO);
            float m = s1 * rc, q = s2 * rc;
            float var = q - m * m;
            if (var < 0.0f) var = 0.0f;
            av = var / (var + epsf);
            bv = (m + c0) * (1.0f - av);
        }
        P0[i] = av;
        P0[NA + i] = bv;
    }
    __syncthreads();
    // 5. x sums of a, b: [Z1][Y1][TX] x 2
    constexpr int NB2 = CT::Z1 * CT::Y1 * TX;
    for (int i = threadIdx.x; i < NB2; i += NT) {
        int x = i % TX, row = i / TX;
        const float *a = P0 + row * CT::X1 + x;
        float s1 = 0.0f, s2 = 0.0f;
#pragma unroll
        for (int d = 0; d <= 2 * RHO; d++) { s1 += a[d]; s2 += a[NA + d]; }
        P1[i] = s1;
        P1[NB2 + i] = s2;
    }
    __syncthreads();
    // 6. y sums: [Z1][TY][TX] x 2
    constexpr int NC = CT::Z1 * TY * TX;
    for (int i = threadIdx.x; i < NC; i += NT) {
        int x = i % TX, t = i / TX, y = t % TY, z = t / TY;
        const float *a = P1 + (z * CT::Y1 + y) * TX + x;
        float s1 = 0.0f, s2 = 0.0f;
#pragma unroll
        for (int d = 0; d <= 2 * RHO; d++) { s1 += a[d * TX]; s2 += a[NB2 + d * TX]; }
        P0[i] = s1;
        P0[NC + i] = s2;
    }
    __syncthreads();
    // 7. z sums -> mean_a, mean_b of the tile
    for (int i = threadIdx.x; i < TZ * TY * TX; i += NT) {
        int x = i % TX, t = i / TX, y = t % TY, z = t / TY;
        int gx = bx + x, gy = by + y, gz = bz + z;
        if (gx >= g.nx || gy >= g.ny || gz >= g.nz) continue;
        const float *a = P0 + (z * TY + y) * TX + x;
        float s1 = 0.0f, s2 = 0.0f;
#pragma unroll
        for (int d = 0; d <= 2 * RHO; d++) { s1 += a[d * TY * TX]; s2 += a[NC + d * TY * TX]; }
        float rc = count_recip(gx, g.nx, RHO) * count_recip(gy, g.ny, RHO) * count_recip(gz, g.nz, RHO);
        size_t o = ((size_t)gz * g.ny + gy) * g.nx + gx;
        p.ma[o] = s1 * rc;
        p.mb[o] = s2 * rc;
    }
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

struct UpArgs {
    int nz, ny, nx, lz, ly, lx, s;
    const float *ma, *mb, *in;
    int z0, y0, x0, dz, dy, dx;      // region of output voxels to produce
    int vin, vout;                   // 16-byte input rows / aligned 4-voxel output groups
    fyg_emit em;
};

constexpr int UPX = 4;               // consecutive x outputs per thread
constexpr int UPK = 64 * UPX / 2 + 4;  // low-res columns one 256-wide chunk touches (s >= 2)

// trilinear upsample of (mean_a, mean_b) from the decimated grid and apply
// out = a*I + b (guided.c:188-228): u = c/s, i0 = min((int)u, l-2), f = min(u-i0, 1).
// As in the reference, the four (z,y) low-res rows are blended once per output
// row -- here into LDS, shared by the 64 threads of the row -- and each voxel
// then needs only the x lerp.  Block (64, 4): 4 rows x 256 voxels.
__global__ __launch_bounds__(NT) void k_upsample(UpArgs p) {
    __shared__ float sa[4][UPK], sb[4][UPK];
    const int ty = threadIdx.y;
    const int y = p.y0 + blockIdx.y * 4 + ty;
    const int z = p.z0 + blockIdx.z;
    const bool row_ok = y < p.y0 + p.dy;
    const float fs = (float)p.s;
    const int xc0 = p.x0 + blockIdx.x * 64 * UPX;
    const int xc1 = min(xc0 + 64 * UPX, p.x0 + p.dx) - 1;
    int kb = (int)((float)xc0 / fs); if (kb > p.lx - 2) kb = p.lx - 2;
    int ke = (int)((float)xc1 / fs); if (ke > p.lx - 2) ke = p.lx - 2;
    const int nk = ke - kb + 2;
    if (row_ok) {
        float uz = (float)z / fs, uy = (float)y / fs;
        int z0 = (int)uz; if (z0 > p.lz - 2) z0 = p.lz - 2;
        int y0 = (int)uy; if (y0 > p.ly - 2) y0 = p.ly - 2;
        float zf = uz - z0; if (zf > 1.0f) zf = 1.0f;
        float yf = uy - y0; if (yf > 1.0f) yf = 1.0f;
        const float w00 = (1 - zf) * (1 - yf), w01 = (1 - zf) * yf, w10 = zf * (1 - yf), w11 = zf * yf;
        const size_t q00 = ((size_t)z0 * p.ly + y0) * p.lx + kb, q01 = q00 + p.lx;
        const size_t q10 = q00 + (size_t)p.ly * p.lx, q11 = q10 + p.lx;
        for (int k = threadIdx.x; k < nk; k += 64) {
            sa[ty][k] = w00 * p.ma[q00 + k] + w01 * p.ma[q01 + k] + w10 * p.ma[q10 + k] + w11 * p.ma[q11 + k];
            sb[ty][k] = w00 * p.mb[q00 + k] + w01 * p.mb[q01 + k] + w10 * p.mb[q10 + k] + w11 * p.mb[q11 + k];
        }
    }
    __syncthreads();
    if (!row_ok) return;
    const int xb = xc0 + threadIdx.x * UPX;
    if (xb > xc1) return;
    const float *irow = p.in + ((size_t)z * p.ny + y) * p.nx;
    const bool full = xb + UPX - 1 <= xc1;
    float iv[UPX], ov[UPX];
    if (full && p.vin) {
        float4 t = *(const float4 *)(irow + xb);
        iv[0] = t.x; iv[1] = t.y; iv[2] = t.z; iv[3] = t.w;
    } else {
        for (int j = 0; j < UPX; j++) iv[j] = xb + j <= xc1 ? irow[xb + j] : 0.0f;
    }
#pragma unroll
    for (int j = 0; j < UPX; j++) {
        int x = xb + j;
        float ux = (float)x / fs;
        int x0 = (int)ux; if (x0 > p.lx - 2) x0 = p.lx - 2;
        float fx = ux - x0; if (fx > 1.0f) fx = 1.0f;
        float gx = 1 - fx;
        int k = x0 - kb;
        float av = gx * sa[ty][k] + fx * sa[ty][k + 1], bv = gx * sb[ty][k] + fx * sb[ty][k + 1];
        ov[j] = av * iv[j] + bv;
    }
    const bool keep = emit_keep(p.em);
    if (full) {
        emit4(p.em, p.ny, p.nx, z, y, xb, ov, p.vout != 0, keep);
    } else {
        for (int j = 0; j < UPX && xb + j <= xc1; j++) emit(p.em, p.ny, p.nx, z, y, xb + j, ov[j]);
    }
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
    int mode;               // 0: (I-c0),(I-c0)^2 from decimated source; 1: moments -> a,b; 2: apply
    Grid g;
    const float *in;
    float c0;
    double eps;
    const float *m0, *m1;
    float *o0, *o1;
    fyg_emit em;
};

__global__ void k_point(PointArgs p) {
    const Grid &g = p.g;
    size_t n = (size_t)g.nz * g.ny * g.nx;
    for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT) {
        int x = (int)(i % g.nx);
        size_t t = i / g.nx;
        int y = (int)(t % g.ny), z = (int)(t / g.ny);
        if (p.mode == 0) {
            float v = p.in[((size_t)z * g.s * g.fny + (size_t)y * g.s) * g.fnx + (size_t)x * g.s] - p.c0;
            p.o0[i] = v;
            p.o1[i] = v * v;
        } else if (p.mode == 1) {
            float m = p.m0[i], q = p.m1[i];
            float var = q - m * m;
            if (var < 0.0f) var = 0.0f;
            float ai = var / (var + (float)p.eps);
            p.o0[i] = ai;
            p.o1[i] = (m + p.c0) * (1.0f - ai);
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

extern "C" size_t fyg_guided_ws_bytes(int nz, int ny, int nx, int s) {
    size_t n = (size_t)nz * ny * nx;
    if (s > 1) {
        size_t lz = (nz + s - 1) / s, ly = (ny + s - 1) / s, lx = (nx + s - 1) / s;
        size_t nl = lz * ly * lx;
        if (lz >= 2 && ly >= 2 && lx >= 2) return sizeof(float) * 8 * nl;
    }
    return sizeof(float) * 6 * n;       // a, b (+ 4 temporaries for the wide-radius fallback)
}

extern "C" int fyg_guided_run(const fyg_guided_args *a, fyg_stream *s) {
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
            PointArgs p0{0, g, a->in, 0.0f, a->eps, nullptr, nullptr, O0, O1, a->em};
            // c0 = first voxel, read on device inside k_point via a one-element copy
            float c0 = 0.0f;
            if (fyg_d2h(&c0, a->in, sizeof(float), s) || fyg_sync(s)) return 1;
            p0.c0 = c0;
            fyg_launch(s, "box_point", k_point, dim3(grid1d(n)), dim3(NT), 0, p0);
            box_global(s, g, r, 2, O0, O1, T0, T1, A, B);   // means in T0, T1
            PointArgs p1{1, g, a->in, c0, a->eps, T0, T1, A, B, a->em};
            fyg_launch(s, "box_point", k_point, dim3(grid1d(n)), dim3(NT), 0, p1);
            box_global(s, g, r, 2, A, B, O0, O1, T0, T1);   // means in O0, O1
            PointArgs p2{2, g, a->in, 0.0f, a->eps, O0, O1, nullptr, nullptr, a->em};
            fyg_launch(s, "box_point", k_point, dim3(grid1d(n)), dim3(NT), 0, p2);
        }
        return fyg_last_error();
    }
    size_t nl = (size_t)lz * ly * lx;
    int rl = r / sd;
    if (rl < 1) rl = 1;
    float *A = ws, *B = ws + nl, *MA = ws + 2 * nl, *MB = ws + 3 * nl;
    Grid gl{lz, ly, lx, sd, ny, nx};
    if (rl == 1) {
        CoefArgs c{gl, a->in, a->eps, MA, MB};
        fyg_launch(s, "guided_coef", k_coef<1>, tile_grid(gl), dim3(NT), CoefTile<1>::bytes, c);
    } else if (rl <= kMaxTileRho) {
        MomArgs m{gl, a->in, a->eps, A, B};
        moments(s, m, rl);
        Box2Args b{gl, A, B, 0, MA, MB, nullptr, a->em};
        box2(s, b, rl);
    } else {
        float *T0 = ws + 4 * nl, *T1 = ws + 5 * nl, *O0 = ws + 6 * nl, *O1 = ws + 7 * nl;
        float c0 = 0.0f;
        if (fyg_d2h(&c0, a->in, sizeof(float), s) || fyg_sync(s)) return 1;
        PointArgs p0{0, gl, a->in, c0, a->eps, nullptr, nullptr, O0, O1, a->em};
        fyg_launch(s, "box_point", k_point, dim3(grid1d(nl)), dim3(NT), 0, p0);
        box_global(s, gl, rl, 2, O0, O1, T0, T1, A, B);
        PointArgs p1{1, gl, a->in, c0, a->eps, T0, T1, A, B, a->em};
        fyg_launch(s, "box_point", k_point, dim3(grid1d(nl)), dim3(NT), 0, p1);
        box_global(s, gl, rl, 2, A, B, O0, O1, T0, T1);
        if (fyg_d2d(MA, O0, sizeof(float) * nl, s) || fyg_d2d(MB, O1, sizeof(float) * nl, s)) return 1;
    }
    UpArgs u{nz, ny, nx, lz, ly, lx, sd, MA, MB, a->in, 0, 0, 0, nz, ny, nx, 0, 0, a->em};
    if (a->em.quant) {
        u.z0 = (int)a->em.iz0; u.y0 = (int)a->em.iy0; u.x0 = (int)a->em.ix0;
        u.dz = (int)a->em.tz; u.dy = (int)a->em.ty; u.dx = (int)a->em.tx;
    }
    u.vin = (nx % 4 == 0) && (u.x0 % 4 == 0) && ((uintptr_t)a->in % 16 == 0);
    u.vout = a->em.quant ? (a->em.tx % 4 == 0 && (uintptr_t)a->em.q % 4 == 0)
                         : (u.vin && (uintptr_t)a->em.out % 16 == 0);
    if ((size_t)u.dz * u.dy * u.dx)
        fyg_launch(s, "guided_upsample", k_upsample, row_grid(u.dx, u.dy, u.dz), dim3(64, 4), 0, u);
    return fyg_last_error();
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
