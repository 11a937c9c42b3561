// fyg_mix.hip -- the deconvolution passes of cubes of mixed shapes, one launch
// per stage for all of them (the batched form's single launch per stage).
//
// A stage's launch is a 1-D grid cut into segments, one per cube: segment i
// covers blocks [start[i], start[i+1]) and names the cube's descriptor and the
// kernel body its length and radix plan select (the same bodies the per-shape
// kernels of fyg_fft_engine.hpp run, so every voxel's arithmetic is theirs).
// A workgroup finds its segment by a binary search over the starts; starts
// and descriptors travel by value in the kernel arguments.  Cubes
// are ordered largest first, so the hardware dispatcher hands the small cubes'
// workgroups to the CUs the large ones leave idle at the end of the stage.
// The dynamic LDS of a launch is the largest body's.
#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "fyg_fft_engine.hpp"

namespace {

// the lengths a mixed launch covers (the sizes of the batched and served
// workloads); a cube with another padded length runs in the grouped form
constexpr int kMixLen[5] = {64, 96, 128, 192, 256};
constexpr int kNMixLen = 5;

int mix_len_index(int L) {
    for (int i = 0; i < kNMixLen; i++)
        if (kMixLen[i] == L) return i;
    return -1;
}

// The stage kernels keep their LDS in dynamic memory, so the compiler cannot
// see that the bodies' ~37 KB per workgroup already limits a CU to four
// workgroups (four waves per SIMD).  FYG_MIX_OCC states that limit, so the
// register allocation of the bodies is planned for the occupancy the launch
// can reach, as it is in the per-shape kernels with their static LDS.
#define FYG_MIX_OCC __attribute__((amdgpu_waves_per_eu(1, 4)))

// body kinds: 2*li + pv for the row, column and Z bodies of length kMixLen[li];
// 2*kNMixLen + 2*li + pv for the plane bodies (li 0: 64, 1: 96)
constexpr int kPlaneKind = 2 * kNMixLen;

#define FYG_MIX_ROWS(BODY)                                                   \
    FYG_MIX_LEN(0, 64, BODY) FYG_MIX_LEN(1, 96, BODY) FYG_MIX_LEN(2, 128, BODY) \
    FYG_MIX_LEN(3, 192, BODY) FYG_MIX_LEN(4, 256, BODY)
#define FYG_MIX_LEN(li, L, BODY)                        \
    case 2 * li: BODY(L, 0); break;                     \
    case 2 * li + 1:                                    \
        if constexpr (plan_alt(L)) BODY(L, 1);          \
        break;
#define FYG_MIX_PLANES(BODY)                                               \
    case kPlaneKind + 0: BODY(64, 0); break;                               \
    case kPlaneKind + 1:                                                   \
        if constexpr (plan_alt(64)) BODY(64, 1);                           \
        break;                                                             \
    case kPlaneKind + 2: BODY(96, 0); break;                               \
    case kPlaneKind + 3:                                                   \
        if constexpr (plan_alt(96)) BODY(96, 1);                           \
        break;

// X forward: rows (u8 gather + prologue, R2C) or fused X+Y planes
template <int MODE, int C>
__global__ __launch_bounds__(256) FYG_MIX_OCC void k_mix_xfwd(MixBatch<XFwd<float>, C> mb) {
    const MixSegs<C> &t = mb.t;
    const int sg = mix_seg(t, blockIdx.x);
    const long b = mix_uniform((int)blockIdx.x - t.start[sg]);
    const XFwd<float> a = mb.a[t.idx[sg]];          // into registers once
#define ROWB(L, PV) rfft_x_body<L, MODE, float, PV>(a, b, mix_lds_as<RowLds<L, float, XLine>>())
#define PLANEB(L, PV) rfft_xy_body<L, MODE, PV>(a, (int)b, mix_lds_as<PlaneLds<L, XLine>>())
    switch (t.kind[sg]) {
        FYG_MIX_ROWS(ROWB)
        FYG_MIX_PLANES(PLANEB)
    }
#undef ROWB
#undef PLANEB
}

// Y forward or inverse: column groups of every batch (z plane) of a cube
template <int C>
__global__ __launch_bounds__(256) FYG_MIX_OCC void k_mix_cols(MixBatch<Cols<float>, C> mb) {
    const MixSegs<C> &t = mb.t;
    const int sg = mix_seg(t, blockIdx.x);
    const int b = mix_uniform((int)blockIdx.x - t.start[sg]), g = mix_uniform(t.nbx[sg]);
    const long bx = mix_uniform(b % g), by = mix_uniform(b / g);
    const Cols<float> &a = mb.a[t.idx[sg]];
#define COLB(L, PV) cols_body<L, float, PV>(a, bx, by, mix_lds_as<ColLds<L, float>>())
    switch (t.kind[sg]) { FYG_MIX_ROWS(COLB) }
#undef COLB
}

// Z forward, weight, Z inverse
template <int C>
__global__ __launch_bounds__(256) FYG_MIX_OCC void k_mix_z(MixBatch<ZFused<float>, C> mb) {
    const MixSegs<C> &t = mb.t;
    const int sg = mix_seg(t, blockIdx.x);
    const long b = mix_uniform((int)blockIdx.x - t.start[sg]);
    const ZFused<float> &a = mb.a[t.idx[sg]];
#define ZB(L, PV) zfused_body<L, float, PV>(a, b, mix_lds_as<ZLds<L, float>>())
    switch (t.kind[sg]) { FYG_MIX_ROWS(ZB) }
#undef ZB
}

// X inverse: rows (C2R + epilogue) or fused Y+X planes
template <int C>
__global__ __launch_bounds__(256) FYG_MIX_OCC void k_mix_xinv(MixBatch<XInv<float>, C> mb) {
    const MixSegs<C> &t = mb.t;
    const int sg = mix_seg(t, blockIdx.x);
    const long b = mix_uniform((int)blockIdx.x - t.start[sg]);
    const XInv<float> a = mb.a[t.idx[sg]];          // into registers once
#define ROWB(L, PV) irfft_x_body<L, float, PV>(a, b, mix_lds_as<RowLds<L, float, XInvLine>>())
#define PLANEB(L, PV) irfft_yx_body<L, PV>(a, (int)b, mix_lds_as<PlaneLds<L, XInvLine>>())
    switch (t.kind[sg]) {
        FYG_MIX_ROWS(ROWB)
        FYG_MIX_PLANES(PLANEB)
    }
#undef ROWB
#undef PLANEB
}

// ring correction images: row y of image si, x block xb (nbx = x blocks per
// row, nby = rows); workgroup 0 steps the generation word when step is set
struct MixRing {
    RingImg r;
    int nxb;
};
template <int C>
__global__ __launch_bounds__(256) void k_mix_ring(MixBatch<MixRing, C> mb, int *step, int nflag) {
    if (step && blockIdx.x == 0) fyg_gen_step_block(step, nflag);
    const MixSegs<C> &t = mb.t;
    const int sg = mix_seg(t, blockIdx.x);
    const int b = mix_uniform((int)blockIdx.x - t.start[sg]);
    const MixRing &m = mb.a[t.idx[sg]];
    const int xb = mix_uniform(b % m.nxb), row = mix_uniform(b / m.nxb);
    ring_image_at(m.r, xb * 256 + threadIdx.x, row % m.r.ny, row / m.r.ny);
}

#undef FYG_MIX_OCC
#undef FYG_MIX_ROWS
#undef FYG_MIX_LEN
#undef FYG_MIX_PLANES

// ---- host side
template <int L>
constexpr size_t row_lds() {
    return sizeof(RowLds<L, float, XLine>) > sizeof(RowLds<L, float, XInvLine>) ? sizeof(RowLds<L, float, XLine>)
                                                                                 : sizeof(RowLds<L, float, XInvLine>);
}
template <int L>
constexpr size_t plane_lds() {
    return sizeof(PlaneLds<L, XLine>) > sizeof(PlaneLds<L, XInvLine>) ? sizeof(PlaneLds<L, XLine>)
                                                                      : sizeof(PlaneLds<L, XInvLine>);
}
// LDS bytes of a stage body of length L: 0 row passes, 1 columns, 2 Z, 3 planes
size_t body_lds(int what, int L) {
    size_t b = 0;
    dispatch(L, [&](auto c) {
        constexpr int LL = decltype(c)::value;
        if constexpr (LL >= 64 && LL <= 256) {
            switch (what) {
            case 0: b = row_lds<LL>(); break;
            case 1: b = sizeof(ColLds<LL, float>); break;
            case 2: b = sizeof(ZLds<LL, float>); break;
            default:
                if constexpr (PlaneShape<LL>::OK) b = plane_lds<LL>();
                break;
            }
        }
    });
    return b;
}

}  // namespace

static_assert(sizeof(MixBatch<XFwd<float>, FYG_MIX_MAX>) <= kMixArgMax, "mixed-launch arguments too large");
static_assert(sizeof(MixBatch<Cols<float>, FYG_MIX_MAX>) <= kMixArgMax, "mixed-launch arguments too large");
static_assert(sizeof(MixBatch<ZFused<float>, FYG_MIX_MAX>) <= kMixArgMax, "mixed-launch arguments too large");
static_assert(sizeof(MixBatch<XInv<float>, FYG_MIX_MAX>) <= kMixArgMax, "mixed-launch arguments too large");
static_assert(sizeof(MixBatch<MixRing, FYG_MIX_MAX>) <= kMixArgMax, "mixed-launch arguments too large");

extern "C" int fyg_mix_len_ok(int L) { return mix_len_index(L) >= 0; }

// n cubes' deconvolutions (fp32, u8 input, one ring mode), one launch per stage
extern "C" int fyg_mix_deconv_run(const fyg_deconv_args *a, int n, fyg_stream *s) {
    if (n < 1) return 0;
    if (n > FYG_MIX_MAX) return 1;
    const int mode = a[0].pro.dering ? 2 : 1;
    for (int i = 0; i < n; i++) {
        const fyg_deconv_args &d = a[i];
        if (d.f64 || !d.in_u8 || (d.pro.dering ? 2 : 1) != mode || mix_len_index(d.px) < 0 ||
            mix_len_index(d.py) < 0 || mix_len_index(d.pz) < 0)
            return 1;
        if (mode == 2 && !d.pro.cimg) return 1;
    }
    const bool may_tune = !s->capturing;
    // per-cube descriptors, built as the grouped form builds them
    std::vector<XFwd<float>> xf(n);
    std::vector<Cols<float>> yf(n), yi(n);
    std::vector<ZFused<float>> zf(n);
    std::vector<XInv<float>> xi(n);
    std::vector<int> planes(n), px(n), py(n), pz(n), pvx(n), pvy(n), pvz(n), pvxy(n);
    std::vector<double> work(n);
    for (int i = 0; i < n; i++) {
        Passes<float> ps;
        if (ps.build(&a[i], 1)) return 1;
        xf[i] = ps.xf[0];
        yf[i] = ps.yf[0];
        yi[i] = ps.yi[0];
        zf[i] = ps.zf[0];
        xi[i] = ps.xi[0];
        const bool tune = may_tune && !a[i].no_tune;
        px[i] = a[i].px; py[i] = a[i].py; pz[i] = a[i].pz;
        planes[i] = use_planes<float>(px[i], py[i]);
        if (planes[i]) {
            pvxy[i] = fyg_radix_plan(FYG_RK_XY, px[i], s, tune);
        } else {
            pvx[i] = fyg_radix_plan(FYG_RK_X, px[i], s, tune);
            pvy[i] = fyg_radix_plan(FYG_RK_Y, py[i], s, tune);
        }
        pvz[i] = fyg_radix_plan(FYG_RK_Z, pz[i], s, tune);
        work[i] = (double)a[i].pz * a[i].py * a[i].hxp;
    }
    // largest spectrum first
    std::vector<int> ord(n);
    for (int i = 0; i < n; i++) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int p, int q) { return work[p] > work[q]; });

    MixStage sx, syf, sz, syi, sxi, sr;
    std::vector<MixRing> rings;
    for (int i : ord) {
        const int lx = mix_len_index(px[i]), ly = mix_len_index(py[i]), lz = mix_len_index(pz[i]);
        if (planes[i]) {
            const int k = kPlaneKind + 2 * lx + pvxy[i];
            sx.add(i, k, a[i].pz, body_lds(3, px[i]));
            sxi.add(i, k, a[i].nz, body_lds(3, px[i]));
        } else {
            dispatch(px[i], [&](auto c) {
                constexpr int L = decltype(c)::value;
                const long g = (xf[i].nlines + Shape<L, float>::NL_RM - 1) / Shape<L, float>::NL_RM;
                const long gi = (xi[i].nlines + Shape<L, float>::NL_RM - 1) / Shape<L, float>::NL_RM;
                sx.add(i, 2 * lx + pvx[i], g, body_lds(0, L));
                sxi.add(i, 2 * lx + pvx[i], gi, body_lds(0, L));
            });
            dispatch(py[i], [&](auto c) {
                constexpr int L = decltype(c)::value;
                const long g = (yf[i].ncols + Shape<L, float>::NL_CM - 1) / Shape<L, float>::NL_CM;
                syf.add(i, 2 * ly + pvy[i], g * a[i].pz, body_lds(1, L), (int)g);
                syi.add(i, 2 * ly + pvy[i], g * a[i].nz, body_lds(1, L), (int)g);
            });
        }
        dispatch(pz[i], [&](auto c) {
            constexpr int L = decltype(c)::value;
            const long nc = (long)zf[i].py * zf[i].hxp;
            sz.add(i, 2 * lz + pvz[i], (nc + Shape<L, float>::NL_ZF - 1) / Shape<L, float>::NL_ZF, body_lds(2, L));
        });
        if (mode == 2) {
            const fyg_prologue &PR = a[i].pro;
            const int s0 = ring_slab(PR, PR.gz0), s1 = ring_slab(PR, PR.gz0 + a[i].nz - 1);
            MixRing m{RingImg{PR, a[i].ny, a[i].nx, s0, s1 - s0 + 1}, (a[i].nx + 255) / 256};
            sr.add((int)rings.size(), 0, (long)m.nxb * a[i].ny * m.r.nimg, 0);
            rings.push_back(m);
        }
    }
    for (const MixStage *st : {&sx, &syf, &sz, &syi, &sxi, &sr})
        if (st->total > 0x7fffffff) return 1;

    // launches of capacity C (the kernel-argument bytes follow the capacity)
    auto run = [&](auto cap) -> int {
        constexpr int C = decltype(cap)::value;
        const fyg_deconv_args &h = a[0];
        if (mode == 2) {
            auto mr = std::make_unique<MixBatch<MixRing, C>>();
            sr.fill(*mr, rings.data(), (int)rings.size());
            fyg_launch(s, "mix_ring", k_mix_ring<C>, dim3((unsigned)sr.total), dim3(256), 0, *mr, h.step_gen,
                       h.step_nflag);
        } else if (h.step_gen && fyg_gen_step(h.step_gen, h.step_nflag, s)) {
            return 1;
        }
        {
            auto mx = std::make_unique<MixBatch<XFwd<float>, C>>();
            sx.fill(*mx, xf.data(), n);
            if (mode == 2)
                fyg_launch(s, "mix_xfwd", k_mix_xfwd<2, C>, dim3((unsigned)sx.total), dim3(256), sx.lds, *mx);
            else
                fyg_launch(s, "mix_xfwd", k_mix_xfwd<1, C>, dim3((unsigned)sx.total), dim3(256), sx.lds, *mx);
        }
        auto mc = std::make_unique<MixBatch<Cols<float>, C>>();
        if (syf.total) {
            syf.fill(*mc, yf.data(), n);
            fyg_launch(s, "mix_cols", k_mix_cols<C>, dim3((unsigned)syf.total), dim3(256), syf.lds, *mc);
        }
        {
            auto mz = std::make_unique<MixBatch<ZFused<float>, C>>();
            sz.fill(*mz, zf.data(), n);
            fyg_launch(s, "mix_z", k_mix_z<C>, dim3((unsigned)sz.total), dim3(256), sz.lds, *mz);
        }
        if (syi.total) {
            syi.fill(*mc, yi.data(), n);
            fyg_launch(s, "mix_cols", k_mix_cols<C>, dim3((unsigned)syi.total), dim3(256), syi.lds, *mc);
        }
        {
            auto mi = std::make_unique<MixBatch<XInv<float>, C>>();
            sxi.fill(*mi, xi.data(), n);
            fyg_launch(s, "mix_xinv", k_mix_xinv<C>, dim3((unsigned)sxi.total), dim3(256), sxi.lds, *mi);
        }
        return 0;
    };
    const int rc = n <= FYG_MIX_SMALL ? run(std::integral_constant<int, FYG_MIX_SMALL>{})
                                      : run(std::integral_constant<int, FYG_MIX_MAX>{});
    if (rc) return 1;
    return fyg_last_error();
}
