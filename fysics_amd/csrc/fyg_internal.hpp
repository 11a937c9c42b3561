// fyg_internal.hpp -- shared by the HIP translation units: the stream object,
// the launch helper that brackets each launch with timing events when
// profiling is on, and small device utilities.
#pragma once
#include <hip/hip_runtime.h>

#include <cstring>

#include "fyg.h"

struct fyg_stream {
    hipStream_t s = nullptr;
    bool owned = false;
    void *ws[FYG_WS_SLOTS] = {};
    size_t wcap[FYG_WS_SLOTS] = {};
    void *pin[2] = {};
    size_t pcap[2] = {};
    int *flag = nullptr;     // [0] the device generation word, then the all-air flags
    int nflag = 0;
    int persist = 0;         // tables from the process-wide cache (stable addresses for graphs)
    int capturing = 0;       // launches are being recorded into a graph: no timing events
    int own_queue = 0;       // created with a hardware queue of its own (a fork's eager helper)
    struct Arena *arena = nullptr;   // per-stream table slots (fyg_stream_table)
    fyg_stream *sub[FYG_SUBS] = {};  // helper streams of a fork (own workspaces), created on first use
    hipEvent_t fork_ev = nullptr;
    hipEvent_t join_ev[FYG_SUBS] = {};
};

namespace fyg_prof {
void begin(fyg_stream *s, const char *name, void **tok);
void end(fyg_stream *s, void *tok);
}  // namespace fyg_prof

// launch k<<<g,b,sh,stream>>>(args...) and time it when profiling is enabled
template <typename... KArgs, typename... Args>
inline void fyg_launch(fyg_stream *s, const char *name, void (*k)(KArgs...), dim3 g, dim3 b,
                       size_t sh, Args... args) {
    void *tok;
    fyg_prof::begin(s, name, &tok);
    k<<<g, b, sh, s->s>>>(static_cast<KArgs>(args)...);
    fyg_prof::end(s, tok);
}

// the step of the device generation word (flags follow it): gen <- gen + 1;
// on reaching the wrap value the flags are cleared first, so no stale
// generation can match a new one.  Run by one whole workgroup (it
// synchronises), before any thread of that workgroup returns.
__device__ __forceinline__ void fyg_gen_step_block(int *gen, int nflag) {
    __shared__ int g;
    if (threadIdx.x == 0) g = *gen;
    __syncthreads();
    const bool wrap = g >= 0x7ffffff0;
    if (wrap)
        for (int i = threadIdx.x; i < nflag; i += blockDim.x) gen[1 + i] = 0;
    __syncthreads();
    if (threadIdx.x == 0) *gen = wrap ? 1 : g + 1;
}

// per-cube descriptors of one batched launch, passed by value in the kernel
// arguments; block index -> cube
template <class T>
struct Batch {
    T a[FYG_BATCH];
};

// segment table of a launch over cubes of mixed shapes, for up to C cubes:
// segment i covers blocks [start[i], start[i+1]) and runs body kind[i] on
// descriptor idx[i]; nbx, nby, nbt carry the cube's grid shape where the body
// needs it
template <int C>
struct MixSegs {
    int n;
    int start[C + 1];
    unsigned short kind[C];
    unsigned short idx[C];
    int nbx[C], nby[C], nbt[C];
};

// the segment of block b: the last one that starts at or before it (the
// starts sit in the kernel arguments; the search is uniform across the block)
template <int C>
__device__ __forceinline__ int mix_seg(const MixSegs<C> &t, int b) {
    int lo = 0, hi = t.n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (t.start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return __builtin_amdgcn_readfirstlane(lo);
}

// a block-uniform integer kept in a scalar register: values derived from the
// segment (block offsets, quotients of the grid shape) would otherwise come
// out of the integer-division expansion in vector registers and take the
// bodies' per-row index arithmetic off the scalar unit
__device__ __forceinline__ int mix_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// the dynamic LDS of a mixed launch, viewed as one body's LDS struct
extern __shared__ __align__(16) unsigned char fyg_mix_lds[];
template <class T>
__device__ __forceinline__ T &mix_lds_as() {
    return *reinterpret_cast<T *>(fyg_mix_lds);
}

// the arguments of a mixed launch: the segment table and every cube's
// descriptor, by value, so the bodies see the descriptors' pointers as global
// ones exactly as in the per-shape kernels and nothing is uploaded before the
// launch.  Two capacities: launches of up to FYG_MIX_SMALL cubes carry a
// quarter of the bytes (the runtime copies every launch's arguments, ~30 KB
// at FYG_MIX_MAX cubes).
#define FYG_MIX_SMALL 32
template <class T, int C>
struct MixBatch {
    MixSegs<C> t;
    T a[C];
};
// the largest kernel arguments a launch on this runtime was measured to take
constexpr size_t kMixArgMax = (size_t)48 << 10;

// host side: appends segments in launch order (up to FYG_MIX_MAX); the
// largest dynamic LDS any body of the launch needs
struct MixStage {
    MixSegs<FYG_MIX_MAX> t;
    size_t lds = 0;
    long total = 0;
    MixStage() { memset(&t, 0, sizeof t); }
    // blocks: the cube's workgroups; align: its first block's index is a
    // multiple of align (the blocks of the gap exit at once)
    void add(int idx, int kind, long blocks, size_t lds_b, int nbx = 0, int nby = 0, int align = 1) {
        if (blocks <= 0) return;
        total = (total + align - 1) / align * align;
        t.start[t.n] = (int)total;
        t.kind[t.n] = (unsigned short)kind;
        t.idx[t.n] = (unsigned short)idx;
        t.nbx[t.n] = nbx;
        t.nby[t.n] = nby;
        t.nbt[t.n] = (int)blocks;
        t.n++;
        total += blocks;
        t.start[t.n] = (int)total;
        if (lds_b > lds) lds = lds_b;
    }
    // the kernel arguments of capacity C for descriptors d[0..nd)
    template <int C, class T>
    void fill(MixBatch<T, C> &mb, const T *d, int nd) const {
        memset(&mb.t, 0, sizeof mb.t);
        mb.t.n = t.n;
        for (int i = 0; i <= t.n; i++) mb.t.start[i] = t.start[i];
        for (int i = 0; i < t.n; i++) {
            mb.t.kind[i] = t.kind[i];
            mb.t.idx[i] = t.idx[i];
            mb.t.nbx[i] = t.nbx[i];
            mb.t.nby[i] = t.nby[i];
            mb.t.nbt[i] = t.nbt[i];
        }
        for (int i = 0; i < nd; i++) mb.a[i] = d[i];
    }
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// A cube's "not all air" flag is set to the call's generation by every
// workgroup that saw a nonzero voxel.  Tens of thousands of workgroups storing
// to one word serialise on the L2 channel that holds it, so a workgroup stores
// only when the word does not hold the generation yet (a stale read at worst
// repeats the store).
__device__ __forceinline__ void fyg_flag_set(int *flag, const int *gen) {
    const int g = *gen;
    if (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) != g) *flag = g;
}
