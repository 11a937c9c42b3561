// fyg_internal.hpp -- shared by the HIP translation units: the stream object,
// the launch helper that brackets each launch with timing events when
// profiling is on, and small device utilities.
#pragma once
#include <hip/hip_runtime.h>

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
