// fyg_runtime.hip -- device runtime behind the fysics drop-in: per-thread
// streams, grow-only device workspaces and pinned staging, a process-wide cache
// of small read-only tables (twiddles, transfer LUTs) and event-based kernel
// timing.  The reference keeps thread-local host scratch that is never freed
// (fft.c:45,109-112; guided.c:49,83; pipeline.c:481-488); the device analogue
// is one stream + workspace set per calling host thread, so concurrent callers
// (OpenMP tile loops, export worker pools) never serialise on a global sync.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <list>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "fyg_internal.hpp"

static thread_local fyg_stream *tls_stream = nullptr;
static thread_local char tls_err[256] = {0};
static std::atomic<int> g_err{0};

static int note(hipError_t e, const char *what) {
    if (e == hipSuccess) return 0;
    snprintf(tls_err, sizeof tls_err, "%s: %s", what, hipGetErrorString(e));
    g_err.store(1);
    return 1;
}

extern "C" const char *fyg_error_string(void) { return tls_err; }

extern "C" void fyg_set_error(const char *msg) {
    snprintf(tls_err, sizeof tls_err, "%s", msg ? msg : "");
    g_err.store(1);
}

extern "C" fyg_stream *fyg_thread_stream(void) {
    if (tls_stream) return tls_stream;
    fyg_stream *st = new fyg_stream();
    if (note(hipStreamCreateWithFlags(&st->s, hipStreamNonBlocking), "hipStreamCreate")) {
        delete st;
        return nullptr;
    }
    st->owned = true;
    tls_stream = st;
    return st;
}

// external streams: one wrapper (with its own workspaces) per hipStream_t
static std::mutex g_wrap_mu;
static std::unordered_map<void *, fyg_stream *> g_wrapped;

extern "C" fyg_stream *fyg_stream_wrap(void *hip_stream) {
    if (!hip_stream) return fyg_thread_stream();
    std::lock_guard<std::mutex> lk(g_wrap_mu);
    auto it = g_wrapped.find(hip_stream);
    if (it != g_wrapped.end()) return it->second;
    fyg_stream *st = new fyg_stream();
    st->s = (hipStream_t)hip_stream;
    g_wrapped[hip_stream] = st;
    return st;
}

extern "C" void *fyg_stream_handle(fyg_stream *s) { return s ? (void *)s->s : nullptr; }

// forget the stream object of a caller's stream (NULL: the calling thread's
// own stream, which is destroyed): waits for its work, then frees its
// workspaces, table slots and helper streams; the next call starts afresh
extern "C" int fyg_stream_release(void *hip_stream) {
    if (!hip_stream) {
        fyg_stream *st = tls_stream;
        tls_stream = nullptr;
        if (st) fyg_stream_free(st);
        return 0;
    }
    fyg_stream *st = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_wrap_mu);
        auto it = g_wrapped.find(hip_stream);
        if (it == g_wrapped.end()) return 0;
        st = it->second;
        g_wrapped.erase(it);
    }
    fyg_stream_private_free(st);
    return 0;
}

// a stream object of its own (workspaces, flags) on an existing hipStream_t, for
// work recorded into a graph: nothing else can grow or free the buffers the
// graph's kernels address.  Tables come from the process-wide cache.
extern "C" fyg_stream *fyg_stream_private(void *hip_stream) {
    fyg_stream *base = hip_stream ? nullptr : fyg_thread_stream();
    if (!hip_stream && !base) return nullptr;
    fyg_stream *st = new fyg_stream();
    st->s = hip_stream ? (hipStream_t)hip_stream : base->s;
    st->persist = 1;
    return st;
}

static void fyg_arena_free(fyg_stream *st);

extern "C" void fyg_stream_private_free(fyg_stream *st) {
    if (!st) return;
    (void)hipStreamSynchronize(st->s);
    for (int i = 0; i < FYG_WS_SLOTS; i++)
        if (st->ws[i]) (void)hipFree(st->ws[i]);
    for (int i = 0; i < 2; i++)
        if (st->pin[i]) (void)hipHostFree(st->pin[i]);
    if (st->flag) (void)hipFree(st->flag);
    fyg_arena_free(st);
    for (int i = 0; i < FYG_SUBS; i++) {
        if (st->sub[i]) fyg_stream_free(st->sub[i]);
        if (st->join_ev[i]) (void)hipEventDestroy(st->join_ev[i]);
    }
    if (st->fork_ev) (void)hipEventDestroy(st->fork_ev);
    delete st;
}

extern "C" int fyg_stream_persistent(fyg_stream *s) { return s && s->persist; }

// A fork's helper overlaps its parent stream only when the two feed different
// hardware queues.  The runtime spreads streams created without a CU mask over
// a few shared queues by use count, so a new helper can land on its parent's
// queue, and the two then run one after the other (the 96-cube mixed batch on
// two lanes: 66 instead of 79 Gvox/s, depending on which streams the process
// had made before).  A stream created with a CU mask gets a queue of its own;
// the mask here enables every CU.  Only eager helpers take one (a graph
// capture records the dependencies, not the queues), and at most kLaneQueues
// per process; beyond that a helper is an ordinary stream.
static std::atomic<int> g_lane_queues{0};
constexpr int kLaneQueues = 12;

static fyg_stream *fyg_stream_new_lane(void) {
    if (g_lane_queues.fetch_add(1) < kLaneQueues) {
        int dev = 0, cu = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cu > 0) {
            std::vector<uint32_t> mask((size_t)(cu + 31) / 32, ~0u);
            if (cu % 32) mask.back() = (1u << (cu % 32)) - 1;
            fyg_stream *st = new fyg_stream();
            if (hipExtStreamCreateWithCUMask(&st->s, (uint32_t)mask.size(), mask.data()) == hipSuccess) {
                st->owned = true;
                st->own_queue = 1;
                return st;
            }
            delete st;
            (void)hipGetLastError();
        }
    }
    g_lane_queues.fetch_sub(1);
    return fyg_stream_new();
}

extern "C" fyg_stream *fyg_stream_sub(fyg_stream *s, int i) {
    if (!s || i < 0 || i >= FYG_SUBS) return nullptr;
    if (s->sub[i]) return s->sub[i];
    fyg_stream *st = s->persist ? fyg_stream_new() : fyg_stream_new_lane();
    if (!st) return nullptr;
    st->persist = s->persist;
    if (note(hipEventCreateWithFlags(&s->join_ev[i], hipEventDisableTiming), "hipEventCreate(join)")) {
        fyg_stream_free(st);
        return nullptr;
    }
    s->sub[i] = st;
    return st;
}

extern "C" int fyg_fork(fyg_stream *s, int n) {
    if (!s || n < 0 || n > FYG_SUBS) return 1;
    if (!s->fork_ev && note(hipEventCreateWithFlags(&s->fork_ev, hipEventDisableTiming), "hipEventCreate(fork)"))
        return 1;
    if (note(hipEventRecord(s->fork_ev, s->s), "hipEventRecord(fork)")) return 1;
    for (int i = 0; i < n; i++) {
        fyg_stream *h = fyg_stream_sub(s, i);
        if (!h) return 1;
        h->capturing = s->capturing;
        if (note(hipStreamWaitEvent(h->s, s->fork_ev, 0), "hipStreamWaitEvent(fork)")) return 1;
    }
    return 0;
}

extern "C" int fyg_join(fyg_stream *s, int n) {
    if (!s || n < 0 || n > FYG_SUBS) return 1;
    for (int i = 0; i < n; i++) {
        fyg_stream *h = s->sub[i];
        if (!h) return 1;
        if (note(hipEventRecord(s->join_ev[i], h->s), "hipEventRecord(join)") ||
            note(hipStreamWaitEvent(s->s, s->join_ev[i], 0), "hipStreamWaitEvent(join)"))
            return 1;
        h->capturing = 0;
    }
    return 0;
}

// waits for the stream, then frees its workspaces (they regrow on demand)
extern "C" int fyg_stream_trim(fyg_stream *s) {
    if (!s) return 1;
    if (note(hipStreamSynchronize(s->s), "hipStreamSynchronize")) return 1;
    for (int i = 0; i < FYG_WS_SLOTS; i++) {
        if (s->ws[i]) (void)hipFree(s->ws[i]);
        s->ws[i] = nullptr;
        s->wcap[i] = 0;
    }
    return 0;
}

// an owned stream object on a new non-blocking hipStream_t (default priority)
extern "C" fyg_stream *fyg_stream_new(void) {
    fyg_stream *st = new fyg_stream();
    if (note(hipStreamCreateWithFlags(&st->s, hipStreamNonBlocking), "hipStreamCreate")) {
        delete st;
        return nullptr;
    }
    st->owned = true;
    return st;
}

extern "C" void fyg_stream_free(fyg_stream *st) {
    if (!st) return;
    const bool own = st->owned, lane = st->own_queue;
    hipStream_t h = st->s;
    fyg_stream_private_free(st);
    if (own) (void)hipStreamDestroy(h);
    if (lane) g_lane_queues.fetch_sub(1);
}

extern "C" int fyg_stream_wait(fyg_stream *s, void *ev) {
    if (!s || !ev) return 1;
    return note(hipStreamWaitEvent(s->s, (hipEvent_t)ev, 0), "hipStreamWaitEvent");
}

extern "C" int fyg_host_alloc(void **p, size_t bytes) {
    *p = nullptr;
    return note(hipHostMalloc(p, bytes ? bytes : 256, hipHostMallocDefault), "hipHostMalloc");
}

extern "C" void fyg_host_free(void *p) {
    if (p) (void)hipHostFree(p);
}

extern "C" int fyg_capture_begin(fyg_stream *s) {
    if (note(hipStreamBeginCapture(s->s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture")) return 1;
    s->capturing = 1;
    return 0;
}

extern "C" int fyg_capture_end(fyg_stream *s, void **exec) {
    s->capturing = 0;
    hipGraph_t g = nullptr;
    if (note(hipStreamEndCapture(s->s, &g), "hipStreamEndCapture")) return 1;
    hipGraphExec_t e = nullptr;
    int rc = note(hipGraphInstantiate(&e, g, nullptr, nullptr, 0), "hipGraphInstantiate");
    (void)hipGraphDestroy(g);
    if (rc) return 1;
    *exec = (void *)e;
    return 0;
}

extern "C" int fyg_graph_launch(void *exec, fyg_stream *s) {
    return note(hipGraphLaunch((hipGraphExec_t)exec, s->s), "hipGraphLaunch");
}

extern "C" void fyg_graph_free(void *exec) {
    if (exec) (void)hipGraphExecDestroy((hipGraphExec_t)exec);
}

extern "C" int fyg_sync(fyg_stream *s) {
    if (!s) return 1;
    return note(hipStreamSynchronize(s->s), "hipStreamSynchronize");
}

extern "C" void *fyg_ws(fyg_stream *s, int slot, size_t bytes) {
    if (!s || slot < 0 || slot >= FYG_WS_SLOTS) return nullptr;
    if (bytes == 0) bytes = 256;
    if (s->wcap[slot] >= bytes) return s->ws[slot];
    // the previous buffer may still be read by queued work on this stream
    if (s->ws[slot]) {
        if (note(hipStreamSynchronize(s->s), "hipStreamSynchronize")) return nullptr;
        (void)hipFree(s->ws[slot]);
        s->ws[slot] = nullptr;
        s->wcap[slot] = 0;
    }
    size_t want = bytes + bytes / 8;      // slack so slowly growing tiles do not realloc each call
    void *p = nullptr;
    if (note(hipMalloc(&p, want), "hipMalloc(workspace)")) return nullptr;
    s->ws[slot] = p;
    s->wcap[slot] = want;
    return p;
}

extern "C" void *fyg_pinned(fyg_stream *s, int slot, size_t bytes) {
    if (!s || slot < 0 || slot > 1) return nullptr;
    if (s->pcap[slot] >= bytes) return s->pin[slot];
    if (s->pin[slot]) {
        if (note(hipStreamSynchronize(s->s), "hipStreamSynchronize")) return nullptr;
        (void)hipHostFree(s->pin[slot]);
        s->pin[slot] = nullptr;
        s->pcap[slot] = 0;
    }
    void *p = nullptr;
    if (note(hipHostMalloc(&p, bytes, hipHostMallocDefault), "hipHostMalloc")) return nullptr;
    s->pin[slot] = p;
    s->pcap[slot] = bytes;
    return p;
}

extern "C" int fyg_malloc(void **p, size_t bytes) {
    return note(hipMalloc(p, bytes ? bytes : 256), "hipMalloc");
}
extern "C" void fyg_free(void *p) {
    if (p) (void)hipFree(p);
}
extern "C" int fyg_h2d(void *dst, const void *src, size_t bytes, fyg_stream *s) {
    if (!bytes) return 0;
    return note(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s->s), "hipMemcpyH2D");
}
extern "C" int fyg_d2h(void *dst, const void *src, size_t bytes, fyg_stream *s) {
    if (!bytes) return 0;
    return note(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s->s), "hipMemcpyD2H");
}
extern "C" int fyg_d2d(void *dst, const void *src, size_t bytes, fyg_stream *s) {
    if (!bytes || dst == src) return 0;
    return note(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s->s), "hipMemcpyD2D");
}

__global__ void k_fill_zero(uint32_t *p, size_t nwords, unsigned char *tail, int ntail) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < nwords; i += stride) p[i] = 0u;
    if (blockIdx.x == 0 && (int)threadIdx.x < ntail) tail[threadIdx.x] = 0;
}

extern "C" int fyg_zero(void *dst, size_t bytes, fyg_stream *s) {
    if (!bytes) return 0;
    size_t nw = bytes / 4;
    int ntail = (int)(bytes - nw * 4);
    unsigned blocks = (unsigned)std::min<size_t>((nw + 255) / 256 + 1, 4096);
    fyg_launch(s, "fill_zero", k_fill_zero, dim3(blocks), dim3(256), 0,
               (uint32_t *)dst, nw, (unsigned char *)dst + nw * 4, ntail);
    return fyg_last_error();
}

__global__ void k_gen_step(int *gen, int nflag) { fyg_gen_step_block(gen, nflag); }

extern "C" int fyg_flags_reserve(fyg_stream *s, int n, int **d_flags, int **d_gen, int *nflag) {
    if (!s || n < 1) return 0;
    if (n > s->nflag) {
        // earlier generations may still be tested by queued work on this stream
        if (s->flag) {
            if (note(hipStreamSynchronize(s->s), "hipStreamSynchronize")) return 0;
            (void)hipFree(s->flag);
            s->flag = nullptr;
            s->nflag = 0;
        }
        int cap = n < 16 ? 16 : n;
        if (note(hipMalloc((void **)&s->flag, sizeof(int) * (cap + 1)), "hipMalloc(flag)")) return 0;
        if (fyg_zero(s->flag, sizeof(int) * (cap + 1), s)) return 0;   // zero never equals a generation
        s->nflag = cap;
    }
    *d_flags = s->flag + 1;
    *d_gen = s->flag;
    *nflag = s->nflag;
    return 1;
}

extern "C" int fyg_gen_step(int *d_gen, int nflag, fyg_stream *s) {
    fyg_launch(s, "gen_step", k_gen_step, dim3(1), dim3(256), 0, d_gen, nflag);
    return fyg_last_error();
}

extern "C" int fyg_flags(fyg_stream *s, int n, int **d_flags, const int **d_gen) {
    int *g = nullptr, nf = 0;
    if (!fyg_flags_reserve(s, n, d_flags, &g, &nf)) return 0;
    if (fyg_gen_step(g, nf, s)) return 0;
    *d_gen = g;
    return 1;
}

extern "C" int fyg_flag(fyg_stream *s, int **d_flag, const int **d_gen) { return fyg_flags(s, 1, d_flag, d_gen); }

extern "C" int fyg_last_error(void) {
    return note(hipGetLastError(), "kernel launch");
}

// ------------------------------------------------------------------ table cache
// Entries are held while a caller may still queue work that reads them: a
// hold (fyg_hold_begin .. fyg_hold_release) collects every table its thread
// finds or makes in between, and a captured graph keeps its hold until it is
// destroyed.  Releasing records an event on the stream the work went to; a
// transfer table ('T' keys, the twiddles, are never evicted) that is neither
// held nor the most recently used leaves the map when the cache is full, and
// its memory is freed once every event of its releases has completed.  No call
// waits for the device.
namespace {
struct Ev {
    hipEvent_t e = nullptr;
    ~Ev() {
        if (e) (void)hipEventDestroy(e);
    }
};
struct TableEntry {
    std::string key;
    void *dev = nullptr;
    size_t bytes = 0;
    int holders = 0;
    std::vector<std::shared_ptr<Ev>> pending;   // releases whose work may still read the table
    std::list<std::shared_ptr<TableEntry>>::iterator lru;
    bool in_lru = false;
};
using EntryP = std::shared_ptr<TableEntry>;
std::mutex g_tab_mu;
std::unordered_map<std::string, EntryP> g_tab;
std::list<EntryP> g_tab_lru;            // transfer tables only, least recently used first
std::vector<EntryP> g_tab_dead;         // evicted, memory freed once their events complete
size_t g_tab_max = 512;
long g_tab_evicted = 0, g_tab_freed = 0;

bool events_done(TableEntry &t) {
    for (auto &ev : t.pending) {
        if (!ev || !ev->e) continue;
        if (hipEventQuery(ev->e) != hipSuccess) return false;
    }
    return true;
}

// frees the evicted entries whose last readers have finished (lock held)
void sweep_dead() {
    size_t k = 0;
    for (size_t i = 0; i < g_tab_dead.size(); i++) {
        EntryP &t = g_tab_dead[i];
        if (t->holders == 0 && events_done(*t)) {
            (void)hipFree(t->dev);
            t->dev = nullptr;
            g_tab_freed++;
        } else {
            g_tab_dead[k++] = t;
        }
    }
    g_tab_dead.resize(k);
}

// evicts least recently used, unheld transfer tables while the cache is full (lock held)
void evict_to_fit() {
    auto it = g_tab_lru.begin();
    while (g_tab_lru.size() >= g_tab_max && it != g_tab_lru.end()) {
        EntryP t = *it;
        if (t->holders > 0) {
            ++it;
            continue;
        }
        it = g_tab_lru.erase(it);
        t->in_lru = false;
        g_tab.erase(t->key);
        g_tab_dead.push_back(t);
        g_tab_evicted++;
    }
    sweep_dead();
}
}  // namespace

struct fyg_hold {
    std::vector<EntryP> held;
    fyg_hold *prev = nullptr;
    bool collecting = true;
};
static thread_local fyg_hold *tls_hold = nullptr;

// the entry of key was found or made: move it to the recent end, hold it (lock held)
static void touch(const EntryP &t) {
    if (t->in_lru) g_tab_lru.splice(g_tab_lru.end(), g_tab_lru, t->lru);
    if (t->key[0] == 'T') return;
    for (fyg_hold *h = tls_hold; h; h = h->prev) {      // nested holds each hold it
        if (!h->collecting) continue;
        bool have = false;
        for (auto &x : h->held)
            if (x == t) {
                have = true;
                break;
            }
        if (!have) {
            t->holders++;
            h->held.push_back(t);
        }
    }
}

extern "C" fyg_hold *fyg_hold_begin(void) {
    fyg_hold *h = new fyg_hold();
    h->prev = tls_hold;
    tls_hold = h;
    return h;
}

extern "C" void fyg_hold_end(fyg_hold *h) {
    if (!h || !h->collecting) return;
    h->collecting = false;
    if (tls_hold == h) tls_hold = h->prev;
}

extern "C" int fyg_hold_release(fyg_hold *h, fyg_stream *s) {
    if (!h) return 0;
    fyg_hold_end(h);
    int rc = 0;
    std::shared_ptr<Ev> ev;
    if (s && !s->capturing && !h->held.empty()) {
        ev = std::make_shared<Ev>();
        rc = note(hipEventCreateWithFlags(&ev->e, hipEventDisableTiming), "hipEventCreate(hold)");
        if (!rc) rc = note(hipEventRecord(ev->e, s->s), "hipEventRecord(hold)");
        if (rc) {
            // the work's completion cannot be tracked: keep the tables held for good
            delete h;
            return 1;
        }
    }
    {
        std::lock_guard<std::mutex> lk(g_tab_mu);
        for (auto &t : h->held) {
            t->holders--;
            if (ev) {
                // drop events of earlier releases that have completed
                size_t k = 0;
                for (size_t i = 0; i < t->pending.size(); i++)
                    if (hipEventQuery(t->pending[i]->e) != hipSuccess) t->pending[k++] = t->pending[i];
                t->pending.resize(k);
                t->pending.push_back(ev);
            }
        }
        sweep_dead();
    }
    delete h;
    return rc;
}

extern "C" const void *fyg_table_find(const void *key, size_t keylen) {
    std::string k((const char *)key, keylen);
    std::lock_guard<std::mutex> lk(g_tab_mu);
    auto it = g_tab.find(k);
    if (it == g_tab.end()) return nullptr;
    touch(it->second);
    return it->second->dev;
}

extern "C" const void *fyg_table(const void *key, size_t keylen, const void *data, size_t bytes) {
    std::string k((const char *)key, keylen);
    std::lock_guard<std::mutex> lk(g_tab_mu);
    auto it = g_tab.find(k);
    if (it != g_tab.end()) {
        touch(it->second);
        return it->second->dev;
    }
    if (!data) return nullptr;
    if (k[0] != 'T') evict_to_fit();
    void *dev = nullptr;
    if (note(hipMalloc(&dev, bytes), "hipMalloc(table)")) return nullptr;
    // synchronous upload: the table must be valid before any stream uses it
    if (note(hipMemcpy(dev, data, bytes, hipMemcpyHostToDevice), "hipMemcpy(table)")) {
        (void)hipFree(dev);
        return nullptr;
    }
    EntryP t = std::make_shared<TableEntry>();
    t->key = k;
    t->dev = dev;
    t->bytes = bytes;
    if (k[0] != 'T') {
        g_tab_lru.push_back(t);
        t->lru = std::prev(g_tab_lru.end());
        t->in_lru = true;
    }
    g_tab[k] = t;
    touch(t);
    return dev;
}

extern "C" void fyg_table_stats(long *st) {
    std::lock_guard<std::mutex> lk(g_tab_mu);
    sweep_dead();
    long held = 0;
    for (auto &kv : g_tab)
        if (kv.second->holders > 0) held++;
    st[0] = (long)g_tab.size();
    st[1] = (long)g_tab_lru.size();
    st[2] = held;
    st[3] = (long)g_tab_dead.size();
    st[4] = g_tab_evicted;
    st[5] = g_tab_freed;
    st[6] = (long)g_tab_max;
}

extern "C" size_t fyg_table_limit(size_t n) {
    std::lock_guard<std::mutex> lk(g_tab_mu);
    size_t old = g_tab_max;
    if (n >= 1) g_tab_max = n;
    return old;
}

// ------------------------------------------------------------------ per-stream table slots
struct Arena {
    char *dev = nullptr;
    char *pin = nullptr;
    std::vector<std::string> key;            // key held by each slot ("" = free)
    std::vector<hipEvent_t> ev;              // recorded after the slot's last upload
    std::vector<long> stamp;                 // last use; the least recent slot is recycled
    std::unordered_map<std::string, int> where;
    long clock = 0;
};

// the stream's table slots (the stream has been synchronised)
static void fyg_arena_free(fyg_stream *st) {
    Arena *a = st->arena;
    if (!a) return;
    for (auto e : a->ev)
        if (e) (void)hipEventDestroy(e);
    if (a->dev) (void)hipFree(a->dev);
    if (a->pin) (void)hipHostFree(a->pin);
    delete a;
    st->arena = nullptr;
}

extern "C" const void *fyg_stream_table(fyg_stream *s, const void *key, size_t keylen, const void *data,
                                        size_t bytes) {
    if (!s || bytes > FYG_ARENA_SLOT) return nullptr;
    if (!s->arena) {
        if (!data) return nullptr;
        Arena *a = new Arena();
        if (note(hipMalloc((void **)&a->dev, FYG_ARENA_SLOT * FYG_ARENA_SLOTS), "hipMalloc(arena)") ||
            note(hipHostMalloc((void **)&a->pin, FYG_ARENA_SLOT * FYG_ARENA_SLOTS, hipHostMallocDefault),
                 "hipHostMalloc(arena)")) {
            if (a->dev) (void)hipFree(a->dev);
            delete a;
            return nullptr;
        }
        a->key.assign(FYG_ARENA_SLOTS, std::string());
        a->ev.assign(FYG_ARENA_SLOTS, nullptr);
        a->stamp.assign(FYG_ARENA_SLOTS, -1);
        s->arena = a;
    }
    Arena *a = s->arena;
    std::string k((const char *)key, keylen);
    auto it = a->where.find(k);
    if (it != a->where.end()) {
        a->stamp[it->second] = ++a->clock;
        return a->dev + FYG_ARENA_SLOT * it->second;
    }
    if (!data) return nullptr;
    int i = 0;
    for (int j = 1; j < FYG_ARENA_SLOTS; j++)
        if (a->stamp[j] < a->stamp[i]) i = j;
    if (!a->key[i].empty()) {
        // the slot's staging memory may still feed its previous upload; its
        // device memory is rewritten in stream order, after earlier readers
        if (note(hipEventSynchronize(a->ev[i]), "hipEventSynchronize(arena)")) return nullptr;
        a->where.erase(a->key[i]);
        a->key[i].clear();
    }
    char *pin = a->pin + FYG_ARENA_SLOT * i, *dev = a->dev + FYG_ARENA_SLOT * i;
    memcpy(pin, data, bytes);
    if (note(hipMemcpyAsync(dev, pin, bytes, hipMemcpyHostToDevice, s->s), "hipMemcpyAsync(arena)")) return nullptr;
    if (!a->ev[i] && note(hipEventCreateWithFlags(&a->ev[i], hipEventDisableTiming), "hipEventCreate")) return nullptr;
    if (note(hipEventRecord(a->ev[i], s->s), "hipEventRecord(arena)")) return nullptr;
    a->key[i] = k;
    a->where[k] = i;
    a->stamp[i] = ++a->clock;
    return dev;
}

extern "C" int fyg_event_record(fyg_stream *s, void **ev) {
    if (!s || !ev) return 1;
    if (!*ev && note(hipEventCreateWithFlags((hipEvent_t *)ev, hipEventDisableTiming), "hipEventCreate")) return 1;
    return note(hipEventRecord((hipEvent_t)*ev, s->s), "hipEventRecord");
}

extern "C" int fyg_event_done(void *ev) {
    hipError_t e = hipEventQuery((hipEvent_t)ev);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    note(e, "hipEventQuery");
    return -1;
}

extern "C" int fyg_event_wait(void *ev) { return note(hipEventSynchronize((hipEvent_t)ev), "hipEventSynchronize"); }

extern "C" void fyg_event_destroy(void *ev) {
    if (ev) (void)hipEventDestroy((hipEvent_t)ev);
}

extern "C" const void *fyg_twiddles(int L) {
    char key[16];
    int kl = snprintf(key, sizeof key, "Ttw%d", L);
    const void *d = fyg_table_find(key, (size_t)kl);
    if (d) return d;
    std::vector<float> tw(2 * (size_t)L);
    for (int m = 0; m < L; m++) {
        double a = 2.0 * M_PI * (double)m / (double)L;
        tw[2 * m] = (float)cos(a);
        tw[2 * m + 1] = (float)sin(a);
    }
    return fyg_table(key, (size_t)kl, tw.data(), sizeof(float) * tw.size());
}

extern "C" const void *fyg_twiddles64(int L) {
    char key[16];
    int kl = snprintf(key, sizeof key, "Tw64%d", L);
    const void *d = fyg_table_find(key, (size_t)kl);
    if (d) return d;
    std::vector<double> tw(2 * (size_t)L);
    for (int m = 0; m < L; m++) {
        double a = 2.0 * M_PI * (double)m / (double)L;
        tw[2 * m] = cos(a);
        tw[2 * m + 1] = sin(a);
    }
    return fyg_table(key, (size_t)kl, tw.data(), sizeof(double) * tw.size());
}

// ------------------------------------------------------------------ profiling
namespace fyg_prof {
std::atomic<int> enabled{0};
std::mutex mu;
struct Rec {
    const char *name;
    hipEvent_t a, b;
};
std::vector<Rec> pending;
struct Acc {
    double ms = 0;
    long n = 0;
};
std::vector<std::pair<std::string, Acc>> acc;

void begin(fyg_stream *s, const char *name, void **tok) {
    *tok = nullptr;
    if (!enabled.load(std::memory_order_relaxed) || s->capturing) return;
    Rec *r = new Rec{name, nullptr, nullptr};
    (void)hipEventCreate(&r->a);
    (void)hipEventCreate(&r->b);
    (void)hipEventRecord(r->a, s->s);
    *tok = r;
}
void end(fyg_stream *s, void *tok) {
    if (!tok) return;
    Rec *r = (Rec *)tok;
    (void)hipEventRecord(r->b, s->s);
    std::lock_guard<std::mutex> lk(mu);
    pending.push_back(*r);
    delete r;
}
}  // namespace fyg_prof

extern "C" void fyg_prof_enable(int on) { fyg_prof::enabled.store(on ? 1 : 0); }

static void prof_drain() {
    std::lock_guard<std::mutex> lk(fyg_prof::mu);
    for (auto &r : fyg_prof::pending) {
        (void)hipEventSynchronize(r.b);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, r.a, r.b);
        bool found = false;
        for (auto &kv : fyg_prof::acc)
            if (kv.first == r.name) {
                kv.second.ms += ms;
                kv.second.n += 1;
                found = true;
                break;
            }
        if (!found) fyg_prof::acc.push_back({std::string(r.name), fyg_prof::Acc{ms, 1}});
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    fyg_prof::pending.clear();
}

extern "C" int fyg_prof_collect(char *names, int names_len, double *ms, long *count, int max) {
    prof_drain();
    std::lock_guard<std::mutex> lk(fyg_prof::mu);
    int n = 0;
    size_t off = 0;
    for (auto &kv : fyg_prof::acc) {
        if (n >= max) break;
        size_t l = kv.first.size() + 1;
        if (off + l > (size_t)names_len) break;
        memcpy(names + off, kv.first.c_str(), l);
        off += l;
        ms[n] = kv.second.ms;
        count[n] = kv.second.n;
        n++;
    }
    return n;
}

extern "C" void fyg_prof_reset(void) {
    prof_drain();
    std::lock_guard<std::mutex> lk(fyg_prof::mu);
    fyg_prof::acc.clear();
}
