// serve.cpp -- device block cache and continuous-batching contraction server.
//
// Block cache (the SIP's BlockManager / CachedBlockMap / LRUArrayPolicy
// role, block_manager.cpp:106-206, cached_block_map.cpp:31-147,
// lru_array_policy.h:37-96): blocks are named by 64-bit keys, live in one
// HBM arena of a fixed budget, and are evicted least-recently-used first
// when a request does not fit.  A block whose device copy is newer than its
// host copy is written back first (spilled) to its host home, or to pinned
// host memory the cache allocates for it.  Eviction is per block (the
// reference evicts whole arrays), and never touches a block in use.
//
// Server: requests (one block contraction or permute each, blocks named by
// cache keys) are queued by any thread; a scheduler thread takes everything
// that has arrived (up to a batch limit and half the cache budget of pinned
// blocks, stopping at the first request that would read a block an earlier
// one in the batch writes, or vice versa; permutes and contractions form
// separate batches, and a permute's destination is touched by no other
// permute of its batch),
// makes its blocks resident, and launches the batch as one grouped launch,
// its work list uploaded through a small ring of descriptor slots (no
// allocation per batch).  Two device streams ("lanes"): large batches run
// in order on lane 0, and a small batch admitted while lane 0 is busy runs on
// lane 1 beside it; a batch that touches a block the other lane's
// unfinished batches write (or writes one they touch) first waits for them
// on the device.  One
// completion thread per lane stamps each batch's requests when its event
// completes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/sipgpu.h"
#include "internal.h"
#include "kernels.h"
#include "plan.h"

using sipgpu::rt::fail;
using sipgpu::rt::hip_fail;
using sipgpu::rt::stream_of;

namespace {

using Clock = std::chrono::steady_clock;

// First-fit sub-allocator over one region (device arena or pinned host
// slab), 256-byte granules.
struct Arena {
    char *base = nullptr;
    size_t size = 0;
    std::map<size_t, size_t> free_ranges; // offset -> length
    size_t used = 0;

    void reset(char *b, size_t n) {
        base = b;
        size = n;
        free_ranges.clear();
        if (n) free_ranges.emplace(0, n);
        used = 0;
    }
    static size_t round(size_t b) { return (std::max<size_t>(b, 1) + 255) & ~(size_t)255; }
    bool alloc(size_t bytes, size_t &off) {
        const size_t need = round(bytes);
        for (auto it = free_ranges.begin(); it != free_ranges.end(); ++it) {
            if (it->second < need) continue;
            off = it->first;
            const size_t rest = it->second - need;
            free_ranges.erase(it);
            if (rest) free_ranges.emplace(off + need, rest);
            used += need;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        const size_t len = round(bytes);
        used -= len;
        auto next = free_ranges.lower_bound(off);
        size_t lo = off, hi = off + len;
        if (next != free_ranges.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == lo) {
                lo = prev->first;
                free_ranges.erase(prev);
            }
        }
        next = free_ranges.lower_bound(hi);
        if (next != free_ranges.end() && next->first == hi) {
            hi += next->second;
            free_ranges.erase(next);
        }
        free_ranges.emplace(lo, hi - lo);
    }
};

constexpr size_t kNotResident = ~(size_t)0;

struct CBlock {
    size_t elems = 0;
    size_t off = kNotResident;     // arena offset while resident
    double *home = nullptr;        // host copy: the caller's, or a pinned spill buffer
    bool home_owned = false;       // home is a range of pinned slab home_slab
    int home_slab = -1;
    bool home_valid = false;       // home holds the block's current value
    int init = SIPGPU_CACHE_UNDEFINED;
    unsigned long long seed = 0, stream = 0;
    double a = 0.0, b = 0.0;
    bool initialised = false;      // contents defined (by upload, fill or a write)
    int pins = 0;
    // Per stream that used the device copy since it became resident, an
    // event recorded at its last release there: the copy is unused once all
    // of them complete (work on several streams may read one block).
    std::vector<std::pair<hipStream_t, hipEvent_t>> uses;
    std::list<unsigned long long>::iterator lru_it;
    bool in_lru = false;
    hipEvent_t spilled = nullptr;  // recorded after the last spill into `home`
};

// An arena range [off, off + len) that becomes reusable once `ev` completes
// (the spill that reads it, or the last kernel that used it).
struct Fence {
    hipEvent_t ev;
    size_t off, len;
};

} // namespace

struct sipgpu_cache_s {
    std::mutex mu;
    Arena arena;
    void *dev = nullptr;
    hipStream_t copy = nullptr;                 // spill (device -> host) copies
    hipStream_t h2d = nullptr;                  // refill (host -> device) copies
    // Spill buffers: ranges of pinned host slabs (one pinned allocation per
    // slab, not per block: pinning is slow, and a stream of first spills
    // would otherwise pay it per block).
    std::vector<Arena> slabs;
    std::unordered_map<unsigned long long, CBlock> blocks;
    std::list<unsigned long long> lru;          // front: least recently released
    std::vector<Fence> fences;                  // freed ranges whose readers may be in flight
    sipgpu_cache_stats st{};
    std::vector<std::vector<unsigned long long>> scopes; // temp blocks of each open scope
    std::unordered_set<unsigned long long> persistent;   // kept by the end of a program
};

namespace {

int make_event(hipEvent_t *e) {
    SIPGPU_HIP_TRY(hipEventCreateWithFlags(e, hipEventDisableTiming));
    return 0;
}

// `s` waits until no launched work uses the block's device copy.
int wait_uses(hipStream_t s, const CBlock &b) {
    for (const auto &u : b.uses) SIPGPU_HIP_TRY(hipStreamWaitEvent(s, u.second, 0));
    return 0;
}

void drop_uses(CBlock &b) {
    for (auto &u : b.uses) (void)hipEventDestroy(u.second);
    b.uses.clear();
}

void drop_spilled(CBlock &b) {
    if (b.spilled) (void)hipEventDestroy(b.spilled);
    b.spilled = nullptr;
}

constexpr size_t kHostSlab = (size_t)256 << 20; // pinned bytes per slab (at least)

// A pinned spill buffer for block b (bytes long) from the first slab with
// room, or from a new slab.
int home_alloc(sipgpu_cache_s &c, CBlock &b, size_t bytes) {
    size_t off = 0;
    for (size_t i = 0; i < c.slabs.size(); ++i)
        if (c.slabs[i].alloc(bytes, off)) {
            b.home = reinterpret_cast<double *>(c.slabs[i].base + off);
            b.home_slab = (int)i;
            b.home_owned = true;
            return 0;
        }
    const size_t n = std::max(Arena::round(bytes), kHostSlab);
    void *p = nullptr;
    if (hipHostMalloc(&p, n, hipHostMallocDefault) != hipSuccess || !p) {
        (void)hipGetLastError();
        return fail(SIPGPU_ENOMEM, "cache: pinned spill buffer");
    }
    Arena a;
    a.reset(static_cast<char *>(p), n);
    a.alloc(bytes, off);
    c.slabs.push_back(std::move(a));
    b.home = reinterpret_cast<double *>(c.slabs.back().base + off);
    b.home_slab = (int)c.slabs.size() - 1;
    b.home_owned = true;
    return 0;
}

// Returns block b's spill buffer to its slab (the caller has made sure no
// copy still reads or writes it).
void home_free(sipgpu_cache_s &c, CBlock &b) {
    if (!b.home_owned) return;
    Arena &a = c.slabs[b.home_slab];
    a.release((size_t)(reinterpret_cast<char *>(b.home) - a.base), b.elems * sizeof(double));
    b.home = nullptr;
    b.home_owned = false;
    b.home_slab = -1;
}

// Cache calls are host-side bookkeeping around uploads, spills, fences and
// synchronisation: none of it may be recorded into a graph (a replay would
// also reuse arena offsets that a later eviction hands to another block).
int refuse_capture(hipStream_t s) {
    if (int rc = sipgpu::rt::refuse_in_capture("a block cache call")) return rc;
    if (s && sipgpu::rt::capturing(s))
        return fail(SIPGPU_ECAPTURE, "a block cache call is not allowed on a capturing stream");
    return 0;
}

// Write back (if needed) and drop the device copy of an unpinned block.
// The arena range is reusable once the returned fence (if any) completes.
int evict_locked(sipgpu_cache_s &c, unsigned long long key, CBlock &b) {
    if (b.in_lru) {
        c.lru.erase(b.lru_it);
        b.in_lru = false;
    }
    const size_t bytes = b.elems * sizeof(double);
    if (b.initialised && !b.home_valid) {
        if (!b.home) {
            if (int rc = home_alloc(c, b, bytes)) return rc;
        }
        if (int rc = wait_uses(c.copy, b)) return rc;
        SIPGPU_HIP_TRY(hipMemcpyAsync(b.home, c.arena.base + b.off, bytes, hipMemcpyDeviceToHost,
                                      c.copy));
        hipEvent_t f;
        if (int rc = make_event(&f)) return rc;
        SIPGPU_HIP_TRY(hipEventRecord(f, c.copy));
        c.fences.push_back(Fence{f, b.off, Arena::round(bytes)});
        // a refill of this block reads `home`: it waits for this copy
        if (!b.spilled) {
            if (int rc = make_event(&b.spilled)) return rc;
        }
        SIPGPU_HIP_TRY(hipEventRecord(b.spilled, c.copy));
        b.home_valid = true;
        c.st.spills += 1;
        c.st.bytes_d2h += bytes;
    } else if (!b.uses.empty()) {
        // a later upload into this range must follow the last kernel reading it
        hipEvent_t f;
        if (int rc = make_event(&f)) return rc;
        if (int rc = wait_uses(c.copy, b)) return rc;
        SIPGPU_HIP_TRY(hipEventRecord(f, c.copy));
        c.fences.push_back(Fence{f, b.off, Arena::round(bytes)});
    }
    c.arena.release(b.off, bytes);
    b.off = kNotResident;
    c.st.evictions += 1;
    (void)key;
    return 0;
}

// Forget a block without writing it back.  Its arena range is fenced like an
// eviction's; an owned spill buffer is freed once the copy stream (which may
// still be filling it) has drained.
int remove_locked(sipgpu_cache_s &c, unsigned long long key) {
    auto it = c.blocks.find(key);
    if (it == c.blocks.end()) return 0;
    CBlock &b = it->second;
    if (b.in_lru) {
        c.lru.erase(b.lru_it);
        b.in_lru = false;
    }
    if (b.off != kNotResident) {
        if (!b.uses.empty()) {
            hipEvent_t f;
            if (int rc = make_event(&f)) return rc;
            if (int rc = wait_uses(c.copy, b)) return rc;
            SIPGPU_HIP_TRY(hipEventRecord(f, c.copy));
            c.fences.push_back(Fence{f, b.off, Arena::round(b.elems * sizeof(double))});
        }
        c.arena.release(b.off, b.elems * sizeof(double));
        b.off = kNotResident;
    }
    if (b.home_owned) {
        // the copy streams may still be filling or reading the spill buffer
        SIPGPU_HIP_TRY(hipStreamSynchronize(c.copy));
        SIPGPU_HIP_TRY(hipStreamSynchronize(c.h2d));
        home_free(c, b);
    }
    drop_uses(b);
    drop_spilled(b);
    c.blocks.erase(it);
    c.persistent.erase(key);
    return 0;
}

// Make `stream` wait for the spills and last uses that freed any part of the
// arena range [off, off + len) (completed fences are dropped on the way).
int fence_locked(sipgpu_cache_s &c, hipStream_t stream, size_t off, size_t len) {
    std::vector<Fence> keep;
    for (const Fence &f : c.fences) {
        if (hipEventQuery(f.ev) == hipSuccess) {
            (void)hipEventDestroy(f.ev);
            continue;
        }
        if (f.off < off + len && off < f.off + f.len)
            SIPGPU_HIP_TRY(hipStreamWaitEvent(stream, f.ev, 0));
        keep.push_back(f);
    }
    c.fences.swap(keep);
    return 0;
}

int reserve_locked(sipgpu_cache_s &c, size_t bytes, size_t &off) {
    while (!c.arena.alloc(bytes, off)) {
        if (c.lru.empty())
            return fail(SIPGPU_EBUDGET, "cache: blocks in use exceed the HBM budget");
        const unsigned long long victim = c.lru.front();
        if (int rc = evict_locked(c, victim, c.blocks[victim])) return rc;
    }
    return 0;
}

int acquire_locked(sipgpu_cache_s &c, hipStream_t stream, unsigned long long key, int mode,
                   double **dev) {
    auto it = c.blocks.find(key);
    if (it == c.blocks.end()) return fail(SIPGPU_ENOTFOUND, "cache: unknown block key");
    CBlock &b = it->second;
    const size_t bytes = b.elems * sizeof(double);
    if (b.off == kNotResident) {
        size_t off;
        if (int rc = reserve_locked(c, bytes, off)) return rc;
        b.off = off;
        double *d = reinterpret_cast<double *>(c.arena.base + off);
        const bool need_value = mode != SIPGPU_CACHE_WRITE;
        if (b.initialised && b.home_valid && need_value) {
            // Refill on the cache's own host-to-device stream, so that it
            // overlaps the kernels already queued on `stream` (and the spills
            // on the device-to-host stream); `stream` waits for the copy.
            if (int rc = fence_locked(c, c.h2d, off, Arena::round(bytes))) return rc;
            if (b.spilled) {
                SIPGPU_HIP_TRY(hipStreamWaitEvent(c.h2d, b.spilled, 0));
                drop_spilled(b);
            }
            SIPGPU_HIP_TRY(hipMemcpyAsync(d, b.home, bytes, hipMemcpyHostToDevice, c.h2d));
            hipEvent_t done;
            if (int rc = make_event(&done)) return rc;
            const hipError_t e1 = hipEventRecord(done, c.h2d);
            const hipError_t e2 = e1 == hipSuccess ? hipStreamWaitEvent(stream, done, 0) : e1;
            (void)hipEventDestroy(done); // released once the recorded copy completes
            SIPGPU_HIP_TRY(e2);
            c.st.bytes_h2d += bytes;
        } else if (int rc = fence_locked(c, stream, off, Arena::round(bytes))) {
            return rc;
        }
        if (!b.initialised) {
            if (b.init == SIPGPU_CACHE_ZERO) {
                if (need_value)
                    SIPGPU_HIP_TRY(sipgpu::launch_block_op(sipgpu::kOpFill, d, nullptr,
                                                           (long long)b.elems, 0.0, stream));
            } else if (b.init == SIPGPU_CACHE_UNIFORM || b.init == SIPGPU_CACHE_NORMAL) {
                SIPGPU_HIP_TRY(sipgpu::launch_fill_random(
                    d, (long long)b.elems, b.seed, b.stream,
                    b.init == SIPGPU_CACHE_UNIFORM ? 0 : 1, b.a, b.b, stream));
            }
            b.initialised = true;
        }
        c.st.misses += 1;
    } else {
        c.st.hits += 1;
    }
    if (b.in_lru) {
        c.lru.erase(b.lru_it);
        b.in_lru = false;
    }
    b.pins += 1;
    if (mode != SIPGPU_CACHE_READ) {
        b.home_valid = false;
        b.initialised = true;
    }
    *dev = reinterpret_cast<double *>(c.arena.base + b.off);
    return 0;
}

int release_locked(sipgpu_cache_s &c, hipStream_t stream, unsigned long long key) {
    auto it = c.blocks.find(key);
    if (it == c.blocks.end()) return fail(SIPGPU_ENOTFOUND, "cache: unknown block key");
    CBlock &b = it->second;
    if (b.pins <= 0) return fail(SIPGPU_EARG, "cache: release without acquire");
    hipEvent_t *ev = nullptr;
    for (auto &u : b.uses)
        if (u.first == stream) ev = &u.second;
    if (!ev) {
        hipEvent_t e;
        if (int rc = make_event(&e)) return rc;
        b.uses.emplace_back(stream, e);
        ev = &b.uses.back().second;
    }
    SIPGPU_HIP_TRY(hipEventRecord(*ev, stream));
    if (--b.pins == 0) {
        c.lru.push_back(key);
        b.lru_it = std::prev(c.lru.end());
        b.in_lru = true;
    }
    return 0;
}

} // namespace

int sipgpu::rt::cache_block_elems(sipgpu_cache_t c, unsigned long long key, size_t *elems) {
    if (!c || !elems) return fail(SIPGPU_EARG, "cache: bad arguments");
    std::lock_guard<std::mutex> g(c->mu);
    auto it = c->blocks.find(key);
    *elems = it == c->blocks.end() ? 0 : it->second.elems;
    return 0;
}

extern "C" {

int sipgpu_cache_create(size_t budget_bytes, sipgpu_cache_t *cache) {
    if (!cache || budget_bytes < 256) return fail(SIPGPU_EARG, "cache: bad arguments");
    if (int rc = sipgpu::rt::ensure_ready()) return rc;
    auto c = std::make_unique<sipgpu_cache_s>();
    const size_t bytes = Arena::round(budget_bytes);
    hipError_t e = hipMalloc(&c->dev, bytes);
    if (e != hipSuccess) return fail(SIPGPU_ENOMEM, "cache: HBM budget not available");
    c->arena.reset(static_cast<char *>(c->dev), bytes);
    e = hipStreamCreateWithFlags(&c->copy, hipStreamNonBlocking);
    if (e != hipSuccess) {
        (void)hipFree(c->dev);
        return hip_fail(e, "cache: copy stream");
    }
    e = hipStreamCreateWithFlags(&c->h2d, hipStreamNonBlocking);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(c->copy);
        (void)hipFree(c->dev);
        return hip_fail(e, "cache: copy stream");
    }
    c->st.budget_bytes = bytes;
    *cache = c.release();
    return 0;
}

int sipgpu_cache_destroy(sipgpu_cache_t c) {
    if (int rc = refuse_capture(nullptr)) return rc;
    if (!c) return fail(SIPGPU_EARG, "null cache");
    SIPGPU_HIP_TRY(hipDeviceSynchronize());
    for (auto &kv : c->blocks) {
        drop_uses(kv.second);
        drop_spilled(kv.second);
    }
    for (Arena &a : c->slabs) (void)hipHostFree(a.base);
    for (const Fence &f : c->fences) (void)hipEventDestroy(f.ev);
    (void)hipStreamDestroy(c->copy);
    (void)hipStreamDestroy(c->h2d);
    (void)hipFree(c->dev);
    delete c;
    return 0;
}

int sipgpu_cache_insert(sipgpu_cache_t c, unsigned long long key, size_t elems,
                        const double *host, int init, unsigned long long seed,
                        unsigned long long stream, double a, double b) {
    if (int rc = refuse_capture(nullptr)) return rc;
    if (!c || elems == 0) return fail(SIPGPU_EARG, "cache: bad arguments");
    if (init < SIPGPU_CACHE_UNDEFINED || init > SIPGPU_CACHE_NORMAL)
        return fail(SIPGPU_EARG, "cache: unknown initialisation");
    std::lock_guard<std::mutex> g(c->mu);
    if (c->blocks.count(key)) return fail(SIPGPU_EARG, "cache: key already present");
    if (Arena::round(elems * sizeof(double)) > c->arena.size)
        return fail(SIPGPU_EBUDGET, "cache: block larger than the HBM budget");
    CBlock blk;
    blk.elems = elems;
    blk.home = const_cast<double *>(host);
    blk.home_valid = host != nullptr;
    blk.initialised = host != nullptr;
    blk.init = init;
    blk.seed = seed;
    blk.stream = stream;
    blk.a = a;
    blk.b = b;
    c->blocks.emplace(key, blk);
    return 0;
}

int sipgpu_cache_acquire(sipgpu_cache_t c, sipgpu_stream_t s, unsigned long long key, int mode,
                         double **dev) {
    if (int rc = refuse_capture(stream_of(s))) return rc;
    if (!c || !dev || mode < SIPGPU_CACHE_READ || mode > SIPGPU_CACHE_READWRITE)
        return fail(SIPGPU_EARG, "cache: bad arguments");
    std::lock_guard<std::mutex> g(c->mu);
    return acquire_locked(*c, stream_of(s), key, mode, dev);
}

int sipgpu_cache_release(sipgpu_cache_t c, sipgpu_stream_t s, unsigned long long key) {
    if (int rc = refuse_capture(stream_of(s))) return rc;
    if (!c) return fail(SIPGPU_EARG, "null cache");
    std::lock_guard<std::mutex> g(c->mu);
    return release_locked(*c, stream_of(s), key);
}

int sipgpu_cache_read(sipgpu_cache_t c, unsigned long long key, double *out) {
    if (int rc = refuse_capture(nullptr)) return rc;
    if (!c || !out) return fail(SIPGPU_EARG, "cache: bad arguments");
    std::lock_guard<std::mutex> g(c->mu);
    auto it = c->blocks.find(key);
    if (it == c->blocks.end()) return fail(SIPGPU_ENOTFOUND, "cache: unknown block key");
    CBlock &b = it->second;
    const size_t bytes = b.elems * sizeof(double);
    SIPGPU_HIP_TRY(hipDeviceSynchronize());
    if (b.off != kNotResident && !(b.initialised && b.home_valid)) {
        SIPGPU_HIP_TRY(hipMemcpy(out, c->arena.base + b.off, bytes, hipMemcpyDeviceToHost));
    } else if (b.initialised && b.home_valid) {
        std::memcpy(out, b.home, bytes);
    } else {
        return fail(SIPGPU_EARG, "cache: block has no defined value yet");
    }
    return 0;
}

int sipgpu_cache_flush(sipgpu_cache_t c, unsigned long long key, int all) {
    if (int rc = refuse_capture(nullptr)) return rc;
    if (!c) return fail(SIPGPU_EARG, "null cache");
    std::lock_guard<std::mutex> g(c->mu);
    SIPGPU_HIP_TRY(hipDeviceSynchronize());
    for (auto &kv : c->blocks) {
        if (!all && kv.first != key) continue;
        CBlock &b = kv.second;
        if (b.off == kNotResident || !b.initialised || b.home_valid) continue;
        const size_t bytes = b.elems * sizeof(double);
        if (!b.home) {
            if (int rc = home_alloc(*c, b, bytes)) return rc;
        }
        SIPGPU_HIP_TRY(hipMemcpy(b.home, c->arena.base + b.off, bytes, hipMemcpyDeviceToHost));
        b.home_valid = true;
        c->st.bytes_d2h += bytes;
    }
    return 0;
}

int sipgpu_cache_evict(sipgpu_cache_t c, unsigned long long key) {
    if (int rc = refuse_capture(nullptr)) return rc;
    if (!c) return fail(SIPGPU_EARG, "null cache");
    std::lock_guard<std::mutex> g(c->mu);
    auto it = c->blocks.find(key);
    if (it == c->blocks.end()) return fail(SIPGPU_ENOTFOUND, "cache: unknown block key");
    if (it->second.pins) return fail(SIPGPU_EARG, "cache: block in use");
    if (it->second.off == kNotResident) return 0;
    return evict_locked(*c, key, it->second);
}

int sipgpu_cache_remove(sipgpu_cache_t c, unsigned long long key) {
    if (int rc = refuse_capture(nullptr)) return rc;
    if (!c) return fail(SIPGPU_EARG, "null cache");
    if (int rc = sipgpu::rt::refuse_in_capture("removing a cached block")) return rc;
    std::lock_guard<std::mutex> g(c->mu);
    auto it = c->blocks.find(key);
    if (it == c->blocks.end()) return fail(SIPGPU_ENOTFOUND, "cache: unknown block key");
    if (it->second.pins) return fail(SIPGPU_EARG, "cache: block in use");
    return remove_locked(*c, key);
}

int sipgpu_cache_scope_enter(sipgpu_cache_t c) {
    if (!c) return fail(SIPGPU_EARG, "null cache");
    std::lock_guard<std::mutex> g(c->mu);
    c->scopes.emplace_back();
    return 0;
}

int sipgpu_cache_scope_add(sipgpu_cache_t c, unsigned long long key) {
    if (!c) return fail(SIPGPU_EARG, "null cache");
    std::lock_guard<std::mutex> g(c->mu);
    if (c->scopes.empty()) return fail(SIPGPU_EARG, "cache: no open scope");
    if (!c->blocks.count(key)) return fail(SIPGPU_ENOTFOUND, "cache: unknown block key");
    c->scopes.back().push_back(key);
    return 0;
}

int sipgpu_cache_scope_leave(sipgpu_cache_t c) {
    if (int rc = refuse_capture(nullptr)) return rc;
    if (!c) return fail(SIPGPU_EARG, "null cache");
    if (int rc = sipgpu::rt::refuse_in_capture("leaving a cache scope")) return rc;
    std::lock_guard<std::mutex> g(c->mu);
    if (c->scopes.empty()) return fail(SIPGPU_EARG, "cache: no open scope");
    // nothing is removed unless everything can be
    for (unsigned long long key : c->scopes.back()) {
        auto it = c->blocks.find(key);
        if (it != c->blocks.end() && it->second.pins)
            return fail(SIPGPU_EARG, "cache: a temp block of the scope is in use");
    }
    for (unsigned long long key : c->scopes.back())
        if (int rc = remove_locked(*c, key)) return rc;
    c->scopes.pop_back();
    return 0;
}

int sipgpu_cache_set_persistent(sipgpu_cache_t c, unsigned long long key, int persistent) {
    if (!c) return fail(SIPGPU_EARG, "null cache");
    std::lock_guard<std::mutex> g(c->mu);
    if (!c->blocks.count(key)) return fail(SIPGPU_ENOTFOUND, "cache: unknown block key");
    if (persistent)
        c->persistent.insert(key);
    else
        c->persistent.erase(key);
    return 0;
}

int sipgpu_cache_end_program(sipgpu_cache_t c, unsigned long long *kept) {
    if (int rc = refuse_capture(nullptr)) return rc;
    if (!c) return fail(SIPGPU_EARG, "null cache");
    if (int rc = sipgpu::rt::refuse_in_capture("ending a program")) return rc;
    std::lock_guard<std::mutex> g(c->mu);
    if (!c->scopes.empty()) return fail(SIPGPU_EARG, "cache: a scope is still open");
    std::vector<unsigned long long> drop;
    for (auto &kv : c->blocks) {
        if (c->persistent.count(kv.first)) continue;
        if (kv.second.pins) return fail(SIPGPU_EARG, "cache: a block is still in use");
        drop.push_back(kv.first);
    }
    for (unsigned long long key : drop)
        if (int rc = remove_locked(*c, key)) return rc;
    // a restored block belongs to the next program: it must be marked again
    // to outlive that one too
    if (kept) *kept = c->persistent.size();
    c->persistent.clear();
    return 0;
}

int sipgpu_cache_get_stats(sipgpu_cache_t c, sipgpu_cache_stats *out) {
    if (!c || !out) return fail(SIPGPU_EARG, "cache: bad arguments");
    std::lock_guard<std::mutex> g(c->mu);
    *out = c->st;
    out->resident_bytes = c->arena.used;
    out->blocks = c->blocks.size();
    return 0;
}

} // extern "C"

// ---------------------------------------------------------------------------
// Server
// ---------------------------------------------------------------------------

namespace {

constexpr int kSlots = 4;
constexpr size_t kSlotBytes = 32u << 20;
constexpr int kLanes = 2; // device streams the batches run on

enum { kKindContract = 0, kKindPermute = 1 };

// A permute request: dims are the source block's extents, transp the
// reference's dim_transp ([0] = 1, then the 1-based new position of each
// old dimension).
struct PermuteReq {
    int rank = 0;
    int dims[SIPGPU_MAX_RANK] = {};
    int transp[SIPGPU_MAX_RANK + 1] = {};
    unsigned long long src = 0, dst = 0;
};

struct Pending {
    unsigned long long ticket;
    int kind;
    sipgpu_request req; // kind == kKindContract
    PermuteReq perm;    // kind == kKindPermute
    Clock::time_point arrival;
};

struct Done {
    Clock::time_point arrival, done;
    int status = 0;
};

// Batches of each lane that last wrote / accessed a block (0: none).
struct KeyUse {
    unsigned long long write[kLanes] = {};
    unsigned long long access[kLanes] = {};
};

} // namespace

struct sipgpu_server_s {
    sipgpu_cache_t cache = nullptr;
    int max_batch = 0;
    int lanes = kLanes;
    hipStream_t stream[kLanes] = {};
    void *slot_dev[kSlots] = {};
    void *slot_host[kSlots] = {};
    hipEvent_t slot_free[kSlots] = {};
    bool slot_used[kSlots] = {};
    int next_slot = 0;

    std::mutex mu;
    std::condition_variable cv_queue, cv_done, cv_inflight;
    std::deque<Pending> queue;
    std::unordered_map<unsigned long long, Done> done;
    unsigned long long next_ticket = 1;
    bool stop = false;
    bool sched_done = false; // the scheduler has pushed its last batch

    struct Flight {
        hipEvent_t ev;
        unsigned long long seq;
        std::vector<std::pair<unsigned long long, Clock::time_point>> reqs;
        int status;
    };
    std::deque<Flight> inflight[kLanes];
    std::vector<hipEvent_t> event_pool;
    // end events of batches not yet stamped complete, by batch number (mu)
    std::unordered_map<unsigned long long, hipEvent_t> live;
    std::thread sched, compl_[kLanes];
    sipgpu_server_stats st{};
    // scheduler thread only
    std::unordered_map<unsigned long long, KeyUse> key_use;
    unsigned long long next_seq = 1;
    unsigned long long cross_waits = 0; // batches that waited on the other lane
    // split batches: planes, result planes and workspace, per lane
    void *scratch[kLanes] = {};
    size_t scratch_bytes[kLanes] = {};
};

namespace {

// A descriptor slot whose previous batch has finished reading it.
int take_slot(sipgpu_server_s &S, int &slot) {
    slot = S.next_slot;
    S.next_slot = (S.next_slot + 1) % kSlots;
    if (S.slot_used[slot]) {
        hipError_t e = hipEventSynchronize(S.slot_free[slot]);
        if (e != hipSuccess) return hip_fail(e, "server: slot");
    }
    return 0;
}

// A panel batch (split precision, or packed FP64): its descriptors through a
// slot, its planes in the lane's scratch buffer.  Batches of a lane run in
// stream order, so one scratch buffer per lane serves them all; it grows
// (after the lane's stream drains) when a batch needs more.
int server_panels(sipgpu_server_s &S, int lane, const sipgpu::rt::PanelPlan &plan,
                  const std::vector<unsigned char> &desc) {
    if (plan.ntiles == 0) return 0;
    const hipStream_t st = S.stream[lane];
    auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
    const size_t need = up(plan.plane_total()) + up(plan.res_total()) + up(plan.work_total());
    if (need > S.scratch_bytes[lane]) {
        SIPGPU_HIP_TRY(hipStreamSynchronize(st));
        if (S.scratch[lane]) SIPGPU_HIP_TRY(hipFree(S.scratch[lane]));
        S.scratch[lane] = nullptr;
        S.scratch_bytes[lane] = 0;
        const size_t bytes = std::max(need, (size_t)256 << 20);
        SIPGPU_HIP_TRY(hipMalloc(&S.scratch[lane], bytes));
        S.scratch_bytes[lane] = bytes;
    }
    int slot = 0;
    if (int rc = take_slot(S, slot)) return rc;
    int8_t *planes = static_cast<int8_t *>(S.scratch[lane]);
    uint8_t *res = reinterpret_cast<uint8_t *>(planes + up(plan.plane_total()));
    double *work = plan.nslots ? reinterpret_cast<double *>(res + up(plan.res_total())) : nullptr;
    std::memcpy(S.slot_host[slot], desc.data(), desc.size());
    hipError_t e = hipMemcpyAsync(S.slot_dev[slot], S.slot_host[slot], desc.size(),
                                  hipMemcpyHostToDevice, st);
    S.slot_used[slot] = true;
    if (e == hipSuccess)
        e = sipgpu::oz_zero_heads(planes, plan.plane_bytes, plan.nmod ? plan.nmod : 1, st);
    if (e != hipSuccess) return hip_fail(e, "server: split batch");
    sipgpu::rt::SplitData sd;
    sipgpu::rt::bind_panel_plan(plan, S.slot_dev[slot], planes, res, work, sd);
    int rc = sipgpu::rt::execute_panels(sd, plan.nmod > 0, SIPGPU_STAGE_ALL, st);
    e = hipEventRecord(S.slot_free[slot], st);
    if (!rc && e != hipSuccess) rc = hip_fail(e, "server: split batch");
    return rc;
}

size_t elems_of(int rank, const int *dims) {
    size_t n = 1;
    for (int i = 0; i < rank; ++i) n *= (size_t)dims[i];
    return n;
}

// The cache's size of a block (0 when the key is unknown).
size_t cache_elems(sipgpu_cache_s *c, unsigned long long key) {
    std::lock_guard<std::mutex> g(c->mu);
    auto it = c->blocks.find(key);
    return it == c->blocks.end() ? 0 : it->second.elems;
}

// Every block a request names holds exactly its extents' elements: a
// mismatch would let a kernel read or write outside the block.
int check_sizes(sipgpu_cache_s *c, const Pending &p) {
    if (p.kind == kKindPermute) {
        const size_t n = elems_of(p.perm.rank, p.perm.dims);
        const size_t a = cache_elems(c, p.perm.src), b = cache_elems(c, p.perm.dst);
        if (!a || !b) return fail(SIPGPU_ENOTFOUND, "server: unknown block key");
        if (a != n || b != n)
            return fail(SIPGPU_EARG, "server: permute extents do not match the cached blocks");
        return 0;
    }
    const sipgpu_request &r = p.req;
    const unsigned long long keys[3] = {r.lkey, r.rkey, r.dkey};
    const size_t want[3] = {elems_of(r.lrank, r.ldims), elems_of(r.rrank, r.rdims),
                            elems_of(r.drank, r.ddims)};
    for (int q = 0; q < 3; ++q) {
        const size_t have = cache_elems(c, keys[q]);
        if (!have) return fail(SIPGPU_ENOTFOUND, "server: unknown block key");
        if (have != want[q])
            return fail(SIPGPU_EARG, "server: request extents do not match the cached blocks");
    }
    return 0;
}

// A batch of permutes (sources distinct from every destination, each
// destination written once: the scheduler's hazard rule) as one launch of
// the batched permute kernel: each block acquired once (sources for
// reading, destinations for writing), the kernel's tables through a
// descriptor slot.
int server_permutes(sipgpu_server_s &S, int lane, const std::vector<Pending> &batch) {
    for (const Pending &p : batch)
        if (int rc = check_sizes(S.cache, p)) return rc;
    const int n = (int)batch.size();
    std::vector<sipgpu::PermutePlan> plans(n);
    for (int e = 0; e < n; ++e) {
        const PermuteReq &q = batch[e].perm;
        if (int rc = sipgpu::plan_permute(q.rank, q.dims, q.transp, plans[e]))
            return fail(rc, "server: invalid permutation");
    }
    const hipStream_t stream = S.stream[lane];
    sipgpu_stream_s sw;
    sw.s = stream;
    std::unordered_map<unsigned long long, double *> ptr;
    std::vector<unsigned long long> acquired;
    int rc = 0;
    for (const Pending &p : batch) {
        const unsigned long long keys[2] = {p.perm.src, p.perm.dst};
        for (int q = 0; q < 2 && !rc; ++q) {
            if (ptr.count(keys[q])) continue;
            double *d = nullptr;
            rc = sipgpu_cache_acquire(S.cache, &sw, keys[q],
                                      q ? SIPGPU_CACHE_WRITE : SIPGPU_CACHE_READ, &d);
            if (!rc) {
                ptr[keys[q]] = d;
                acquired.push_back(keys[q]);
            }
        }
        if (rc) break;
    }
    if (!rc) {
        std::vector<const double *> in(n);
        std::vector<double *> out(n);
        for (int e = 0; e < n; ++e) {
            in[e] = ptr[batch[e].perm.src];
            out[e] = ptr[batch[e].perm.dst];
        }
        auto round256 = [](size_t x) { return (x + 255) & ~(size_t)255; };
        const size_t first_off = round256((size_t)n * sizeof(sipgpu::PermArgs));
        const size_t entry_off = first_off + round256((size_t)(n + 1) * sizeof(int));
        std::vector<unsigned char> host(entry_off);
        int tile_cap = 0, tv_cap = 1, items = 16;
        const int64_t grid = sipgpu::permute_batch_args(
            plans.data(), in.data(), out.data(), n,
            reinterpret_cast<sipgpu::PermArgs *>(host.data()),
            reinterpret_cast<int *>(host.data() + first_off), &tile_cap, &tv_cap, &items);
        const size_t bytes = entry_off + (size_t)std::max<int64_t>(grid, 0) * sizeof(int);
        if (grid >= ((int64_t)1 << 31) || bytes > kSlotBytes) {
            // tables beyond a descriptor slot: one launch per permute
            for (int e = 0; e < n && !rc; ++e) {
                const hipError_t he = sipgpu::launch_permute(plans[e], in[e], out[e], stream);
                if (he != hipSuccess) rc = hip_fail(he, "server: permute");
            }
        } else if (grid > 0) {
            host.resize(bytes);
            const int *first = reinterpret_cast<const int *>(host.data() + first_off);
            int *entry = reinterpret_cast<int *>(host.data() + entry_off);
            for (int e = 0; e < n; ++e)
                for (int w = first[e]; w < first[e + 1]; ++w) entry[w] = e;
            int slot = 0;
            rc = take_slot(S, slot);
            if (!rc) {
                std::memcpy(S.slot_host[slot], host.data(), bytes);
                const unsigned char *dev = static_cast<const unsigned char *>(S.slot_dev[slot]);
                hipError_t e = hipMemcpyAsync(S.slot_dev[slot], S.slot_host[slot], bytes,
                                              hipMemcpyHostToDevice, stream);
                if (e == hipSuccess)
                    e = sipgpu::launch_permute_batch(
                        reinterpret_cast<const sipgpu::PermArgs *>(dev),
                        reinterpret_cast<const int *>(dev + first_off),
                        reinterpret_cast<const int *>(dev + entry_off), (int)grid, tile_cap,
                        tv_cap, items, stream);
                if (e == hipSuccess) e = hipEventRecord(S.slot_free[slot], stream);
                if (e != hipSuccess) rc = hip_fail(e, "server: permute launch");
                S.slot_used[slot] = true;
            }
        }
    }
    for (unsigned long long k : acquired) {
        const int r2 = sipgpu_cache_release(S.cache, &sw, k);
        if (!rc) rc = r2;
    }
    if (!rc) {
        S.st.batches += 1;
        S.st.requests += batch.size();
        S.st.max_batch = std::max<unsigned long long>(S.st.max_batch, batch.size());
    }
    return rc;
}

int server_batch(sipgpu_server_s &S, int lane, std::vector<Pending> &batch) {
    if (!batch.empty() && batch[0].kind == kKindPermute) return server_permutes(S, lane, batch);
    for (const Pending &p : batch)
        if (int rc = check_sizes(S.cache, p)) return rc;
    const hipStream_t stream = S.stream[lane];
    // blocks of the batch, each acquired once: destinations for write
    // (read-write if their first entry accumulates), operands for read
    std::unordered_map<unsigned long long, double *> dptr;
    std::vector<unsigned long long> acquired;
    std::unordered_map<unsigned long long, int> dmode;
    for (const Pending &p : batch)
        if (!dmode.count(p.req.dkey))
            dmode[p.req.dkey] = p.req.accumulate ? SIPGPU_CACHE_READWRITE : SIPGPU_CACHE_WRITE;
    int rc = 0;
    sipgpu_stream_s sw;
    sw.s = stream;
    for (const Pending &p : batch) {
        const unsigned long long keys[3] = {p.req.lkey, p.req.rkey, p.req.dkey};
        for (int q = 0; q < 3 && !rc; ++q) {
            if (dptr.count(keys[q])) continue;
            const int mode = q == 2 ? dmode[keys[q]] : SIPGPU_CACHE_READ;
            double *d = nullptr;
            rc = sipgpu_cache_acquire(S.cache, &sw, keys[q], mode, &d);
            if (!rc) {
                dptr[keys[q]] = d;
                acquired.push_back(keys[q]);
            }
        }
        if (rc) break;
    }
    sipgpu::rt::BatchImage img;
    sipgpu::rt::PanelPlan plan;
    const int nmod = batch.empty() ? 0 : batch[0].req.moduli;
    bool panels = nmod > 0; // split, or packed FP64
    if (!rc) {
        std::vector<sipgpu_contraction> list(batch.size());
        for (size_t i = 0; i < batch.size(); ++i) {
            const sipgpu_request &r = batch[i].req;
            sipgpu_contraction &e = list[i];
            std::memcpy(e.ptrn, r.ptrn, sizeof(e.ptrn));
            e.l = dptr[r.lkey];
            e.r = dptr[r.rkey];
            e.d = dptr[r.dkey];
            e.lrank = r.lrank;
            e.rrank = r.rrank;
            e.drank = r.drank;
            std::memcpy(e.ldims, r.ldims, sizeof(e.ldims));
            std::memcpy(e.rdims, r.rdims, sizeof(e.rdims));
            std::memcpy(e.ddims, r.ddims, sizeof(e.ddims));
            e.accumulate = r.accumulate;
        }
        if (nmod) {
            rc = sipgpu::rt::build_panel_plan(list.data(), (int)list.size(), nmod, plan);
        } else {
            rc = sipgpu::rt::build_batch_image(list.data(), (int)list.size(), img);
            // large FP64 batches run on packed panels, as sipgpu_batch_plan
            // plans them (the table kernel gathers per tile)
            if (!rc && img.ntasks > 0 && img.flops >= 1e9 && img.flops / img.ntasks >= 5e7 &&
                sipgpu::rt::build_panel_plan(list.data(), (int)list.size(), 0, plan) == 0)
                panels = true;
        }
    }
    const std::vector<unsigned char> &desc = panels ? plan.host : img.host;
    if (!rc && desc.size() > kSlotBytes)
        rc = fail(SIPGPU_ESIZE, "server: batch work list exceeds its descriptor slot");
    if (!rc && panels) rc = server_panels(S, lane, plan, desc);
    if (!rc && !panels && img.ntiles > 0) {
        int slot = 0;
        rc = take_slot(S, slot);
        if (!rc) {
            std::memcpy(S.slot_host[slot], img.host.data(), img.host.size());
            hipError_t e = hipMemcpyAsync(S.slot_dev[slot], S.slot_host[slot], img.host.size(),
                                          hipMemcpyHostToDevice, stream);
            if (e == hipSuccess)
                e = sipgpu::launch_gemm(img.cfg, img.args_at(S.slot_dev[slot]), img.ntiles,
                                        stream);
            if (e == hipSuccess) e = hipEventRecord(S.slot_free[slot], stream);
            if (e != hipSuccess) rc = hip_fail(e, "server: launch");
            S.slot_used[slot] = true;
        }
    }
    for (unsigned long long k : acquired) {
        const int r2 = sipgpu_cache_release(S.cache, &sw, k);
        if (!rc) rc = r2;
    }
    if (!rc) {
        S.st.batches += 1;
        S.st.requests += batch.size();
        S.st.flops += panels ? plan.flops : img.flops;
        S.st.max_batch = std::max<unsigned long long>(S.st.max_batch, batch.size());
    }
    return rc;
}

// Blocks a batch reads and writes.
void batch_keys(const std::vector<Pending> &batch, std::vector<unsigned long long> &reads,
                std::vector<unsigned long long> &writes) {
    for (const Pending &p : batch) {
        if (p.kind == kKindPermute) {
            reads.push_back(p.perm.src);
            writes.push_back(p.perm.dst);
        } else {
            reads.push_back(p.req.lkey);
            reads.push_back(p.req.rkey);
            writes.push_back(p.req.dkey);
            if (p.req.accumulate) reads.push_back(p.req.dkey);
        }
    }
}

// Orders a batch about to run on `lane` after the other lanes' batches that
// it conflicts with: a block it reads after their last write of it, a block
// it writes after their last access.  Batches of one lane are ordered by
// their stream.  Returns whether it had to wait.
bool order_after_other_lanes(sipgpu_server_s &S, int lane,
                             const std::vector<unsigned long long> &reads,
                             const std::vector<unsigned long long> &writes) {
    std::vector<unsigned long long> after;
    for (unsigned long long k : reads) {
        auto it = S.key_use.find(k);
        if (it == S.key_use.end()) continue;
        for (int o = 0; o < S.lanes; ++o)
            if (o != lane && it->second.write[o]) after.push_back(it->second.write[o]);
    }
    for (unsigned long long k : writes) {
        auto it = S.key_use.find(k);
        if (it == S.key_use.end()) continue;
        for (int o = 0; o < S.lanes; ++o)
            if (o != lane && it->second.access[o]) after.push_back(it->second.access[o]);
    }
    if (after.empty()) return false;
    std::sort(after.begin(), after.end());
    after.erase(std::unique(after.begin(), after.end()), after.end());
    bool waited = false;
    // Only this thread records the pooled events, so an event found live
    // here still marks the end of that batch.
    std::lock_guard<std::mutex> lk(S.mu);
    for (unsigned long long seq : after) {
        auto it = S.live.find(seq);
        if (it == S.live.end()) continue; // already complete
        (void)hipStreamWaitEvent(S.stream[lane], it->second, 0);
        waited = true;
    }
    return waited;
}

void note_use(sipgpu_server_s &S, int lane, unsigned long long seq,
              const std::vector<unsigned long long> &reads,
              const std::vector<unsigned long long> &writes) {
    for (unsigned long long k : reads) S.key_use[k].access[lane] = seq;
    for (unsigned long long k : writes) {
        KeyUse &u = S.key_use[k];
        u.write[lane] = seq;
        u.access[lane] = seq;
    }
}

// Work of a batch, 2 M N K summed over its contractions (|L| |R| / |D| =
// K^2 for a block contraction).
double batch_flops(const std::vector<Pending> &batch) {
    double f = 0.0;
    for (const Pending &p : batch) {
        if (p.kind != kKindContract) continue;
        const double l = (double)elems_of(p.req.lrank, p.req.ldims);
        const double r = (double)elems_of(p.req.rrank, p.req.rdims);
        const double d = (double)elems_of(p.req.drank, p.req.ddims);
        f += 2.0 * d * std::sqrt(l * r / d);
    }
    return f;
}

// Large batches run one after another on lane 0: each is a persistent
// grouped launch sized for the whole GPU, and two of them side by side only
// split the CUs (measured: 43 -> 27 TFLOP/s at saturation).  A small batch
// admitted while lane 0 is busy takes lane 1 and starts as CUs free up.
constexpr double kLargeBatchFlops = 1e9;

int pick_lane(sipgpu_server_s &S, const std::vector<Pending> &batch) {
    if (S.lanes < 2 || batch_flops(batch) >= kLargeBatchFlops) return 0;
    std::lock_guard<std::mutex> lk(S.mu);
    return S.inflight[0].empty() ? 0 : 1;
}

void scheduler_loop(sipgpu_server_s *S) {
    (void)hipSetDevice(sipgpu::rt::device_id());
    for (;;) {
        std::vector<Pending> batch;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_queue.wait(lk, [&] { return S->stop || !S->queue.empty(); });
            if (S->queue.empty() && S->stop) return;
            std::unordered_set<unsigned long long> reads, writes;
            // the blocks a batch pins must fit beside the ones it evicts
            const size_t pin_limit = S->cache->arena.size / 2;
            size_t pinned = 0;
            while (!S->queue.empty() && (int)batch.size() < S->max_batch) {
                const Pending &p = S->queue.front();
                // permutes batch with permutes (one launch), contractions
                // with contractions
                if (!batch.empty() && (p.kind == kKindPermute) != (batch[0].kind == kKindPermute))
                    break;
                if (p.kind == kKindPermute) {
                    const PermuteReq &q = p.perm;
                    // program order: a destination is neither read nor
                    // written by another permute of the batch
                    if (writes.count(q.src) || reads.count(q.dst) || writes.count(q.dst)) break;
                    const size_t add = (reads.count(q.src) ? 0 : 8 * elems_of(q.rank, q.dims)) +
                                       8 * elems_of(q.rank, q.dims);
                    if (!batch.empty() && pinned + add > pin_limit) break;
                    pinned += add;
                    reads.insert(q.src);
                    writes.insert(q.dst);
                    batch.push_back(p);
                    S->queue.pop_front();
                    continue;
                }
                const sipgpu_request &r = p.req;
                // keep program order: no block both written and read in one batch
                if (writes.count(r.lkey) || writes.count(r.rkey) || reads.count(r.dkey)) break;
                // one precision per batch
                if (!batch.empty() && r.moduli != batch[0].req.moduli) break;
                size_t add = 0;
                if (!reads.count(r.lkey)) add += 8 * elems_of(r.lrank, r.ldims);
                if (!reads.count(r.rkey)) add += 8 * elems_of(r.rrank, r.rdims);
                if (!writes.count(r.dkey)) add += 8 * elems_of(r.drank, r.ddims);
                if (!batch.empty() && pinned + add > pin_limit) break;
                pinned += add;
                reads.insert(r.lkey);
                reads.insert(r.rkey);
                writes.insert(r.dkey);
                batch.push_back(p);
                S->queue.pop_front();
            }
        }
        std::vector<unsigned long long> rkeys, wkeys;
        batch_keys(batch, rkeys, wkeys);
        const int lane = pick_lane(*S, batch);
        const unsigned long long seq = S->next_seq++;
        if (order_after_other_lanes(*S, lane, rkeys, wkeys)) S->cross_waits += 1;
        const int rc = server_batch(*S, lane, batch);
        hipEvent_t ev = nullptr;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            if (!S->event_pool.empty()) {
                ev = S->event_pool.back();
                S->event_pool.pop_back();
            }
        }
        if (!ev) (void)hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        (void)hipEventRecord(ev, S->stream[lane]);
        note_use(*S, lane, seq, rkeys, wkeys);
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->live[seq] = ev;
            sipgpu_server_s::Flight f;
            f.ev = ev;
            f.seq = seq;
            f.status = rc;
            for (const Pending &p : batch) f.reqs.emplace_back(p.ticket, p.arrival);
            S->inflight[lane].push_back(std::move(f));
        }
        S->cv_inflight.notify_all();
    }
}

// One completion thread per lane: a lane's batches end in stream order, so
// each is stamped as soon as its own event completes.
void completion_loop(sipgpu_server_s *S, int lane) {
    (void)hipSetDevice(sipgpu::rt::device_id());
    for (;;) {
        sipgpu_server_s::Flight f;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_inflight.wait(lk, [&] { return !S->inflight[lane].empty() || S->sched_done; });
            if (S->inflight[lane].empty()) return; // scheduler finished and all batches stamped
            f = std::move(S->inflight[lane].front());
            S->inflight[lane].pop_front();
        }
        const hipError_t e = hipEventSynchronize(f.ev);
        const Clock::time_point t = Clock::now();
        std::lock_guard<std::mutex> lk(S->mu);
        for (auto &r : f.reqs) {
            Done d;
            d.arrival = r.second;
            d.done = t;
            d.status = f.status ? f.status : (e == hipSuccess ? 0 : SIPGPU_EDEVICE);
            S->done[r.first] = d;
        }
        S->live.erase(f.seq);
        S->event_pool.push_back(f.ev);
        S->cv_done.notify_all();
    }
}

int check_request(const sipgpu_request &r) {
    if (r.lrank < 1 || r.lrank > SIPGPU_MAX_RANK || r.rrank < 1 || r.rrank > SIPGPU_MAX_RANK ||
        r.drank < 1 || r.drank > SIPGPU_MAX_RANK)
        return fail(SIPGPU_ERANK, "server: requests are block = block * block, rank 1..6");
    if (r.dkey == r.lkey || r.dkey == r.rkey)
        return fail(SIPGPU_EHAZARD, "server: a request writes a block it reads");
    if (r.moduli != 0 && (r.moduli < 6 || r.moduli > 14))
        return fail(SIPGPU_EARG, "server: moduli must be 0 (FP64) or 6..14");
    return 0;
}

int enqueue(sipgpu_server_s *S, Pending p, unsigned long long *ticket) {
    if (int rc = check_sizes(S->cache, p)) return rc;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stop) return fail(SIPGPU_EARG, "server: stopped");
        p.ticket = S->next_ticket++;
        if (ticket) *ticket = p.ticket;
        S->queue.push_back(p);
    }
    S->cv_queue.notify_one();
    return 0;
}

int submit_at(sipgpu_server_s *S, const sipgpu_request *r, Clock::time_point arrival,
              unsigned long long *ticket) {
    if (int rc = check_request(*r)) return rc;
    Pending p{};
    p.kind = kKindContract;
    p.req = *r;
    p.arrival = arrival;
    return enqueue(S, p, ticket);
}

int wait_ticket(sipgpu_server_s *S, unsigned long long ticket, double *latency_ms) {
    std::unique_lock<std::mutex> lk(S->mu);
    S->cv_done.wait(lk, [&] { return S->done.count(ticket) != 0; });
    const Done d = S->done[ticket];
    S->done.erase(ticket);
    if (latency_ms)
        *latency_ms = std::chrono::duration<double, std::milli>(d.done - d.arrival).count();
    if (d.status) return fail(d.status, "server: request failed (" + std::to_string(d.status) + ")");
    return 0;
}

} // namespace

extern "C" {

int sipgpu_server_create(sipgpu_cache_t cache, int max_batch, sipgpu_server_t *server) {
    if (int rc = sipgpu::rt::refuse_in_capture("creating a server")) return rc;
    if (!cache || !server || max_batch < 1) return fail(SIPGPU_EARG, "server: bad arguments");
    if (int rc = sipgpu::rt::ensure_ready()) return rc;
    auto S = std::make_unique<sipgpu_server_s>();
    S->cache = cache;
    S->max_batch = max_batch;
    // SIPGPU_SERVER_LANES=1 runs every batch on one stream (experiments)
    if (const char *e = std::getenv("SIPGPU_SERVER_LANES"))
        S->lanes = std::max(1, std::min(kLanes, std::atoi(e)));
    for (int l = 0; l < S->lanes; ++l)
        SIPGPU_HIP_TRY(hipStreamCreateWithFlags(&S->stream[l], hipStreamNonBlocking));
    for (int i = 0; i < kSlots; ++i) {
        SIPGPU_HIP_TRY(hipMalloc(&S->slot_dev[i], kSlotBytes));
        SIPGPU_HIP_TRY(hipHostMalloc(&S->slot_host[i], kSlotBytes, hipHostMallocDefault));
        SIPGPU_HIP_TRY(hipEventCreateWithFlags(&S->slot_free[i], hipEventDisableTiming));
    }
    S->sched = std::thread(scheduler_loop, S.get());
    for (int l = 0; l < S->lanes; ++l) S->compl_[l] = std::thread(completion_loop, S.get(), l);
    *server = S.release();
    return 0;
}

int sipgpu_server_submit(sipgpu_server_t S, const sipgpu_request *r, unsigned long long *ticket) {
    if (!S || !r) return fail(SIPGPU_EARG, "server: bad arguments");
    return submit_at(S, r, Clock::now(), ticket);
}

int sipgpu_server_submit_permute(sipgpu_server_t S, int rank, const int *dims,
                                 const int *dim_transp, unsigned long long src_key,
                                 unsigned long long dst_key, unsigned long long *ticket) {
    if (!S || !dims || !dim_transp) return fail(SIPGPU_EARG, "server: bad arguments");
    if (rank < 1 || rank > SIPGPU_MAX_RANK)
        return fail(SIPGPU_ERANK, "server: permutes of rank 1..6");
    if (src_key == dst_key) return fail(SIPGPU_EHAZARD, "server: a permute writes the block it reads");
    Pending p{};
    p.kind = kKindPermute;
    p.perm.rank = rank;
    p.perm.src = src_key;
    p.perm.dst = dst_key;
    std::memcpy(p.perm.dims, dims, sizeof(int) * rank);
    std::memcpy(p.perm.transp, dim_transp, sizeof(int) * (rank + 1));
    // a permutation of 1..rank (the kernel's plan checks it again)
    unsigned seen = 0;
    for (int i = 1; i <= rank; ++i) {
        const int t = dim_transp[i];
        if (t < 1 || t > rank || (seen >> t) & 1u)
            return fail(SIPGPU_EPERM, "server: dim_transp is not a permutation");
        seen |= 1u << t;
    }
    for (int i = 0; i < rank; ++i)
        if (dims[i] < 1) return fail(SIPGPU_EARG, "server: permute extents must be positive");
    p.arrival = Clock::now();
    return enqueue(S, p, ticket);
}

int sipgpu_server_wait(sipgpu_server_t S, unsigned long long ticket, double *latency_ms) {
    if (!S) return fail(SIPGPU_EARG, "null server");
    return wait_ticket(S, ticket, latency_ms);
}

int sipgpu_server_replay(sipgpu_server_t S, const sipgpu_request *reqs, const double *arrival_s,
                         int n, double *latency_ms, double *wall_s) {
    if (!S || n < 0 || (n > 0 && (!reqs || !arrival_s || !latency_ms)))
        return fail(SIPGPU_EARG, "server: bad arguments");
    std::vector<unsigned long long> tickets(n);
    const Clock::time_point t0 = Clock::now();
    int rc = 0;
    for (int i = 0; i < n && !rc; ++i) {
        const Clock::time_point at =
            t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(arrival_s[i]));
        while (Clock::now() < at) {
            // spin for the last 50 us, sleep before that
            if (at - Clock::now() > std::chrono::microseconds(200))
                std::this_thread::sleep_for(std::chrono::microseconds(100));
        }
        rc = submit_at(S, &reqs[i], at, &tickets[i]);
    }
    for (int i = 0; i < n; ++i) {
        if (!tickets[i]) continue;
        const int r2 = wait_ticket(S, tickets[i], &latency_ms[i]);
        if (!rc) rc = r2;
    }
    if (wall_s) *wall_s = std::chrono::duration<double>(Clock::now() - t0).count();
    return rc;
}

int sipgpu_server_get_stats(sipgpu_server_t S, sipgpu_server_stats *out) {
    if (!S || !out) return fail(SIPGPU_EARG, "server: bad arguments");
    std::lock_guard<std::mutex> lk(S->mu);
    *out = S->st;
    return 0;
}

int sipgpu_server_lanes(sipgpu_server_t S, int *lanes, unsigned long long *cross_waits) {
    if (!S) return fail(SIPGPU_EARG, "null server");
    if (lanes) *lanes = S->lanes;
    if (cross_waits) *cross_waits = S->cross_waits;
    return 0;
}

int sipgpu_server_destroy(sipgpu_server_t S) {
    if (int rc = sipgpu::rt::refuse_in_capture("destroying a server")) return rc;
    if (!S) return fail(SIPGPU_EARG, "null server");
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_queue.notify_all();
    S->sched.join();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->sched_done = true;
    }
    S->cv_inflight.notify_all();
    for (int l = 0; l < S->lanes; ++l) S->compl_[l].join();
    for (int l = 0; l < S->lanes; ++l) (void)hipStreamSynchronize(S->stream[l]);
    for (int i = 0; i < kSlots; ++i) {
        (void)hipFree(S->slot_dev[i]);
        (void)hipHostFree(S->slot_host[i]);
        (void)hipEventDestroy(S->slot_free[i]);
    }
    for (hipEvent_t e : S->event_pool) (void)hipEventDestroy(e);
    for (int l = 0; l < S->lanes; ++l) {
        if (S->scratch[l]) (void)hipFree(S->scratch[l]);
        (void)hipStreamDestroy(S->stream[l]);
    }
    delete S;
    return 0;
}

} // extern "C"
