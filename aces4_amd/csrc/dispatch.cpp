// dispatch.cpp -- the SIP worker's dispatch of contraction and permute
// instructions onto the device request queue.
//
// The interpreter keeps a stack of block selectors (array, index ids, and
// through the SipTables the block's extents).  For a contraction it pops
// the RIGHT operand first, then the left, and builds the ACES label pattern
// [destination ids | left ids | right ids] for get_contraction_ptrn_
// (Interpreter::handle_contraction, interpreter.cpp:1210-1262, reached from
// the block_contract_op / block_contract_to_scalar_op cases, :734-757).  For
// a permute it pops the LEFT-hand side first, then the right-hand side, and
// turns the two index-id lists into a 0-based old-to-new vector with -1 in
// unused slots (Interpreter::permute_rhs_to_lhs, :2049-2084, reached from
// block_permute_op, :656-679), which Block::transpose_copy hands to the
// permute kernel as [1, perm + 1] with the SOURCE block's extents
// (block.cpp:216-255).
//
// Here the same steps produce requests for the continuous-batching server
// (serve.cpp) instead of synchronous kernel calls: contractions and
// permutes are queued and return at once.  The server keeps submission
// order between requests that share a block.  Host-visible uses of a block
// (a print, a super-instruction's argument, a scalar pushed onto the
// expression stack) first wait for the queued requests that write it; the
// worker tracks, per block, the last queued writer and the queued readers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/sipgpu.h"
#include "internal.h"
#include "plan.h"

using sipgpu::rt::fail;

namespace {

struct Selector {
    unsigned long long key = 0;
    int rank = 0;
    int ids[SIPGPU_MAX_RANK] = {};
    int ext[SIPGPU_MAX_RANK] = {};
};

size_t elems_of(int rank, const int *ext) {
    size_t n = 1;
    for (int i = 0; i < rank; ++i) n *= (size_t)ext[i];
    return n;
}

} // namespace

struct sipgpu_worker_s {
    std::mutex mu;
    sipgpu_cache_t cache = nullptr;
    sipgpu_server_t server = nullptr;
    sipgpu_stream_t stream = nullptr; // synchronous work: scalar results, host writes
    double *scalar = nullptr;         // device result of a contraction to a scalar
    std::vector<Selector> stack;
    // queued requests not yet waited for, and per block the last queued
    // writer and the queued readers since it
    std::set<unsigned long long> pending;
    std::unordered_map<unsigned long long, unsigned long long> writer;
    std::unordered_map<unsigned long long, std::vector<unsigned long long>> readers;
    int first_error = 0;
    std::string first_message;
    sipgpu_worker_stats st{};
};

namespace {

// Waits for one queued request (once: the server forgets a ticket when it
// has been waited for).  A failed request's status is kept as the worker's
// first error and returned.
int wait_one(sipgpu_worker_s &W, unsigned long long t) {
    if (!W.pending.count(t)) return 0;
    W.pending.erase(t);
    const int rc = sipgpu_server_wait(W.server, t, nullptr);
    if (rc && !W.first_error) {
        W.first_error = rc;
        W.first_message = sipgpu_last_error();
    }
    W.st.waits += 1;
    return rc;
}

// Read-after-write: the queued request that last writes `key`.
int wait_writer(sipgpu_worker_s &W, unsigned long long key) {
    auto it = W.writer.find(key);
    if (it == W.writer.end()) return 0;
    const int rc = wait_one(W, it->second);
    W.writer.erase(it);
    return rc;
}

// Write-after-read / write-after-write: every queued request on `key`.
int wait_all_on(sipgpu_worker_s &W, unsigned long long key) {
    int rc = wait_writer(W, key);
    auto it = W.readers.find(key);
    if (it != W.readers.end()) {
        for (unsigned long long t : it->second) {
            const int r2 = wait_one(W, t);
            if (!rc) rc = r2;
        }
        W.readers.erase(it);
    }
    return rc;
}

void note_queued(sipgpu_worker_s &W, unsigned long long t,
                 std::initializer_list<unsigned long long> reads, unsigned long long write) {
    W.pending.insert(t);
    for (unsigned long long k : reads) {
        std::vector<unsigned long long> &v = W.readers[k];
        if (v.size() >= 64) // forget readers already waited for
            v.erase(std::remove_if(v.begin(), v.end(),
                                   [&](unsigned long long x) { return !W.pending.count(x); }),
                    v.end());
        v.push_back(t);
    }
    // the server runs the new writer after the queued readers of its block,
    // so waiting for the writer covers them
    W.writer[write] = t;
    W.readers.erase(write);
}

// A destination the SIP would create on first write
// (BlockManager::get_block_for_writing): registered in the cache if absent.
int ensure_block(sipgpu_worker_s &W, unsigned long long key, int rank, const int *ext, int init) {
    const size_t n = elems_of(rank, ext);
    size_t have = 0;
    if (int rc = sipgpu::rt::cache_block_elems(W.cache, key, &have)) return rc;
    if (have == 0) return sipgpu_cache_insert(W.cache, key, n, nullptr, init, 0, 0, 0.0, 0.0);
    if (have != n) return fail(SIPGPU_EARG, "worker: block extents differ from the cached block");
    return 0;
}

int pop(sipgpu_worker_s &W, Selector &out) {
    if (W.stack.empty()) return fail(SIPGPU_EARG, "worker: block selector stack is empty");
    out = W.stack.back();
    W.stack.pop_back();
    return 0;
}

// The contraction pattern of handle_contraction: labels of the destination,
// then the left, then the right operand.
int pattern_of(int drank, const int *dids, const Selector &l, const Selector &r, int *ptrn) {
    int aces[3 * SIPGPU_MAX_RANK];
    int n = 0;
    for (int i = 0; i < drank; ++i) aces[n++] = dids[i];
    for (int i = 0; i < l.rank; ++i) aces[n++] = l.ids[i];
    for (int i = 0; i < r.rank; ++i) aces[n++] = r.ids[i];
    const int ierr = sipgpu::convert_aces_pattern(drank, l.rank, r.rank, aces, ptrn);
    if (ierr)
        return fail(SIPGPU_EPATTERN,
                    "worker: get_contraction_ptrn_ returned ierr " + std::to_string(ierr));
    return 0;
}

} // namespace

extern "C" {

int sipgpu_permute_rhs_to_lhs(int lrank, const int *lhs_ids, const int *rhs_ids, int *perm) {
    if (lrank < 0 || lrank > SIPGPU_MAX_RANK || !lhs_ids || !rhs_ids || !perm)
        return fail(SIPGPU_EARG, "permute_rhs_to_lhs: bad arguments");
    for (int i = 0; i < SIPGPU_MAX_RANK; ++i) perm[i] = 0;
    for (int i = 0; i < lrank; ++i) {
        int j = 0;
        while (j < SIPGPU_MAX_RANK && rhs_ids[j] != lhs_ids[i]) ++j;
        if (j >= lrank) return fail(SIPGPU_EPERM, "permute_rhs_to_lhs: illegal transpose");
        perm[j] = i;
    }
    for (int i = lrank; i < SIPGPU_MAX_RANK; ++i) perm[i] = -1;
    return 0;
}

int sipgpu_worker_create(sipgpu_cache_t cache, sipgpu_server_t server, sipgpu_worker_t *worker) {
    if (int rc = sipgpu::rt::refuse_in_capture("creating a worker")) return rc;
    if (!cache || !server || !worker) return fail(SIPGPU_EARG, "worker: bad arguments");
    auto W = std::make_unique<sipgpu_worker_s>();
    W->cache = cache;
    W->server = server;
    if (int rc = sipgpu_stream_create(&W->stream)) return rc;
    void *p = nullptr;
    if (int rc = sipgpu_malloc(&p, sizeof(double))) {
        (void)sipgpu_stream_destroy(W->stream);
        return rc;
    }
    W->scalar = static_cast<double *>(p);
    *worker = W.release();
    return 0;
}

int sipgpu_worker_destroy(sipgpu_worker_t W) {
    if (!W) return fail(SIPGPU_EARG, "null worker");
    const int rc = sipgpu_worker_flush(W, 0);
    (void)sipgpu_free(W->scalar);
    (void)sipgpu_stream_destroy(W->stream);
    delete W;
    return rc;
}

int sipgpu_worker_push_selector(sipgpu_worker_t W, unsigned long long key, int rank,
                                const int *index_ids, const int *extents) {
    if (!W || rank < 0 || rank > SIPGPU_MAX_RANK || (rank > 0 && (!index_ids || !extents)))
        return fail(SIPGPU_EARG, "worker: bad selector");
    Selector s;
    s.key = key;
    s.rank = rank;
    for (int i = 0; i < SIPGPU_MAX_RANK; ++i) {
        s.ids[i] = i < rank ? index_ids[i] : -1; // unused_index_slot
        s.ext[i] = i < rank ? extents[i] : 1;
        if (i < rank && extents[i] < 1) return fail(SIPGPU_EARG, "worker: extents must be positive");
    }
    std::lock_guard<std::mutex> g(W->mu);
    W->stack.push_back(s);
    return 0;
}

int sipgpu_worker_stack_depth(sipgpu_worker_t W, int *depth) {
    if (!W || !depth) return fail(SIPGPU_EARG, "worker: bad arguments");
    std::lock_guard<std::mutex> g(W->mu);
    *depth = (int)W->stack.size();
    return 0;
}

int sipgpu_worker_contract(sipgpu_worker_t W, unsigned long long dkey, int drank,
                           const int *dindex_ids, const int *dextents, int accumulate,
                           int moduli, unsigned long long *ticket) {
    if (!W || drank < 1 || drank > SIPGPU_MAX_RANK || !dindex_ids || !dextents)
        return fail(SIPGPU_EARG, "worker: contraction into a block of rank 1..6");
    std::lock_guard<std::mutex> g(W->mu);
    Selector r, l;
    if (int rc = pop(*W, r)) return rc; // the right operand is on top
    if (int rc = pop(*W, l)) return rc;
    int ptrn[2 * SIPGPU_MAX_RANK] = {};
    if (int rc = pattern_of(drank, dindex_ids, l, r, ptrn)) return rc;
    if (!sipgpu::pattern_valid(ptrn, l.rank, l.ext, r.rank, r.ext, drank, dextents))
        return fail(SIPGPU_EPATTERN, "worker: pattern does not fit the block extents");
    if (int rc = ensure_block(*W, dkey, drank, dextents, SIPGPU_CACHE_ZERO)) return rc;
    sipgpu_request q{};
    std::memcpy(q.ptrn, ptrn, sizeof(q.ptrn));
    q.lrank = l.rank;
    q.rrank = r.rank;
    q.drank = drank;
    std::memcpy(q.ldims, l.ext, sizeof(q.ldims));
    std::memcpy(q.rdims, r.ext, sizeof(q.rdims));
    for (int i = 0; i < SIPGPU_MAX_RANK; ++i) q.ddims[i] = i < drank ? dextents[i] : 1;
    q.lkey = l.key;
    q.rkey = r.key;
    q.dkey = dkey;
    q.accumulate = accumulate ? 1 : 0;
    q.moduli = moduli;
    unsigned long long t = 0;
    if (int rc = sipgpu_server_submit(W->server, &q, &t)) return rc;
    note_queued(*W, t, {l.key, r.key}, dkey);
    W->st.contractions += 1;
    if (ticket) *ticket = t;
    return 0;
}

int sipgpu_worker_permute(sipgpu_worker_t W, const int *compiler_perm, unsigned long long *ticket) {
    if (!W) return fail(SIPGPU_EARG, "null worker");
    std::lock_guard<std::mutex> g(W->mu);
    Selector lhs, rhs;
    if (int rc = pop(*W, lhs)) return rc; // the left-hand side is on top
    if (int rc = pop(*W, rhs)) return rc;
    if (lhs.rank < 1 || rhs.rank != lhs.rank)
        return fail(SIPGPU_EARG, "worker: a permute's sides must have the same rank >= 1");
    int perm[SIPGPU_MAX_RANK];
    if (int rc = sipgpu_permute_rhs_to_lhs(lhs.rank, lhs.ids, rhs.ids, perm)) return rc;
    if (compiler_perm)
        for (int i = 0; i < SIPGPU_MAX_RANK; ++i)
            if (perm[i] != compiler_perm[i])
                return fail(SIPGPU_EPERM, "worker: permutation vector from compiler differs");
    // Block::transpose_copy: [1, perm + 1] over the source's extents
    int transp[SIPGPU_MAX_RANK + 1];
    transp[0] = 1;
    for (int i = 0; i < lhs.rank; ++i) transp[i + 1] = perm[i] + 1;
    for (int j = 0; j < rhs.rank; ++j)
        if (lhs.ext[perm[j]] != rhs.ext[j])
            return fail(SIPGPU_EARG, "worker: permute extents do not match");
    if (int rc = ensure_block(*W, lhs.key, lhs.rank, lhs.ext, SIPGPU_CACHE_UNDEFINED)) return rc;
    unsigned long long t = 0;
    if (int rc = sipgpu_server_submit_permute(W->server, rhs.rank, rhs.ext, transp, rhs.key,
                                              lhs.key, &t))
        return rc;
    note_queued(*W, t, {rhs.key}, lhs.key);
    W->st.permutes += 1;
    if (ticket) *ticket = t;
    return 0;
}

int sipgpu_worker_contract_to_scalar(sipgpu_worker_t W, double *value) {
    if (!W || !value) return fail(SIPGPU_EARG, "worker: bad arguments");
    std::lock_guard<std::mutex> g(W->mu);
    Selector r, l;
    if (int rc = pop(*W, r)) return rc;
    if (int rc = pop(*W, l)) return rc;
    int ptrn[2 * SIPGPU_MAX_RANK] = {};
    if (int rc = pattern_of(0, nullptr, l, r, ptrn)) return rc;
    // the scalar goes onto the expression stack: a host-visible result
    if (int rc = wait_writer(*W, l.key)) return rc;
    if (int rc = wait_writer(*W, r.key)) return rc;
    double *dl = nullptr, *dr = nullptr;
    int rc = sipgpu_cache_acquire(W->cache, W->stream, l.key, SIPGPU_CACHE_READ, &dl);
    if (rc) return rc;
    rc = sipgpu_cache_acquire(W->cache, W->stream, r.key, SIPGPU_CACHE_READ, &dr);
    if (!rc) {
        rc = sipgpu_contract(W->stream, ptrn, dl, l.rank, l.ext, dr, r.rank, r.ext, W->scalar, 0,
                             nullptr, 0);
        const int r2 = sipgpu_cache_release(W->cache, W->stream, r.key);
        if (!rc) rc = r2;
    }
    const int r3 = sipgpu_cache_release(W->cache, W->stream, l.key);
    if (!rc) rc = r3;
    if (!rc) rc = sipgpu_stream_synchronize(W->stream);
    if (!rc) rc = sipgpu_memcpy_d2h(value, W->scalar, sizeof(double));
    if (!rc) W->st.scalars += 1;
    return rc;
}

int sipgpu_worker_read(sipgpu_worker_t W, unsigned long long key, double *out) {
    if (!W || !out) return fail(SIPGPU_EARG, "worker: bad arguments");
    std::lock_guard<std::mutex> g(W->mu);
    if (int rc = wait_writer(*W, key)) return rc;
    W->st.reads += 1;
    return sipgpu_cache_read(W->cache, key, out);
}

int sipgpu_worker_write(sipgpu_worker_t W, unsigned long long key, int rank, const int *extents,
                        const double *host) {
    if (!W || !host || rank < 0 || rank > SIPGPU_MAX_RANK || (rank > 0 && !extents))
        return fail(SIPGPU_EARG, "worker: bad arguments");
    std::lock_guard<std::mutex> g(W->mu);
    // a super-instruction writes the block: every queued use of it first
    if (int rc = wait_all_on(*W, key)) return rc;
    if (int rc = ensure_block(*W, key, rank, extents, SIPGPU_CACHE_UNDEFINED)) return rc;
    double *d = nullptr;
    int rc = sipgpu_cache_acquire(W->cache, W->stream, key, SIPGPU_CACHE_WRITE, &d);
    if (rc) return rc;
    // on the worker's stream: the acquire may have made it wait for a spill
    // out of the range the block now occupies
    const hipStream_t hs = W->stream->s;
    hipError_t e = hipMemcpyAsync(d, host, elems_of(rank, extents) * sizeof(double),
                                  hipMemcpyHostToDevice, hs);
    if (e == hipSuccess) e = hipStreamSynchronize(hs);
    if (e != hipSuccess) rc = sipgpu::rt::hip_fail(e, "worker: block upload");
    const int r2 = sipgpu_cache_release(W->cache, W->stream, key);
    if (!rc) rc = r2;
    if (!rc) W->st.writes += 1;
    return rc;
}

int sipgpu_worker_flush(sipgpu_worker_t W, int writeback) {
    if (!W) return fail(SIPGPU_EARG, "null worker");
    std::lock_guard<std::mutex> g(W->mu);
    const std::vector<unsigned long long> all(W->pending.begin(), W->pending.end());
    for (unsigned long long t : all) (void)wait_one(*W, t);
    W->writer.clear();
    W->readers.clear();
    W->st.flushes += 1;
    int rc = W->first_error;
    if (rc) {
        fail(rc, "worker: a queued request failed: " + W->first_message);
        W->first_error = 0;
        W->first_message.clear();
        return rc;
    }
    if (writeback) rc = sipgpu_cache_flush(W->cache, 0, 1);
    return rc;
}

int sipgpu_worker_get_stats(sipgpu_worker_t W, sipgpu_worker_stats *out) {
    if (!W || !out) return fail(SIPGPU_EARG, "worker: bad arguments");
    std::lock_guard<std::mutex> g(W->mu);
    *out = W->st;
    out->queued = W->pending.size();
    return 0;
}

} // extern "C"
