// runtime.cpp -- libsipgpu's host runtime and C ABI.
//
// Drop-in layer: the reference's tensor-kernel symbols (host pointers,
// synchronous, assign semantics, the reference's ierr codes), staged
// through a grow-only device arena so no call allocates once warm.
// Native layer: device blocks, streams, cached contraction plans and
// planned batches (one grouped FP64-MFMA launch per batch).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <climits>
#include <numeric>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <tuple>
#include <mutex>
#include <queue>
#include <set>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/sipgpu.h"
#include "internal.h"
#include "kernels.h"
#include "ozaki.h"
#include "plan.h"

using namespace sipgpu;

struct sipgpu_event_s {
    hipEvent_t e = nullptr;
};

namespace {
thread_local std::string g_last_error;
} // namespace

namespace sipgpu {
namespace rt {

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    return fail(SIPGPU_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

} // namespace rt
} // namespace sipgpu

namespace {

using sipgpu::rt::BatchImage;
using sipgpu::rt::fail;
using sipgpu::rt::hip_fail;
double g_last_clock_mhz = 0.0; // shader clock seen by the last MFMA calibration
// Compute units of the library's device (256 on MI355X), read at init.
int64_t g_device_cus = 256;

#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) return hip_fail(e_, #expr);                                 \
    } while (0)

// Grow-only device buffer.  Freed only at process exit by the driver.
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    int reserve(size_t bytes) {
        if (bytes <= cap) return 0;
        if (p) {
            hipError_t e = hipFree(p);
            if (e != hipSuccess) return hip_fail(e, "hipFree");
            p = nullptr;
            cap = 0;
        }
        size_t want = std::max<size_t>(bytes, 1 << 20);
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) return fail(SIPGPU_ENOMEM, "device allocation failed");
        cap = want;
        return 0;
    }
};

// Device copy of one contraction shape's offset tables.
struct DeviceShape {
    ContractionShape shape;
    int32_t *tab = nullptr; // koffL | moffL | koffR | noffR | moffD | noffD
    int kL = 0, mL = 0, kR = 0, nR = 0, mD = 0, nD = 0;
};

struct Context {
    std::mutex mu;
    bool ready = false;
    int device = 0;
    sipgpu_stream_s stream;
    DevBuf stage_l, stage_r, stage_d;
    std::unordered_map<std::string, std::unique_ptr<DeviceShape>> shapes;
};

Context &ctx() {
    static Context *c = new Context(); // never destroyed: HIP may be gone at exit
    return *c;
}

// Device memory of sipgpu_contract_noalloc: the K-segment workspace of a
// split single contraction, then the contract-to-scalar partials.  One copy
// per device, held until the process ends.  On the host `scratch` names a
// stand-in, not device memory: init resolves each device's address with
// hipGetSymbolAddress, and host code only ever passes that address on.
// A split contraction has at most g_device_cus / 2 tiles of 64 x 64 and
// four K segments, so its three workspace segments hold at most
// 3 * 128 * 4096 doubles (12 MiB) on 256 CUs.
constexpr size_t kNoAllocWsDoubles = (size_t)2 << 20; // 16 MiB
constexpr size_t kNoAllocPartDoubles = (size_t)std::max(kDotBlocks, kEnergyBlocks);
static __device__ __align__(256) double scratch[kNoAllocWsDoubles + kNoAllocPartDoubles];
constexpr int kMaxDevices = 64;
double *g_scratch_dev[kMaxDevices] = {}; // hipGetSymbolAddress of `scratch`, per device

// Workspace and partials a contraction uses instead of its stream's own.
struct FixedScratch {
    double *ws;      // K segments 1.. of a split contraction
    size_t ws_bytes;
    double *part;    // contract-to-scalar partials (kDotBlocks doubles)
};

// the stream scratch holds the partial sums of dots, energies and norms
static_assert(kNormBlocks <= std::max(kDotBlocks, kEnergyBlocks), "scratch too small for norm2");

int ensure_stream_scratch(sipgpu_stream_s &s) {
    if (s.scratch) return 0;
    HIP_TRY(hipMalloc(&s.scratch, std::max(kDotBlocks, kEnergyBlocks) * sizeof(double)));
    return 0;
}

int init_locked(Context &c, int device) {
    if (c.ready) return 0;
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(SIPGPU_EARG, "no such device");
    HIP_TRY(hipSetDevice(device));
    int cus = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    if (cus > 0) g_device_cus = cus;
    if (device < kMaxDevices && !g_scratch_dev[device]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_dev[device] = static_cast<double *>(p);
    }
    HIP_TRY(hipStreamCreateWithFlags(&c.stream.s, hipStreamNonBlocking));
    int rc = ensure_stream_scratch(c.stream);
    if (rc) return rc;
    c.device = device;
    c.ready = true;
    return 0;
}

int ensure_init(Context &c) { return init_locked(c, c.device); }

sipgpu_stream_s &resolve(Context &c, sipgpu_stream_t s) { return s ? *s : c.stream; }

std::string shape_key(const int *ptrn, int lr, const int *ld, int rr, const int *rd, int dr,
                      const int *dd) {
    std::string k;
    auto put = [&](int v) { k.append(reinterpret_cast<const char *>(&v), sizeof(v)); };
    put(lr); put(rr); put(dr);
    for (int i = 0; i < lr + rr; ++i) put(ptrn[i]);
    for (int i = 0; i < lr; ++i) put(ld[i]);
    for (int i = 0; i < rr; ++i) put(rd[i]);
    for (int i = 0; i < dr; ++i) put(dd[i]);
    return k;
}

// Looks up or builds the device tables of a contraction shape.  Returns the
// reference's ierr (1, -1) or a SIPGPU_E* code.
int get_shape(Context &c, const int *ptrn, int lr, const int *ld, int rr, const int *rd,
              int dr, const int *dd, DeviceShape **out, bool may_plan = true) {
    const std::string key = shape_key(ptrn, lr, ld, rr, rd, dr, dd);
    auto it = c.shapes.find(key);
    if (it != c.shapes.end()) { *out = it->second.get(); return 0; }
    if (!may_plan)
        return fail(SIPGPU_ENOTFOUND, "no-allocation contraction of a shape that is not planned "
                                      "(run sipgpu_contract with it once)");
    // a new shape uploads its tables: run it once before capturing
    if (int rc = sipgpu::rt::refuse_in_capture("a contraction of a new shape")) return rc;
    auto ds = std::make_unique<DeviceShape>();
    int rc = plan_contraction(ptrn, lr, ld, rr, rd, dr, dd, ds->shape, true);
    if (rc) return rc;
    const ContractionShape &s = ds->shape;
    std::vector<int32_t> pool;
    auto add = [&](const std::vector<int32_t> &t) {
        const int at = (int)pool.size();
        pool.insert(pool.end(), t.begin(), t.end());
        return at;
    };
    ds->kL = add(s.koffL); ds->mL = add(s.moffL);
    ds->kR = add(s.koffR); ds->nR = add(s.noffR);
    ds->mD = add(s.moffD); ds->nD = add(s.noffD);
    HIP_TRY(hipMalloc(&ds->tab, pool.size() * sizeof(int32_t)));
    HIP_TRY(hipMemcpy(ds->tab, pool.data(), pool.size() * sizeof(int32_t),
                      hipMemcpyHostToDevice));
    *out = ds.get();
    c.shapes.emplace(key, std::move(ds));
    return 0;
}

// Kernel-side roles of one block pair: D(m, n) = sum_k A(k, m) B(k, n) with
// (A, B) = (L, R), or (R, L) on the transposed problem when the destination's
// fastest dim is on the left side.
void make_task_pair(const ContractionShape &s, int kL, int mL, int kR, int nR, int mD, int nD,
                    const double *l, const double *r, double *d, int beta, GemmTask &task,
                    GemmPair &pair) {
    const int fl = (s.l_kfast ? 1 : 0) | (s.r_kfast ? 2 : 0);
    if (!s.swap) {
        task = GemmTask{d, (int)s.M, (int)s.N, mD, nD, 0, 0, beta, 0};
        pair = GemmPair{l, r, (int)s.K, kL, mL, kR, nR, fl};
    } else {
        task = GemmTask{d, (int)s.N, (int)s.M, nD, mD, 0, 0, beta, 0};
        pair = GemmPair{r, l, (int)s.K, kR, nR, kL, mL, ((fl & 1) << 1) | ((fl & 2) >> 1)};
    }
}

int pick_config(int64_t M, int64_t N) {
    // SIPGPU_GEMM_CONFIG=<n> pins the tile configuration (tuning runs).
    static const int forced = [] {
        const char *e = std::getenv("SIPGPU_GEMM_CONFIG");
        return e ? std::atoi(e) : -1;
    }();
    if (forced >= 0 && forced < kGemmConfigs) return forced;
    // Config 0 (64x64 tile, four waves of 32x32, no occupancy hint) measured
    // fastest on every block shape of the benchmark terms; the larger tiles
    // fit fewer workgroups per CU.
    (void)M;
    (void)N;
    return kGemm64x64;
}

// One contraction on device pointers (no locking; caller holds the lock).
// `fx` (may be null) replaces the stream's workspace and partials; with it
// the call allocates nothing and refuses shapes it would have to plan.
int contract_device(Context &c, sipgpu_stream_s &st, const int *ptrn, const double *l, int lr,
                    const int *ld, const double *r, int rr, const int *rd, double *d, int dr,
                    const int *dd, int accumulate, const FixedScratch *fx = nullptr) {
    DeviceShape *ds = nullptr;
    int rc = get_shape(c, ptrn, lr, ld, rr, rd, dr, dd, &ds, fx == nullptr);
    if (rc) return rc;
    const ContractionShape &s = ds->shape;
    if (s.dsize == 0) return 0;
    const int32_t *tab = ds->tab;
    if (dr > 0 && lr > 0 && rr > 0) {
        GemmTask task;
        GemmPair pair;
        make_task_pair(s, ds->kL, ds->mL, ds->kR, ds->nR, ds->mD, ds->nD, l, r, d,
                       accumulate ? 1 : 0, task, pair);
        const int cfg = pick_config(task.M, task.N);
        const int bm = gemm_tile_m(cfg), bn = gemm_tile_n(cfg);
        GemmArgs a{};
        a.tab = tab;
        a.tilesM = (task.M + bm - 1) / bm;
        const int64_t ntiles = (int64_t)a.tilesM * ((task.N + bn - 1) / bn);
        if (ntiles >= ((int64_t)1 << 31)) return fail(SIPGPU_ESIZE, "too many tiles");
        a.ntiles = (int)ntiles;
        task.first = 0;
        task.npairs = 1;
        a.task0 = task;
        a.pair0 = pair;
        // Too few tiles to fill the CUs and a long K: split K into segments
        // that run side by side, segment s > 0 into the stream's workspace,
        // then add the segments to D in order (a fixed summation order).
        // Up to 8 segments of at least 64 k, two workgroups per CU (a ring
        // block pair: p50 47.6 -> 44.5 us against 4 segments of 128, one per
        // CU), and no more than the 16 MiB of the no-allocation workspace
        // holds, so both entry points split every shape the same way.
        const int64_t K = pair.K;
        int nseg = 1, kseg = (int)K;
        if (cfg == kGemm64x64 && ntiles * 2 <= g_device_cus && K >= 256) {
            const int64_t fit = 1 + (int64_t)(kNoAllocWsDoubles / std::max<int64_t>(s.dsize, 1));
            const int want = (int)std::max<int64_t>(
                1, std::min<int64_t>({8, 2 * g_device_cus / ntiles, K / 64, fit}));
            kseg = (int)(((K + want - 1) / want + 15) / 16 * 16);
            nseg = (int)((K + kseg - 1) / kseg);
        }
        if (nseg > 1) {
            const size_t need = (size_t)(nseg - 1) * (size_t)s.dsize * sizeof(double);
            if (fx) {
                if (need > fx->ws_bytes)
                    return fail(SIPGPU_ESIZE, "no-allocation contraction: its K segments need " +
                                                  std::to_string(need) + " bytes, more than the " +
                                                  std::to_string(fx->ws_bytes) +
                                                  " of the static scratch");
            } else if (need > st.ws_bytes) {
                // No allocation while this thread captures (on any stream)
                // or this stream is captured.  Running unsplit instead would
                // change the summation order, so a replay would differ in
                // bits from the eager result: refuse, and the caller runs the
                // contraction once on this stream before capturing.
                if (sipgpu::rt::thread_captures() > 0 || sipgpu::rt::capturing(st.s))
                    return fail(SIPGPU_ECAPTURE,
                                "a contraction whose split workspace must grow (run it once "
                                "on this stream before capturing)");
                // at least 32 MiB, doubling: a stream of uneven blocks
                // reallocates a few times at most.  The old workspace is
                // kept until the stream is destroyed: graphs captured on
                // this stream address it.
                const size_t grow = std::max({need, 2 * st.ws_bytes, (size_t)32 << 20});
                double *fresh = nullptr;
                if (hipMalloc(&fresh, grow) != hipSuccess) {
                    // HBM taken (by a block cache, say): the unsplit
                    // contraction needs no workspace, so run it that way
                    (void)hipGetLastError();
                    nseg = 1;
                } else {
                    if (st.ws) st.retired.push_back(st.ws);
                    st.ws = fresh;
                    st.ws_bytes = grow;
                }
            }
        }
        if (nseg > 1) {
            double *ws = fx ? fx->ws : st.ws;
            a.ksplit = nseg;
            a.kseg = kseg;
            a.ws = ws;
            a.ws_stride = s.dsize;
            a.ntiles = (int)ntiles * nseg;
            HIP_TRY(launch_gemm(cfg, a, a.ntiles, st.s));
            AddManyArgs am{};
            am.n = nseg - 1;
            for (int q = 0; q < am.n; ++q) am.p[q] = ws + (size_t)q * s.dsize;
            HIP_TRY(launch_block_add_many(d, am, s.dsize, st.s));
            return 0;
        }
        HIP_TRY(launch_gemm(cfg, a, a.ntiles, st.s));
        return 0;
    }
    if (accumulate) return fail(SIPGPU_EARG, "accumulate is only defined for block results");
    if (dr == 0 && lr > 0 && rr > 0) {
        if (!fx) {
            rc = ensure_stream_scratch(st);
            if (rc) return rc;
        }
        HIP_TRY(launch_dot(d, l, tab + ds->kL, r, tab + ds->kR, (int)s.K,
                           fx ? fx->part : st.scratch, st.s));
    } else if (dr > 0 && lr > 0 && rr == 0) {
        HIP_TRY(launch_scale_scatter(d, tab + ds->mD, l, tab + ds->mL, r, (int)s.M, st.s));
    } else if (dr > 0 && lr == 0 && rr > 0) {
        HIP_TRY(launch_scale_scatter(d, tab + ds->nD, r, tab + ds->nR, l, (int)s.N, st.s));
    } else if (dr == 0 && lr == 0 && rr == 0) {
        HIP_TRY(launch_scalar_mul(d, l, r, st.s));
    }
    return 0;
}

int64_t prod(int n, const int *dims) {
    int64_t p = 1;
    for (int i = 0; i < n; ++i) p *= dims[i];
    return p;
}

} // namespace

// ---------------------------------------------------------------------------
// Batches
// ---------------------------------------------------------------------------

using sipgpu::rt::SplitData;
using sipgpu::rt::PanelPlan;

struct sipgpu_batch_s {
    int cfg = kGemm64x64;
    int ntiles = 0;     // including XCD padding slots
    int ntasks = 0, npairs = 0, real_tiles = 0;
    int ngroups = 0;    // panel paths: GEMM tasks after merging result blocks
    double flops = 0.0;
    void *dev = nullptr; // tiles | tasks | pairs | tab | parts
    GemmArgs args{};
    // table kernel, result blocks cut into pair groups: the groups' workspace
    // copies and the entries that add them into the blocks
    double *ws = nullptr;
    const AddPartsEntry *parts = nullptr;
    int nparts = 0;
    long long parts_max_n = 0;
    bool split = false;
    bool packed = false; // FP64 on packed operand panels
    SplitData sd;
};

namespace {

constexpr int kXcds = 8; // MI355X: workgroups go round-robin to 8 XCDs

// Offset tables of a batch, pooled in one device array; tables with equal
// contents are stored once (shapes of uneven segments share most of them).
struct TablePool {
    std::vector<int32_t> pool;
    std::unordered_map<std::string, int> index;
    int add(const std::vector<int32_t> &t) {
        std::string key(reinterpret_cast<const char *>(t.data()), t.size() * sizeof(int32_t));
        auto it = index.find(key);
        if (it != index.end()) return it->second;
        const int off = (int)pool.size();
        pool.insert(pool.end(), t.begin(), t.end());
        index.emplace(std::move(key), off);
        return off;
    }
    std::array<int, 6> add_shape(const ContractionShape &s) {
        return {add(s.koffL), add(s.moffL), add(s.koffR), add(s.noffR), add(s.moffD),
                add(s.noffD)};
    }
    // The same tables with rows (columns) renumbered in order of their
    // destination offsets: a relabeling of the product's rows that changes no
    // sum, after which the destination's unit-stride index, wherever it sits
    // among the row (column) indices, runs along consecutive rows (columns).
    std::array<int, 6> add_shape_dest_ordered(const ContractionShape &s) {
        auto order = [](const std::vector<int32_t> &d) {
            std::vector<int> p(d.size());
            std::iota(p.begin(), p.end(), 0);
            std::stable_sort(p.begin(), p.end(), [&](int a, int b) { return d[a] < d[b]; });
            return p;
        };
        auto take = [](const std::vector<int32_t> &t, const std::vector<int> &p) {
            std::vector<int32_t> out(p.size());
            for (size_t i = 0; i < p.size(); ++i) out[i] = t[p[i]];
            return out;
        };
        const std::vector<int> pm = order(s.moffD), pn = order(s.noffD);
        return {add(s.koffL), add(take(s.moffL, pm)), add(s.koffR), add(take(s.noffR, pn)),
                add(take(s.moffD, pm)), add(take(s.noffD, pn))};
    }
};

// Static XCD schedule of a grouped launch.  Slot t of the launch runs on
// XCD t % 8, so the tile list is laid out XCD-major.  Whole tasks go to one
// XCD (their tiles then share its L2), heaviest first onto the least loaded
// XCD; a task with more than 1/16 of all tiles is first cut into rectangular
// blocks of tiles of at most that size, so a batch of few large tasks still
// spans the chip.  Empty slots are task -1.
template <class Tile>
std::vector<std::vector<Tile>> xcd_lists(const std::vector<int> &tm, const std::vector<int> &tn,
                                         const std::vector<double> &cost, int &real_tiles) {
    const int ntasks = (int)tm.size();
    long long total = 0;
    for (int t = 0; t < ntasks; ++t) total += (long long)tm[t] * tn[t];
    const long long cap = std::max<long long>(1, (total + 2 * kXcds - 1) / (2 * kXcds));
    // A unit is a rectangle of one task's tiles: the whole task, or for a
    // task above `cap` tiles one block of a pr x pc cut chosen as square as
    // possible, so the tiles an XCD runs together share few operand panels
    // (a 4500 x 4500 grid cut in 9 x 9-tile blocks: 18 panels per 81 tiles
    // instead of 38 for a column-major range).
    struct Unit { int task, m0, m1, n0, n1; double cost; };
    std::vector<Unit> units;
    for (int t = 0; t < ntasks; ++t) {
        const long long nt = (long long)tm[t] * tn[t];
        const long long pieces = std::max<long long>(1, (nt + cap - 1) / cap);
        int pr = (int)std::lround(std::sqrt((double)pieces * tm[t] / std::max(tn[t], 1)));
        pr = std::max(1, std::min(pr, tm[t]));
        int pc = (int)std::min<long long>(tn[t], (pieces + pr - 1) / pr);
        pc = std::max(pc, 1);
        const int bm = (tm[t] + pr - 1) / pr, bn = (tn[t] + pc - 1) / pc;
        for (int n0 = 0; n0 < tn[t]; n0 += bn)
            for (int m0 = 0; m0 < tm[t]; m0 += bm) {
                const int m1 = std::min(tm[t], m0 + bm), n1 = std::min(tn[t], n0 + bn);
                const double c = cost[t] * (double)(m1 - m0) * (n1 - n0) / (double)std::max<long long>(nt, 1);
                units.push_back(Unit{t, m0, m1, n0, n1, c});
            }
    }
    std::stable_sort(units.begin(), units.end(),
                     [](const Unit &a, const Unit &b) { return a.cost > b.cost; });
    std::vector<std::vector<Tile>> per_xcd(kXcds);
    std::vector<double> load(kXcds, 0.0);
    real_tiles = 0;
    for (const Unit &u : units) {
        const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        for (int q = u.n0; q < u.n1; ++q)
            for (int p = u.m0; p < u.m1; ++p) {
                Tile tl{};
                tl.task = u.task;
                tl.mt = p;
                tl.nt = q;
                per_xcd[x].push_back(tl);
            }
        load[x] += u.cost;
        real_tiles += (u.m1 - u.m0) * (u.n1 - u.n0);
    }
    return per_xcd;
}

// The launch order that sends list x to XCD x: slot j * kXcds + x.
template <class Tile>
std::vector<Tile> interleave(const std::vector<std::vector<Tile>> &per_xcd) {
    size_t per = 0;
    for (auto &v : per_xcd) per = std::max(per, v.size());
    Tile empty{};
    empty.task = -1;
    std::vector<Tile> tiles(per * kXcds, empty);
    for (int x = 0; x < kXcds; ++x)
        for (size_t j = 0; j < per_xcd[x].size(); ++j) tiles[j * kXcds + x] = per_xcd[x][j];
    return tiles;
}

template <class Tile>
std::vector<Tile> xcd_schedule(const std::vector<int> &tm, const std::vector<int> &tn,
                               const std::vector<double> &cost, int &real_tiles) {
    return interleave(xcd_lists<Tile>(tm, tn, cost, real_tiles));
}

// Makespan of a list dispatched in order onto `slots` identical workgroup
// slots (each unit starts on the slot that frees first).
double list_makespan(const std::vector<double> &cost, int slots) {
    std::priority_queue<double, std::vector<double>, std::greater<double>> q;
    for (int i = 0; i < slots; ++i) q.push(0.0);
    double end = 0.0;
    for (double c : cost) {
        const double t = q.top() + c;
        q.pop();
        q.push(t);
        end = std::max(end, t);
    }
    return end;
}

// Persistent schedule of the packed FP64 GEMM: `slots` workgroups per XCD
// (two per CU), each running its own list of work units.  Per XCD, the
// tiles of whole rounds go in list order to the least loaded workgroup (the
// order the hardware would dispatch them in, so concurrent tiles share
// panels in L2); the last, partly filled round is water-filled: its stages,
// concatenated in list order, are cut into one contiguous piece per
// workgroup such that all workgroups finish together.  A tile cut into
// several K segments takes consecutive workspace slots (one per segment, in
// K order) and one fixup entry.  Units of workgroup j of XCD x go to
// workgroup j * kXcds + x, so they run on XCD x.
void persistent_schedule(const std::vector<std::vector<OzTile>> &lists,
                         const std::vector<std::vector<double>> &costs, int slots,
                         std::vector<OzTile> &units, std::vector<int> &first,
                         std::vector<OzTile> &fix, int &nslots) {
    std::vector<std::vector<OzTile>> wg((size_t)slots * kXcds);
    constexpr int kMinSeg = 4; // stages
    for (int x = 0; x < kXcds; ++x) {
        const auto &L = lists[x];
        const auto &c = costs[x];
        const int n = (int)L.size();
        const int r = n % slots;
        const int full = n - r;
        std::vector<double> load(slots, 0.0);
        using Slot = std::pair<double, int>;
        std::priority_queue<Slot, std::vector<Slot>, std::greater<Slot>> q;
        for (int j = 0; j < slots; ++j) q.push({0.0, j});
        for (int i = 0; i < full; ++i) {
            const Slot sl = q.top();
            q.pop();
            wg[(size_t)sl.second * kXcds + x].push_back(L[i]);
            load[sl.second] = sl.first + c[i];
            q.push({load[sl.second], sl.second});
        }
        if (r == 0) continue;
        // water level F: sum over workgroups of max(0, F - load) = tail work
        double W = 0.0;
        for (int i = full; i < n; ++i) W += c[i];
        std::vector<double> sorted = load;
        std::sort(sorted.begin(), sorted.end());
        double F = sorted.back() + W / slots, acc = 0.0;
        for (int m = 1; m <= slots; ++m) {
            acc += sorted[m - 1];
            const double f = (W + acc) / m;
            if (m == slots || f <= sorted[m]) {
                F = f;
                break;
            }
        }
        // cut the tail into pieces, workgroups in index order: piece j ends
        // at the stage boundary (a multiple of kMinSeg into its tile)
        // nearest the cumulative target, so rounding does not accumulate
        struct Seg { int j, s0, s1; };
        std::vector<std::vector<Seg>> segs(n - full);
        int i = full, st = L[full].s0;
        double pos = 0.0, target = 0.0;
        for (int j = 0; j < slots && i < n; ++j) {
            target += std::max(0.0, F - load[j]);
            if (j == slots - 1) target = 1e300; // the last takes the rest
            while (i < n) {
                const int nst = L[i].s1 - L[i].s0;
                const double cs = c[i] / nst; // cost per stage
                const int rem = L[i].s1 - st;
                if (pos + rem * cs <= target + 0.5 * kMinSeg * cs) {
                    segs[i - full].push_back({j, st, L[i].s1});
                    pos += rem * cs;
                    if (++i < n) st = L[i].s0;
                    continue;
                }
                int k = (int)std::lround((target - pos) / cs / kMinSeg) * kMinSeg;
                if (k > 0 && L[i].s1 - (st + k) < kMinSeg) k = rem;
                if (k > 0) {
                    segs[i - full].push_back({j, st, st + k});
                    st += k;
                    pos += k * cs;
                    if (st == L[i].s1 && ++i < n) st = L[i].s0;
                }
                break;
            }
        }
        for (int t = 0; t < n - full; ++t) {
            const OzTile &u = L[full + t];
            const auto &sg = segs[t];
            if (sg.size() == 1) { // whole tile
                wg[(size_t)sg[0].j * kXcds + x].push_back(u);
                continue;
            }
            OzTile f = u;
            f.part = nslots;
            f.s0 = (int)sg.size();
            fix.push_back(f);
            for (const Seg &e : sg) {
                OzTile v = u;
                v.s0 = e.s0;
                v.s1 = e.s1;
                v.part = nslots++;
                wg[(size_t)e.j * kXcds + x].push_back(v);
            }
        }
    }
    first.assign(1, 0);
    for (const auto &v : wg) {
        units.insert(units.end(), v.begin(), v.end());
        first.push_back((int)units.size());
    }
}

// No destination may overlap an operand or another destination (partially:
// entries naming the same block accumulate into it).
template <class ShapeVec>
int check_hazards(const sipgpu_contraction *list, int n, const ShapeVec &shapes,
                  const std::vector<int> &shape_of) {
    struct Range { uintptr_t lo, hi; int kind, who; };
    std::vector<Range> rs;
    for (int i = 0; i < n; ++i) {
        const ContractionShape &s = *shapes[shape_of[i]];
        auto rng = [](const double *p, int64_t cnt) {
            return std::make_pair((uintptr_t)p, (uintptr_t)(p + cnt));
        };
        auto L = rng(list[i].l, s.lsize), R = rng(list[i].r, s.rsize), D = rng(list[i].d, s.dsize);
        rs.push_back({L.first, L.second, 0, i});
        rs.push_back({R.first, R.second, 0, i});
        rs.push_back({D.first, D.second, 1, i});
    }
    {
        std::vector<Range> ds;
        for (auto &r : rs) if (r.kind == 1) ds.push_back(r);
        std::sort(ds.begin(), ds.end(), [](const Range &a, const Range &b) { return a.lo < b.lo; });
        for (size_t i = 1; i < ds.size(); ++i) {
            if (ds[i].lo < ds[i - 1].hi && !(ds[i].lo == ds[i - 1].lo && ds[i].hi == ds[i - 1].hi))
                return fail(SIPGPU_EHAZARD, "destination blocks partially overlap");
        }
        std::vector<Range> ops;
        for (auto &r : rs) if (r.kind == 0) ops.push_back(r);
        std::sort(ops.begin(), ops.end(), [](const Range &a, const Range &b) { return a.lo < b.lo; });
        // prefix maximum of operand ends: a destination [lo, hi) overlaps an
        // operand iff some operand starting before hi ends after lo
        std::vector<uintptr_t> maxhi(ops.size());
        uintptr_t m = 0;
        for (size_t i = 0; i < ops.size(); ++i) maxhi[i] = m = std::max(m, ops[i].hi);
        for (auto &d : ds) {
            auto it = std::lower_bound(ops.begin(), ops.end(), d.hi,
                                       [](const Range &a, uintptr_t v) { return a.lo < v; });
            if (it != ops.begin() && maxhi[(it - ops.begin()) - 1] > d.lo)
                return fail(SIPGPU_EHAZARD, "a batch reads a block that it writes");
        }
    }

    return 0;
}

} // namespace

int sipgpu::rt::build_batch_image(const sipgpu_contraction *list, int n, BatchImage &b,
                                  int cfg, int split_cus) {
    struct Entry {
        const sipgpu_contraction *e;
        std::string key;
    };
    // Shapes of this batch, deduplicated; tables pooled.
    std::unordered_map<std::string, int> shape_index;
    std::vector<const ContractionShape *> shapes;
    std::vector<std::unique_ptr<ContractionShape>> owned;
    std::vector<int> shape_of(n);
    for (int i = 0; i < n; ++i) {
        const sipgpu_contraction &e = list[i];
        if (!e.l || !e.r || !e.d) return fail(SIPGPU_EARG, "null block pointer in batch");
        if (e.lrank < 1 || e.lrank > SIPGPU_MAX_RANK || e.rrank < 1 ||
            e.rrank > SIPGPU_MAX_RANK || e.drank < 1 || e.drank > SIPGPU_MAX_RANK)
            return fail(SIPGPU_ERANK, "batch entries must be block = block * block, rank 1..6");
        std::string key = shape_key(e.ptrn, e.lrank, e.ldims, e.rrank, e.rdims, e.drank, e.ddims);
        auto it = shape_index.find(key);
        if (it == shape_index.end()) {
            auto s = std::make_unique<ContractionShape>();
            int rc = plan_contraction(e.ptrn, e.lrank, e.ldims, e.rrank, e.rdims, e.drank,
                                      e.ddims, *s, true);
            if (rc == 1 || rc == -1)
                return fail(SIPGPU_EPATTERN, "batch entry " + std::to_string(i) +
                                                 ": contraction pattern rejected");
            if (rc) return fail(rc, "batch entry " + std::to_string(i) + ": plan failed");
            it = shape_index.emplace(key, (int)shapes.size()).first;
            shapes.push_back(s.get());
            owned.push_back(std::move(s));
        }
        shape_of[i] = it->second;
    }

    if (int rc = check_hazards(list, n, shapes, shape_of)) return rc;

    // Tasks: entries grouped by destination, in list order.  An assign entry
    // discards everything before it; the task keeps entries from the last
    // assign on (beta = 0), or all of them with beta = 1.
    std::unordered_map<const double *, std::vector<int>> by_dest;
    std::vector<const double *> dest_order;
    for (int i = 0; i < n; ++i) {
        auto &v = by_dest[list[i].d];
        if (v.empty()) dest_order.push_back(list[i].d);
        v.push_back(i);
    }

    TablePool tp;
    std::vector<int32_t> &pool = tp.pool;
    std::unordered_map<int, std::array<int, 6>> shape_tab;
    auto tabs_of = [&](int si) -> std::array<int, 6> {
        auto it = shape_tab.find(si);
        if (it != shape_tab.end()) return it->second;
        const std::array<int, 6> at = tp.add_shape(*shapes[si]);
        shape_tab.emplace(si, at);
        return at;
    };

    std::vector<GemmTask> tasks;
    std::vector<GemmPair> pairs;
    std::vector<double> task_cost;
    int64_t maxM = 0, maxN = 0, sumM = 0, sumN = 0;
    b.flops = 0.0;
    for (const double *dp : dest_order) {
        const auto &idx = by_dest[dp];
        size_t start = 0;
        for (size_t j = 0; j < idx.size(); ++j)
            if (!list[idx[j]].accumulate) start = j;
        const int beta = list[idx[start]].accumulate ? 1 : 0;
        const ContractionShape &s0 = *shapes[shape_of[idx[start]]];
        const auto t0 = tabs_of(shape_of[idx[start]]);
        GemmTask task;
        GemmPair unused;
        make_task_pair(s0, t0[0], t0[1], t0[2], t0[3], t0[4], t0[5], nullptr, nullptr,
                       const_cast<double *>(dp), beta, task, unused);
        task.first = (int)pairs.size();
        double cost = 0.0;
        for (size_t j = start; j < idx.size(); ++j) {
            const int i = idx[j];
            const ContractionShape &s = *shapes[shape_of[i]];
            if (s.M != s0.M || s.N != s0.N || s.moffD != s0.moffD || s.noffD != s0.noffD ||
                s.swap != s0.swap)
                return fail(SIPGPU_EARG, "entries writing one block disagree on its layout");
            const auto t = tabs_of(shape_of[i]);
            GemmTask tk;
            GemmPair pr;
            make_task_pair(s, t[0], t[1], t[2], t[3], t[4], t[5], list[i].l, list[i].r,
                           const_cast<double *>(dp), beta, tk, pr);
            pairs.push_back(pr);
            cost += 2.0 * (double)s.M * (double)s.N * (double)s.K;
        }
        task.npairs = (int)pairs.size() - task.first;
        b.flops += cost;
        tasks.push_back(task);
        task_cost.push_back(cost);
        maxM = std::max<int64_t>(maxM, task.M);
        maxN = std::max<int64_t>(maxN, task.N);
        sumM += task.M;
        sumN += task.N;
    }
    const int nblocks = (int)tasks.size();
    b.cfg = cfg >= 0 && cfg < kGemmConfigs
                ? cfg
                : pick_config(nblocks ? sumM / nblocks : 0, nblocks ? sumN / nblocks : 0);
    const int bm = gemm_tile_m(b.cfg), bn = gemm_tile_n(b.cfg);

    // Too few tiles to fill the CUs: each result block's pairs are cut into
    // up to 16 contiguous groups, about four tiles per CU in all (the ladder
    // of 16 blocks of 100 x 100 x 16 pairs of K 100: 0.171 ms unsplit, 0.044
    // with two tiles per CU, 0.037 with four).  Group 0 is
    // the block's own task; group q > 0 assigns into a workspace copy of the
    // block, and after the GEMM the copies are added into the block in group
    // order (a fixed summation order).
    std::vector<AddPartsEntry> parts;
    if (split_cus > 0) {
        long long tiles0 = 0;
        for (const GemmTask &t : tasks)
            tiles0 += (long long)((t.M + bm - 1) / bm) * ((t.N + bn - 1) / bn);
        if (tiles0 > 0 && tiles0 * 2 <= split_cus) {
            const int G = (int)std::min<long long>(16, 4 * split_cus / tiles0);
            auto group_cost = [&](const GemmTask &t, int p0, int p1) {
                double c = 0.0;
                for (int p = p0; p < p1; ++p) c += 2.0 * (double)t.M * t.N * pairs[p].K;
                return c;
            };
            for (int t = 0; t < nblocks; ++t) {
                const int np = tasks[t].npairs, first = tasks[t].first;
                const int g = std::min(G, np);
                if (g < 2) continue;
                const long long dsize = (long long)tasks[t].M * tasks[t].N;
                AddPartsEntry pe{tasks[t].D, nullptr, dsize, g - 1, 0};
                b.parts_off.push_back(b.ws_doubles);
                for (int q = 1; q < g; ++q) {
                    GemmTask sub = tasks[t];
                    sub.first = first + (int)((long long)q * np / g);
                    sub.npairs = first + (int)((long long)(q + 1) * np / g) - sub.first;
                    sub.beta = 0;
                    sub.D = nullptr; // bound to the workspace
                    b.ws_task.push_back((int)tasks.size());
                    b.ws_off.push_back(b.ws_doubles);
                    b.ws_doubles += dsize;
                    task_cost.push_back(group_cost(sub, sub.first, sub.first + sub.npairs));
                    tasks.push_back(sub);
                }
                tasks[t].npairs = (int)((long long)np / g);
                task_cost[t] = group_cost(tasks[t], first, first + tasks[t].npairs);
                parts.push_back(pe);
                b.parts_max_n = std::max(b.parts_max_n, dsize);
            }
        }
    }
    const int ntasks = (int)tasks.size();

    std::vector<int> tms(ntasks), tns(ntasks);
    for (int t = 0; t < ntasks; ++t) {
        tms[t] = (tasks[t].M + bm - 1) / bm;
        tns[t] = (tasks[t].N + bn - 1) / bn;
    }
    int real_tiles = 0;
    const std::vector<GemmTile> tiles = xcd_schedule<GemmTile>(tms, tns, task_cost, real_tiles);
    b.real_tiles = real_tiles;

    // Upload: tiles | tasks | pairs | tables | parts, each 256-byte aligned.
    auto align = [](size_t v) { return (v + 255) & ~(size_t)255; };
    const size_t o_tiles = 0;
    const size_t o_tasks = align(o_tiles + tiles.size() * sizeof(GemmTile));
    const size_t o_pairs = align(o_tasks + tasks.size() * sizeof(GemmTask));
    const size_t o_tab = align(o_pairs + pairs.size() * sizeof(GemmPair));
    const size_t o_parts = align(o_tab + pool.size() * sizeof(int32_t));
    const size_t total = align(o_parts + parts.size() * sizeof(AddPartsEntry)) + 256;
    std::vector<unsigned char> host(total, 0);
    std::memcpy(host.data() + o_tiles, tiles.data(), tiles.size() * sizeof(GemmTile));
    std::memcpy(host.data() + o_tasks, tasks.data(), tasks.size() * sizeof(GemmTask));
    std::memcpy(host.data() + o_pairs, pairs.data(), pairs.size() * sizeof(GemmPair));
    std::memcpy(host.data() + o_tab, pool.data(), pool.size() * sizeof(int32_t));
    if (!parts.empty())
        std::memcpy(host.data() + o_parts, parts.data(), parts.size() * sizeof(AddPartsEntry));
    b.host = std::move(host);
    b.o_tiles = o_tiles;
    b.o_tasks = o_tasks;
    b.o_pairs = o_pairs;
    b.o_tab = o_tab;
    b.o_parts = o_parts;
    b.nparts = (int)parts.size();
    b.ntiles = (int)tiles.size();
    b.ntasks = nblocks;
    b.npairs = (int)pairs.size();
    (void)maxM;
    (void)maxN;
    return 0;
}

namespace {

int plan_batch_locked(Context &, const sipgpu_contraction *list, int n, sipgpu_batch_s &b,
                      int cfg = -1, bool groups = false) {
    BatchImage img;
    if (int rc = sipgpu::rt::build_batch_image(list, n, img, cfg, groups ? (int)g_device_cus : 0))
        return rc;
    if (img.ws_doubles) {
        HIP_TRY(hipMalloc(&b.ws, (size_t)img.ws_doubles * sizeof(double)));
        img.bind_workspace(b.ws);
    }
    HIP_TRY(hipMalloc(&b.dev, img.host.size()));
    HIP_TRY(hipMemcpy(b.dev, img.host.data(), img.host.size(), hipMemcpyHostToDevice));
    b.args = img.args_at(b.dev);
    b.parts = img.nparts ? reinterpret_cast<const AddPartsEntry *>(
                               static_cast<unsigned char *>(b.dev) + img.o_parts)
                         : nullptr;
    b.nparts = img.nparts;
    b.parts_max_n = img.parts_max_n;
    b.cfg = img.cfg;
    b.flops = img.flops;
    b.ntiles = img.ntiles;
    b.ntasks = img.ntasks;
    b.npairs = img.npairs;
    b.real_tiles = img.real_tiles;
    return 0;
}

} // namespace

namespace {

// Pairwise coprime moduli <= 255, largest first.
constexpr int kModuli[kOzMaxModuli] = {255, 254, 253, 251, 247, 241, 239, 233,
                                       229, 227, 223, 211, 199, 197, 193, 191};
constexpr int kOzMinModuli = 6, kOzMaxUsable = 14; // M < 2^112: exact in 128-bit two's complement

int inv_mod(int a, int p) { // a^-1 mod p (gcd(a, p) = 1)
    int t = 0, nt = 1, r = p, nr = ((a % p) + p) % p;
    while (nr) {
        const int q = r / nr;
        int tmp = t - q * nt; t = nt; nt = tmp;
        tmp = r - q * nr; r = nr; nr = tmp;
    }
    return t < 0 ? t + p : t;
}

// Moduli constants for n moduli and the largest task K.  Returns P (bits).
int oz_constants(int n, long long kmax, OzConst &c) {
    c = OzConst{};
    c.n = n;
    unsigned __int128 M = 1;
    double log2M = 0.0;
    for (int i = 0; i < n; ++i) {
        c.p[i] = kModuli[i];
        M *= (unsigned __int128)kModuli[i];
        log2M += std::log2((double)kModuli[i]);
    }
    for (int i = 0; i < n; ++i) {
        const int p = c.p[i];
        long long mi = 1; // (M / p_i) mod p_i
        for (int j = 0; j < n; ++j)
            if (j != i) mi = (mi * (c.p[j] % p)) % p;
        c.u[i] = inv_mod((int)mi, p);
        c.c17[i] = (int)((1LL << 17) % p);
        c.c34[i] = (int)((1LL << 34) % p);
        c.off[i] = p << 19;
        c.bar[i] = (uint32_t)((1ULL << 32) / (unsigned long long)p);
        c.rp[i] = 1.0 / p;
        const unsigned __int128 Mi = M / (unsigned __int128)p;
        c.mi[i][0] = (uint64_t)Mi;
        c.mi[i][1] = (uint64_t)(Mi >> 64);
    }
    c.m[0] = (uint64_t)M;
    c.m[1] = (uint64_t)(M >> 64);
    // |C| <= K 2^(2P-2) <= M / 4
    const double lk = std::log2((double)std::max<long long>(kmax, 1));
    int P = (int)std::floor((log2M - 2.0 - lk) / 2.0) + 1;
    P = std::min(P, 52);
    c.P = P;
    return P;
}

// Panel planner: nmod > 0 plans the split-precision path (1-byte residue
// panels, nmod moduli); nmod == 0 the packed FP64 path (panels of doubles,
// products scattered straight into the destinations).
} // namespace

int sipgpu::rt::build_panel_plan(const sipgpu_contraction *list, int n, int nmod, PanelPlan &P,
                                 const PanelOptions &opt) {
    if (nmod != 0 && (nmod < kOzMinModuli || nmod > kOzMaxUsable))
        return fail(SIPGPU_EARG, "split path: nmoduli must be in 6..14");
    const int eb = nmod ? 1 : 8; // bytes per panel element
    // shapes
    std::unordered_map<std::string, int> shape_index;
    std::vector<std::unique_ptr<ContractionShape>> shapes;
    std::vector<int> shape_of(n);
    for (int i = 0; i < n; ++i) {
        const sipgpu_contraction &e = list[i];
        if (!e.l || !e.r || !e.d) return fail(SIPGPU_EARG, "null block pointer in batch");
        if (e.lrank < 1 || e.lrank > SIPGPU_MAX_RANK || e.rrank < 1 ||
            e.rrank > SIPGPU_MAX_RANK || e.drank < 1 || e.drank > SIPGPU_MAX_RANK)
            return fail(SIPGPU_ERANK, "batch entries must be block = block * block, rank 1..6");
        const std::string key =
            shape_key(e.ptrn, e.lrank, e.ldims, e.rrank, e.rdims, e.drank, e.ddims);
        auto it = shape_index.find(key);
        if (it == shape_index.end()) {
            auto sh = std::make_unique<ContractionShape>();
            const int rc = plan_contraction(e.ptrn, e.lrank, e.ldims, e.rrank, e.rdims, e.drank,
                                            e.ddims, *sh, true);
            if (rc == 1 || rc == -1)
                return fail(SIPGPU_EPATTERN, "batch entry " + std::to_string(i) +
                                                 ": contraction pattern rejected");
            if (rc) return fail(rc, "batch entry " + std::to_string(i) + ": plan failed");
            it = shape_index.emplace(key, (int)shapes.size()).first;
            shapes.push_back(std::move(sh));
        }
        shape_of[i] = it->second;
    }
    if (int rc = check_hazards(list, n, shapes, shape_of)) return rc;
    // table pool (per shape, identical tables shared)
    TablePool tp;
    std::vector<int32_t> &pool = tp.pool;
    std::vector<std::array<int, 6>> tabs(shapes.size());
    for (size_t si = 0; si < shapes.size(); ++si) tabs[si] = tp.add_shape_dest_ordered(*shapes[si]);
    auto unit_stride = [&](int off, int len) { return len > 1 && pool[off + 1] - pool[off] == 1; };
    // operand instances: one per (block, k table, x table)
    struct Inst { const double *ptr; int K, X, Kp, ktab, xtab, xfast; };
    std::map<std::tuple<const double *, int, int, int, int>, int> inst_index;
    std::vector<Inst> insts;
    const int kq = kOzChunk / eb; // panel elements per 16-byte chunk
    auto instance = [&](const double *ptr, int si, int side) {
        const ContractionShape &sh = *shapes[si];
        const int ktab = tabs[si][side == 0 ? 0 : 2], xtab = tabs[si][side == 0 ? 1 : 3];
        const int X = side == 0 ? (int)sh.M : (int)sh.N;
        const auto key = std::make_tuple(ptr, ktab, xtab, (int)sh.K, X);
        auto it = inst_index.find(key);
        if (it != inst_index.end()) return it->second;
        Inst in{ptr, (int)sh.K, X, (int)((sh.K + kq - 1) / kq * kq), ktab, xtab,
                unit_stride(xtab, X) ? 1 : 0};
        insts.push_back(in);
        inst_index.emplace(key, (int)insts.size() - 1);
        return (int)insts.size() - 1;
    };
    // result blocks, one per destination (same rules as the FP64 batch):
    // the pairs from its last assigning entry on, summed in list order
    struct Dest {
        OzSub s;
        int M, N;
        std::vector<int> a, b; // left / right instance of each pair
        double cost;
    };
    std::unordered_map<const double *, std::vector<int>> by_dest;
    std::vector<const double *> dest_order;
    for (int i = 0; i < n; ++i) {
        auto &v = by_dest[list[i].d];
        if (v.empty()) dest_order.push_back(list[i].d);
        v.push_back(i);
    }
    std::vector<Dest> dests;
    long long kmax = 0;
    int npairs = 0;
    P.flops = 0.0;
    for (const double *dp : dest_order) {
        const auto &idx = by_dest[dp];
        size_t start = 0;
        for (size_t j = 0; j < idx.size(); ++j)
            if (!list[idx[j]].accumulate) start = j;
        const int si0 = shape_of[idx[start]];
        const ContractionShape &s0 = *shapes[si0];
        Dest d{};
        d.M = (int)s0.M;
        d.N = (int)s0.N;
        d.s.D = const_cast<double *>(dp);
        d.s.mD = tabs[si0][4];
        d.s.nD = tabs[si0][5];
        d.s.beta = list[idx[start]].accumulate ? 1 : 0;
        d.s.nfast = (unit_stride(d.s.nD, d.N) || d.M == 1) ? 1 : 0;
        long long ktot = 0, kptot = 0;
        for (size_t j = start; j < idx.size(); ++j) {
            const int i = idx[j];
            const int si = shape_of[i];
            const ContractionShape &sh = *shapes[si];
            // rows and columns are enumerated in destination order, so entries
            // agree when they split the block's indices into the same row and
            // column sets, whatever order their operands hold them in
            if (sh.M != s0.M || sh.N != s0.N || tabs[si][4] != tabs[si0][4] ||
                tabs[si][5] != tabs[si0][5])
                return fail(SIPGPU_EARG, "entries writing one block disagree on its layout");
            d.a.push_back(instance(list[i].l, si, 0));
            d.b.push_back(instance(list[i].r, si, 1));
            ktot += sh.K;
            kptot += insts[d.a.back()].Kp;
            d.cost += 2.0 * (double)sh.M * (double)sh.N * (double)sh.K;
        }
        // exact int32 accumulation: sum of |a b| <= Kvirtual * 127^2 < 2^31
        if (nmod && kptot + kOzStage > 131072)
            return fail(SIPGPU_ESIZE, "split path: a result block sums more than 131072 k");
        if ((long long)d.M * d.N >= INT_MAX)
            return fail(SIPGPU_ESIZE, "split path: result block too large");
        npairs += (int)d.a.size();
        kmax = std::max(kmax, ktot);
        P.flops += d.cost;
        dests.push_back(std::move(d));
    }
    const int ndest = (int)dests.size();
    // Orientation: the GEMM's output lanes run along its columns, so a block
    // whose unit-stride index is among its rows is computed transposed
    // (D^T = R^T L: the same sums with the operands' roles swapped) and its
    // stores stay contiguous.
    for (Dest &d : dests) {
        const bool mfast = d.M > 1 && unit_stride(d.s.mD, d.M);
        const bool nfast = d.N > 1 && unit_stride(d.s.nD, d.N);
        if (mfast && !nfast) {
            std::swap(d.a, d.b);
            std::swap(d.M, d.N);
            std::swap(d.s.mD, d.s.nD);
            d.s.nfast = 1;
        }
    }

    // Merged tasks.  A row key is a result block's list of left instances,
    // a column key its list of right instances.  Result blocks whose column
    // keys pair with the same set of row keys form a grid (rows x columns)
    // over one virtual K: per pair, the grid's left instances are stacked
    // row after row in one panel and its right ones likewise, and the grid
    // runs as one GEMM whose tiles cross block boundaries (a 450-row block
    // no longer wastes a quarter of its last 128-row tile).  Grids are cut to
    // at most kMaxGrid rows and columns.
    std::map<std::vector<int>, int> rkey, ckey;
    std::vector<std::vector<int>> rvec, cvec; // instance lists by key
    std::vector<int> d_r(ndest), d_c(ndest);
    std::set<std::pair<int, int>> seen;
    for (int i = 0; i < ndest; ++i) {
        auto r = rkey.emplace(dests[i].a, (int)rvec.size());
        if (r.second) rvec.push_back(dests[i].a);
        auto c = ckey.emplace(dests[i].b, (int)cvec.size());
        if (c.second) cvec.push_back(dests[i].b);
        d_r[i] = r.first->second;
        d_c[i] = c.first->second;
        if (!seen.insert({d_r[i], d_c[i]}).second) { // the same product into two blocks
            d_c[i] = (int)cvec.size();
            cvec.push_back(dests[i].b);
        }
    }
    auto rows_ext = [&](int key) { return insts[rvec[key][0]].X; };
    auto cols_ext = [&](int key) { return insts[cvec[key][0]].X; };
    std::vector<std::vector<int>> rows_of(cvec.size());
    std::map<std::pair<int, int>, int> dest_at;
    for (int i = 0; i < ndest; ++i) {
        rows_of[d_c[i]].push_back(d_r[i]);
        dest_at[{d_r[i], d_c[i]}] = i;
    }
    std::map<std::vector<int>, int> grid_of;
    std::vector<std::pair<std::vector<int>, std::vector<int>>> grids; // (row keys, col keys)
    const char *merge_env = std::getenv("SIPGPU_MERGE"); // "0": one task per result block
    const bool merge = opt.merge >= 0 ? opt.merge != 0 : !(merge_env && !std::strcmp(merge_env, "0"));
    if (!merge) {
        for (int i = 0; i < ndest; ++i) grids.push_back({{d_r[i]}, {d_c[i]}});
    } else {
        for (int c = 0; c < (int)cvec.size(); ++c) {
            if (rows_of[c].empty()) continue;
            std::sort(rows_of[c].begin(), rows_of[c].end());
            auto it = grid_of.emplace(rows_of[c], (int)grids.size());
            if (it.second) grids.push_back({rows_of[c], {}});
            grids[it.first->second].second.push_back(c);
        }
    }
    constexpr int kMaxGrid = 32768;
    auto cut = [](const std::vector<int> &keys, const std::function<int(int)> &ext) {
        std::vector<std::vector<int>> out(1);
        long long s = 0;
        for (int k : keys) {
            if (!out.back().empty() && s + ext(k) > kMaxGrid) {
                out.emplace_back();
                s = 0;
            }
            out.back().push_back(k);
            s += ext(k);
        }
        return out;
    };
    std::vector<OzTask> tasks;
    std::vector<double> task_cost;
    std::vector<std::vector<int>> task_rows, task_cols; // row / column keys of each task
    std::vector<OzSub> subs;
    for (const auto &gr : grids)
        for (const auto &rs : cut(gr.first, rows_ext))
            for (const auto &cs : cut(gr.second, cols_ext)) {
                OzTask t{};
                const Dest &d0 = dests[dest_at.at({rs[0], cs[0]})];
                double cost = 0.0;
                for (int r : rs) t.M += rows_ext(r);
                for (int c : cs) t.N += cols_ext(c);
                if ((long long)t.M * t.N >= INT_MAX)
                    return fail(SIPGPU_ESIZE, "split path: result block too large");
                t.nfast = d0.s.nfast;
                t.rmap = t.cmap = -1;
                if (rs.size() == 1 && cs.size() == 1) {
                    t.D = d0.s.D;
                    t.mD = d0.s.mD;
                    t.nD = d0.s.nD;
                    t.beta = d0.s.beta;
                    cost = d0.cost;
                } else {
                    // row map: (row key ordinal, row inside its block) per row
                    t.rmap = (int)pool.size();
                    for (size_t g = 0; g < rs.size(); ++g)
                        for (int m = 0; m < rows_ext(rs[g]); ++m) {
                            pool.push_back((int32_t)g);
                            pool.push_back(m);
                        }
                    t.cmap = (int)pool.size();
                    for (size_t h = 0; h < cs.size(); ++h)
                        for (int m = 0; m < cols_ext(cs[h]); ++m) {
                            pool.push_back((int32_t)h);
                            pool.push_back(m);
                        }
                    t.sub0 = (int)subs.size();
                    t.ncp = (int)cs.size();
                    for (int r : rs)
                        for (int c : cs) {
                            const Dest &d = dests[dest_at.at({r, c})];
                            subs.push_back(d.s);
                            cost += d.cost;
                        }
                }
                tasks.push_back(t);
                task_cost.push_back(cost);
                task_rows.push_back(rs);
                task_cols.push_back(cs);
            }
    const int ntasks = (int)tasks.size();
    if (pool.size() >= (size_t)INT_MAX) return fail(SIPGPU_ESIZE, "batch tables too large");

    // Operand panels (after the plane's zero bytes).  A merged task's stack
    // for one pair holds a copy of each of its keys' instances, row after
    // row, with exponents of its own.  An unstacked operand is converted
    // once and shared by every task that uses it unstacked; on the split
    // path those tasks then share its row scales, so the instances of one
    // task side (all scaled alike: its rows share one scale over all its
    // pairs) are united into components with one exponent slot each.
    std::vector<OzBlock> blocks;
    std::map<std::pair<int, int>, int> shared_copy; // (instance, side) -> block
    long long panel = kOzZero;
    int nexpo = 0;
    std::vector<int> uf(2 * insts.size());
    std::iota(uf.begin(), uf.end(), 0);
    auto find = [&](int x) {
        while (uf[x] != x) x = uf[x] = uf[uf[x]];
        return x;
    };
    for (int t = 0; t < ntasks && nmod; ++t)
        for (int side = 0; side < 2; ++side) {
            const auto &keys = side == 0 ? task_rows[t] : task_cols[t];
            if (keys.size() != 1) continue;
            const auto &v = (side == 0 ? rvec : cvec)[keys[0]];
            for (int ii : v) uf[find(2 * ii + side)] = find(2 * v[0] + side);
        }
    std::unordered_map<int, int> comp_slot;
    auto slot_of = [&](int ii, int side) {
        auto it = comp_slot.emplace(find(2 * ii + side), nexpo);
        if (it.second) nexpo += insts[ii].X;
        return it.first->second;
    };
    auto place = [&](int ii, int exp_off) {
        const Inst &in = insts[ii];
        OzBlock ob{};
        ob.data = in.ptr;
        ob.K = in.K;
        ob.X = in.X;
        ob.Kp = in.Kp;
        ob.ktab = in.ktab;
        ob.xtab = in.xtab;
        ob.xfast = in.xfast;
        ob.exp_off = exp_off;
        ob.panel_off = panel;
        panel += (long long)in.X * in.Kp * eb;
        blocks.push_back(ob);
        return (int)blocks.size() - 1;
    };
    auto align_panel = [&]() { panel = (panel + 255) / 256 * 256; };
    // chunk offsets: bytes in a residue plane, 16-byte units in an FP64 one
    const long long unit = nmod ? 1 : kOzChunk;
    std::vector<OzChunkRef> chunks;
    for (int t = 0; t < ntasks; ++t) {
        const auto &rs = task_rows[t], &cs = task_cols[t];
        if (nmod) {
            for (int side = 0; side < 2; ++side) {
                const auto &keys = side == 0 ? rs : cs;
                int &e = side == 0 ? tasks[t].expA : tasks[t].expB;
                if (keys.size() == 1) {
                    e = slot_of((side == 0 ? rvec : cvec)[keys[0]][0], side);
                } else {
                    e = nexpo;
                    nexpo += side == 0 ? tasks[t].M : tasks[t].N;
                }
            }
        }
        auto stack = [&](int side, int p) {
            const auto &keys = side == 0 ? rs : cs;
            const auto &vec = side == 0 ? rvec : cvec;
            const int exp0 = side == 0 ? tasks[t].expA : tasks[t].expB;
            if (keys.size() == 1) {
                const int ii = vec[keys[0]][p];
                // the FP64 panel of an instance is the same on either side
                const auto key = std::make_pair(ii, nmod ? side : 0);
                auto it = shared_copy.find(key);
                if (it == shared_copy.end()) {
                    it = shared_copy.emplace(key, place(ii, exp0)).first;
                    align_panel();
                }
                return blocks[it->second].panel_off;
            }
            const long long base = panel;
            int off = 0;
            for (int key : keys) {
                const int ii = vec[key][p];
                place(ii, exp0 + off);
                off += insts[ii].X;
            }
            align_panel();
            return base;
        };
        // virtual K: the pairs' panel chunks, padded with chunks of the zero
        // bytes to whole stages (FP64) or whole pairs of stages (the INT8
        // GEMM runs two stages per loop trip)
        tasks[t].chunk0 = (int)chunks.size();
        const int np = (int)rvec[rs[0]].size();
        for (int p = 0; p < np; ++p) {
            const long long la = stack(0, p), lb = stack(1, p);
            const int Kp = insts[rvec[rs[0]][p]].Kp;
            for (int k = 0; k < Kp; k += kq)
                chunks.push_back(OzChunkRef{(uint32_t)((la + (long long)k * eb) / unit),
                                            (uint32_t)((lb + (long long)k * eb) / unit), Kp * eb,
                                            0});
            if (panel >= (1LL << 32) * unit - 65536)
                return fail(SIPGPU_ESIZE, nmod ? "split path: operand panels exceed 4 GiB per modulus"
                                               : "packed path: operand panels exceed 64 GiB");
        }
        while ((chunks.size() - tasks[t].chunk0) % ((nmod ? 2 : 1) * kOzStageChunks))
            chunks.push_back(OzChunkRef{0u, 0u, 0, 0});
        tasks[t].nchunks = (int)chunks.size() - tasks[t].chunk0;
    }
    // conversion work
    std::vector<OzConvTile> conv, rows;
    for (int i = 0; i < (int)blocks.size(); ++i)
        for (int x0 = 0; x0 < blocks[i].X; x0 += kOzTileX) {
            rows.push_back(OzConvTile{i, x0, 0, 0});
            for (int k0 = 0; k0 < blocks[i].Kp; k0 += kOzTileK)
                conv.push_back(OzConvTile{i, x0, k0, 0});
        }
    // result planes, reconstruction work
    long long res_elems = 0;
    std::vector<OzRecItem> rec;
    for (int t = 0; t < ntasks && nmod; ++t) {
        tasks[t].res_off = res_elems; // 16-byte aligned: the reconstruction reads words
        const long long mn = (long long)tasks[t].M * tasks[t].N;
        for (long long e0 = 0; e0 < mn; e0 += kOzRecChunk) rec.push_back(OzRecItem{t, (int)e0});
        res_elems += (mn + 15) / 16 * 16;
    }
    // GEMM tile: 256 x 256 with eight waves of 64 x 128 (two per SIMD, two
    // thirds of the LDS traffic per MFMA of the 128 x 128 tile) when the batch has at
    // least one such tile per CU over all moduli, else 128 x 128.  Measured
    // with 14 moduli: ring 1.98 ms against 2.24 (128) and 2.97 (256, four
    // waves of 128 x 128, one per SIMD); the block-sparse ladder likewise.
    // SIPGPU_OZ_TILE=128|256|257 or the tuner pins it.
    long long big_tiles = 0;
    for (const OzTask &t : tasks)
        big_tiles += (long long)((t.M + kOzBigTile - 1) / kOzBigTile) *
                     ((t.N + kOzBigTile - 1) / kOzBigTile);
    int tile = (nmod && big_tiles * nmod >= 256) ? kOzWideTile : kOzBM;
    if (const char *e = std::getenv("SIPGPU_OZ_TILE")) {
        const int v = std::atoi(e);
        if (nmod && (v == kOzBM || v == kOzBigTile || v == kOzWideTile)) tile = v;
    }
    if (nmod && (opt.tile == kOzBM || opt.tile == kOzBigTile || opt.tile == kOzWideTile))
        tile = opt.tile;
    // the packed FP64 GEMM: 128 x 128 tiles, or 64 x 64 when chosen (small
    // batches cut into K segments write a quarter of the partial sums)
    if (!nmod && opt.tile == kOzSmallTile) tile = kOzSmallTile;
    const int tsize = oz_tile_size(tile);
    // tiles: the same XCD schedule as the FP64 batch
    const char *tail_env = std::getenv("SIPGPU_SPLIT_TAIL"); // "0": one tile per workgroup
    std::vector<int> tms(ntasks), tns(ntasks);
    for (int t = 0; t < ntasks; ++t) {
        tms[t] = (tasks[t].M + tsize - 1) / tsize;
        tns[t] = (tasks[t].N + tsize - 1) / tsize;
    }
    int real_tiles = 0;
    std::vector<std::vector<OzTile>> lists = xcd_lists<OzTile>(tms, tns, task_cost, real_tiles);
    std::vector<std::vector<double>> costs(kXcds);
    for (int x = 0; x < kXcds; ++x)
        for (OzTile &u : lists[x]) {
            u.part = -1;
            u.s0 = 0;
            u.s1 = tasks[u.task].nchunks / kOzStageChunks;
            costs[x].push_back(task_cost[u.task] / ((double)tms[u.task] * tns[u.task]));
        }
    std::vector<OzTile> tiles, fix;
    std::vector<int> wg_first;
    int nslots = 0;
    const bool tail = opt.tail >= 0 ? opt.tail != 0 : !(tail_env && !std::strcmp(tail_env, "0"));
    if (!nmod && tail) {
        // the packed FP64 GEMM runs two workgroups per CU, 32 CUs per XCD;
        // tail = 2, 3, 4 run 32, 16, 8 workgroups per XCD, which cuts a small
        // batch's tiles into fewer K segments: less partial-sum traffic
        // through the fixup pass where the segments are short
        const int slots = 64 >> (std::min(std::max(opt.tail, 1), 4) - 1);
        persistent_schedule(lists, costs, slots, tiles, wg_first, fix, nslots);
    } else {
        tiles = interleave(lists);
    }
    if (tiles.size() >= (size_t)INT_MAX) return fail(SIPGPU_ESIZE, "split path: too many tiles");

    OzConst oc{};
    if (nmod) {
        const int P = oz_constants(nmod, kmax, oc);
        if (P < 8) return fail(SIPGPU_EARG, "split path: too few moduli for this K");
    }

    // the descriptor image
    auto align = [](size_t v) { return (v + 255) & ~(size_t)255; };
    size_t off = 0;
    P.o_blocks = off; off = align(off + blocks.size() * sizeof(OzBlock));
    P.o_tasks = off; off = align(off + tasks.size() * sizeof(OzTask));
    P.o_tiles = off; off = align(off + tiles.size() * sizeof(OzTile));
    P.o_chunks = off; off = align(off + chunks.size() * sizeof(OzChunkRef));
    P.o_tab = off; off = align(off + pool.size() * sizeof(int32_t));
    P.o_conv = off; off = align(off + conv.size() * sizeof(OzConvTile));
    P.o_rows = off; off = align(off + rows.size() * sizeof(OzConvTile));
    P.o_rec = off; off = align(off + rec.size() * sizeof(OzRecItem));
    P.o_subs = off; off = align(off + subs.size() * sizeof(OzSub));
    P.o_fix = off; off = align(off + fix.size() * sizeof(OzTile));
    P.o_wg = off; off = align(off + wg_first.size() * sizeof(int));
    P.o_expo = off; off = align(off + (size_t)nexpo * sizeof(int));
    P.host.assign(off + 256, 0);
    auto put = [&](size_t o, const void *src, size_t bytes) {
        if (bytes) std::memcpy(P.host.data() + o, src, bytes);
    };
    put(P.o_blocks, blocks.data(), blocks.size() * sizeof(OzBlock));
    put(P.o_tasks, tasks.data(), tasks.size() * sizeof(OzTask));
    put(P.o_tiles, tiles.data(), tiles.size() * sizeof(OzTile));
    put(P.o_chunks, chunks.data(), chunks.size() * sizeof(OzChunkRef));
    put(P.o_tab, pool.data(), pool.size() * sizeof(int32_t));
    put(P.o_conv, conv.data(), conv.size() * sizeof(OzConvTile));
    put(P.o_rows, rows.data(), rows.size() * sizeof(OzConvTile));
    put(P.o_rec, rec.data(), rec.size() * sizeof(OzRecItem));
    put(P.o_subs, subs.data(), subs.size() * sizeof(OzSub));
    put(P.o_fix, fix.data(), fix.size() * sizeof(OzTile));
    put(P.o_wg, wg_first.data(), wg_first.size() * sizeof(int));
    P.nmod = nmod;
    P.c = oc;
    P.plane_bytes = (panel + 255) / 256 * 256;
    P.res_plane_bytes = nmod ? (res_elems + 16 + 255) / 256 * 256 : 0; // a word past the last region
    P.nslots = nslots;
    P.tile = tile;
    P.ntiles = (int)tiles.size();
    P.nconv = (int)conv.size();
    P.nrows = (int)rows.size();
    P.nrec = (int)rec.size();
    P.nfix = (int)fix.size();
    P.nwg = wg_first.empty() ? 0 : (int)wg_first.size() - 1;
    P.nexpo = nexpo;
    P.ndest = ndest;
    P.ngroups = ntasks;
    P.npairs = npairs;
    P.real_tiles = real_tiles;
    return 0;
}

void sipgpu::rt::bind_panel_plan(const PanelPlan &P, void *desc, int8_t *planes,
                                 uint8_t *res_planes, double *work, SplitData &sd) {
    unsigned char *base = static_cast<unsigned char *>(desc);
    sd.c = P.c;
    sd.planes = planes;
    sd.plane_bytes = P.plane_bytes;
    sd.res_planes = res_planes;
    sd.res_plane_bytes = P.res_plane_bytes;
    sd.work = work;
    sd.expo = reinterpret_cast<int *>(base + P.o_expo);
    sd.nexpo = P.nexpo;
    sd.row_items = reinterpret_cast<const OzConvTile *>(base + P.o_rows);
    sd.nrow_items = P.nrows;
    const int32_t *dtab = reinterpret_cast<const int32_t *>(base + P.o_tab);
    const OzTask *dtasks = reinterpret_cast<const OzTask *>(base + P.o_tasks);
    const OzSub *dsubs = reinterpret_cast<const OzSub *>(base + P.o_subs);
    OzConvArgs &ca = sd.cargs;
    ca = OzConvArgs{};
    ca.blocks = reinterpret_cast<const OzBlock *>(base + P.o_blocks);
    ca.items = reinterpret_cast<const OzConvTile *>(base + P.o_conv);
    ca.nitems = P.nconv;
    ca.tab = dtab;
    ca.expo = sd.expo;
    ca.planes = planes;
    ca.plane_bytes = P.plane_bytes;
    ca.c = P.c;
    OzGemmArgs &g = sd.gargs;
    g = OzGemmArgs{};
    g.tiles = reinterpret_cast<const OzTile *>(base + P.o_tiles);
    g.tasks = dtasks;
    g.chunks = reinterpret_cast<const OzChunkRef *>(base + P.o_chunks);
    g.planes = planes;
    g.plane_bytes = P.plane_bytes;
    g.res_planes = res_planes;
    g.res_plane_bytes = P.res_plane_bytes;
    g.ntiles = P.ntiles;
    g.tile = P.tile;
    g.tab = dtab;
    g.subs = dsubs;
    g.work = work;
    g.fix = reinterpret_cast<const OzTile *>(base + P.o_fix);
    g.nfix = P.nfix;
    g.wg_first = P.nwg ? reinterpret_cast<const int *>(base + P.o_wg) : nullptr;
    g.nwg = P.nwg;
    g.c = P.c;
    OzRecArgs &r = sd.rargs;
    r = OzRecArgs{};
    r.tasks = dtasks;
    r.items = reinterpret_cast<const OzRecItem *>(base + P.o_rec);
    r.nitems = P.nrec;
    r.tab = dtab;
    r.subs = dsubs;
    r.expo = sd.expo;
    r.res_planes = res_planes;
    r.res_plane_bytes = P.res_plane_bytes;
    r.c = P.c;
}

int sipgpu::rt::execute_panels(SplitData &sd, bool split, int stages, hipStream_t s) {
    if (!split) {
        if (stages & SIPGPU_STAGE_PREPARE) HIP_TRY(oz_pack(sd.cargs, s));
        if (stages & SIPGPU_STAGE_PRODUCT) {
            HIP_TRY(oz_gemm_f64(sd.gargs, s));
            HIP_TRY(oz_fixup_f64(sd.gargs, s));
        }
        return 0;
    }
    if (stages & SIPGPU_STAGE_PREPARE) {
        HIP_TRY(oz_fill_int(sd.expo, sd.nexpo, INT_MIN, s));
        HIP_TRY(oz_row_exponents(sd.cargs, sd.row_items, sd.nrow_items, s));
        HIP_TRY(oz_to_residues(sd.cargs, s));
    }
    if (stages & SIPGPU_STAGE_PRODUCT) HIP_TRY(oz_gemm(sd.gargs, s));
    if (stages & SIPGPU_STAGE_FINISH) HIP_TRY(oz_reconstruct(sd.rargs, s));
    return 0;
}

namespace {

// A panel batch with memory of its own: descriptors, planes (zeroed once:
// the zero bytes at the start of each plane are never written again),
// result planes and the split-tile workspace.
int plan_panels_locked(const sipgpu_contraction *list, int n, int nmod, sipgpu_batch_s &b,
                       const sipgpu::rt::PanelOptions &opt = sipgpu::rt::PanelOptions()) {
    PanelPlan P;
    if (int rc = sipgpu::rt::build_panel_plan(list, n, nmod, P, opt)) return rc;
    SplitData &sd = b.sd;
    HIP_TRY(hipMalloc(&b.dev, P.host.size()));
    HIP_TRY(hipMemcpy(b.dev, P.host.data(), P.host.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&sd.planes, P.plane_total()));
    HIP_TRY(hipMemset(sd.planes, 0, P.plane_total()));
    HIP_TRY(hipDeviceSynchronize());
    if (P.res_total()) HIP_TRY(hipMalloc(&sd.res_planes, P.res_total()));
    if (P.work_total()) HIP_TRY(hipMalloc(&sd.work, P.work_total()));
    sipgpu::rt::bind_panel_plan(P, b.dev, sd.planes, sd.res_planes, sd.work, sd);
    b.split = nmod > 0;
    b.packed = nmod == 0;
    b.ntiles = P.ntiles;
    b.ntasks = P.ndest;
    b.ngroups = P.ngroups;
    b.npairs = P.npairs;
    b.real_tiles = P.real_tiles;
    b.flops = P.flops;
    return 0;
}

void free_batch_memory(sipgpu_batch_s &b) {
    if (b.dev) (void)hipFree(b.dev);
    if (b.ws) (void)hipFree(b.ws);
    b.ws = nullptr;
    b.parts = nullptr;
    b.nparts = 0;
    if (b.sd.planes) (void)hipFree(b.sd.planes);
    if (b.sd.res_planes) (void)hipFree(b.sd.res_planes);
    if (b.sd.work) (void)hipFree(b.sd.work);
    b.dev = nullptr;
    b.sd.planes = nullptr;
    b.sd.res_planes = nullptr;
    b.sd.work = nullptr;
}

int execute_batch(Context &c, sipgpu_batch_s &b, int stages, sipgpu_stream_t s) {
    if (b.ntiles == 0) return 0;
    const hipStream_t hs = resolve(c, s).s;
    if (b.split || b.packed) return sipgpu::rt::execute_panels(b.sd, b.split, stages, hs);
    if (stages & SIPGPU_STAGE_PRODUCT) {
        HIP_TRY(launch_gemm(b.cfg, b.args, b.ntiles, hs));
        if (b.nparts) HIP_TRY(launch_add_parts(b.parts, b.nparts, b.parts_max_n, hs));
    }
    return 0;
}

} // namespace

namespace {

// How a batch is planned: path 0 = table kernel (tile configuration cfg),
// 1 = packed FP64 panels, 2 = split precision; merge / tail / tile as in
// PanelOptions (-1: default).
struct PlanChoice {
    int path = -1, cfg = -1, merge = -1, tail = -1, tile = -1;
    bool operator==(const PlanChoice &o) const {
        return path == o.path && cfg == o.cfg && merge == o.merge && tail == o.tail &&
               tile == o.tile;
    }
};

// Tuned choices by batch signature (guarded by the context lock).
std::unordered_map<uint64_t, PlanChoice> &tuned_table() {
    static auto *t = new std::unordered_map<uint64_t, PlanChoice>();
    return *t;
}

// Signature of a batch: its shapes, patterns, accumulate flags and which
// entries share operand / destination blocks (blocks numbered in order of
// first appearance), and the precision.  Two batches with equal signatures
// plan identically up to the block addresses.
uint64_t batch_signature(const sipgpu_contraction *list, int n, int nmod) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](uint64_t v) {
        for (int i = 0; i < 8; ++i) {
            h ^= (v >> (8 * i)) & 0xff;
            h *= 1099511628211ull;
        }
    };
    std::unordered_map<const double *, int> id;
    auto block = [&](const double *p) {
        auto it = id.emplace(p, (int)id.size());
        return it.first->second;
    };
    mix((uint64_t)nmod);
    mix((uint64_t)n);
    for (int i = 0; i < n; ++i) {
        const sipgpu_contraction &e = list[i];
        mix(((uint64_t)e.lrank << 16) | ((uint64_t)e.rrank << 8) | (uint64_t)e.drank);
        for (int k = 0; k < e.lrank + e.rrank && k < 2 * SIPGPU_MAX_RANK; ++k) mix((uint64_t)(int64_t)e.ptrn[k]);
        for (int k = 0; k < e.lrank; ++k) mix((uint64_t)e.ldims[k]);
        for (int k = 0; k < e.rrank; ++k) mix((uint64_t)e.rdims[k]);
        for (int k = 0; k < e.drank; ++k) mix((uint64_t)e.ddims[k]);
        mix((uint64_t)e.accumulate);
        mix((uint64_t)block(e.l));
        mix((uint64_t)block(e.r));
        mix((uint64_t)block(e.d));
    }
    return h;
}

bool tuned_choice(Context &, const sipgpu_contraction *list, int n, int nmod, PlanChoice &ch) {
    auto &t = tuned_table();
    if (t.empty()) return false;
    auto it = t.find(batch_signature(list, n, nmod));
    if (it == t.end()) return false;
    ch = it->second;
    return true;
}

int plan_choice_locked(Context &c, const sipgpu_contraction *list, int n, int nmod,
                       const PlanChoice &ch, std::unique_ptr<sipgpu_batch_s> &out) {
    auto b = std::make_unique<sipgpu_batch_s>();
    sipgpu::rt::PanelOptions opt;
    opt.merge = ch.merge;
    opt.tail = ch.tail;
    opt.tile = ch.tile;
    int rc;
    if (ch.path == 0)
        rc = plan_batch_locked(c, list, n, *b, ch.cfg, ch.tail == 1);
    else
        rc = plan_panels_locked(list, n, ch.path == 2 ? nmod : 0, *b, opt);
    if (rc) {
        free_batch_memory(*b);
        return rc;
    }
    out = std::move(b);
    return 0;
}

// The plan sipgpu_batch_plan makes without a tuned choice.  Large batches
// run on packed operand panels: the permutations are applied once per
// block (one streaming pass) instead of per tile, and the GEMM runs over
// whole 16-k stages of each task's virtual K.  Small batches (and single
// contractions) keep the table-driven kernel: no extra pass, no panel
// memory.  SIPGPU_FP64_PATH=tables|packed pins it.
int plan_default_locked(Context &c, const sipgpu_contraction *list, int n,
                        std::unique_ptr<sipgpu_batch_s> &out, PlanChoice *chosen) {
    std::unique_ptr<sipgpu_batch_s> b;
    PlanChoice tab;
    tab.path = 0;
    if (int rc = plan_choice_locked(c, list, n, 0, tab, b)) {
        // the table kernel takes one pattern per result block; a batch that
        // writes a block through several patterns runs on panels
        PlanChoice pk;
        pk.path = 1;
        if (plan_choice_locked(c, list, n, 0, pk, b)) return rc;
        if (chosen) *chosen = pk;
        out = std::move(b);
        return 0;
    }
    bool packed = b->ntasks > 0 && b->flops >= 1e9 && b->flops / b->ntasks >= 5e7;
    if (const char *e = std::getenv("SIPGPU_FP64_PATH")) {
        if (!std::strcmp(e, "tables")) packed = false;
        if (!std::strcmp(e, "packed")) packed = b->ntasks > 0;
    }
    PlanChoice used = tab;
    used.cfg = b->cfg;
    if (packed) {
        PlanChoice pk;
        pk.path = 1;
        std::unique_ptr<sipgpu_batch_s> pb;
        if (plan_choice_locked(c, list, n, 0, pk, pb) == 0) {
            free_batch_memory(*b);
            b = std::move(pb);
            used = pk;
        }
    }
    if (chosen) *chosen = used;
    out = std::move(b);
    return 0;
}

// Elements of a block for the elementwise operations: rank 0 is one element
// and a product that is not positive is none, as the reference's loops run
// over 0 .. product - 1 (tensor_dil_omp.F90:172-187).  A product beyond 2^40
// elements is rejected.
int block_elems(int rank, const int *dims, long long &n) {
    n = 1;
    if (rank < 0) return SIPGPU_ERANK;
    if (rank > 0 && !dims) return SIPGPU_EARG;
    for (int i = 0; i < rank; ++i) {
        n *= dims[i];
        if (n == 0) break;
        if (n > (1LL << 40) || n < -(1LL << 40)) return SIPGPU_ESIZE;
    }
    if (n < 0) n = 0;
    return 0;
}

// tensor_block_add_ applies the factor unless it lies within 1e-13 of one,
// except for scalars, which are always multiplied (tensor_dil_omp.F90:417-421).
int add_op(int rank, double factor) {
    return rank == 0 || std::fabs(factor - 1.0) > 1e-13 ? sipgpu::kOpAddScaled : sipgpu::kOpAdd;
}

} // namespace

int sipgpu::rt::ensure_ready() {
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    return ensure_init(c);
}

int sipgpu::rt::device_id() { return ctx().device; }

hipStream_t sipgpu::rt::stream_of(sipgpu_stream_t s) { return s ? s->s : ctx().stream.s; }

extern "C" {

// ---------------------------------------------------------------------------
// Drop-in layer
// ---------------------------------------------------------------------------

long long tensor_size_by_shape_(int *num_dim, int *dims, int *ierr) {
    long long s = 1;
    *ierr = 0;
    if (*num_dim > 0) {
        for (int i = 0; i < *num_dim; ++i) s *= dims[i];
        if (s <= 0) *ierr = 2;
    } else if (*num_dim < 0) {
        *ierr = 1;
        s = 0;
    }
    return s;
}

void get_contraction_ptrn_(int *drank, int *lrank, int *rrank, int *aces_ptrn, int *my_ptrn,
                           int *ierr) {
    *ierr = convert_aces_pattern(*drank, *lrank, *rrank, aces_ptrn, my_ptrn);
}

void tensor_block_copy__(int *nthreads, int *rank, int *dims, int *dim_transp, double *in,
                         double *out, int *ierr) {
    (void)nthreads;
    *ierr = 0;
    if (*rank < 0) { *ierr = *rank; return; }
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    int rc = ensure_init(c);
    if (rc) { *ierr = rc; return; }
    PermutePlan p;
    if (*rank == 0) {
        p.size = 1;
        p.identity = true;
    } else {
        rc = plan_permute(*rank, dims, dim_transp, p);
        if (rc) { *ierr = rc; return; }
    }
    if (p.size <= 0) return;
    const size_t bytes = (size_t)p.size * sizeof(double);
    if ((rc = c.stage_l.reserve(bytes)) || (rc = c.stage_d.reserve(bytes))) { *ierr = rc; return; }
    hipStream_t s = c.stream.s;
    hipError_t e = hipMemcpyAsync(c.stage_l.p, in, bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = launch_permute(p, static_cast<double *>(c.stage_l.p), static_cast<double *>(c.stage_d.p), s);
    if (e == hipSuccess) e = hipMemcpyAsync(out, c.stage_d.p, bytes, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) *ierr = hip_fail(e, "tensor_block_copy__");
}

void tensor_block_contract__(int *nthreads, int *ptrn, double *l, int *lrank, int *ldims,
                             double *r, int *rrank, int *rdims, double *d, int *drank,
                             int *ddims, int *ierr) {
    (void)nthreads;
    *ierr = 0;
    const int lr = *lrank, rr = *rrank, dr = *drank;
    if (lr < 0 || lr > kRefMaxRank || rr < 0 || rr > kRefMaxRank || dr < 0 || dr > kRefMaxRank) {
        *ierr = -1;
        return;
    }
    if (!pattern_valid(ptrn, lr, ldims, rr, rdims, dr, ddims)) { *ierr = 1; return; }
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    int rc = ensure_init(c);
    if (rc) { *ierr = rc; return; }
    const int64_t ls = prod(lr, ldims), rs = prod(rr, rdims), dsz = prod(dr, ddims);
    if (ls <= 0 || rs <= 0 || dsz <= 0) return; // empty block: nothing to do
    if ((rc = c.stage_l.reserve(ls * 8)) || (rc = c.stage_r.reserve(rs * 8)) ||
        (rc = c.stage_d.reserve(dsz * 8))) {
        *ierr = rc;
        return;
    }
    hipStream_t s = c.stream.s;
    hipError_t e = hipMemcpyAsync(c.stage_l.p, l, ls * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(c.stage_r.p, r, rs * 8, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) { *ierr = hip_fail(e, "tensor_block_contract__ upload"); return; }
    rc = contract_device(c, c.stream, ptrn, static_cast<double *>(c.stage_l.p), lr, ldims,
                         static_cast<double *>(c.stage_r.p), rr, rdims,
                         static_cast<double *>(c.stage_d.p), dr, ddims, 0);
    if (rc) { *ierr = rc; return; }
    e = hipMemcpyAsync(d, c.stage_d.p, dsz * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) *ierr = hip_fail(e, "tensor_block_contract__ download");
}


// The remaining kernels of the reference module.  Host blocks are staged
// through the library's device buffers; every call is synchronous.

void tensor_block_init__(int *nthreads, double *tens, int *tens_rank, int *dims, double *sval,
                         int *ierr) {
    (void)nthreads;
    *ierr = 0;
    if (*tens_rank < 0) { *ierr = -1; return; }
    long long n = 0;
    if (int rc = block_elems(*tens_rank, dims, n)) { *ierr = rc; return; }
    if (n == 0) return;
    if (int rc = sipgpu::rt::refuse_in_capture("a host block operation")) { *ierr = rc; return; }
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    int rc = ensure_init(c);
    if (!rc) rc = c.stage_d.reserve(n * 8);
    if (rc) { *ierr = rc; return; }
    hipStream_t s = c.stream.s;
    double *d = static_cast<double *>(c.stage_d.p);
    hipError_t e = launch_block_op(kOpFill, d, d, n, *sval, s);
    if (e == hipSuccess) e = hipMemcpyAsync(tens, d, n * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) *ierr = hip_fail(e, "tensor_block_init__");
}

void tensor_block_scale__(int *nthreads, double *tens, int *tens_rank, int *dims,
                          double *scale_fac, int *ierr) {
    (void)nthreads;
    *ierr = 0;
    if (*tens_rank < 0) { *ierr = -1; return; }
    long long n = 0;
    if (int rc = block_elems(*tens_rank, dims, n)) { *ierr = rc; return; }
    if (n == 0) return;
    if (int rc = sipgpu::rt::refuse_in_capture("a host block operation")) { *ierr = rc; return; }
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    int rc = ensure_init(c);
    if (!rc) rc = c.stage_d.reserve(n * 8);
    if (rc) { *ierr = rc; return; }
    hipStream_t s = c.stream.s;
    double *d = static_cast<double *>(c.stage_d.p);
    hipError_t e = hipMemcpyAsync(d, tens, n * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = launch_block_op(kOpScale, d, d, n, *scale_fac, s);
    if (e == hipSuccess) e = hipMemcpyAsync(tens, d, n * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) *ierr = hip_fail(e, "tensor_block_scale__");
}

double tensor_block_norm2__(int *nthreads, double *tens, int *tens_rank, int *dims, int *ierr) {
    (void)nthreads;
    *ierr = 0;
    if (*tens_rank < 0) { *ierr = -1; return 0.0; }
    long long n = 0;
    if (int rc = block_elems(*tens_rank, dims, n)) { *ierr = rc; return 0.0; }
    if (n == 0) return 0.0;
    if (int rc = sipgpu::rt::refuse_in_capture("a host block operation")) { *ierr = rc; return 0.0; }
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    int rc = ensure_init(c);
    if (!rc) rc = c.stage_l.reserve(n * 8);
    if (!rc) rc = c.stage_r.reserve(8);
    if (rc) { *ierr = rc; return 0.0; }
    hipStream_t s = c.stream.s;
    double *x = static_cast<double *>(c.stage_l.p), *r = static_cast<double *>(c.stage_r.p);
    double v = 0.0;
    hipError_t e = hipMemcpyAsync(x, tens, n * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = launch_block_norm2(x, n, r, c.stream.scratch, s);
    if (e == hipSuccess) e = hipMemcpyAsync(&v, r, 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) { *ierr = hip_fail(e, "tensor_block_norm2__"); return 0.0; }
    return v;
}

} // extern "C"

namespace {

// Shared body of tensor_block_slice__ / tensor_block_insert__: the span of
// the tensor block the slice touches is staged, never the whole block.
int host_slice(bool insert, int rank, double *tens, const int *tens_ext, double *slc,
               const int *slice_ext, const int *ext_beg) {
    if (rank < 0) return 1; // invalid number of dimensions (tensor_dil_omp.F90:325)
    SlicePlan p;
    if (int rc = plan_slice(rank, tens_ext, slice_ext, ext_beg, p))
        return fail(rc, "slice outside the tensor block or too large");
    // an empty slice: the reference's divide_segment_ rejects the empty range
    // (tensor_dil_omp.F90:1248-1254) and its ierr is returned
    if (p.count == 0) return -1;
    if (int rc = sipgpu::rt::refuse_in_capture("a host block operation")) return rc;
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    int rc = ensure_init(c);
    if (!rc) rc = c.stage_l.reserve(p.span * 8);
    if (!rc) rc = c.stage_d.reserve(p.count * 8);
    if (rc) return rc;
    hipStream_t s = c.stream.s;
    double *t = static_cast<double *>(c.stage_l.p), *x = static_cast<double *>(c.stage_d.p);
    const SliceArgs a = make_slice_args(p);
    hipError_t e = hipMemcpyAsync(t, tens + p.base, p.span * 8, hipMemcpyHostToDevice, s);
    if (insert) {
        if (e == hipSuccess) e = hipMemcpyAsync(x, slc, p.count * 8, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = launch_slice(a, true, t, x, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(tens + p.base, t, p.span * 8, hipMemcpyDeviceToHost, s);
    } else {
        if (e == hipSuccess) e = launch_slice(a, false, t, x, s);
        if (e == hipSuccess) e = hipMemcpyAsync(slc, x, p.count * 8, hipMemcpyDeviceToHost, s);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, insert ? "tensor_block_insert__" : "tensor_block_slice__");
    return 0;
}

} // namespace

extern "C" {

void tensor_block_slice__(int *nthreads, int *dim_num, double *tens, int *tens_ext, double *slice,
                          int *slice_ext, int *ext_beg, int *ierr) {
    (void)nthreads;
    *ierr = host_slice(false, *dim_num, tens, tens_ext, slice, slice_ext, ext_beg);
}

void tensor_block_insert__(int *nthreads, int *dim_num, double *tens, int *tens_ext,
                           double *slice, int *slice_ext, int *ext_beg, int *ierr) {
    (void)nthreads;
    *ierr = host_slice(true, *dim_num, tens, tens_ext, slice, slice_ext, ext_beg);
}

void tensor_block_add__(int *nthreads, int *dim_num, int *dim_extents, double *tens0,
                        double *tens1, double *scale_fac, int *ierr) {
    (void)nthreads;
    *ierr = 0;
    if (*dim_num < 0) { *ierr = -1; return; }
    long long n = 0;
    if (int rc = block_elems(*dim_num, dim_extents, n)) { *ierr = rc; return; }
    if (n == 0) return;
    if (int rc = sipgpu::rt::refuse_in_capture("a host block operation")) { *ierr = rc; return; }
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    int rc = ensure_init(c);
    if (!rc) rc = c.stage_d.reserve(n * 8);
    if (!rc) rc = c.stage_l.reserve(n * 8);
    if (rc) { *ierr = rc; return; }
    hipStream_t s = c.stream.s;
    double *d = static_cast<double *>(c.stage_d.p), *x = static_cast<double *>(c.stage_l.p);
    hipError_t e = hipMemcpyAsync(d, tens0, n * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(x, tens1, n * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = launch_block_op(add_op(*dim_num, *scale_fac), d, x, n, *scale_fac, s);
    if (e == hipSuccess) e = hipMemcpyAsync(tens0, d, n * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) *ierr = hip_fail(e, "tensor_block_add__");
}

// ---------------------------------------------------------------------------
// Native layer
// ---------------------------------------------------------------------------

const char *sipgpu_last_error(void) { return g_last_error.c_str(); }
const char *sipgpu_version(void) { return "sipgpu 0.1 gfx950"; }

int sipgpu_init(int device) {
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    if (c.ready && c.device != device) return fail(SIPGPU_EARG, "already initialised on another device");
    return init_locked(c, device);
}

int sipgpu_device_count(int *count) {
    if (!count) return fail(SIPGPU_EARG, "null count");
    hipError_t e = hipGetDeviceCount(count);
    if (e == hipErrorNoDevice) { *count = 0; return 0; }
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
    return 0;
}

int sipgpu_device_synchronize(void) {
    if (int rc = sipgpu::rt::refuse_in_capture("device synchronisation")) return rc;
    HIP_TRY(hipDeviceSynchronize());
    return 0;
}

int sipgpu_malloc(void **dptr, size_t bytes) {
    if (int rc = sipgpu::rt::refuse_in_capture("allocation")) return rc;
    if (!dptr) return fail(SIPGPU_EARG, "null pointer");
    Context &c = ctx();
    {
        std::lock_guard<std::mutex> g(c.mu);
        int rc = ensure_init(c);
        if (rc) return rc;
    }
    hipError_t e = hipMalloc(dptr, bytes ? bytes : 1);
    if (e != hipSuccess) return fail(SIPGPU_ENOMEM, "hipMalloc failed");
    return 0;
}

int sipgpu_free(void *dptr) {
    if (int rc = sipgpu::rt::refuse_in_capture("freeing")) return rc;
    HIP_TRY(hipDeviceSynchronize()); // no launch in flight may still use it
    HIP_TRY(hipFree(dptr));
    return 0;
}

// The plain copies are ordered after all work enqueued so far on every
// stream (the library's streams do not synchronise with the null stream).
int sipgpu_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    if (int rc = sipgpu::rt::refuse_in_capture("a host copy")) return rc;
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return 0;
}
int sipgpu_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    if (int rc = sipgpu::rt::refuse_in_capture("a host copy")) return rc;
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return 0;
}
int sipgpu_memcpy_d2d(void *dst, const void *src, size_t bytes) {
    if (int rc = sipgpu::rt::refuse_in_capture("a blocking copy")) return rc;
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    HIP_TRY(hipDeviceSynchronize());
    return 0;
}
int sipgpu_memset(void *dptr, int value, size_t bytes) {
    if (int rc = sipgpu::rt::refuse_in_capture("a blocking memset")) return rc;
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemset(dptr, value, bytes));
    HIP_TRY(hipDeviceSynchronize());
    return 0;
}

int sipgpu_stream_create(sipgpu_stream_t *s) {
    if (!s) return fail(SIPGPU_EARG, "null stream");
    Context &c = ctx();
    {
        std::lock_guard<std::mutex> g(c.mu);
        int rc = ensure_init(c);
        if (rc) return rc;
    }
    auto *st = new sipgpu_stream_s();
    hipError_t e = hipStreamCreateWithFlags(&st->s, hipStreamNonBlocking);
    if (e != hipSuccess) { delete st; return hip_fail(e, "hipStreamCreate"); }
    int rc = ensure_stream_scratch(*st);
    if (rc) { (void)hipStreamDestroy(st->s); delete st; return rc; }
    *s = st;
    return 0;
}

int sipgpu_stream_destroy(sipgpu_stream_t s) {
    if (!s) return fail(SIPGPU_EARG, "null stream");
    hipError_t e = hipStreamSynchronize(s->s);
    if (s->scratch) (void)hipFree(s->scratch);
    if (s->ws) (void)hipFree(s->ws);
    for (double *w : s->retired) (void)hipFree(w);
    if (e == hipSuccess) e = hipStreamDestroy(s->s);
    delete s;
    if (e != hipSuccess) return hip_fail(e, "hipStreamDestroy");
    return 0;
}

int sipgpu_stream_synchronize(sipgpu_stream_t s) {
    if (int rc = sipgpu::rt::refuse_in_capture("stream synchronisation")) return rc;
    Context &c = ctx();
    if (!s && !c.ready) return 0;
    HIP_TRY(hipStreamSynchronize(resolve(c, s).s));
    return 0;
}

int sipgpu_event_create(sipgpu_event_t *e) {
    if (!e) return fail(SIPGPU_EARG, "null event");
    auto *ev = new sipgpu_event_s();
    hipError_t r = hipEventCreate(&ev->e);
    if (r != hipSuccess) { delete ev; return hip_fail(r, "hipEventCreate"); }
    *e = ev;
    return 0;
}
int sipgpu_event_destroy(sipgpu_event_t e) {
    if (!e) return fail(SIPGPU_EARG, "null event");
    hipError_t r = hipEventDestroy(e->e);
    delete e;
    if (r != hipSuccess) return hip_fail(r, "hipEventDestroy");
    return 0;
}
int sipgpu_event_record(sipgpu_event_t e, sipgpu_stream_t s) {
    if (!e) return fail(SIPGPU_EARG, "null event");
    Context &c = ctx();
    {
        std::lock_guard<std::mutex> g(c.mu);
        int rc = ensure_init(c);
        if (rc) return rc;
    }
    HIP_TRY(hipEventRecord(e->e, resolve(c, s).s));
    return 0;
}
int sipgpu_stream_wait_event(sipgpu_stream_t s, sipgpu_event_t e) {
    if (!e) return fail(SIPGPU_EARG, "null event");
    Context &c = ctx();
    {
        std::lock_guard<std::mutex> g(c.mu);
        int rc = ensure_init(c);
        if (rc) return rc;
    }
    HIP_TRY(hipStreamWaitEvent(resolve(c, s).s, e->e, 0));
    return 0;
}
int sipgpu_event_synchronize(sipgpu_event_t e) {
    if (!e) return fail(SIPGPU_EARG, "null event");
    HIP_TRY(hipEventSynchronize(e->e));
    return 0;
}
int sipgpu_event_elapsed_ms(sipgpu_event_t a, sipgpu_event_t b, float *ms) {
    if (!a || !b || !ms) return fail(SIPGPU_EARG, "null argument");
    HIP_TRY(hipEventElapsedTime(ms, a->e, b->e));
    return 0;
}

int sipgpu_permute(sipgpu_stream_t s, int rank, const int *dims, const int *dim_transp,
                   const double *in, double *out) {
    if (rank < 0) return fail(SIPGPU_ERANK, "negative rank");
    if (!in || !out) return fail(SIPGPU_EARG, "null block pointer");
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    int rc = ensure_init(c);
    if (rc) return rc;
    PermutePlan p;
    if (rank == 0) {
        p.size = 1;
        p.identity = true;
    } else {
        rc = plan_permute(rank, dims, dim_transp, p);
        if (rc) return fail(rc, "invalid permutation");
    }
    HIP_TRY(launch_permute(p, in, out, resolve(c, s).s));
    return 0;
}

struct sipgpu_permute_batch_s {
    void *dev = nullptr; // PermArgs[n] | int first[n + 1] | int entry[grid]
    size_t first_off = 0, entry_off = 0;
    int n = 0, grid = 0;
    int tile_cap = 0, tv_cap = 1; // LDS of the launch
    int items = 16;               // elements per thread
    long long elements = 0;
};

int sipgpu_permute_batch_plan(const sipgpu_permutation *list, int n,
                              sipgpu_permute_batch_t *batch) {
    if (int rc = sipgpu::rt::refuse_in_capture("permute batch planning")) return rc;
    if (!batch || n < 0 || (n > 0 && !list)) return fail(SIPGPU_EARG, "bad permute batch arguments");
    *batch = nullptr;
    std::vector<PermutePlan> plans(n);
    std::vector<const double *> in(n);
    std::vector<double *> out(n);
    long long elements = 0;
    for (int e = 0; e < n; ++e) {
        const sipgpu_permutation &q = list[e];
        if (!q.in || !q.out) return fail(SIPGPU_EARG, "null block pointer in a permute batch");
        if (q.rank < 0 || q.rank > SIPGPU_MAX_RANK)
            return fail(SIPGPU_ERANK, "permute batch: rank outside 0..6");
        if (q.rank == 0) {
            plans[e].size = 1;
            plans[e].identity = true;
        } else if (int rc = plan_permute(q.rank, q.dims, q.dim_transp, plans[e])) {
            return fail(rc, "invalid permutation in a permute batch");
        }
        in[e] = q.in;
        out[e] = q.out;
        elements += plans[e].size;
    }
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    if (int rc = ensure_init(c)) return rc;
    auto round256 = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t first_off = round256((size_t)n * sizeof(PermArgs));
    const size_t entry_off = first_off + round256((size_t)(n + 1) * sizeof(int));
    std::vector<unsigned char> host(entry_off);
    int tile_cap = 0, tv_cap = 1, items = 16;
    const int64_t grid = permute_batch_args(
        plans.data(), in.data(), out.data(), n, reinterpret_cast<PermArgs *>(host.data()),
        reinterpret_cast<int *>(host.data() + first_off), &tile_cap, &tv_cap, &items);
    if (grid >= ((int64_t)1 << 31)) return fail(SIPGPU_ESIZE, "permute batch: too many tiles");
    // the entry of every workgroup
    host.resize(entry_off + (size_t)grid * sizeof(int));
    {
        const int *first = reinterpret_cast<const int *>(host.data() + first_off);
        int *entry = reinterpret_cast<int *>(host.data() + entry_off);
        for (int e = 0; e < n; ++e)
            for (int w = first[e]; w < first[e + 1]; ++w) entry[w] = e;
    }
    auto b = std::make_unique<sipgpu_permute_batch_s>();
    b->first_off = first_off;
    b->entry_off = entry_off;
    b->n = n;
    b->tile_cap = tile_cap;
    b->tv_cap = tv_cap;
    b->items = items;
    b->grid = (int)grid;
    b->elements = elements;
    if (grid > 0) {
        HIP_TRY(hipMalloc(&b->dev, host.size()));
        const hipError_t e = hipMemcpy(b->dev, host.data(), host.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(b->dev);
            return hip_fail(e, "permute batch upload");
        }
    }
    *batch = b.release();
    return 0;
}

int sipgpu_permute_batch_execute(sipgpu_permute_batch_t b, sipgpu_stream_t s) {
    if (!b) return fail(SIPGPU_EARG, "null permute batch");
    if (b->grid == 0) return 0;
    Context &c = ctx();
    const unsigned char *base = static_cast<const unsigned char *>(b->dev);
    HIP_TRY(launch_permute_batch(reinterpret_cast<const PermArgs *>(base),
                                 reinterpret_cast<const int *>(base + b->first_off),
                                 reinterpret_cast<const int *>(base + b->entry_off), b->grid,
                                 b->tile_cap, b->tv_cap, b->items, resolve(c, s).s));
    return 0;
}

int sipgpu_permute_batch_info(sipgpu_permute_batch_t b, long long *elements, int *workgroups) {
    if (!b) return fail(SIPGPU_EARG, "null permute batch");
    if (elements) *elements = b->elements;
    if (workgroups) *workgroups = b->grid;
    return 0;
}

int sipgpu_permute_batch_destroy(sipgpu_permute_batch_t b) {
    if (int rc = sipgpu::rt::refuse_in_capture("destroying a permute batch")) return rc;
    if (!b) return fail(SIPGPU_EARG, "null permute batch");
    // the tables may still be read by a launch in flight on any stream
    const hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return hip_fail(e, "sipgpu_permute_batch_destroy");
    if (b->dev) (void)hipFree(b->dev);
    delete b;
    return 0;
}

int sipgpu_contract(sipgpu_stream_t s, const int *ptrn, const double *l, int lrank,
                    const int *ldims, const double *r, int rrank, const int *rdims, double *d,
                    int drank, const int *ddims, int accumulate) {
    if (!l || !r || !d || (lrank + rrank > 0 && !ptrn)) return fail(SIPGPU_EARG, "null argument");
    if (lrank < 0 || lrank > SIPGPU_MAX_RANK || rrank < 0 || rrank > SIPGPU_MAX_RANK ||
        drank < 0 || drank > SIPGPU_MAX_RANK)
        return fail(SIPGPU_ERANK, "rank outside 0..6");
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    int rc = ensure_init(c);
    if (rc) return rc;
    rc = contract_device(c, resolve(c, s), ptrn, l, lrank, ldims, r, rrank, rdims, d, drank,
                         ddims, accumulate);
    if (rc == 1 || rc == -1) return fail(SIPGPU_EPATTERN, "contraction pattern rejected");
    return rc;
}

int sipgpu_contract_noalloc(sipgpu_stream_t s, const int *ptrn, const double *l, int lrank,
                            const int *ldims, const double *r, int rrank, const int *rdims,
                            double *d, int drank, const int *ddims, int accumulate) {
    if (!l || !r || !d || (lrank + rrank > 0 && !ptrn)) return fail(SIPGPU_EARG, "null argument");
    if (lrank < 0 || lrank > SIPGPU_MAX_RANK || rrank < 0 || rrank > SIPGPU_MAX_RANK ||
        drank < 0 || drank > SIPGPU_MAX_RANK)
        return fail(SIPGPU_ERANK, "rank outside 0..6");
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    // initialisation creates the library's stream and scratch: set-up, done
    // by sipgpu_init (or any other call) beforehand
    if (!c.ready)
        return fail(SIPGPU_EARG, "no-allocation contraction before the library is initialised "
                                 "(call sipgpu_init)");
    if (c.device >= kMaxDevices || !g_scratch_dev[c.device])
        return fail(SIPGPU_EDEVICE, "no static scratch on this device");
    double *base = g_scratch_dev[c.device];
    const FixedScratch fx{base, kNoAllocWsDoubles * sizeof(double), base + kNoAllocWsDoubles};
    int rc = contract_device(c, resolve(c, s), ptrn, l, lrank, ldims, r, rrank, rdims, d, drank,
                             ddims, accumulate, &fx);
    if (rc == 1 || rc == -1) return fail(SIPGPU_EPATTERN, "contraction pattern rejected");
    return rc;
}

int sipgpu_batch_plan(const sipgpu_contraction *list, int n, sipgpu_batch_t *batch) {
    if (int rc = sipgpu::rt::refuse_in_capture("batch planning")) return rc;
    if (!batch || (n > 0 && !list) || n < 0) return fail(SIPGPU_EARG, "bad batch arguments");
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    int rc = ensure_init(c);
    if (rc) return rc;
    std::unique_ptr<sipgpu_batch_s> b;
    PlanChoice ch;
    if (tuned_choice(c, list, n, 0, ch)) {
        rc = plan_choice_locked(c, list, n, 0, ch, b);
        if (rc == 0) {
            *batch = b.release();
            return 0;
        }
    }
    rc = plan_default_locked(c, list, n, b, nullptr);
    if (rc) return rc;
    *batch = b.release();
    return 0;
}

int sipgpu_batch_execute(sipgpu_batch_t b, sipgpu_stream_t s) {
    if (!b) return fail(SIPGPU_EARG, "null batch");
    return execute_batch(ctx(), *b, SIPGPU_STAGE_ALL, s);
}

int sipgpu_batch_execute_stages(sipgpu_batch_t b, int stages, sipgpu_stream_t s) {
    if (!b) return fail(SIPGPU_EARG, "null batch");
    if (stages & ~SIPGPU_STAGE_ALL) return fail(SIPGPU_EARG, "unknown stage bits");
    return execute_batch(ctx(), *b, stages, s);
}

int sipgpu_batch_destroy(sipgpu_batch_t b) {
    if (int rc = sipgpu::rt::refuse_in_capture("destroying a batch")) return rc;
    if (!b) return fail(SIPGPU_EARG, "null batch");
    // the work list may still be read by a launch in flight on any stream
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return hip_fail(e, "sipgpu_batch_destroy");
    free_batch_memory(*b);
    delete b;
    return 0;
}

int sipgpu_batch_plan_split(const sipgpu_contraction *list, int n, int nmoduli,
                            sipgpu_batch_t *batch) {
    if (int rc = sipgpu::rt::refuse_in_capture("batch planning")) return rc;
    if (!batch || (n > 0 && !list) || n < 0) return fail(SIPGPU_EARG, "bad batch arguments");
    if (nmoduli < kOzMinModuli || nmoduli > kOzMaxUsable)
        return fail(SIPGPU_EARG, "split path: nmoduli must be in 6..14");
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    int rc = ensure_init(c);
    if (rc) return rc;
    std::unique_ptr<sipgpu_batch_s> b;
    PlanChoice ch;
    ch.path = 2;
    (void)tuned_choice(c, list, n, nmoduli, ch);
    rc = plan_choice_locked(c, list, n, nmoduli, ch, b);
    if (rc) return rc;
    *batch = b.release();
    return 0;
}

int sipgpu_batch_tune(const sipgpu_contraction *list, int n, int nmoduli, sipgpu_batch_t *batch,
                      sipgpu_tune_result *result) {
    if (int rc = sipgpu::rt::refuse_in_capture("tuning")) return rc;
    if (!batch || (n > 0 && !list) || n < 1) return fail(SIPGPU_EARG, "bad batch arguments");
    if (nmoduli != 0 && (nmoduli < 6 || nmoduli > 14))
        return fail(SIPGPU_EARG, "split path: nmoduli must be in 6..14");
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    int rc = ensure_init(c);
    if (rc) return rc;
    // scratch destinations, one per distinct result block
    std::unordered_map<const double *, size_t> dest_off;
    size_t total = 0;
    for (int i = 0; i < n; ++i) {
        if (!list[i].d) return fail(SIPGPU_EARG, "null block pointer in batch");
        if (dest_off.count(list[i].d)) continue;
        dest_off[list[i].d] = total;
        total += ((size_t)std::max<int64_t>(prod(list[i].drank, list[i].ddims), 1) + 31) / 32 * 32;
    }
    double *scratch = nullptr;
    HIP_TRY(hipMalloc(&scratch, total * sizeof(double)));
    std::vector<sipgpu_contraction> tl(list, list + n);
    for (auto &e : tl) e.d = scratch + dest_off[e.d];
    hipError_t he = hipMemset(scratch, 0, total * sizeof(double));
    std::vector<PlanChoice> cands;
    if (nmoduli == 0) {
        for (int cfg : {0, 1, 2, 3, 5, 6, 11, 12, 14}) {
            PlanChoice ch;
            ch.path = 0;
            ch.cfg = cfg;
            cands.push_back(ch);
        }
        // the table kernel with each result block's pairs cut into groups
        // (a batch whose tiles fill at most half of the CUs)
        for (int cfg : {0, 11}) {
            PlanChoice ch;
            ch.path = 0;
            ch.cfg = cfg;
            ch.tail = 1;
            cands.push_back(ch);
        }
        for (int merge : {1, 0})
            for (int tail : {1, 0, 2, 3, 4}) {
                PlanChoice ch;
                ch.path = 1;
                ch.merge = merge;
                ch.tail = tail;
                cands.push_back(ch);
            }
        for (int merge : {1, 0})
            for (int tail : {1, 2}) {
                PlanChoice ch;
                ch.path = 1;
                ch.merge = merge;
                ch.tail = tail;
                ch.tile = kOzSmallTile;
                cands.push_back(ch);
            }
    } else {
        for (int tile : {kOzBM, kOzBigTile, kOzWideTile})
            for (int merge : {1, 0}) {
                PlanChoice ch;
                ch.path = 2;
                ch.merge = merge;
                ch.tile = tile;
                cands.push_back(ch);
            }
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (he == hipSuccess) he = hipEventCreate(&e0);
    if (he == hipSuccess) he = hipEventCreate(&e1);
    const hipStream_t hs = c.stream.s;
    auto time_plan = [&](sipgpu_batch_s &b, double &ms) -> int {
        if (int r = execute_batch(c, b, SIPGPU_STAGE_ALL, nullptr)) return r;
        HIP_TRY(hipStreamSynchronize(hs));
        ms = 1e300;
        for (int rep = 0; rep < 5; ++rep) {
            HIP_TRY(hipEventRecord(e0, hs));
            if (int r = execute_batch(c, b, SIPGPU_STAGE_ALL, nullptr)) return r;
            HIP_TRY(hipEventRecord(e1, hs));
            HIP_TRY(hipEventSynchronize(e1));
            float t = 0.f;
            HIP_TRY(hipEventElapsedTime(&t, e0, e1));
            ms = std::min(ms, (double)t);
        }
        return 0;
    };
    // the untuned plan
    PlanChoice def;
    double def_ms = 0.0;
    if (he == hipSuccess) {
        std::unique_ptr<sipgpu_batch_s> b;
        int r = 0;
        if (nmoduli == 0) {
            r = plan_default_locked(c, tl.data(), n, b, &def);
        } else {
            def.path = 2;
            r = plan_choice_locked(c, tl.data(), n, nmoduli, def, b);
        }
        if (r == 0) {
            if (time_plan(*b, def_ms)) def_ms = 0.0;
            (void)hipDeviceSynchronize();
            free_batch_memory(*b);
        }
    }
    PlanChoice best;
    double best_ms = 1e300;
    int timed = 0;
    for (const PlanChoice &ch : cands) {
        if (he != hipSuccess) break;
        std::unique_ptr<sipgpu_batch_s> b;
        if (plan_choice_locked(c, tl.data(), n, nmoduli, ch, b)) continue; // not plannable
        double ms = 0.0;
        const int r = time_plan(*b, ms);
        (void)hipDeviceSynchronize();
        free_batch_memory(*b);
        if (r) {
            rc = r;
            break;
        }
        ++timed;
        if (ms < best_ms) {
            best_ms = ms;
            best = ch;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(scratch);
    if (he != hipSuccess) return hip_fail(he, "tuning");
    if (rc) return rc;
    if (!timed) return fail(SIPGPU_EARG, "tuning: no candidate plan could be made");
    // the untuned plan stays unless a candidate beats it by more than the
    // timing noise (2%): a best-of-five minimum over scratch destinations
    // otherwise picks plans that are no faster in the caller's run
    if (def_ms > 0.0 && best_ms > 0.98 * def_ms) {
        best = def;
        best_ms = def_ms;
    }
    tuned_table()[batch_signature(list, n, nmoduli)] = best;
    std::unique_ptr<sipgpu_batch_s> b;
    if ((rc = plan_choice_locked(c, list, n, nmoduli, best, b))) return rc;
    *batch = b.release();
    if (result) {
        result->candidates = timed;
        result->best_ms = best_ms;
        result->default_ms = def_ms;
        result->path = best.path;
        result->config = best.cfg;
        result->merge = best.merge;
        result->tail = best.tail;
        result->tile = best.tile;
    }
    return 0;
}

int sipgpu_batch_plan_choice(const sipgpu_contraction *list, int n, int nmoduli, int path,
                             int config, int merge, int tail, int tile, sipgpu_batch_t *batch) {
    if (int rc = sipgpu::rt::refuse_in_capture("batch planning")) return rc;
    if (!batch || (n > 0 && !list) || n < 0) return fail(SIPGPU_EARG, "bad batch arguments");
    if (path < 0 || path > 2 || (path == 2) != (nmoduli != 0) || config < -1 ||
        config >= kGemmConfigs || tail < -1 || tail > 4 ||
        (path == 2 && tile != -1 && tile != kOzBM && tile != kOzBigTile && tile != kOzWideTile) ||
        (path != 2 && tile != -1 && !(path == 1 && tile == kOzSmallTile)))
        return fail(SIPGPU_EARG, "bad plan choice");
    if (nmoduli != 0 && (nmoduli < 6 || nmoduli > 14))
        return fail(SIPGPU_EARG, "split path: nmoduli must be in 6..14");
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    int rc = ensure_init(c);
    if (rc) return rc;
    PlanChoice ch;
    ch.path = path;
    ch.cfg = config;
    ch.merge = merge;
    ch.tail = tail;
    ch.tile = tile;
    std::unique_ptr<sipgpu_batch_s> b;
    if ((rc = plan_choice_locked(c, list, n, nmoduli, ch, b))) return rc;
    *batch = b.release();
    return 0;
}

int sipgpu_tune_save(const char *path) {
    if (!path) return fail(SIPGPU_EARG, "null path");
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    FILE *f = std::fopen(path, "w");
    if (!f) return fail(SIPGPU_EARG, std::string("cannot write ") + path);
    std::fprintf(f, "# sipgpu tuning table: signature path config merge tail tile\n");
    for (const auto &kv : tuned_table())
        std::fprintf(f, "%016llx %d %d %d %d %d\n", (unsigned long long)kv.first, kv.second.path,
                     kv.second.cfg, kv.second.merge, kv.second.tail, kv.second.tile);
    const bool ok = std::fclose(f) == 0;
    return ok ? 0 : fail(SIPGPU_EARG, std::string("cannot write ") + path);
}

int sipgpu_tune_load(const char *path) {
    if (!path) return fail(SIPGPU_EARG, "null path");
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    FILE *f = std::fopen(path, "r");
    if (!f) return fail(SIPGPU_ENOTFOUND, std::string("cannot read ") + path);
    char line[256];
    int bad = 0;
    while (std::fgets(line, sizeof line, f)) {
        if (line[0] == '#' || line[0] == '\n') continue;
        unsigned long long sig = 0;
        PlanChoice ch;
        if (std::sscanf(line, "%llx %d %d %d %d %d", &sig, &ch.path, &ch.cfg, &ch.merge, &ch.tail,
                        &ch.tile) != 6 ||
            ch.path < 0 || ch.path > 2 || ch.cfg < -1 || ch.cfg >= kGemmConfigs ||
            ch.merge < -1 || ch.merge > 1 || ch.tail < -1 || ch.tail > 4 ||
            (ch.tile != -1 && !(ch.path == 1 && ch.tile == kOzSmallTile) &&
             !(ch.path == 2 && (ch.tile == kOzBM || ch.tile == kOzBigTile ||
                                ch.tile == kOzWideTile)))) {
            ++bad;
            continue;
        }
        tuned_table()[(uint64_t)sig] = ch;
    }
    std::fclose(f);
    return bad ? fail(SIPGPU_EARG, std::string("malformed lines in ") + path) : 0;
}

int sipgpu_tune_clear(void) {
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    tuned_table().clear();
    return 0;
}

int sipgpu_batch_plan_split_tol(const sipgpu_contraction *list, int n, double rel_err,
                                sipgpu_batch_t *batch) {
    if (int rc = sipgpu::rt::refuse_in_capture("batch planning")) return rc;
    if (!batch || (n > 0 && !list) || n < 0 || !(rel_err > 0.0))
        return fail(SIPGPU_EARG, "bad batch arguments");
    // longest summed K over the result blocks (all entries counted: an
    // upper bound, so the width chosen is never too wide)
    std::unordered_map<const double *, long long> ksum;
    long long kmax = 1;
    for (int i = 0; i < n; ++i) {
        const sipgpu_contraction &e = list[i];
        if (e.lrank < 1 || e.lrank > SIPGPU_MAX_RANK || e.rrank < 0 ||
            e.rrank > SIPGPU_MAX_RANK)
            return fail(SIPGPU_ERANK, "batch entries must be block = block * block, rank 1..6");
        long long k = 1;
        for (int d = 0; d < e.lrank; ++d)
            if (e.ptrn[d] < 0) k *= std::max(e.ldims[d], 1);
        kmax = std::max(kmax, ksum[e.d] += k);
    }
    int nmod = 0;
    for (int m = kOzMinModuli; m <= kOzMaxUsable && !nmod; ++m) {
        OzConst oc;
        const int P = oz_constants(m, kmax, oc);
        if (P >= 8 && 16.0 * std::ldexp(1.0, -P) <= rel_err) nmod = m;
    }
    if (!nmod)
        return fail(SIPGPU_EARG, "split path: 14 moduli do not meet this error target");
    return sipgpu_batch_plan_split(list, n, nmod, batch);
}

int sipgpu_batch_split_info(sipgpu_batch_t b, int *nmoduli, int *bits) {
    if (!b) return fail(SIPGPU_EARG, "null batch");
    if (nmoduli) *nmoduli = b->split ? b->sd.c.n : 0;
    if (bits) *bits = b->split ? b->sd.c.P : 53;
    return 0;
}

int sipgpu_batch_kind(sipgpu_batch_t b, int *kind) {
    if (!b || !kind) return fail(SIPGPU_EARG, "null batch");
    *kind = b->split ? SIPGPU_BATCH_SPLIT : b->packed ? SIPGPU_BATCH_PACKED : SIPGPU_BATCH_TABLES;
    return 0;
}

int sipgpu_batch_info(sipgpu_batch_t b, double *flops, int *ntasks, int *ntiles, int *npairs) {
    if (!b) return fail(SIPGPU_EARG, "null batch");
    if (flops) *flops = b->flops;
    if (ntasks) *ntasks = b->ntasks;
    if (ntiles) *ntiles = b->real_tiles;
    if (npairs) *npairs = b->npairs;
    return 0;
}

namespace {
int energy_args(int rank, const int *dims, const double *const *eps, const int *signs,
                EnergyArgs &a) {
    if (rank < 1 || rank > kEnergyMaxRank || !dims || !eps || !signs)
        return fail(SIPGPU_EARG, "energy: rank must be 1..6 with dims, energies and signs");
    a = EnergyArgs{};
    a.rank = rank;
    a.n = 1;
    for (int d = 0; d < rank; ++d) {
        if (dims[d] < 1) return fail(SIPGPU_EARG, "energy: extents must be positive");
        if (signs[d] != 1 && signs[d] != -1) return fail(SIPGPU_EARG, "energy: signs are +1/-1");
        if (!eps[d]) return fail(SIPGPU_EARG, "energy: null energy vector");
        a.dims[d] = dims[d];
        a.eps[d] = eps[d];
        a.sign[d] = signs[d];
        a.n *= dims[d];
    }
    return 0;
}
} // namespace

int sipgpu_energy_denominator(sipgpu_stream_t s, int rank, const int *dims, double *x,
                              const double *const *eps, const int *signs) {
    if (!x) return fail(SIPGPU_EARG, "energy: null block");
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    int rc = ensure_init(c);
    if (rc) return rc;
    EnergyArgs a;
    if ((rc = energy_args(rank, dims, eps, signs, a))) return rc;
    HIP_TRY(launch_denominator_divide(a, x, resolve(c, s).s));
    return 0;
}

int sipgpu_energy_sum(sipgpu_stream_t s, int rank, const int *dims, const double *x,
                      const double *y, const double *const *eps, const int *signs,
                      double factor, double *out) {
    if (!x || !y || !out) return fail(SIPGPU_EARG, "energy: null block or result");
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    int rc = ensure_init(c);
    if (rc) return rc;
    EnergyArgs a;
    if ((rc = energy_args(rank, dims, eps, signs, a))) return rc;
    sipgpu_stream_s &st = resolve(c, s);
    if ((rc = ensure_stream_scratch(st))) return rc;
    HIP_TRY(launch_energy_sum(a, x, y, factor, out, st.scratch, st.s));
    return 0;
}

int sipgpu_energy_sum_parts(sipgpu_stream_t s, int rank, const int *dims, double *t,
                            const double *const *parts, int nparts, const double *const *eps,
                            const int *signs, double factor, double *out) {
    if (!t || !out || (nparts > 0 && !parts)) return fail(SIPGPU_EARG, "energy: null pointer");
    if (nparts < 0 || nparts > kAddManyMax) return fail(SIPGPU_EARG, "energy: 0..8 partial blocks");
    AddManyArgs ps{};
    ps.n = nparts;
    for (int q = 0; q < nparts; ++q) {
        if (!parts[q]) return fail(SIPGPU_EARG, "energy: null partial block");
        ps.p[q] = parts[q];
    }
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    int rc = ensure_init(c);
    if (rc) return rc;
    EnergyArgs a;
    if ((rc = energy_args(rank, dims, eps, signs, a))) return rc;
    sipgpu_stream_s &st = resolve(c, s);
    if ((rc = ensure_stream_scratch(st))) return rc;
    HIP_TRY(launch_energy_sum_parts(a, t, ps, factor, out, st.scratch, st.s));
    return 0;
}

int sipgpu_fill_random(sipgpu_stream_t s, double *out, long long n, unsigned long long seed,
                       unsigned long long stream, int dist, double a, double b) {
    if (n < 0 || (n > 0 && !out) || (dist != 0 && dist != 1))
        return fail(SIPGPU_EARG, "fill_random: bad arguments");
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    int rc = ensure_init(c);
    if (rc) return rc;
    HIP_TRY(launch_fill_random(out, n, seed, stream, dist, a, b, resolve(c, s).s));
    return 0;
}

int sipgpu_block_init(sipgpu_stream_t s, double *t, int rank, const int *dims, double value) {
    long long n = 0;
    if (int rc = block_elems(rank, dims, n)) return fail(rc, "block_init: bad rank or extents");
    if (n == 0) return 0;
    if (!t) return fail(SIPGPU_EARG, "block_init: null block");
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    if (int rc = ensure_init(c)) return rc;
    HIP_TRY(launch_block_op(kOpFill, t, t, n, value, resolve(c, s).s));
    return 0;
}

int sipgpu_block_scale(sipgpu_stream_t s, double *t, int rank, const int *dims, double factor) {
    long long n = 0;
    if (int rc = block_elems(rank, dims, n)) return fail(rc, "block_scale: bad rank or extents");
    if (n == 0) return 0;
    if (!t) return fail(SIPGPU_EARG, "block_scale: null block");
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    if (int rc = ensure_init(c)) return rc;
    HIP_TRY(launch_block_op(kOpScale, t, t, n, factor, resolve(c, s).s));
    return 0;
}

int sipgpu_block_add(sipgpu_stream_t s, double *t0, const double *t1, int rank, const int *dims,
                     double factor) {
    long long n = 0;
    if (int rc = block_elems(rank, dims, n)) return fail(rc, "block_add: bad rank or extents");
    if (n == 0) return 0;
    if (!t0 || !t1) return fail(SIPGPU_EARG, "block_add: null block");
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    if (int rc = ensure_init(c)) return rc;
    HIP_TRY(launch_block_op(add_op(rank, factor), t0, t1, n, factor, resolve(c, s).s));
    return 0;
}

int sipgpu_block_add_many(sipgpu_stream_t s, double *t, const double *const *parts, int nparts,
                          int rank, const int *dims) {
    long long n = 0;
    if (int rc = block_elems(rank, dims, n)) return fail(rc, "block_add_many: bad rank or extents");
    if (nparts < 0 || nparts > kAddManyMax)
        return fail(SIPGPU_EARG, "block_add_many: 0..8 sources");
    if (n == 0 || nparts == 0) return 0;
    if (!t || !parts) return fail(SIPGPU_EARG, "block_add_many: null pointer");
    AddManyArgs a{};
    a.n = nparts;
    for (int q = 0; q < nparts; ++q) {
        if (!parts[q]) return fail(SIPGPU_EARG, "block_add_many: null source");
        a.p[q] = parts[q];
    }
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    if (int rc = ensure_init(c)) return rc;
    HIP_TRY(launch_block_add_many(t, a, n, resolve(c, s).s));
    return 0;
}

int sipgpu_block_norm2(sipgpu_stream_t s, const double *t, int rank, const int *dims,
                       double *result) {
    long long n = 0;
    if (int rc = block_elems(rank, dims, n)) return fail(rc, "block_norm2: bad rank or extents");
    if (!result || (n > 0 && !t)) return fail(SIPGPU_EARG, "block_norm2: null pointer");
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    if (int rc = ensure_init(c)) return rc;
    sipgpu_stream_s &st = resolve(c, s);
    if (int rc = ensure_stream_scratch(st)) return rc;
    HIP_TRY(launch_block_norm2(t, n, result, st.scratch, st.s));
    return 0;
}

} // extern "C"

namespace {

int device_slice(sipgpu_stream_t s, bool insert, int rank, double *tens, const int *tens_ext,
                 double *slc, const int *slice_ext, const int *ext_beg) {
    if (rank < 0) return fail(SIPGPU_ERANK, "slice: negative rank");
    if (rank > 0 && (!tens_ext || !slice_ext || !ext_beg))
        return fail(SIPGPU_EARG, "slice: null extents");
    SlicePlan p;
    if (int rc = plan_slice(rank, tens_ext, slice_ext, ext_beg, p))
        return fail(rc, "slice outside the tensor block or too large");
    if (p.count == 0) return 0;
    if (!tens || !slc) return fail(SIPGPU_EARG, "slice: null block");
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    if (int rc = ensure_init(c)) return rc;
    HIP_TRY(launch_slice(make_slice_args(p), insert, tens + p.base, slc, resolve(c, s).s));
    return 0;
}

} // namespace

extern "C" {

int sipgpu_block_slice(sipgpu_stream_t s, int rank, const double *tens, const int *tens_ext,
                       double *slice, const int *slice_ext, const int *ext_beg) {
    return device_slice(s, false, rank, const_cast<double *>(tens), tens_ext, slice, slice_ext,
                        ext_beg);
}

int sipgpu_block_insert(sipgpu_stream_t s, int rank, double *tens, const int *tens_ext,
                        const double *slice, const int *slice_ext, const int *ext_beg) {
    return device_slice(s, true, rank, tens, tens_ext, const_cast<double *>(slice), slice_ext,
                        ext_beg);
}

int sipgpu_plan_slice(int rank, const int *tens_ext, const int *slice_ext, const int *ext_beg,
                      int *canon_rank, int *ext, long long *tstride, long long *info) {
    if (rank < 0) return fail(SIPGPU_ERANK, "slice: negative rank");
    if (!canon_rank || !ext || !tstride || !info ||
        (rank > 0 && (!tens_ext || !slice_ext || !ext_beg)))
        return fail(SIPGPU_EARG, "null argument");
    SlicePlan p;
    if (int rc = plan_slice(rank, tens_ext, slice_ext, ext_beg, p))
        return fail(rc, "slice outside the tensor block or too large");
    *canon_rank = p.rank;
    for (int i = 0; i < p.rank; ++i) { ext[i] = p.ext[i]; tstride[i] = p.tstride[i]; }
    info[0] = p.base;
    info[1] = p.count;
    info[2] = p.span;
    info[3] = p.tsize;
    return 0;
}

int sipgpu_measure_peaks(double *mfma_f64_tflops, double *copy_gbs, size_t copy_bytes) {
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    int rc = ensure_init(c);
    if (rc) return rc;
    hipStream_t s = c.stream.s;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, c.device));
    if (mfma_f64_tflops) {
        // one to eight workgroups of four waves per CU (one to eight waves per
        // SIMD): the best rate counts, with the clock it ran at
        const int max_blocks = prop.multiProcessorCount * 8, iters = 4000;
        double *out = nullptr;
        long long *ticks = nullptr;
        HIP_TRY(hipMalloc(&out, (size_t)max_blocks * 256 * sizeof(double)));
        HIP_TRY(hipMalloc(&ticks, (size_t)max_blocks * 2 * sizeof(long long)));
        HIP_TRY(hipMemsetAsync(out, 0, (size_t)max_blocks * 256 * sizeof(double), s));
        HIP_TRY(launch_mfma_peak(out, ticks, max_blocks, 10, s)); // warm-up
        int wall_khz = 0;
        HIP_TRY(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, c.device));
        double best = 0.0;
        for (int per_cu : {1, 2, 4, 8}) {
            const int blocks = prop.multiProcessorCount * per_cu;
            HIP_TRY(hipEventRecord(e0, s));
            HIP_TRY(launch_mfma_peak(out, ticks, blocks, iters, s));
            HIP_TRY(hipEventRecord(e1, s));
            HIP_TRY(hipEventSynchronize(e1));
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
            const double flops = (double)blocks * 4 /*waves*/ * iters * 8 * 2.0 * 16 * 16 * 4;
            const double tf = flops / (ms * 1e-3) / 1e12;
            if (tf <= best) continue;
            best = tf;
            std::vector<long long> t((size_t)blocks * 2);
            HIP_TRY(hipMemcpy(t.data(), ticks, t.size() * sizeof(long long),
                              hipMemcpyDeviceToHost));
            double cyc = 0, wall = 0;
            for (int i = 0; i < blocks; ++i) { cyc += (double)t[2 * i]; wall += (double)t[2 * i + 1]; }
            g_last_clock_mhz = wall > 0 ? cyc / wall * wall_khz / 1e3 : 0.0;
        }
        *mfma_f64_tflops = best;
        HIP_TRY(hipFree(out));
        HIP_TRY(hipFree(ticks));
    }
    if (copy_gbs) {
        const size_t n = std::max<size_t>(copy_bytes / sizeof(double), 1 << 20) & ~(size_t)1;
        double *a = nullptr, *b = nullptr;
        HIP_TRY(hipMalloc(&a, n * sizeof(double)));
        HIP_TRY(hipMalloc(&b, n * sizeof(double)));
        HIP_TRY(hipMemsetAsync(a, 0, n * sizeof(double), s));
        HIP_TRY(launch_copy(a, b, n, s)); // warm-up
        float best = 1e30f;
        for (int r = 0; r < 5; ++r) {
            HIP_TRY(hipEventRecord(e0, s));
            HIP_TRY(launch_copy(a, b, n, s));
            HIP_TRY(hipEventRecord(e1, s));
            HIP_TRY(hipEventSynchronize(e1));
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
            best = std::min(best, ms);
        }
        *copy_gbs = 2.0 * n * sizeof(double) / (best * 1e-3) / 1e9;
        HIP_TRY(hipFree(a));
        HIP_TRY(hipFree(b));
    }
    HIP_TRY(hipEventDestroy(e0));
    HIP_TRY(hipEventDestroy(e1));
    return 0;
}

double sipgpu_measured_clock_mhz(void) { return g_last_clock_mhz; }

int sipgpu_measure_lds(double *bytes_per_clock_per_cu) {
    if (!bytes_per_clock_per_cu) return fail(SIPGPU_EARG, "null argument");
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    if (int rc = ensure_init(c)) return rc;
    hipStream_t s = c.stream.s;
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, c.device));
    int wall_khz = 0;
    HIP_TRY(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, c.device));
    const int blocks = prop.multiProcessorCount * 2, iters = 4096;
    int *out = nullptr;
    long long *ticks = nullptr;
    HIP_TRY(hipMalloc(&out, (size_t)blocks * sizeof(int)));
    HIP_TRY(hipMalloc(&ticks, (size_t)blocks * 2 * sizeof(long long)));
    double best = 0.0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = launch_lds_read(8, blocks, out, ticks, 16, s); // warm-up
    for (int waves : {8, 16}) {
        if (e != hipSuccess) break;
        e = hipEventRecord(e0, s);
        if (e == hipSuccess) e = launch_lds_read(waves, blocks, out, ticks, iters, s);
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        std::vector<long long> t((size_t)blocks * 2);
        if (e == hipSuccess)
            e = hipMemcpy(t.data(), ticks, t.size() * sizeof(long long), hipMemcpyDeviceToHost);
        if (e != hipSuccess || ms <= 0.f) break;
        // the shader clock the loops ran at (their clock and wall-clock
        // ticks), then all bytes over the kernel's duration per CU and clock
        double cyc = 0.0, wall = 0.0;
        for (int b = 0; b < blocks; ++b) {
            cyc += (double)t[2 * b];
            wall += (double)t[2 * b + 1];
        }
        const double hz = wall > 0.0 ? cyc / wall * wall_khz * 1e3 : prop.clockRate * 1e3;
        const double bytes = (double)blocks * waves * 64 * iters * 8 * 16;
        best = std::max(best, bytes / (ms * 1e-3) / prop.multiProcessorCount / hz);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(out);
    (void)hipFree(ticks);
    if (e != hipSuccess) return hip_fail(e, "sipgpu_measure_lds");
    *bytes_per_clock_per_cu = best;
    return 0;
}

int sipgpu_plan_tables(const int *ptrn, int lrank, const int *ldims, int rrank, const int *rdims,
                       int drank, const int *ddims, long long *mnk, int *o2n, int *koffL,
                       int *moffL, int *koffR, int *noffR, int *moffD, int *noffD) {
    if (lrank < 0 || lrank > kRefMaxRank || rrank < 0 || rrank > kRefMaxRank || drank < 0 ||
        drank > kRefMaxRank)
        return fail(SIPGPU_ERANK, "rank outside 0..32");
    ContractionShape s;
    const bool tables = koffL || moffL || koffR || noffR || moffD || noffD;
    int rc = plan_contraction(ptrn, lrank, ldims, rrank, rdims, drank, ddims, s, tables);
    if (rc == 1 || rc == -1) return fail(SIPGPU_EPATTERN, "contraction pattern rejected");
    if (rc) return fail(rc, "plan failed");
    if (mnk) { mnk[0] = s.M; mnk[1] = s.N; mnk[2] = s.K; }
    if (o2n) {
        int j = 0;
        for (int i = 1; i <= lrank; ++i) o2n[j++] = s.lo2n[i];
        for (int i = 1; i <= rrank; ++i) o2n[j++] = s.ro2n[i];
        for (int i = 1; i <= drank; ++i) o2n[j++] = s.do2n[i];
    }
    auto put = [](int *dst, const std::vector<int32_t> &t) {
        if (dst) std::copy(t.begin(), t.end(), dst);
    };
    put(koffL, s.koffL); put(moffL, s.moffL); put(koffR, s.koffR);
    put(noffR, s.noffR); put(moffD, s.moffD); put(noffD, s.noffD);
    return 0;
}

int sipgpu_plan_permute(int rank, const int *dims, const int *dim_transp, int *info) {
    if (!info || (rank > 0 && (!dims || !dim_transp))) return fail(SIPGPU_EARG, "null argument");
    PermutePlan p;
    if (rank == 0) {
        p.identity = true;
    } else {
        int rc = plan_permute(rank, dims, dim_transp, p);
        if (rc) return fail(rc, "invalid permutation");
    }
    info[0] = p.rank;
    info[1] = p.nU;
    info[2] = p.row_mode ? 1 : 0;
    info[3] = p.identity ? 1 : 0;
    return 0;
}

} // extern "C"
