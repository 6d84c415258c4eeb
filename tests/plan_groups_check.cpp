// Host-only check of the table planner's pair groups (no device calls):
// sipgpu::rt::build_batch_image on ladder- and ring-shaped batches, with and
// without the split.  Every pair lands in exactly one GEMM task, the groups
// after the first assign into disjoint workspace copies of their block, the
// add-parts entries cover exactly those copies, the tiles cover every task's
// grid once, and binding a workspace points the tasks and entries into it.
// Block addresses are stand-in integers (the planner only compares them).
#include <cstdio>
#include <cstdlib>
#include <vector>
#include <set>
#include "internal.h"
using namespace sipgpu;

static int fails = 0;
#define CHECK(c) do { if (!(c)) { std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c); ++fails; } } while (0)

static void run(int nv_seg, int seg, int no_seg, int oseg, int split_cus, int accumulate_mix) {
    int dr = 4, lr = 4, rr = 4;
    int aces[12] = {1, 2, 3, 4, 1, 2, 5, 6, 5, 6, 3, 4};
    int ptrn[8], ierr = 0;
    get_contraction_ptrn_(&dr, &lr, &rr, aces, ptrn, &ierr);
    CHECK(ierr == 0);
    std::vector<sipgpu_contraction> list;
    uintptr_t base = 1 << 20;
    auto ptr = [&](int id) { return (double *)(base + (uintptr_t)id * 67108864); };
    // D(a,b,i,j) += V(a,b,c,d) T(c,d,i,j)
    for (int a = 0; a < nv_seg; ++a) for (int b = 0; b < nv_seg; ++b)
    for (int i = 0; i < no_seg; ++i) for (int j = 0; j < no_seg; ++j) {
        int did = 100000 + ((a * nv_seg + b) * no_seg + i) * no_seg + j;
        int q = 0;
        for (int c = 0; c < nv_seg; ++c) for (int d = 0; d < nv_seg; ++d) {
            sipgpu_contraction e{};
            e.l = ptr(((a * nv_seg + b) * nv_seg + c) * nv_seg + d);
            e.r = ptr(50000 + ((c * nv_seg + d) * no_seg + i) * no_seg + j);
            e.d = ptr(did);
            e.lrank = e.rrank = e.drank = 4;
            int ld[4] = {seg, seg, seg, seg}, rd[4] = {seg, seg, oseg, oseg}, ddm[4] = {seg, seg, oseg, oseg};
            for (int k = 0; k < 4; ++k) { e.ldims[k] = ld[k]; e.rdims[k] = rd[k]; e.ddims[k] = ddm[k]; }
            for (int k = 0; k < 8; ++k) e.ptrn[k] = ptrn[k];
            e.accumulate = accumulate_mix ? (q == 2 ? 0 : 1) : 1;
            list.push_back(e);
            ++q;
        }
    }
    sipgpu::rt::BatchImage img;
    int rc = sipgpu::rt::build_batch_image(list.data(), (int)list.size(), img, 0, split_cus);
    CHECK(rc == 0);
    const GemmTask *tasks = (const GemmTask *)(img.host.data() + img.o_tasks);
    const GemmTile *tiles = (const GemmTile *)(img.host.data() + img.o_tiles);
    const AddPartsEntry *parts = (const AddPartsEntry *)(img.host.data() + img.o_parts);
    const int nsub = (int)img.ws_task.size();
    const int ntot = img.ntasks + nsub;
    std::printf("blocks %d pairs %d subtasks %d parts %d ws_doubles %lld tiles %d\n", img.ntasks,
                img.npairs, nsub, img.nparts, img.ws_doubles, img.ntiles);
    // every pair index in exactly one task, groups contiguous
    std::vector<int> owner(img.npairs, -1);
    for (int t = 0; t < ntot; ++t)
        for (int p = tasks[t].first; p < tasks[t].first + tasks[t].npairs; ++p) {
            CHECK(p >= 0 && p < img.npairs);
            CHECK(owner[p] < 0);
            owner[p] = t;
        }
    for (int p = 0; p < img.npairs; ++p) CHECK(owner[p] >= 0);
    // sub-tasks: beta 0, null D, offsets disjoint; parts cover them
    std::set<long long> offs;
    for (int i = 0; i < nsub; ++i) {
        const GemmTask &s = tasks[img.ws_task[i]];
        CHECK(img.ws_task[i] >= img.ntasks);
        CHECK(s.beta == 0 && s.D == nullptr);
        CHECK(img.ws_off[i] + (long long)s.M * s.N <= img.ws_doubles);
        CHECK(offs.insert(img.ws_off[i]).second);
    }
    long long covered = 0;
    for (int i = 0; i < img.nparts; ++i) {
        CHECK(parts[i].p == nullptr && parts[i].nparts >= 1);
        CHECK(img.parts_off[i] + parts[i].nparts * parts[i].n <= img.ws_doubles);
        covered += parts[i].nparts * parts[i].n;
        // the entry's block is a real block task's destination
        bool found = false;
        for (int t = 0; t < img.ntasks; ++t) if (tasks[t].D == parts[i].t) { found = true; CHECK((long long)tasks[t].M * tasks[t].N == parts[i].n); }
        CHECK(found);
    }
    CHECK(covered == img.ws_doubles);
    // tiles cover every task's grid exactly once
    const int bm = 64, bn = 64;
    std::vector<std::vector<int>> seen(ntot);
    for (int t = 0; t < ntot; ++t) seen[t].assign(((tasks[t].M + bm - 1) / bm) * ((tasks[t].N + bn - 1) / bn), 0);
    for (int i = 0; i < img.ntiles; ++i) {
        if (tiles[i].task < 0) continue;
        CHECK(tiles[i].task < ntot);
        const GemmTask &t = tasks[tiles[i].task];
        const int tm = (t.M + bm - 1) / bm;
        CHECK(tiles[i].mt < tm && tiles[i].nt < (t.N + bn - 1) / bn);
        seen[tiles[i].task][tiles[i].nt * tm + tiles[i].mt]++;
    }
    for (int t = 0; t < ntot; ++t) for (int v : seen[t]) CHECK(v == 1);
    // binding
    double *ws = (double *)(uintptr_t)0x700000000000ull;
    img.bind_workspace(ws);
    for (int i = 0; i < nsub; ++i) CHECK(tasks[img.ws_task[i]].D == ws + img.ws_off[i]);
    for (int i = 0; i < img.nparts; ++i) CHECK(parts[i].p == ws + img.parts_off[i]);
}

int main() {
    run(4, 10, 1, 10, 256, 0);   // the ladder: 16 blocks x 16 pairs
    run(4, 10, 1, 10, 256, 1);   // with an assigning entry in each block
    run(4, 10, 1, 10, 0, 0);     // split off
    run(3, 7, 2, 5, 256, 0);
    run(2, 40, 1, 20, 256, 0);   // large blocks: no split
    std::printf(fails ? "FAILED %d\n" : "all ok\n", fails);
    return fails != 0;
}
