// triples.cpp -- host side of the on-chip-W (T) energy (triples.hip).
//
// A plan holds the 5-block tables of the virtual and occupied index
// ranges, the two packed operand panel arrays (PA over row sets
// (x, y1, y2), PB over column sets (x2, x3, y3)), and one energy partial
// per box slot.  sipgpu_triples_pack fills the panels from the T, V1 and V2
// blocks; sipgpu_triples_run computes the partials of a range of box
// slots (one workgroup per box); sipgpu_triples_energy sums all partials in
// slot order into the energy (deterministic).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/sipgpu.h"
#include "internal.h"
#include "kernels.h"

using sipgpu::TriplesBlk;
using sipgpu::rt::fail;
using sipgpu::rt::hip_fail;

struct sipgpu_triples_s {
    int nv = 0, no = 0;
    std::vector<int> vseg, oseg;
    std::vector<TriplesBlk> vblk, oblk;
    int kd = 0, kl = 0, nkp = 0;
    // device memory
    void *tables = nullptr; // TriplesBlk vblk | oblk, ints, pointer tables
    const TriplesBlk *d_vblk = nullptr, *d_oblk = nullptr;
    const int *d_vext = nullptr, *d_oext = nullptr, *d_vseg_of = nullptr, *d_oseg_of = nullptr;
    const int *d_vseg_start = nullptr, *d_oseg_start = nullptr;
    const double *const *d_T = nullptr, *const *d_V1 = nullptr, *const *d_V2 = nullptr;
    const double *e_occ = nullptr, *e_vir = nullptr;
    double *PA = nullptr, *PB = nullptr;
    long long pa_doubles = 0, pb_doubles = 0;
    double *partial = nullptr, *sums = nullptr, *tuple_partial = nullptr;
    long long slots = 0;   // box slots of the super-box order
    int super[6] = {}, nblk[6] = {};
    int max_tuple_boxes = 0;
    bool packed = false;
};

namespace {

void cut_blocks(const std::vector<int> &seg, std::vector<TriplesBlk> &out) {
    int g = 0;
    for (int s = 0; s < (int)seg.size(); ++s) {
        for (int off = 0; off < seg[s]; off += sipgpu::kTriplesBox)
            out.push_back(TriplesBlk{s, off, std::min(sipgpu::kTriplesBox, seg[s] - off), g + off});
        g += seg[s];
    }
}

// first 5-block of each segment, and the count
void seg_blocks(const std::vector<TriplesBlk> &blk, int s, int &first, int &count) {
    first = -1;
    count = 0;
    for (int i = 0; i < (int)blk.size(); ++i)
        if (blk[i].seg == s) {
            if (first < 0) first = i;
            ++count;
        }
}

} // namespace

extern "C" {

int sipgpu_triples_create(int nv, const int *v_seg, int no, const int *o_seg,
                          const double *const *T, const double *const *V1,
                          const double *const *V2, const double *e_occ, const double *e_vir,
                          sipgpu_triples_t *plan) {
    if (int rc = sipgpu::rt::refuse_in_capture("planning the triples")) return rc;
    if (nv < 1 || no < 1 || !v_seg || !o_seg || !T || !V1 || !V2 || !e_occ || !e_vir || !plan)
        return fail(SIPGPU_EARG, "triples: bad arguments");
    if (int rc = sipgpu::rt::ensure_ready()) return rc;
    auto P = std::make_unique<sipgpu_triples_s>();
    P->nv = nv;
    P->no = no;
    P->vseg.assign(v_seg, v_seg + nv);
    P->oseg.assign(o_seg, o_seg + no);
    for (int x : P->vseg)
        if (x < 1) return fail(SIPGPU_EARG, "triples: segment extents must be positive");
    for (int x : P->oseg)
        if (x < 1) return fail(SIPGPU_EARG, "triples: segment extents must be positive");
    cut_blocks(P->vseg, P->vblk);
    cut_blocks(P->oseg, P->oblk);
    for (int x : P->vseg) P->kd += x;
    for (int x : P->oseg) P->kl += x;
    // K padded to whole pairs of k-pairs (the kernel loads two per trip)
    P->nkp = ((P->kd + P->kl + 15) / 16) * 2;
    const long long nvb = (long long)P->vblk.size(), nob = (long long)P->oblk.size();
    const long long per_set = 8LL * P->nkp * 128;
    P->pa_doubles = nvb * nob * nob * per_set;
    P->pb_doubles = nvb * nvb * nob * per_set;
    const int nblk[6] = {(int)nvb, (int)nvb, (int)nvb, (int)nob, (int)nob, (int)nob};
    P->slots = 64;
    for (int d = 0; d < 6; ++d) {
        P->nblk[d] = nblk[d];
        P->super[d] = (nblk[d] + 1) / 2;
        P->slots *= P->super[d];
    }
    int vmax = 0, omax = 0;
    for (int s = 0; s < nv; ++s) vmax = std::max(vmax, (P->vseg[s] + 4) / 5);
    for (int s = 0; s < no; ++s) omax = std::max(omax, (P->oseg[s] + 4) / 5);
    P->max_tuple_boxes = vmax * vmax * vmax * omax * omax * omax;

    // tables: blocks, then int arrays, then pointer tables (8-byte aligned)
    std::vector<int> vseg_of(P->kd), oseg_of(P->kl), vstart(nv), ostart(no);
    for (int s = 0, g = 0; s < nv; g += P->vseg[s], ++s) {
        vstart[s] = g;
        for (int q = 0; q < P->vseg[s]; ++q) vseg_of[g + q] = s;
    }
    for (int s = 0, g = 0; s < no; g += P->oseg[s], ++s) {
        ostart[s] = g;
        for (int q = 0; q < P->oseg[s]; ++q) oseg_of[g + q] = s;
    }
    const size_t nT = (size_t)nv * nv * no * no, nV1 = (size_t)nv * nv * nv * no,
                 nV2 = (size_t)no * nv * no * no;
    std::vector<unsigned char> img;
    auto put = [&](const void *p, size_t bytes) {
        const size_t at = (img.size() + 15) & ~(size_t)15;
        img.resize(at + bytes);
        std::memcpy(img.data() + at, p, bytes);
        return at;
    };
    const size_t o_vblk = put(P->vblk.data(), P->vblk.size() * sizeof(TriplesBlk));
    const size_t o_oblk = put(P->oblk.data(), P->oblk.size() * sizeof(TriplesBlk));
    const size_t o_vext = put(P->vseg.data(), nv * sizeof(int));
    const size_t o_oext = put(P->oseg.data(), no * sizeof(int));
    const size_t o_vso = put(vseg_of.data(), vseg_of.size() * sizeof(int));
    const size_t o_oso = put(oseg_of.data(), oseg_of.size() * sizeof(int));
    const size_t o_vst = put(vstart.data(), nv * sizeof(int));
    const size_t o_ost = put(ostart.data(), no * sizeof(int));
    const size_t o_T = put(T, nT * sizeof(double *));
    const size_t o_V1 = put(V1, nV1 * sizeof(double *));
    const size_t o_V2 = put(V2, nV2 * sizeof(double *));
    for (size_t i = 0; i < nT; ++i)
        if (!T[i]) return fail(SIPGPU_EARG, "triples: null T block");
    for (size_t i = 0; i < nV1; ++i)
        if (!V1[i]) return fail(SIPGPU_EARG, "triples: null V1 block");
    for (size_t i = 0; i < nV2; ++i)
        if (!V2[i]) return fail(SIPGPU_EARG, "triples: null V2 block");

    auto cleanup = [&]() {
        if (P->tables) (void)hipFree(P->tables);
        if (P->PA) (void)hipFree(P->PA);
        if (P->PB) (void)hipFree(P->PB);
        if (P->partial) (void)hipFree(P->partial);
        if (P->sums) (void)hipFree(P->sums);
        if (P->tuple_partial) (void)hipFree(P->tuple_partial);
    };
    hipError_t e = hipMalloc(&P->tables, img.size());
    if (e == hipSuccess)
        e = hipMemcpy(P->tables, img.data(), img.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&P->PA, (size_t)P->pa_doubles * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&P->PB, (size_t)P->pb_doubles * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&P->partial, (size_t)P->slots * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&P->sums, (size_t)sipgpu::kTriplesSums * sizeof(double));
    if (e == hipSuccess)
        e = hipMalloc(&P->tuple_partial, (size_t)P->max_tuple_boxes * sizeof(double));
    if (e != hipSuccess) {
        cleanup();
        (void)hipGetLastError();
        return fail(SIPGPU_ENOMEM, "triples: device memory for the packed operands");
    }
    const unsigned char *base = static_cast<const unsigned char *>(P->tables);
    P->d_vblk = reinterpret_cast<const TriplesBlk *>(base + o_vblk);
    P->d_oblk = reinterpret_cast<const TriplesBlk *>(base + o_oblk);
    P->d_vext = reinterpret_cast<const int *>(base + o_vext);
    P->d_oext = reinterpret_cast<const int *>(base + o_oext);
    P->d_vseg_of = reinterpret_cast<const int *>(base + o_vso);
    P->d_oseg_of = reinterpret_cast<const int *>(base + o_oso);
    P->d_vseg_start = reinterpret_cast<const int *>(base + o_vst);
    P->d_oseg_start = reinterpret_cast<const int *>(base + o_ost);
    P->d_T = reinterpret_cast<const double *const *>(base + o_T);
    P->d_V1 = reinterpret_cast<const double *const *>(base + o_V1);
    P->d_V2 = reinterpret_cast<const double *const *>(base + o_V2);
    P->e_occ = e_occ;
    P->e_vir = e_vir;
    *plan = P.release();
    return 0;
}

int sipgpu_triples_destroy(sipgpu_triples_t P) {
    if (int rc = sipgpu::rt::refuse_in_capture("freeing the triples")) return rc;
    if (!P) return fail(SIPGPU_EARG, "null triples plan");
    SIPGPU_HIP_TRY(hipDeviceSynchronize());
    for (void *p : {(void *)P->tables, (void *)P->PA, (void *)P->PB, (void *)P->partial,
                    (void *)P->sums, (void *)P->tuple_partial})
        if (p) (void)hipFree(p);
    delete P;
    return 0;
}

int sipgpu_triples_info(sipgpu_triples_t P, long long *slots, double *flops_per_box,
                        size_t *panel_bytes) {
    if (!P) return fail(SIPGPU_EARG, "null triples plan");
    if (slots) *slots = P->slots;
    // the MFMA work a box issues: six 128 x 128 tiles over the padded K
    if (flops_per_box) *flops_per_box = 6.0 * 2.0 * 128.0 * 128.0 * (8.0 * P->nkp);
    if (panel_bytes) *panel_bytes = (size_t)(P->pa_doubles + P->pb_doubles) * sizeof(double);
    return 0;
}

int sipgpu_triples_pack(sipgpu_triples_t P, sipgpu_stream_t s) {
    if (!P) return fail(SIPGPU_EARG, "null triples plan");
    const hipStream_t hs = sipgpu::rt::stream_of(s);
    for (int side = 0; side < 2; ++side) {
        sipgpu::TriplesPackArgs a{};
        a.out = side == 0 ? P->PA : P->PB;
        a.total = side == 0 ? P->pa_doubles : P->pb_doubles;
        a.side = side;
        a.nkp = P->nkp;
        a.kd = P->kd;
        a.kl = P->kl;
        a.n1 = side == 0 ? (int)P->oblk.size() : (int)P->vblk.size();
        a.n2 = (int)P->oblk.size();
        a.blk0 = P->d_vblk;
        a.blk1 = side == 0 ? P->d_oblk : P->d_vblk;
        a.blk2 = P->d_oblk;
        a.nv = P->nv;
        a.no = P->no;
        a.vext = P->d_vext;
        a.oext = P->d_oext;
        a.vseg_of = P->d_vseg_of;
        a.oseg_of = P->d_oseg_of;
        a.vseg_start = P->d_vseg_start;
        a.oseg_start = P->d_oseg_start;
        a.T = P->d_T;
        a.V1 = P->d_V1;
        a.V2 = P->d_V2;
        SIPGPU_HIP_TRY(sipgpu::launch_triples_pack(a, hs));
    }
    // partials of slots no run has computed count as zero
    SIPGPU_HIP_TRY(sipgpu::launch_block_op(sipgpu::kOpFill, P->partial, nullptr, P->slots, 0.0, hs));
    P->packed = true;
    return 0;
}

namespace {

sipgpu::TriplesArgs box_args(const sipgpu_triples_s &P) {
    sipgpu::TriplesArgs a{};
    a.PA = P.PA;
    a.PB = P.PB;
    a.nkp = P.nkp;
    a.nvb = (int)P.vblk.size();
    a.nob = (int)P.oblk.size();
    a.vblk = P.d_vblk;
    a.oblk = P.d_oblk;
    a.e_vir = P.e_vir;
    a.e_occ = P.e_occ;
    for (int d = 0; d < 6; ++d) {
        a.super[d] = P.super[d];
        a.nblk[d] = P.nblk[d];
    }
    return a;
}

} // namespace

int sipgpu_triples_run(sipgpu_triples_t P, long long box0, long long nbox, sipgpu_stream_t s) {
    if (!P) return fail(SIPGPU_EARG, "null triples plan");
    if (!P->packed) return fail(SIPGPU_EARG, "triples: pack the operands first");
    if (box0 < 0 || nbox < 0 || box0 > P->slots || nbox > P->slots - box0)
        return fail(SIPGPU_EARG, "triples: box slots outside the plan");
    if (nbox == 0) return 0;
    sipgpu::TriplesArgs a = box_args(*P);
    a.partial = P->partial;
    const hipStream_t hs = sipgpu::rt::stream_of(s);
    // one launch per SIPGPU_TRIPLES_LAUNCH_BOXES slots: a dispatch holds
    // fewer than 2^32 work-items (8.4M boxes of 512), and a launch stays
    // near a second of the whole sum's ~20 s on one GPU
    for (long long b = box0; b < box0 + nbox; b += SIPGPU_TRIPLES_LAUNCH_BOXES) {
        a.box0 = b;
        a.nbox = std::min<long long>(SIPGPU_TRIPLES_LAUNCH_BOXES, box0 + nbox - b);
        a.per_xcd = (a.nbox + 7) / 8;
        SIPGPU_HIP_TRY(sipgpu::launch_triples_boxes(a, 8 * a.per_xcd, hs));
    }
    return 0;
}

int sipgpu_triples_energy(sipgpu_triples_t P, double factor, double *out, sipgpu_stream_t s) {
    if (!P || !out) return fail(SIPGPU_EARG, "triples: bad arguments");
    SIPGPU_HIP_TRY(sipgpu::launch_triples_reduce(P->partial, P->slots, P->sums, factor, out,
                                                 sipgpu::rt::stream_of(s)));
    return 0;
}

int sipgpu_triples_tuple(sipgpu_triples_t P, const int *tuple, double factor, double *w_out,
                         double *out, sipgpu_stream_t s) {
    if (!P || !tuple || !out) return fail(SIPGPU_EARG, "triples: bad arguments");
    if (!P->packed) return fail(SIPGPU_EARG, "triples: pack the operands first");
    sipgpu::TriplesArgs a = box_args(*P);
    a.tuple_mode = 1;
    a.partial = P->tuple_partial;
    a.wout = w_out;
    long long grid = 1;
    for (int d = 0; d < 6; ++d) {
        const bool virt = d < 3;
        const int nseg = virt ? P->nv : P->no;
        if (tuple[d] < 0 || tuple[d] >= nseg) return fail(SIPGPU_EARG, "triples: tuple outside the segments");
        seg_blocks(virt ? P->vblk : P->oblk, tuple[d], a.first[d], a.count[d]);
        a.wdims[d] = (virt ? P->vseg : P->oseg)[tuple[d]];
        grid *= a.count[d];
    }
    const hipStream_t hs = sipgpu::rt::stream_of(s);
    SIPGPU_HIP_TRY(sipgpu::launch_triples_boxes(a, grid, hs));
    SIPGPU_HIP_TRY(sipgpu::launch_triples_reduce(P->tuple_partial, grid, P->sums, factor, out, hs));
    return 0;
}

} // extern "C"
