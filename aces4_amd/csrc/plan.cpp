// plan.cpp -- contraction-pattern logic and index-offset tables (host).
//
// The integer logic restates the reference's pattern handling
// (tensor_dil_omp.F90:84-142 get_contraction_ptrn, :799-859
// determine_index_permutations, :861-896 contr_ptrn_ok): its outputs are
// integer maps and must match the reference bit for bit.  The offset tables
// are this library's own device-side representation of those maps.
#include "plan.h"

#include <algorithm>
#include <cstring>
#include <numeric>

#include "../../include/sipgpu.h"

namespace sipgpu {

int convert_aces_pattern(int drank, int lrank, int rrank, const int *aces,
                         int *my_ptrn) {
    if (drank < 0 || lrank < 0 || rrank < 0) return 1;
    const int m = lrank + rrank, n = drank + m;
    if (n % 2 != 0) return 2;
    if (m == 0) return drank > 0 ? 3 : 0;
    // Positions 1..n stably ordered by label (the reference's merge sort is
    // stable, so equal labels keep their original order).
    std::vector<int> pos(n);
    std::iota(pos.begin(), pos.end(), 1);
    std::stable_sort(pos.begin(), pos.end(),
                     [&](int a, int b) { return aces[a - 1] < aces[b - 1]; });
    for (int i = 0; i + 1 < n; i += 2) {
        const int lab = aces[pos[i] - 1];
        if (lab != aces[pos[i + 1] - 1]) return 4;     // unpaired label
        if (i > 0 && lab == aces[pos[i - 1] - 1]) return 5; // label used 3+ times
    }
    for (int i = 0; i + 1 < n; i += 2) {
        const int j = pos[i], k = pos[i + 1];
        if (j <= drank && k > drank) {
            my_ptrn[k - drank - 1] = j;
        } else if (j > drank && j <= drank + lrank && k > drank + lrank) {
            my_ptrn[j - drank - 1] = -(k - (drank + lrank));
            my_ptrn[k - drank - 1] = -(j - drank);
        } else {
            return 6;
        }
    }
    return 0;
}

bool pattern_valid(const int *ptrn, int lr, const int *ldims, int rr,
                   const int *rdims, int dr, const int *ddims) {
    const int total = dr + lr + rr;
    if (total <= 0) return true;
    std::vector<int> hits(total, 0);
    for (int j = 0; j < lr; ++j) {
        const int p = ptrn[j];
        if (p > 0 && p <= dr) {
            ++hits[p - 1];
            ++hits[dr + j];
            if (ldims[j] != ddims[p - 1]) return false;
        } else if (p < 0 && -p <= rr) {
            ++hits[dr + lr - p - 1];
            if (ldims[j] != rdims[-p - 1]) return false;
        } else {
            return false;
        }
    }
    for (int j = 0; j < rr; ++j) {
        const int p = ptrn[lr + j];
        if (p > 0 && p <= dr) {
            ++hits[p - 1];
            ++hits[dr + lr + j];
            if (rdims[j] != ddims[p - 1]) return false;
        } else if (p < 0 && -p <= lr) {
            ++hits[dr - p - 1];
            if (rdims[j] != ldims[-p - 1]) return false;
        } else {
            return false;
        }
    }
    return std::all_of(hits.begin(), hits.end(), [](int h) { return h == 1; });
}

namespace {

bool is_identity(int n, const int *o2n) {
    for (int j = 1; j <= n; ++j)
        if (o2n[j] != j) return false;
    return true;
}

// Offsets of a column-major walk over `ext` (first fastest), where step i
// moves by `stride[i]` elements.
void build_table(const std::vector<int> &ext, const std::vector<int64_t> &stride,
                 std::vector<int32_t> &out) {
    int64_t n = 1;
    for (int e : ext) n *= e;
    out.assign((size_t)n, 0);
    if (ext.empty()) return;
    std::vector<int> idx(ext.size(), 0);
    int64_t off = 0;
    for (int64_t i = 0; i < n; ++i) {
        out[(size_t)i] = (int32_t)off;
        for (size_t d = 0; d < ext.size(); ++d) {
            if (++idx[d] < ext[d]) { off += stride[d]; break; }
            off -= (int64_t)(ext[d] - 1) * stride[d];
            idx[d] = 0;
        }
    }
}

std::vector<int64_t> colmajor_strides(int rank, const int *dims) {
    std::vector<int64_t> s(rank > 0 ? rank : 0);
    int64_t acc = 1;
    for (int i = 0; i < rank; ++i) { s[i] = acc; acc *= dims[i]; }
    return s;
}

} // namespace

int plan_contraction(const int *ptrn, int lr, const int *ldims, int rr,
                     const int *rdims, int dr, const int *ddims,
                     ContractionShape &c, bool build_tables) {
    if (lr < 0 || lr > kRefMaxRank || rr < 0 || rr > kRefMaxRank || dr < 0 ||
        dr > kRefMaxRank)
        return -1;
    if (!pattern_valid(ptrn, lr, ldims, rr, rdims, dr, ddims)) return 1;
    c = ContractionShape();
    c.lrank = lr; c.rrank = rr; c.drank = dr;

    // Destination: the product matrix holds the left-uncontracted dims then
    // the right-uncontracted dims, each in its operand's order; do2n maps
    // matrix dim j to destination dim do2n[j].
    for (int j = 0; j < dr; ++j) c.dsize *= ddims[j];
    int nmat = 0;
    c.do2n[0] = 1;
    for (int j = 0; j < lr + rr; ++j)
        if (ptrn[j] > 0) c.do2n[++nmat] = ptrn[j];
    c.dtransp = dr > 0 && !is_identity(nmat, c.do2n);

    // Right operand: contracted dims first in their own order, then the
    // uncontracted ones.
    int ncd = 0;
    for (int j = 0; j < rr; ++j) c.rsize *= rdims[j];
    if (rr > 0) {
        c.ro2n[0] = 1;
        for (int j = 0; j < rr; ++j)
            if (ptrn[lr + j] < 0) { c.ro2n[j + 1] = ++ncd; c.K *= rdims[j]; }
        int pos = ncd;
        for (int j = 0; j < rr; ++j)
            if (ptrn[lr + j] > 0) { c.ro2n[j + 1] = ++pos; c.N *= rdims[j]; }
        c.rtransp = !is_identity(rr, c.ro2n);
    }

    // Left operand: contracted dims in the order of their partners in the
    // right operand, then its uncontracted dims in its own order.
    for (int j = 0; j < lr; ++j) c.lsize *= ldims[j];
    int ncl = 0;
    if (lr > 0) {
        c.lo2n[0] = 1;
        std::vector<int> cdims; // left dims (0-based) that are contracted
        for (int j = 0; j < lr; ++j)
            if (ptrn[j] < 0) cdims.push_back(j);
        ncl = (int)cdims.size();
        std::stable_sort(cdims.begin(), cdims.end(),
                         [&](int a, int b) { return -ptrn[a] < -ptrn[b]; });
        for (int i = 0; i < ncl; ++i) c.lo2n[cdims[i] + 1] = i + 1;
        int pos = ncl;
        for (int j = 0; j < lr; ++j)
            if (ptrn[j] > 0) { c.lo2n[j + 1] = ++pos; c.M *= ldims[j]; }
        c.ltransp = !is_identity(lr, c.lo2n);
    }

    const int64_t lim = (int64_t)1 << 31;
    if (c.lsize >= lim || c.rsize >= lim || c.dsize >= lim) return SIPGPU_ESIZE;
    if (!build_tables) return 0;

    // Offset tables: walk each matrix index over the dims it spans, in the
    // order the reference's permuted copies would store them.
    const auto sl = colmajor_strides(lr, ldims);
    const auto sr = colmajor_strides(rr, rdims);
    const auto sd = colmajor_strides(dr, ddims);
    {
        // Contracted dims in the right operand's order: (extent, L stride, R
        // stride).  A right dim's partner is the left dim whose pattern entry
        // points at it (the reference pairs by the left entries, :840-846);
        // pattern_valid guarantees exactly one such left dim.
        struct KDim { int ext; int64_t sL, sR; };
        std::vector<KDim> kd;
        for (int j = 0; j < rr; ++j) {
            if (ptrn[lr + j] >= 0) continue;
            int partner = -1;
            for (int x = 0; x < lr; ++x)
                if (ptrn[x] == -(j + 1)) partner = x;
            if (partner < 0) return 1;
            kd.push_back({rdims[j], sl[partner], sr[j]});
        }
        auto unit = [&](bool left) {
            for (size_t i = 0; i < kd.size(); ++i)
                if (kd[i].ext > 1 && (left ? kd[i].sL : kd[i].sR) == 1) return (int)i;
            return -1;
        };
        int first = unit(true);
        if (first < 0) first = unit(false);
        if (first > 0) std::rotate(kd.begin(), kd.begin() + first, kd.begin() + first + 1);
        std::vector<int> ke;
        std::vector<int64_t> ksl, ksr;
        for (const KDim &k : kd) { ke.push_back(k.ext); ksl.push_back(k.sL); ksr.push_back(k.sR); }
        build_table(ke, ksl, c.koffL);
        build_table(ke, ksr, c.koffR);
        std::vector<int> me, ne;
        std::vector<int64_t> ms, ns;
        for (int j = 0; j < lr; ++j)
            if (ptrn[j] > 0) { me.push_back(ldims[j]); ms.push_back(sl[j]); }
        for (int j = 0; j < rr; ++j)
            if (ptrn[lr + j] > 0) { ne.push_back(rdims[j]); ns.push_back(sr[j]); }
        build_table(me, ms, c.moffL);
        build_table(ne, ns, c.noffR);
    }
    {
        std::vector<int> me, ne;
        std::vector<int64_t> ms, ns;
        int j = 0;
        for (int i = 0; i < lr + rr; ++i) {
            if (ptrn[i] <= 0) continue;
            const int dd = c.do2n[++j] - 1; // destination dim of matrix dim j
            if (i < lr) { me.push_back(ddims[dd]); ms.push_back(sd[dd]); }
            else        { ne.push_back(ddims[dd]); ns.push_back(sd[dd]); }
        }
        build_table(me, ms, c.moffD);
        build_table(ne, ns, c.noffD);
    }
    if (c.koffL.size() != (size_t)c.K || c.koffR.size() != (size_t)c.K ||
        c.moffL.size() != (size_t)c.M || c.noffR.size() != (size_t)c.N ||
        c.moffD.size() != (size_t)c.M || c.noffD.size() != (size_t)c.N)
        return 1; // inconsistent extents: cannot happen after pattern_valid
    c.l_kfast = c.K > 1 && c.koffL[1] == 1;
    c.r_kfast = c.K > 1 && c.koffR[1] == 1;
    const bool d_mfast = c.M > 1 && c.moffD[1] == 1;
    const bool d_nfast = c.N > 1 && c.noffD[1] == 1;
    c.swap = d_mfast && !d_nfast;
    return 0;
}

int plan_permute(int rank, const int *dims, const int *dim_transp,
                 PermutePlan &p) {
    p = PermutePlan();
    if (rank < 0 || rank > kRefMaxRank) return SIPGPU_ERANK;
    std::vector<int> seen(rank + 1, 0);
    for (int i = 1; i <= rank; ++i) {
        const int t = dim_transp[i];
        if (t < 1 || t > rank || seen[t]) return SIPGPU_EPERM;
        seen[t] = 1;
    }
    int64_t size = 1;
    for (int i = 0; i < rank; ++i) {
        if (dims[i] < 0) return SIPGPU_EARG;
        size *= dims[i];
    }
    p.size = size;
    if (size >= ((int64_t)1 << 31)) return SIPGPU_ESIZE;
    if (size <= 1) { p.identity = true; return 0; }

    // Output stride of each input dim: product of the extents placed before
    // it in the output order.
    std::vector<int> n2o(rank + 1);
    for (int i = 1; i <= rank; ++i) n2o[dim_transp[i]] = i;
    std::vector<int64_t> sout(rank);
    int64_t acc = 1;
    for (int q = 1; q <= rank; ++q) { sout[n2o[q] - 1] = acc; acc *= dims[n2o[q] - 1]; }

    // Canonical dims: drop extent-1 dims, merge an input dim into its
    // predecessor when the two stay adjacent and in order in the output.
    std::vector<int64_t> ce, cs;
    for (int i = 0; i < rank; ++i) {
        if (dims[i] == 1) continue;
        if (!ce.empty() && cs.back() * ce.back() == sout[i]) {
            ce.back() *= dims[i];
        } else {
            ce.push_back(dims[i]);
            cs.push_back(sout[i]);
        }
    }
    const int cr = (int)ce.size();
    if (cr <= 1) { p.identity = true; return 0; }
    if (cr > 8) return SIPGPU_ERANK; // needs rank > 8; the SIP stops at 6
    p.rank = cr;
    for (int i = 0; i < cr; ++i) { p.ext[i] = (int)ce[i]; p.sout[i] = cs[i]; }

    int dout = 0; // canonical input dim that is fastest in the output
    for (int i = 0; i < cr; ++i)
        if (cs[i] == 1) dout = i;
    if (dout == 0) {
        p.row_mode = true;
        p.nU = 1;
    } else {
        // Tile rows: leading input dims (contiguous in the input), never the
        // output-fastest dim, grown until a row holds 64 elements.
        int a = 1;
        int64_t su = ce[0];
        while (a < dout && su < 64) su *= ce[a++];
        p.nU = a;
    }
    p.sizeU = 1;
    for (int i = 0; i < p.nU; ++i) p.sizeU *= ce[i];
    std::vector<int> rest;
    for (int i = p.nU; i < cr; ++i) rest.push_back(i);
    std::sort(rest.begin(), rest.end(), [&](int a, int b) { return cs[a] < cs[b]; });
    p.nV = (int)rest.size();
    p.sizeV = 1;
    for (int i = 0; i < p.nV; ++i) { p.vdim[i] = rest[i]; p.sizeV *= ce[rest[i]]; }
    return 0;
}

int plan_slice(int rank, const int *tens_ext, const int *slice_ext, const int *ext_beg,
               SlicePlan &p) {
    p = SlicePlan();
    if (rank < 0 || rank > kRefMaxRank) return SIPGPU_ERANK;
    const int64_t kLimit = int64_t(1) << 31;
    int64_t tsize = 1, count = 1, tstride = 1;
    for (int i = 0; i < rank; ++i) {
        if (ext_beg[i] < 0 || slice_ext[i] < 0 || tens_ext[i] < 0 ||
            int64_t(ext_beg[i]) + slice_ext[i] > tens_ext[i])
            return SIPGPU_EARG;
        tsize *= tens_ext[i];
        count *= slice_ext[i];
        if (tsize >= kLimit || count >= kLimit) return SIPGPU_ESIZE;
    }
    p.tsize = tsize;
    p.count = count;
    if (count == 0) return 0;
    // Canonical dims: element offsets grow with the slice's digits, so
    // element 0 is the lowest tensor offset touched and the last the highest.
    int cr = 0;
    int64_t ce[kRefMaxRank], cs[kRefMaxRank];
    for (int i = 0; i < rank; ++i) {
        p.base += int64_t(ext_beg[i]) * tstride;
        if (slice_ext[i] > 1) {
            if (cr > 0 && cs[cr - 1] * ce[cr - 1] == tstride) {
                ce[cr - 1] *= slice_ext[i]; // continues the previous dim's run
            } else {
                ce[cr] = slice_ext[i];
                cs[cr] = tstride;
                ++cr;
            }
        }
        tstride *= tens_ext[i];
    }
    if (cr > kSliceMaxRank) return SIPGPU_ERANK; // needs rank > 8; the SIP stops at 6
    p.rank = cr;
    int64_t last = 0;
    for (int i = 0; i < cr; ++i) {
        p.ext[i] = (int)ce[i];
        p.tstride[i] = cs[i];
        last += (ce[i] - 1) * cs[i];
    }
    p.span = last + 1;
    return 0;
}

} // namespace sipgpu
