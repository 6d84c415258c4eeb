/*
 * oracle.c -- CPU restatement of the SIP contract / permute kernels.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Reference: the Fortran module
 * src/sip/tensor_algebra/tensor_dil_omp.F90.  Each function below names the
 * lines it follows.  It is a restatement, not a translation: the
 * reference's cache blocking, OpenMP splitting and loop-priority machinery
 * only change speed, never the values, so they are left out.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

/* tensor_dil_omp.F90:66-82 */
long long orc_tensor_size_by_shape(int num_dim, const int *dims, int *ierr) {
    long long s = 1;
    *ierr = 0;
    if (num_dim > 0) {
        for (int i = 0; i < num_dim; ++i) s *= dims[i];
        if (s <= 0) *ierr = 2;
    } else if (num_dim < 0) {
        *ierr = 1;
        s = 0;
    }
    return s;
}

/* tensor_dil_omp.F90:1183-1236 (stable merge sort by key, sign in trn[0]).
 * A stable sort has a unique result, so insertion sort gives the same
 * permutation; the sign is tracked as the parity of adjacent swaps. */
void orc_merge_sort_key_int(int ni, const int *key, int *trn) {
    for (int i = 2; i <= ni; ++i) {
        int item = trn[i];
        int j = i - 1;
        while (j >= 1 && key[trn[j] - 1] > key[item - 1]) {
            trn[j + 1] = trn[j];
            trn[0] = -trn[0];
            --j;
        }
        trn[j + 1] = item;
    }
}

/* tensor_dil_omp.F90:84-142 */
void orc_get_contraction_ptrn(int drank, int lrank, int rrank,
                              const int *aces_ptrn, int *my_ptrn, int *ierr) {
    *ierr = 0;
    if (drank < 0 || lrank < 0 || rrank < 0) { *ierr = 1; return; }
    int m = lrank + rrank, n = drank + m;
    if (n % 2 != 0) { *ierr = 2; return; }
    if (m == 0) { if (drank > 0) *ierr = 3; return; }
    int *trn = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    trn[0] = 1;
    for (int i = 1; i <= n; ++i) trn[i] = i;
    orc_merge_sort_key_int(n, aces_ptrn, trn);
    /* Labels must come in pairs, never three of a kind. */
    for (int i = 1; i <= n - 1; i += 2) {
        if (aces_ptrn[trn[i] - 1] == aces_ptrn[trn[i + 1] - 1]) {
            if (i > 1 && aces_ptrn[trn[i] - 1] == aces_ptrn[trn[i - 1] - 1]) {
                *ierr = 5; free(trn); return;
            }
        } else {
            *ierr = 4; free(trn); return;
        }
    }
    for (int i = 1; i <= n - 1; i += 2) {
        int j = trn[i], k = trn[i + 1]; /* j < k: the sort is stable */
        if (j <= drank && k > drank) {
            my_ptrn[k - drank - 1] = j; /* uncontracted: goes to dest dim j */
        } else if (j > drank && j <= drank + lrank && k > drank + lrank) {
            my_ptrn[j - drank - 1] = -(k - (drank + lrank));
            my_ptrn[k - drank - 1] = -(j - drank);
        } else {
            *ierr = 6; free(trn); return;
        }
    }
    free(trn);
}

/* tensor_dil_omp.F90:861-896 */
int orc_contr_ptrn_ok(const int *ptrn, int lr, const int *ldims, int rr,
                      const int *rdims, int dr, const int *ddims) {
    int jl = dr + lr + rr;
    if (jl <= 0) return 1;
    int *bus = (int *)calloc((size_t)jl, sizeof(int));
    int ok = 1;
    for (int j0 = 1; j0 <= lr && ok; ++j0) {
        int j1 = ptrn[j0 - 1];
        if (j1 > 0 && j1 <= dr) {
            bus[j1 - 1]++; bus[dr + j0 - 1]++;
            if (ldims[j0 - 1] != ddims[j1 - 1]) ok = 0;
        } else if (j1 < 0 && -j1 <= rr) {
            bus[dr + lr - j1 - 1]++;
            if (ldims[j0 - 1] != rdims[-j1 - 1]) ok = 0;
        } else {
            ok = 0;
        }
    }
    for (int j0 = lr + 1; j0 <= lr + rr && ok; ++j0) {
        int j1 = ptrn[j0 - 1];
        if (j1 > 0 && j1 <= dr) {
            bus[j1 - 1]++; bus[dr + j0 - 1]++;
            if (rdims[j0 - lr - 1] != ddims[j1 - 1]) ok = 0;
        } else if (j1 < 0 && -j1 <= lr) {
            bus[dr - j1 - 1]++;
            if (rdims[j0 - lr - 1] != ldims[-j1 - 1]) ok = 0;
        } else {
            ok = 0;
        }
    }
    for (int j0 = 0; j0 < jl && ok; ++j0)
        if (bus[j0] != 1) ok = 0;
    free(bus);
    return ok;
}

static int perm_trivial(int ni, const int *trn) {
    for (int j = 1; j <= ni; ++j)
        if (trn[j] != j) return 0;
    return 1;
}

/* tensor_dil_omp.F90:799-859 */
void orc_determine_index_permutations(const int *ptrn, int lr, const int *ldims,
                                      int rr, const int *rdims, int dr,
                                      const int *ddims, orc_perms *p) {
    memset(p, 0, sizeof(*p));
    p->lld = p->lrd = p->lcd = 1;
    /* Destination: the GEMM result holds left-uncontracted then
     * right-uncontracted dims; do2n maps them to destination dims. */
    p->dsize = 1;
    if (dr > 0) {
        for (int j = 0; j < dr; ++j) p->dsize *= ddims[j];
        p->do2n[0] = 1;
        int j1 = 0;
        for (int j0 = 1; j0 <= lr + rr; ++j0)
            if (ptrn[j0 - 1] > 0) p->do2n[++j1] = ptrn[j0 - 1];
        p->dtransp = !perm_trivial(j1, p->do2n);
    }
    /* Right operand: contracted dims first (original order), then the rest. */
    p->rsize = 1;
    if (rr > 0) {
        for (int j = 0; j < rr; ++j) p->rsize *= rdims[j];
        p->ro2n[0] = 1;
        int j1 = 0;
        for (int j0 = 1; j0 <= rr; ++j0)
            if (ptrn[lr + j0 - 1] < 0) { p->ro2n[j0] = ++j1; p->lcd *= rdims[j0 - 1]; }
        p->ncd = j1;
        for (int j0 = 1; j0 <= rr; ++j0)
            if (ptrn[lr + j0 - 1] > 0) { p->ro2n[j0] = ++j1; p->lrd *= rdims[j0 - 1]; }
        p->nru = j1 - p->ncd;
        p->rtransp = !perm_trivial(j1, p->ro2n);
    }
    /* Left operand: contracted dims sorted into the right operand's order,
     * then the uncontracted dims in their own order. */
    p->lsize = 1;
    if (lr > 0) {
        int jkey[ORC_MAX_RANK], jtrn1[ORC_MAX_RANK + 1], jtrn0[ORC_MAX_RANK + 1];
        for (int j = 0; j < lr; ++j) p->lsize *= ldims[j];
        p->lo2n[0] = 1;
        int j1 = 0;
        for (int j0 = 1; j0 <= lr; ++j0)
            if (ptrn[j0 - 1] < 0) { ++j1; jtrn1[j1] = j0; jkey[j1 - 1] = -ptrn[j0 - 1]; }
        p->ncd = j1;
        jtrn0[0] = 1;
        for (int jj = 1; jj <= j1; ++jj) jtrn0[jj] = jj;
        orc_merge_sort_key_int(j1, jkey, jtrn0);
        for (int j0 = 1; j0 <= j1; ++j0) p->lo2n[jtrn1[jtrn0[j0]]] = j0;
        for (int j0 = 1; j0 <= lr; ++j0)
            if (ptrn[j0 - 1] > 0) { p->lo2n[j0] = ++j1; p->lld *= ldims[j0 - 1]; }
        p->nlu = j1 - p->ncd;
        p->ltransp = !perm_trivial(j1, p->lo2n);
    }
}

/* tensor_dil_omp.F90:438-660 (semantics only): old dim i goes to new
 * position dim_transp[i]; the output is column-major in the new order. */
void orc_tensor_block_copy(int rank, const int *dims, const int *dim_transp,
                           const double *in, double *out, int *ierr) {
    *ierr = 0;
    if (rank < 0) { *ierr = rank; return; }
    if (rank == 0) { out[0] = in[0]; return; }
    long long n2o_ext[ORC_MAX_RANK];
    int n2o[ORC_MAX_RANK + 1];
    for (int i = 1; i <= rank; ++i) n2o[dim_transp[i]] = i;
    long long bases_out[ORC_MAX_RANK + 1];
    long long bs = 1;
    for (int i = 1; i <= rank; ++i) {
        n2o_ext[i - 1] = dims[n2o[i] - 1];
        bases_out[n2o[i]] = bs;
        bs *= n2o_ext[i - 1];
    }
    long long idx[ORC_MAX_RANK];
    for (int i = 0; i < rank; ++i) idx[i] = 0;
    long long lout = 0;
    for (long long lin = 0; lin < bs; ++lin) {
        out[lout] = in[lin];
        for (int i = 0; i < rank; ++i) { /* odometer over input dims */
            if (++idx[i] < dims[i]) { lout += bases_out[i + 1]; break; }
            lout -= (idx[i] - 1) * bases_out[i + 1];
            idx[i] = 0;
        }
    }
}

/* The reference's GEMM call, dgemm('T','N',M,N,K,1,L,K,R,K,1,D,M) on a
 * zeroed D (tensor_dil_omp.F90:752, :762), with sequential-k summation. */
void orc_gemm_tn(long long m, long long n, long long k, const double *l,
                 const double *r, double *d) {
    for (long long j = 0; j < n; ++j) {
        const double *rc = r + j * k;
        for (long long i = 0; i < m; ++i) {
            const double *lc = l + i * k;
            double v = 0.0;
            for (long long c = 0; c < k; ++c) v += lc[c] * rc[c];
            d[j * m + i] = v;
        }
    }
}

/* tensor_dil_omp.F90:394-436 */
void orc_tensor_block_add(int rank, const int *dims, double *t0,
                          const double *t1, double scale, int *ierr) {
    *ierr = 0;
    if (rank == 0) { t0[0] = t0[0] + t1[0] * scale; return; }
    if (rank < 0) { *ierr = -1; return; }
    long long n = 1;
    for (int i = 0; i < rank; ++i) n *= dims[i];
    double ds = scale - 1.0;
    if (ds < 0) ds = -ds;
    if (ds > 1e-13) {
        for (long long i = 0; i < n; ++i) t0[i] = t0[i] + t1[i] * scale;
    } else {
        for (long long i = 0; i < n; ++i) t0[i] = t0[i] + t1[i];
    }
}

/* tensor_dil_omp.F90:662-797: permute operands to contracted-first order,
 * zero the matrix result (assign semantics), multiply, permute back. */
void orc_tensor_block_contract(const int *ptrn, const double *l, int lr,
                               const int *ldims, const double *r, int rr,
                               const int *rdims, double *d, int dr,
                               const int *ddims, int *ierr) {
    *ierr = 0;
    if (lr < 0 || lr > ORC_MAX_RANK || rr < 0 || rr > ORC_MAX_RANK || dr < 0 ||
        dr > ORC_MAX_RANK) {
        *ierr = -1;
        return;
    }
    if (!orc_contr_ptrn_ok(ptrn, lr, ldims, rr, rdims, dr, ddims)) { *ierr = 1; return; }
    orc_perms p;
    orc_determine_index_permutations(ptrn, lr, ldims, rr, rdims, dr, ddims, &p);
    const double *ltp = l, *rtp = r;
    double *ltmp = NULL, *rtmp = NULL, *dtmp = NULL, *dtp = d;
    if (p.ltransp) {
        ltmp = (double *)malloc(sizeof(double) * (size_t)p.lsize);
        orc_tensor_block_copy(lr, ldims, p.lo2n, l, ltmp, ierr);
        ltp = ltmp;
    }
    if (p.rtransp) {
        rtmp = (double *)malloc(sizeof(double) * (size_t)p.rsize);
        orc_tensor_block_copy(rr, rdims, p.ro2n, r, rtmp, ierr);
        rtp = rtmp;
    }
    if (p.dtransp) {
        dtmp = (double *)malloc(sizeof(double) * (size_t)p.dsize);
        dtp = dtmp;
    }
    for (long long i = 0; i < p.dsize; ++i) dtp[i] = 0.0;
    if (dr > 0 && lr > 0 && rr > 0) {
        orc_gemm_tn(p.lld, p.lrd, p.lcd, ltp, rtp, dtp);
    } else if (dr == 0 && lr > 0 && rr > 0) { /* full contraction */
        double v = 0.0;
        for (long long i = 0; i < p.lcd; ++i) v += ltp[i] * rtp[i];
        dtp[0] = 0.0 + v;
    } else if (dr > 0 && lr > 0 && rr == 0) {
        orc_tensor_block_add(dr, ddims, dtp, ltp, rtp[0], ierr);
    } else if (dr > 0 && lr == 0 && rr > 0) {
        orc_tensor_block_add(dr, ddims, dtp, rtp, ltp[0], ierr);
    } else if (dr == 0 && lr == 0 && rr == 0) {
        dtp[0] = ltp[0] * rtp[0];
    }
    if (p.dtransp) {
        int mdims[ORC_MAX_RANK];
        for (int j = 0; j < dr; ++j) mdims[j] = ddims[p.do2n[j + 1] - 1];
        orc_tensor_block_copy(dr, mdims, p.do2n, dtp, d, ierr);
    }
    free(ltmp);
    free(rtmp);
    free(dtmp);
}

/* energy_denominator_rhf.F:399-444 (do_rhf_den6; do_rhf_den4 :310-352 and
 * do_rhf_den2 :255-300 are the same rule at lower rank): loop over the
 * block in memory order, add the occupied energies left to right, subtract
 * the virtual ones left to right, divide. */
void orc_energy_denominator(int rank, const int *dims, double *x, const double *const *eps,
                            const int *signs) {
    long long n = 1;
    for (int d = 0; d < rank; ++d) n *= dims[d];
    int idx[ORC_MAX_RANK] = {0};
    for (long long e = 0; e < n; ++e) {
        double den = 0.0;
        int first = 1;
        for (int d = 0; d < rank; ++d)
            if (signs[d] > 0) {
                den = first ? eps[d][idx[d]] : den + eps[d][idx[d]];
                first = 0;
            }
        for (int d = 0; d < rank; ++d)
            if (signs[d] < 0) {
                den = first ? -eps[d][idx[d]] : den - eps[d][idx[d]];
                first = 0;
            }
        x[e] = x[e] / den;
        for (int d = 0; d < rank; ++d) { /* odometer, first index fastest */
            if (++idx[d] < dims[d]) break;
            idx[d] = 0;
        }
    }
}

/* tensor_dil_omp.F90:910-938 */
double orc_fcontract(long long n, const double *l, const double *r, double d) {
    double val = 0.0;
    for (long long i = 0; i < n; ++i) val = val + l[i] * r[i];
    return d + val;
}

/* Elements of a block of rank > 0; 0 when the product is not positive. */
static long long orc_elems(int rank, const int *dims) {
    long long n = 1;
    for (int i = 0; i < rank; ++i) n *= dims[i];
    return n > 0 ? n : 0;
}

/* tensor_dil_omp.F90:144-191 */
void orc_block_init(int rank, const int *dims, double *t, double sval, int *ierr) {
    *ierr = 0;
    if (rank < 0) { *ierr = -1; return; }
    if (rank == 0) { t[0] = sval; return; }
    long long n = orc_elems(rank, dims);
    for (long long i = 0; i < n; ++i) t[i] = sval;
}

/* tensor_dil_omp.F90:193-228 */
void orc_block_scale(int rank, const int *dims, double *t, double s, int *ierr) {
    *ierr = 0;
    if (rank < 0) { *ierr = -1; return; }
    if (rank == 0) { t[0] = t[0] * s; return; }
    long long n = orc_elems(rank, dims);
    for (long long i = 0; i < n; ++i) t[i] = t[i] * s;
}

/* tensor_dil_omp.F90:230-269 */
double orc_block_norm2(int rank, const int *dims, const double *t, int *ierr) {
    *ierr = 0;
    if (rank < 0) { *ierr = -1; return 0.0; }
    if (rank == 0) return t[0] * t[0];
    long long n = orc_elems(rank, dims);
    double v = 0.0;
    for (long long i = 0; i < n; ++i) v = v + t[i] * t[i];
    return v;
}

/* Column-major offset in `tens` of element `l` of the slice: the slice's
 * multi-index shifted by ext_beg (tensor_dil_omp.F90:308, :370). */
static long long orc_slice_offset(int rank, const int *tens_ext, const int *slice_ext,
                                  const int *ext_beg, long long l) {
    long long off = 0, stride = 1;
    for (int i = 0; i < rank; ++i) {
        long long im = l % slice_ext[i];
        l /= slice_ext[i];
        off += (ext_beg[i] + im) * stride;
        stride *= tens_ext[i];
    }
    return off;
}

/* Shared rank / size rules of slice and insert (tensor_dil_omp.F90:293,
 * :322-326 and divide_segment_ :1248-1254); returns the slice's element
 * count, or 0 when there is nothing to copy. */
static long long orc_slice_count(int rank, const int *slice_ext, int *ierr) {
    *ierr = 0;
    if (rank < 0) { *ierr = 1; return 0; }
    if (rank == 0) return 1;
    long long n = 1;
    for (int i = 0; i < rank; ++i) n *= slice_ext[i];
    if (n <= 0) { *ierr = -1; return 0; }
    return n;
}

/* tensor_dil_omp.F90:271-330 */
void orc_block_slice(int rank, const double *tens, const int *tens_ext, double *slice,
                     const int *slice_ext, const int *ext_beg, int *ierr) {
    long long n = orc_slice_count(rank, slice_ext, ierr);
    if (rank == 0 && !*ierr) { slice[0] = tens[0]; return; }
    for (long long l = 0; l < n; ++l)
        slice[l] = tens[orc_slice_offset(rank, tens_ext, slice_ext, ext_beg, l)];
}

/* tensor_dil_omp.F90:332-392 */
void orc_block_insert(int rank, double *tens, const int *tens_ext, const double *slice,
                      const int *slice_ext, const int *ext_beg, int *ierr) {
    long long n = orc_slice_count(rank, slice_ext, ierr);
    if (rank == 0 && !*ierr) { tens[0] = slice[0]; return; }
    for (long long l = 0; l < n; ++l)
        tens[orc_slice_offset(rank, tens_ext, slice_ext, ext_beg, l)] = slice[l];
}
