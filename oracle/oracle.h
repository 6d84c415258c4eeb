/*
 * oracle.h -- CPU restatement of the SIP block-contraction algorithm.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product library (libsipgpu.so) never links or calls it.
 *
 * Every routine restates the reference's Fortran kernel module
 * src/sip/tensor_algebra/tensor_dil_omp.F90 (cited per function in
 * oracle.c).  Conventions follow the reference exactly:
 *   - blocks are dense FP64, column-major (first index fastest);
 *   - a permutation vector `dim_transp[0..rank]` has dim_transp[0] = sign
 *     (ignored) and dim_transp[i] = 1-based new position of old dim i;
 *   - a contraction pattern `ptrn[0..lrank+rrank-1]` holds +k (uncontracted,
 *     goes to destination dim k) or -k (contracted with dim k of the other
 *     operand), 1-based;
 *   - error codes are the reference's.
 *
 * The GEMM inside the contraction sums each output element over k in
 * ascending order starting from 0.0, with no fused multiply-add.  That is
 * the summation order of the reference's own BLAS-free GEMM
 * (tensor_dil_omp.F90:940-1179) for every shape class except its
 * "contracted dimension only is large" class, which uses an OpenMP
 * reduction (tensor_dil_omp.F90:1097-1127).
 */
#ifndef SIPGPU_ORACLE_H
#define SIPGPU_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_RANK 32

/* Result of determine_index_permutations (tensor_dil_omp.F90:799-859). */
typedef struct {
    int lo2n[ORC_MAX_RANK + 1]; /* [0]=sign, [1..lrank] 1-based new position */
    int ro2n[ORC_MAX_RANK + 1];
    int do2n[ORC_MAX_RANK + 1]; /* matrix-order dim j -> destination dim */
    int ltransp, rtransp, dtransp;
    int nlu, nru, ncd;
    long long lld, lrd, lcd; /* M, N, K */
    long long lsize, rsize, dsize;
} orc_perms;

long long orc_tensor_size_by_shape(int num_dim, const int *dims, int *ierr);

/* Stable sort of trn[1..ni] (item numbers, 1-based) by key[item-1];
 * trn[0] is a sign flipped per odd transposition, as in the reference. */
void orc_merge_sort_key_int(int ni, const int *key, int *trn);

void orc_get_contraction_ptrn(int drank, int lrank, int rrank,
                              const int *aces_ptrn, int *my_ptrn, int *ierr);

int orc_contr_ptrn_ok(const int *ptrn, int lrank, const int *ldims,
                      int rrank, const int *rdims, int drank,
                      const int *ddims);

void orc_determine_index_permutations(const int *ptrn, int lrank,
                                      const int *ldims, int rrank,
                                      const int *rdims, int drank,
                                      const int *ddims, orc_perms *p);

void orc_tensor_block_copy(int rank, const int *dims, const int *dim_transp,
                           const double *in, double *out, int *ierr);

/* D(M,N) = sum_k L(k,m) R(k,n), L stored K x M, R stored K x N. */
void orc_gemm_tn(long long m, long long n, long long k, const double *l,
                 const double *r, double *d);

void orc_tensor_block_contract(const int *ptrn, const double *l, int lrank,
                               const int *ldims, const double *r, int rrank,
                               const int *rdims, double *d, int drank,
                               const int *ddims, int *ierr);

void orc_tensor_block_add(int rank, const int *dims, double *t0,
                          const double *t1, double scale, int *ierr);

/* tensor_block_init_ (tensor_dil_omp.F90:144-191): rank 0 sets t[0]; a
 * positive rank sets all prod(dims) elements (none when the product is not
 * positive); a negative rank sets ierr = -1 and writes nothing. */
void orc_block_init(int rank, const int *dims, double *t, double sval, int *ierr);

/* tensor_block_scale_ (tensor_dil_omp.F90:193-228): t *= s elementwise, same
 * rank rules as orc_block_init. */
void orc_block_scale(int rank, const int *dims, double *t, double s, int *ierr);

/* tensor_block_norm2_ (tensor_dil_omp.F90:230-269): sum of t[i]^2 in
 * ascending i (the reference reduces with OpenMP; one thread gives this
 * order); rank 0 gives t[0]^2, a negative rank 0 with ierr = -1. */
double orc_block_norm2(int rank, const int *dims, const double *t, int *ierr);

/* tensor_block_slice_ (tensor_dil_omp.F90:271-330): slice(i) =
 * tens(ext_beg + i) for every multi-index i of the slice, both column-major.
 * rank 0 copies the scalar; a negative rank gives ierr = 1; an empty slice
 * (a non-positive element count) gives ierr = -1 and writes nothing (the
 * reference's divide_segment_, :1238-1256, rejects a zero-length range).
 * Like the reference, no bounds are checked. */
void orc_block_slice(int rank, const double *tens, const int *tens_ext, double *slice,
                     const int *slice_ext, const int *ext_beg, int *ierr);

/* tensor_block_insert_ (tensor_dil_omp.F90:332-392): tens(ext_beg + i) =
 * slice(i); error rules as orc_block_slice. */
void orc_block_insert(int rank, double *tens, const int *tens_ext, const double *slice,
                      const int *slice_ext, const int *ext_beg, int *ierr);

/* energy_denominator_rhf (src/sip/super_instructions/qm/qm-generic/
 * energy_denominator_rhf.F:15-352): x /= eps in place, where for element
 * (i_0 .. i_{rank-1}) eps adds eps_d[i_d] over the occupied indices
 * (sign_d = +1) in index order, then subtracts eps_d[i_d] over the virtual
 * ones (sign_d = -1) in index order (do_rhf_den6 :399-444: eps = epsb +
 * epsd + epsf - epsa - epsc - epse). */
void orc_energy_denominator(int rank, const int *dims, double *x, const double *const *eps,
                            const int *signs);

/* tensor_block_fcontract (tensor_dil_omp.F90:910-938): returns
 * d + sum_l l[l] * r[l] over n elements, summed in ascending order (the
 * reference reduces with OpenMP, whose order is unspecified). */
double orc_fcontract(long long n, const double *l, const double *r, double d);

#ifdef __cplusplus
}
#endif
#endif
