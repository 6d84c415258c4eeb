/*
 * sipgpu.h -- C ABI of libsipgpu.so, the MI355X (gfx950) implementation of the
 * SIP block-contraction hot path: the `contract` and `permute`
 * super-instructions that the SIAL interpreter issues on segmented blocks.
 *
 * Two layers share one library:
 *
 * (1) Drop-in layer.  The exact symbols, argument meaning, assign semantics
 *     and error codes of the reference's tensor kernel module, so a SIP build
 *     links libsipgpu.so in place of its tensor_dil_omp library.  All scalars
 *     are passed by pointer (binary-identical to the reference's C++
 *     by-reference prototypes), `int` is 32-bit, blocks are dense FP64
 *     column-major host arrays, and every call is synchronous.
 *
 * (2) Native layer (`sipgpu_*`).  Device-resident blocks, HIP streams,
 *     planned batches of contractions executed as one grouped launch, and the
 *     device block cache.  All functions return 0 on success or a negative
 *     SIPGPU_E* code; sipgpu_last_error() gives the message.
 *
 * Conventions shared with the reference:
 *   - a permutation `dim_transp[0..rank]` has dim_transp[0] = sign (ignored)
 *     and dim_transp[i] = 1-based new position of old dimension i;
 *   - a contraction pattern `ptrn[0..lrank+rrank-1]` holds +k (uncontracted,
 *     becomes destination dimension k) or -k (contracted with dimension k of
 *     the other operand), 1-based, as produced by get_contraction_ptrn_.
 */
#ifndef SIPGPU_H
#define SIPGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Drop-in layer                                                             */
/* ------------------------------------------------------------------------ */

/* Replaces tensor_size_by_shape_ (tensor_dil_omp.F90:66-82; prototype
 * tensor_ops_c_prototypes.h:41).  ierr: 1 negative rank, 2 non-positive size. */
long long tensor_size_by_shape_(int *num_dim, int *dims, int *ierr);

/* Replaces get_contraction_ptrn_ (tensor_dil_omp.F90:84-142; prototype
 * tensor_ops_c_prototypes.h:55).  aces_ptrn holds drank+lrank+rrank integer
 * labels (destination, left, right), each label exactly twice.  ierr codes
 * 1..6 as in the reference. */
void get_contraction_ptrn_(int *drank, int *lrank, int *rrank, int *aces_ptrn,
                           int *my_ptrn, int *ierr);

/* Replaces tensor_block_copy__ (tensor_dil_omp.F90:438-660; prototype
 * tensor_ops_c_prototypes.h:152), the body of the permute super-instruction
 * (Block::transpose_copy, block.cpp:216-255).  `dims` are the SOURCE
 * extents.  ierr = rank for a negative rank; SIPGPU_E* for an invalid
 * permutation or a device failure.  nthreads is accepted and ignored. */
void tensor_block_copy__(int *nthreads, int *rank, int *dims, int *dim_transp,
                         double *in, double *out, int *ierr);

/* Replaces tensor_block_contract__ (tensor_dil_omp.F90:662-906; prototype
 * tensor_ops_c_prototypes.h:178), called by Interpreter::handle_contraction
 * (interpreter.cpp:1210-1262).  D is overwritten (assign semantics,
 * tensor_dil_omp.F90:752).  ierr: 1 bad pattern, -1 bad ranks, SIPGPU_E* for
 * device failures.  nthreads is accepted and ignored. */
void tensor_block_contract__(int *nthreads, int *ptrn, double *l, int *lrank,
                             int *ldims, double *r, int *rrank, int *rdims,
                             double *d, int *drank, int *ddims, int *ierr);

/* The rest of the reference module's C interface (prototypes
 * tensor_ops_c_prototypes.h:70-133).  Same conventions as above: host blocks,
 * synchronous, nthreads accepted and ignored; device failures give a
 * negative SIPGPU_E* ierr.  Element counts follow the reference: rank 0 is
 * one element, and a product of extents that is not positive is no element. */

/* Replaces tensor_block_init__ (tensor_dil_omp.F90:144-191; prototype
 * tensor_ops_c_prototypes.h:70): every element = *sval.  ierr -1 for a
 * negative rank. */
void tensor_block_init__(int *nthreads, double *tens, int *tens_rank, int *dims, double *sval,
                         int *ierr);

/* Replaces tensor_block_scale__ (tensor_dil_omp.F90:193-228; prototype
 * tensor_ops_c_prototypes.h:84): every element *= *scale_fac.  ierr -1 for
 * a negative rank. */
void tensor_block_scale__(int *nthreads, double *tens, int *tens_rank, int *dims,
                          double *scale_fac, int *ierr);

/* Replaces tensor_block_norm2__ (tensor_dil_omp.F90:230-269; prototype
 * tensor_ops_c_prototypes.h:86): the sum of squares.  The reference's
 * OpenMP reduction has no fixed order; this one is fixed for a given size.
 * ierr -1 (result 0) for a negative rank. */
double tensor_block_norm2__(int *nthreads, double *tens, int *tens_rank, int *dims, int *ierr);

/* Replaces tensor_block_slice__ (tensor_dil_omp.F90:271-330; prototype
 * tensor_ops_c_prototypes.h:100), called by Block::extract_slice
 * (dynamic_data/block.cpp:294): slice(i) = tens(ext_beg + i) for every
 * column-major multi-index i of the slice.  ierr: 1 negative dim_num, -1
 * empty slice (as the reference's segment division reports it),
 * SIPGPU_EARG for a slice that leaves the tensor block (which the reference
 * does not check), SIPGPU_ESIZE for blocks of 2^31 elements or more.  Only
 * the span of `tens` the slice touches is staged to the device. */
void tensor_block_slice__(int *nthreads, int *dim_num, double *tens, int *tens_ext, double *slice,
                          int *slice_ext, int *ext_beg, int *ierr);

/* Replaces tensor_block_insert__ (tensor_dil_omp.F90:332-392; prototype
 * tensor_ops_c_prototypes.h:117), called by Block::insert_slice
 * (dynamic_data/block.cpp:320): tens(ext_beg + i) = slice(i).  Errors as
 * tensor_block_slice__. */
void tensor_block_insert__(int *nthreads, int *dim_num, double *tens, int *tens_ext,
                           double *slice, int *slice_ext, int *ext_beg, int *ierr);

/* Replaces tensor_block_add__ (tensor_dil_omp.F90:394-436; prototype
 * tensor_ops_c_prototypes.h:133): tens0 += tens1 * s, the product and the
 * sum rounded separately; for rank > 0 a factor within 1e-13 of one is not
 * applied (tens0 += tens1).  ierr -1 for a negative rank. */
void tensor_block_add__(int *nthreads, int *dim_num, int *dim_extents, double *tens0,
                        double *tens1, double *scale_fac, int *ierr);

/* ------------------------------------------------------------------------ */
/* Native layer                                                              */
/* ------------------------------------------------------------------------ */

#define SIPGPU_MAX_RANK 6

enum {
    SIPGPU_OK = 0,
    SIPGPU_EPATTERN = -101,   /* contraction pattern rejected */
    SIPGPU_ERANK = -102,      /* rank outside 0..SIPGPU_MAX_RANK */
    SIPGPU_EPERM = -103,      /* dim_transp is not a permutation */
    SIPGPU_ESIZE = -104,      /* block too large (>= 2^31 elements) */
    SIPGPU_EDEVICE = -105,    /* HIP runtime error */
    SIPGPU_ENOMEM = -106,     /* device or host allocation failed */
    SIPGPU_EARG = -107,       /* invalid argument (null pointer, bad handle) */
    SIPGPU_EHAZARD = -108,    /* batch writes a block it also reads */
    SIPGPU_EBUDGET = -109,    /* block cache budget cannot hold the request */
    SIPGPU_ENOTFOUND = -110,  /* block id not present in the cache */
    SIPGPU_ECAPTURE = -111    /* not allowed while the stream is being captured */
};

typedef struct sipgpu_stream_s *sipgpu_stream_t;   /* wraps a hipStream_t */
typedef struct sipgpu_event_s *sipgpu_event_t;     /* wraps a hipEvent_t */
typedef struct sipgpu_batch_s *sipgpu_batch_t;     /* a planned batch */

const char *sipgpu_last_error(void);
const char *sipgpu_version(void);

/* Selects the device (default 0) and creates the library's own stream. */
int sipgpu_init(int device);
int sipgpu_device_count(int *count);
int sipgpu_device_synchronize(void);

/* Device memory (plain HBM allocations, no caching).  The copies and the
 * memset are synchronous and ordered after all work already enqueued on any
 * stream; sipgpu_free waits for that work too. */
int sipgpu_malloc(void **dptr, size_t bytes);
int sipgpu_free(void *dptr);
int sipgpu_memcpy_h2d(void *dst, const void *src, size_t bytes);
int sipgpu_memcpy_d2h(void *dst, const void *src, size_t bytes);
int sipgpu_memcpy_d2d(void *dst, const void *src, size_t bytes);
int sipgpu_memset(void *dptr, int value, size_t bytes);

/* Streams (default priority only) and events. A NULL stream means the
 * library's own stream. */
int sipgpu_stream_create(sipgpu_stream_t *s);
int sipgpu_stream_destroy(sipgpu_stream_t s);
int sipgpu_stream_synchronize(sipgpu_stream_t s);
int sipgpu_event_create(sipgpu_event_t *e);
int sipgpu_event_destroy(sipgpu_event_t e);
int sipgpu_event_record(sipgpu_event_t e, sipgpu_stream_t s);
/* Work enqueued on `s` after this call waits for the work recorded by `e`
 * (no host wait): the join of work split over several streams. */
int sipgpu_stream_wait_event(sipgpu_stream_t s, sipgpu_event_t e);
int sipgpu_event_synchronize(sipgpu_event_t e);
int sipgpu_event_elapsed_ms(sipgpu_event_t start, sipgpu_event_t stop, float *ms);

/* Permute a device block, asynchronously on `s`.  Same meaning as
 * tensor_block_copy__; `in` and `out` must not overlap. */
int sipgpu_permute(sipgpu_stream_t s, int rank, const int *dims,
                   const int *dim_transp, const double *in, double *out);

/* A batch of permutes run as one launch.  The reference permutes one block
 * per call (tensor_block_copy__, tensor_ops_c_prototypes.h:152, from the
 * permute super-instruction and Block::transpose_copy, block.cpp:216-232);
 * a batch is that call repeated over its entries, so that many small blocks
 * (whose single launches are latency-bound) fill the device together.
 * Each entry has sipgpu_permute's meaning: `in` (rank, dims) into `out` with
 * dim_transp[0..rank] (sign slot first, then the 1-based new position of
 * each old dimension).  Entries run concurrently: no entry's output may
 * overlap another entry's input or output. */
typedef struct {
    const double *in;
    double *out;
    int rank;
    int dims[SIPGPU_MAX_RANK];
    int dim_transp[SIPGPU_MAX_RANK + 1];
} sipgpu_permutation;
typedef struct sipgpu_permute_batch_s *sipgpu_permute_batch_t;
/* Validates every entry (SIPGPU_EARG null pointer, SIPGPU_ERANK rank outside
 * 0..6, SIPGPU_EPERM bad dim_transp, SIPGPU_ESIZE block of 2^31 elements or
 * more) and uploads the batch's tile table; the blocks stay the caller's. */
int sipgpu_permute_batch_plan(const sipgpu_permutation *list, int n,
                              sipgpu_permute_batch_t *batch);
/* One launch over every entry's tiles, asynchronously on `s`. */
int sipgpu_permute_batch_execute(sipgpu_permute_batch_t batch, sipgpu_stream_t s);
/* Elements moved per execution (reads = writes) and workgroups launched. */
int sipgpu_permute_batch_info(sipgpu_permute_batch_t batch, long long *elements,
                              int *workgroups);
int sipgpu_permute_batch_destroy(sipgpu_permute_batch_t batch);

/* Contract device blocks, asynchronously on `s`.  accumulate = 0 gives the
 * reference's assign semantics (D = L*R); accumulate = 1 gives D += L*R
 * (the block_contract_accumulate_op semantics of opcode.h:117, which the
 * reference interpreter leaves unimplemented). */
int sipgpu_contract(sipgpu_stream_t s, const int *ptrn, const double *l,
                    int lrank, const int *ldims, const double *r, int rrank,
                    const int *rdims, double *d, int drank, const int *ddims,
                    int accumulate);

/* sipgpu_contract that allocates nothing.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under graph
 * capture, tight serving loops).  Same arguments and results, bit for bit;
 * the K-segment workspace of a split contraction and the contract-to-scalar
 * partials live in a static device array of 16 MiB + 32 KiB per device,
 * held until the process ends.  Preconditions: the library is initialised
 * (SIPGPU_EARG otherwise) and the shape (pattern and extents) has been run
 * once through sipgpu_contract, which uploads its offset tables
 * (SIPGPU_ENOTFOUND otherwise).  Both entry points cap a contraction's K
 * split so its segments fit the 16 MiB (a larger need would be refused with
 * SIPGPU_ESIZE).  All calls on one device share that array: calls that
 * overlap on the device (other streams, other threads) race. */
int sipgpu_contract_noalloc(sipgpu_stream_t s, const int *ptrn, const double *l,
                            int lrank, const int *ldims, const double *r,
                            int rrank, const int *rdims, double *d, int drank,
                            const int *ddims, int accumulate);

/* One contraction of a batch.  Entries that name the same destination block
 * (same pointer, extents and pattern destination order) form one task: their
 * products are summed in list order inside one accumulator and written once;
 * if the first such entry has accumulate = 0 the block is overwritten,
 * otherwise the sum is added to it. */
typedef struct {
    const double *l;
    const double *r;
    double *d;
    int lrank, rrank, drank;
    int ldims[SIPGPU_MAX_RANK];
    int rdims[SIPGPU_MAX_RANK];
    int ddims[SIPGPU_MAX_RANK];
    int ptrn[2 * SIPGPU_MAX_RANK];
    int accumulate;
} sipgpu_contraction;

/* Plans a batch: validates every entry, builds the index-offset tables
 * (permutations fused into the operand loads and the result scatter),
 * groups entries into tasks and tiles, and uploads the work list.  The plan
 * owns only its work list; blocks stay owned by the caller. */
int sipgpu_batch_plan(const sipgpu_contraction *list, int n, sipgpu_batch_t *batch);
/* Runs a planned batch as one grouped FP64-MFMA launch on `s`. */
int sipgpu_batch_execute(sipgpu_batch_t batch, sipgpu_stream_t s);
int sipgpu_batch_destroy(sipgpu_batch_t batch);
/* Plans the same batch for the split-precision path: every FP64 product is
 * emulated with INT8 MFMA (Ozaki scheme II): rows of the left operand and
 * columns of the right are scaled to P-bit integers, their product is formed
 * exactly modulo `nmoduli` (6..14) pairwise-coprime moduli <= 255 with INT32
 * accumulation, and recovered by the Chinese remainder theorem in FP64.  P
 * is the largest width for which the exact product fits (sipgpu_batch_split_info);
 * the result's relative error is of order 2^-P.  Executed with
 * sipgpu_batch_execute; each execution rescales and re-slices its operands.
 * Blocks that share rows through a result block share their row scales. */
int sipgpu_batch_plan_split(const sipgpu_contraction *list, int n, int nmoduli,
                            sipgpu_batch_t *batch);
/* The split path with the moduli count set by an error target: the fewest
 * moduli (6..14) whose operand width P, for this batch's longest K, meets
 * 16 * 2^-P <= rel_err.  The split products' relative error (Frobenius, per
 * result block) is about kappa 2^-P with kappa 4-5 on the benchmark blocks
 * (Gaussian and uniform data); 16 leaves room for rows with a wider spread.
 * Fails with SIPGPU_EARG when 14 moduli do not meet the target (run the
 * plain FP64 path instead).  sipgpu_batch_split_info reports the choice. */
int sipgpu_batch_plan_split_tol(const sipgpu_contraction *list, int n, double rel_err,
                                sipgpu_batch_t *batch);
/* Moduli count and operand bits P of a split batch (0 and 53 for FP64 batches). */
int sipgpu_batch_split_info(sipgpu_batch_t batch, int *nmoduli, int *bits);
/* How a planned batch runs.  sipgpu_batch_plan picks the packed path for
 * batches of at least 1e9 flop averaging 5e7 flop per result block: each
 * operand block is first packed once into K-contiguous panels, then one
 * grouped GEMM runs over whole 16-k stages of every task's concatenated K
 * (SIPGPU_FP64_PATH=tables|packed in the environment pins the choice). */
enum { SIPGPU_BATCH_TABLES = 0, SIPGPU_BATCH_PACKED = 1, SIPGPU_BATCH_SPLIT = 2 };
int sipgpu_batch_kind(sipgpu_batch_t batch, int *kind);
/* Runs only some stages of a planned batch on `s` (no reference
 * counterpart: the SIP issues a contraction as one call).  PREPARE builds
 * the operand panels (packed and split paths: packing, row / column scales
 * and residues), PRODUCT runs the grouped GEMM, FINISH the CRT
 * reconstruction of a split batch.  A stage absent from a batch's path is a
 * no-op.  Running PREPARE, PRODUCT, FINISH in that order on one stream is
 * sipgpu_batch_execute; callers use it to time the GEMM by itself or to
 * overlap one batch's preparation with another's product on two streams
 * (each stage reads what the previous stage of the same batch wrote, so the
 * caller orders them with events). */
enum { SIPGPU_STAGE_PREPARE = 1, SIPGPU_STAGE_PRODUCT = 2, SIPGPU_STAGE_FINISH = 4,
       SIPGPU_STAGE_ALL = 7 };
int sipgpu_batch_execute_stages(sipgpu_batch_t batch, int stages, sipgpu_stream_t s);
/* Autotuner (no reference counterpart).  Plans the batch every candidate
 * way -- the table kernel in each of its tile configurations (and, for
 * configurations 0 and 11, with each result block's pairs cut into groups
 * that run side by side when the tiles fill at most half of the CUs: tail 1
 * on path 0; the groups' workspace copies are added into the block after the
 * GEMM in group order), packed FP64
 * panels with and without merged result blocks, without the persistent
 * water-filled schedule (tail 0) and with it over 64, 32, 16 or 8
 * workgroups per XCD (tail 1, 2, 3, 4: fewer workgroups cut a small batch's
 * tiles into fewer K segments), in 128 x 128 tiles and in 64 x 64 tiles
 * (tile 64: a small batch's K segments write a quarter of the partial
 * sums); for the split path (nmoduli > 0) the
 * three INT8 GEMMs (tile 128: 128 x 128, four waves; 256: 256 x 256, four
 * waves of 128 x 128; 257: 256 x 256, eight waves of 64 x 128) with and
 * without merging -- times each (one warm-up and five
 * timed executions on the library stream, into scratch copies of the
 * destinations, so the caller's blocks are not touched) and returns the
 * fastest plan.  The untuned plan is kept (its fields -1, "default") unless
 * the fastest candidate beats it by more than 2%.  The choice is remembered under the batch's signature (its
 * shapes, patterns, accumulate flags and which entries share blocks, not the
 * addresses), so later sipgpu_batch_plan / _plan_split calls on batches with
 * that signature plan the same way.  Synchronises the device. */
typedef struct {
    int candidates;     /* plans timed */
    double best_ms;     /* per execution, the chosen plan */
    double default_ms;  /* per execution, the untuned plan (0 if it failed) */
    int path;           /* 0 table kernel, 1 packed FP64, 2 split precision */
    int config, merge, tail, tile;
} sipgpu_tune_result;
int sipgpu_batch_tune(const sipgpu_contraction *list, int n, int nmoduli, sipgpu_batch_t *batch,
                      sipgpu_tune_result *result);
/* Plans the batch one given way (the fields of sipgpu_tune_result: path,
 * config and tail (1: pair groups) for the table kernel, merge / tail / tile
 * for the panel paths, -1 for a default; tile 64 or -1 for packed FP64,
 * 128 / 256 / 257 or -1 for the split path): pins a plan, e.g. one found by
 * tuning elsewhere. */
int sipgpu_batch_plan_choice(const sipgpu_contraction *list, int n, int nmoduli, int path,
                             int config, int merge, int tail, int tile, sipgpu_batch_t *batch);
/* The tuning table as text (one "signature path config merge tail tile" line
 * per entry): save writes it, load merges a file into it, clear empties it. */
int sipgpu_tune_save(const char *path);
int sipgpu_tune_load(const char *path);
int sipgpu_tune_clear(void);
/* Sum of 2*M*N*K over the batch, and its task / tile / pair counts. */
int sipgpu_batch_info(sipgpu_batch_t batch, double *flops, int *ntasks,
                      int *ntiles, int *npairs);

/* ------------------------------------------------------------------------
 * Orbital-energy denominators (energy_denominator_rhf.F:15-352, the SIP's
 * energy_denominator_rhf super-instruction, generalised to any rank <= 6).
 * Block x (device, column-major, extents dims) has one orbital-energy vector
 * per index, eps[d] (device, dims[d] entries: that index's segment), and a
 * sign per index, +1 occupied / -1 virtual.  The denominator of an element
 * is its occupied energies added in index order, then its virtual energies
 * subtracted in index order (e_b + e_d + e_f - e_a - e_c - e_e for the
 * reference's (a,i,b,j,c,k) blocks, do_rhf_den6 :399-444). */
int sipgpu_energy_denominator(sipgpu_stream_t s, int rank, const int *dims, double *x,
                              const double *const *eps, const int *signs);
/* *out += factor * sum_e (x[e] / D(e)) * y[e], enqueued on `s`; `out` is a
 * device double.  With y = x this is the (T) energy of a W block, i.e. the
 * denominator division then the contract-to-scalar of rccsdpt_aaa.sialx
 * :2166-2174 in one pass, without writing the divided block.  Fixed
 * summation order (deterministic). */
int sipgpu_energy_sum(sipgpu_stream_t s, int rank, const int *dims, const double *x,
                      const double *y, const double *const *eps, const int *signs,
                      double factor, double *out);
/* t = ((t + parts[0]) + parts[1]) + ... (0..8 partial blocks, stored in t),
 * then *out += factor * sum_e (t[e] / D(e)) * t[e], in one pass: the result
 * of sipgpu_block_add_many followed by sipgpu_energy_sum(x = y = t), bit for
 * bit.  A W block whose terms were written as partial blocks is summed and
 * reduced without reading it back. */
int sipgpu_energy_sum_parts(sipgpu_stream_t s, int rank, const int *dims, double *t,
                            const double *const *parts, int nparts, const double *const *eps,
                            const int *signs, double factor, double *out);

/* ------------------------------------------------------------------------
 * Block operations on device blocks, enqueued on `s` (asynchronous): the
 * drop-in layer's init / scale / add / norm2 / slice / insert with device
 * pointers, same element rules and arithmetic (tensor_dil_omp.F90:144-436).
 * Errors: SIPGPU_ERANK for a negative rank, SIPGPU_EARG for null pointers or
 * a slice outside the tensor block, SIPGPU_ESIZE for oversized blocks.  An
 * empty block or slice is a successful no-op. */
int sipgpu_block_init(sipgpu_stream_t s, double *t, int rank, const int *dims, double value);
int sipgpu_block_scale(sipgpu_stream_t s, double *t, int rank, const int *dims, double factor);
/* t0 += t1 * factor (a factor within 1e-13 of one is not applied for rank > 0). */
int sipgpu_block_add(sipgpu_stream_t s, double *t0, const double *t1, int rank, const int *dims,
                     double factor);
/* t = ((t + parts[0]) + parts[1]) + ... elementwise, 0..8 sources added in
 * order: the sequence of roundings of the same tensor_block_add__ calls with
 * factor one, in one pass over HBM.  No reference counterpart as one call;
 * a block accumulated from several producers (partial results written by
 * separate batches) is summed with it. */
int sipgpu_block_add_many(sipgpu_stream_t s, double *t, const double *const *parts, int nparts,
                          int rank, const int *dims);
/* *result (a device double) = sum of squares, in an order fixed by the
 * element count (per-workgroup partials, then one workgroup). */
int sipgpu_block_norm2(sipgpu_stream_t s, const double *t, int rank, const int *dims,
                       double *result);
int sipgpu_block_slice(sipgpu_stream_t s, int rank, const double *tens, const int *tens_ext,
                       double *slice, const int *slice_ext, const int *ext_beg);
int sipgpu_block_insert(sipgpu_stream_t s, int rank, double *tens, const int *tens_ext,
                        const double *slice, const int *slice_ext, const int *ext_beg);

/* ------------------------------------------------------------------------
 * Device block cache.  Replaces the SIP's BlockManager / CachedBlockMap /
 * LRUArrayPolicy residency logic (block_manager.cpp:106-206,
 * cached_block_map.cpp:31-147, lru_array_policy.h:37-96) for blocks in
 * HBM: blocks are named by 64-bit keys and held in one arena of a fixed
 * budget; when a block does not fit, unpinned blocks are evicted
 * least-recently-released first, and a block whose device copy is newer
 * than its host copy is first written back (spilled) to its host home, or
 * to pinned host memory the cache allocates.  The reference evicts whole
 * arrays; this cache evicts single blocks.  Work that uses a block must be
 * enqueued between sipgpu_cache_acquire and sipgpu_cache_release on the
 * same stream; ordering between streams is the caller's. */
typedef struct sipgpu_cache_s *sipgpu_cache_t;

enum { /* initial value of a block with no host copy */
    SIPGPU_CACHE_UNDEFINED = 0,
    SIPGPU_CACHE_ZERO = 1,
    SIPGPU_CACHE_UNIFORM = 2, /* sipgpu_fill_random(seed, stream, dist 0, a, b) */
    SIPGPU_CACHE_NORMAL = 3   /* sipgpu_fill_random(seed, stream, dist 1, a) */
};
enum { SIPGPU_CACHE_READ = 1, SIPGPU_CACHE_WRITE = 2, SIPGPU_CACHE_READWRITE = 3 };

typedef struct {
    unsigned long long hits, misses, evictions, spills;
    unsigned long long bytes_h2d, bytes_d2h;
    unsigned long long resident_bytes, budget_bytes, blocks;
} sipgpu_cache_stats;

int sipgpu_cache_create(size_t budget_bytes, sipgpu_cache_t *cache);
int sipgpu_cache_destroy(sipgpu_cache_t cache);
/* Registers a block of `elems` doubles.  `host` (may be NULL) is its home
 * copy, read on a miss and written by spills; it must outlive the block.
 * Without a home, `init` defines its first value. */
int sipgpu_cache_insert(sipgpu_cache_t cache, unsigned long long key, size_t elems,
                        const double *host, int init, unsigned long long seed,
                        unsigned long long stream, double a, double b);
/* Makes the block resident (on a miss: evicting as needed, then uploading
 * on the cache's host-to-device stream, which `s` waits for), pins it and
 * returns its device address.  Work enqueued on `s` after the call sees the
 * block's value; other streams must be ordered after `s`.  WRITE marks the
 * device copy as the current value without uploading it. */
int sipgpu_cache_acquire(sipgpu_cache_t cache, sipgpu_stream_t s, unsigned long long key,
                         int mode, double **dev);
/* Unpins the block after the work enqueued on `s` that uses it. */
int sipgpu_cache_release(sipgpu_cache_t cache, sipgpu_stream_t s, unsigned long long key);
/* Current value of a block into host memory (synchronises the device). */
int sipgpu_cache_read(sipgpu_cache_t cache, unsigned long long key, double *out);
/* Writes back dirty resident blocks (one key, or all when `all`) to their
 * homes: the SIP's flush before super-instructions and checkpoints. */
int sipgpu_cache_flush(sipgpu_cache_t cache, unsigned long long key, int all);
int sipgpu_cache_evict(sipgpu_cache_t cache, unsigned long long key);
int sipgpu_cache_get_stats(sipgpu_cache_t cache, sipgpu_cache_stats *stats);

/* Block lifetimes of a SIAL program.
 *
 * sipgpu_cache_remove forgets a block without writing it back: the
 * BlockManager's delete of a block whose value is no longer needed
 * (block_manager.cpp:91-101).  SIPGPU_EARG while the block is acquired.
 *
 * Scopes replace BlockManager::enter_scope / leave_scope
 * (block_manager.cpp:182-206): sipgpu_cache_scope_add makes a block a temp
 * of the innermost open scope (what get_block_for_writing does for a block
 * created inside a scope, :105-115), and sipgpu_cache_scope_leave removes
 * all of that scope's temps (none if any of them is still acquired) and
 * closes it.  Scopes nest.
 *
 * Persistence replaces set_persistent / save_marked_arrays /
 * restore_persistent (worker_persistent_array_manager.h:20-57):
 * sipgpu_cache_end_program removes every block not marked persistent and
 * keeps the marked ones with their values, resident or spilled, for the next
 * program; the marks are cleared, so a block outlives the next program only
 * if that program marks it again (an object "can only be restored once").
 * `kept` (may be NULL) receives the number of blocks kept.  It fails while a
 * scope is open or an unmarked block is acquired. */
int sipgpu_cache_remove(sipgpu_cache_t cache, unsigned long long key);
int sipgpu_cache_scope_enter(sipgpu_cache_t cache);
int sipgpu_cache_scope_add(sipgpu_cache_t cache, unsigned long long key);
int sipgpu_cache_scope_leave(sipgpu_cache_t cache);
int sipgpu_cache_set_persistent(sipgpu_cache_t cache, unsigned long long key, int persistent);
int sipgpu_cache_end_program(sipgpu_cache_t cache, unsigned long long *kept);

/* ------------------------------------------------------------------------
 * Contraction server: continuous batching of single block contractions
 * (and permutes) whose blocks live in a cache.  A scheduler thread takes
 * every request that has arrived (up to `max_batch`, stopping at the first
 * one that would read a block an earlier request of the batch writes, or the
 * reverse, and at a change of precision; consecutive permutes form a batch
 * of their own, in which no permute's destination is another's source or
 * destination), makes its blocks resident and launches the batch: one
 * batched permute launch, one grouped FP64
 * launch, or the split-precision stages with their planes in a server-owned
 * scratch buffer per stream (grown, after that stream drains, to the
 * largest batch's need).  Two device streams: large batches (>= 1e9 flop)
 * run in order on the first, and a small batch admitted while the first is
 * busy runs on the second, beside it on the device; a batch that reads a
 * block an unfinished batch of the other stream writes (or writes one it
 * reads or writes) waits for it on the device first.  One
 * completion thread per stream stamps each request's completion.  A batch
 * pins at most half the cache budget (one request beyond that alone).
 * Requests on the same blocks take effect in submission order.  Every block
 * a request names must be in the cache with exactly its extents' elements
 * (SIPGPU_ENOTFOUND / SIPGPU_EARG at submission otherwise).  No reference
 * counterpart: the SIP's worker runs one contraction at a time
 * (interpreter.cpp:1210-1262). */
typedef struct sipgpu_server_s *sipgpu_server_t;

typedef struct {
    int ptrn[2 * SIPGPU_MAX_RANK];
    int lrank, rrank, drank;
    int ldims[SIPGPU_MAX_RANK], rdims[SIPGPU_MAX_RANK], ddims[SIPGPU_MAX_RANK];
    unsigned long long lkey, rkey, dkey;
    int accumulate;
    int moduli; /* 0: plain FP64; 6..14: split precision over that many moduli */
} sipgpu_request;

typedef struct {
    unsigned long long batches, requests, max_batch;
    double flops;
} sipgpu_server_stats;

int sipgpu_server_create(sipgpu_cache_t cache, int max_batch, sipgpu_server_t *server);
int sipgpu_server_submit(sipgpu_server_t server, const sipgpu_request *req,
                         unsigned long long *ticket);
/* Queues a permute of block `src_key` into block `dst_key` (the permute
 * super-instruction, tensor_block_copy__ semantics: `dims` are the source's
 * extents, dim_transp[0] = 1 and dim_transp[i] the new position of old
 * dimension i).  It runs as a batch of its own, in submission order with
 * the contractions around it. */
int sipgpu_server_submit_permute(sipgpu_server_t server, int rank, const int *dims,
                                 const int *dim_transp, unsigned long long src_key,
                                 unsigned long long dst_key, unsigned long long *ticket);
/* Device streams the server's batches alternate over, and how many batches
 * had to wait on the device for a conflicting batch of another stream. */
int sipgpu_server_lanes(sipgpu_server_t server, int *lanes, unsigned long long *cross_waits);
/* Waits for a request; latency = completion - submission, milliseconds. */
int sipgpu_server_wait(sipgpu_server_t server, unsigned long long ticket, double *latency_ms);
/* Submits reqs[i] at arrival_s[i] seconds after the call (an arrival
 * schedule), waits for all, and returns each latency (completion - arrival)
 * and the wall time of the whole replay. */
int sipgpu_server_replay(sipgpu_server_t server, const sipgpu_request *reqs,
                         const double *arrival_s, int n, double *latency_ms, double *wall_s);
int sipgpu_server_get_stats(sipgpu_server_t server, sipgpu_server_stats *stats);
int sipgpu_server_destroy(sipgpu_server_t server);

/* ------------------------------------------------------------------------
 * Worker dispatch: the SIP interpreter's contraction and permute
 * instructions, taken in the interpreter's own terms and queued on a server.
 *
 * sipgpu_worker_push_selector pushes a block selector (the block's cache key,
 * its rank, the SIAL index ids of the selector and the block's extents) as
 * the interpreter's block_selector_stack_ does.  Then:
 *  - sipgpu_worker_contract replaces the block_contract_op case and
 *    Interpreter::handle_contraction (interpreter.cpp:734-741, :1210-1262):
 *    it pops the RIGHT operand first, then the left, builds the ACES pattern
 *    [destination ids | left ids | right ids], converts it with
 *    get_contraction_ptrn_ (SIPGPU_EPATTERN on a non-zero ierr, which the
 *    interpreter turns into a failure at :1243) and queues the contraction
 *    (accumulate = 1: the block_contract_accumulate_op semantics; moduli 0 or
 *    6..14 as in sipgpu_request).  A destination missing from the cache is
 *    created zeroed, as get_block_for_writing creates it;
 *  - sipgpu_worker_permute replaces block_permute_op, permute_rhs_to_lhs and
 *    Block::transpose_copy (interpreter.cpp:656-679, :2049-2084;
 *    block.cpp:216-255): it pops the LEFT-hand side first, builds the 0-based
 *    old-to-new vector (-1 in unused slots), compares it with the compiler's
 *    vector when one is given (SIPGPU_EPERM if they differ, as the
 *    interpreter's check fails), and queues the permute of the right-hand
 *    block, [1, perm + 1] over its extents, into the left-hand block;
 *  - sipgpu_worker_contract_to_scalar replaces block_contract_to_scalar_op
 *    (interpreter.cpp:748-757): pops right then left, waits for the queued
 *    writers of both blocks, and returns the scalar the interpreter would
 *    push onto its expression stack.
 * Host-visible uses: sipgpu_worker_read waits for the queued writer of a
 * block, then reads it; sipgpu_worker_write (a super-instruction's output)
 * waits for every queued request on the block, then stores the host data;
 * sipgpu_worker_flush waits for everything queued (a barrier, a checkpoint,
 * the end of a program) and, with `writeback`, writes dirty blocks back to
 * their host homes.  A queued request's failure is reported by the flush.
 * One thread drives a worker, as one interpreter thread drives the SIP. */
typedef struct sipgpu_worker_s *sipgpu_worker_t;

typedef struct {
    unsigned long long contractions, permutes, scalars, reads, writes;
    unsigned long long waits, flushes, queued;
} sipgpu_worker_stats;

int sipgpu_worker_create(sipgpu_cache_t cache, sipgpu_server_t server, sipgpu_worker_t *worker);
int sipgpu_worker_destroy(sipgpu_worker_t worker);
int sipgpu_worker_push_selector(sipgpu_worker_t worker, unsigned long long key, int rank,
                                const int *index_ids, const int *extents);
int sipgpu_worker_stack_depth(sipgpu_worker_t worker, int *depth);
int sipgpu_worker_contract(sipgpu_worker_t worker, unsigned long long dkey, int drank,
                           const int *dindex_ids, const int *dextents, int accumulate,
                           int moduli, unsigned long long *ticket);
int sipgpu_worker_permute(sipgpu_worker_t worker, const int *compiler_perm,
                          unsigned long long *ticket);
int sipgpu_worker_contract_to_scalar(sipgpu_worker_t worker, double *value);
int sipgpu_worker_read(sipgpu_worker_t worker, unsigned long long key, double *out);
int sipgpu_worker_write(sipgpu_worker_t worker, unsigned long long key, int rank,
                        const int *extents, const double *host);
int sipgpu_worker_flush(sipgpu_worker_t worker, int writeback);
int sipgpu_worker_get_stats(sipgpu_worker_t worker, sipgpu_worker_stats *stats);

/* Interpreter::permute_rhs_to_lhs's vector (interpreter.cpp:2049-2084), host
 * only: perm[j] = i where the right-hand side's index id j (of
 * SIPGPU_MAX_RANK, unused ids -1) is the left-hand side's id i; slots from
 * lrank on are -1.  SIPGPU_EPERM for an id of the left-hand side that the
 * right-hand side's first lrank ids lack ("illegal transpose"). */
int sipgpu_permute_rhs_to_lhs(int lrank, const int *lhs_ids, const int *rhs_ids, int *perm);

/* ------------------------------------------------------------------------
 * The (T) triples energy with W kept on chip.  Replaces the per-tuple
 * sequence of rccsdpt_aaa.sialx (:2164-2174: the W block built from its
 * twelve contractions, energy_denominator_rhf, then the contraction of the
 * divided block with the undivided one into esum) for every block tuple
 * (A, B, C, I, J, K) of the unrestricted sum:
 *   W(a,b,c,i,j,k) = sum over the 6 permutations P of the pairs (a,i),
 *     (b,j), (c,k) of  sum_d V1(x2,x3,d,y3) T(x1,d,y1,y2)
 *                    + sum_l T(x3,x2,y3,l) V2(l,x1,y2,y1)
 *   E += factor * sum W * (W / D),  D = e_i + e_j + e_k - e_a - e_b - e_c.
 * T(a,d,i,j), V1(b,c,d,k) and V2(l,c,j,k) (V2 holding -V) are given as
 * tables of device block pointers in segment order, first index slowest
 * (T: nv*nv*no*no entries; V1: nv*nv*nv*no; V2: no*nv*no*no); e_occ and
 * e_vir are device arrays of all occupied / virtual orbital energies.  W is
 * computed in 5^6 boxes held in LDS and is never written to HBM.
 *
 * sipgpu_triples_create plans and allocates the packed operand panels
 * (sipgpu_triples_info: their bytes, ~4.3 GB for o = 30, v = 240) and one
 * energy partial per box slot; sipgpu_triples_pack (re)packs them from the
 * blocks and zeroes the partials; sipgpu_triples_run computes the partials of
 * box slots [box0, box0 + nbox) (all `slots` slots cover the whole sum), in
 * one kernel launch per SIPGPU_TRIPLES_LAUNCH_BOXES slots (a launch of
 * 512-thread workgroups holds fewer than 2^32 work-items; 2^21 boxes of
 * o = 30, v = 240 run ~1.8 s on one MI355X);
 * sipgpu_triples_energy adds factor * (sum of all partials, in slot order)
 * to *out (a device double).  sipgpu_triples_tuple computes one tuple's
 * energy into *out (+=) and, if w_out is not null, writes its W block
 * (column-major a, b, c, i, j, k).  All calls on one plan must be ordered
 * (one stream). */
#define SIPGPU_TRIPLES_LAUNCH_BOXES (1LL << 21)
typedef struct sipgpu_triples_s *sipgpu_triples_t;
int sipgpu_triples_create(int nv, const int *v_seg, int no, const int *o_seg,
                          const double *const *T, const double *const *V1,
                          const double *const *V2, const double *e_occ, const double *e_vir,
                          sipgpu_triples_t *plan);
int sipgpu_triples_info(sipgpu_triples_t plan, long long *slots, double *flops_per_box,
                        size_t *panel_bytes);
int sipgpu_triples_pack(sipgpu_triples_t plan, sipgpu_stream_t s);
int sipgpu_triples_run(sipgpu_triples_t plan, long long box0, long long nbox, sipgpu_stream_t s);
int sipgpu_triples_energy(sipgpu_triples_t plan, double factor, double *out, sipgpu_stream_t s);
int sipgpu_triples_tuple(sipgpu_triples_t plan, const int *tuple, double factor, double *w_out,
                         double *out, sipgpu_stream_t s);
int sipgpu_triples_destroy(sipgpu_triples_t plan);

/* ------------------------------------------------------------------------
 * Graphs of CC iterations (no reference counterpart: the SIP re-issues every
 * super-instruction each iteration).  Everything the library enqueues on `s`
 * between begin and end -- batch executions and their stages, single
 * contractions and permutes of device blocks, energy denominators and sums,
 * seeded fills -- is captured into one graph and replayed by launch with
 * the same pointers and sizes; the blocks' contents may change between
 * replays.  Batches must be planned and single-contraction shapes run once
 * before capture: planning during capture fails with SIPGPU_ECAPTURE and
 * leaves the capture intact.  The captured work clears memory with kernels
 * only (no memset nodes).  Single contractions and energy sums in a graph use
 * the capture stream's scratch and K-split workspace: destroy the graph before
 * that stream, and do not replay it on another stream while single
 * contractions run on the capture stream. */
typedef struct sipgpu_graph_s *sipgpu_graph_t;
int sipgpu_graph_begin(sipgpu_stream_t s);
int sipgpu_graph_end(sipgpu_stream_t s, sipgpu_graph_t *graph);
int sipgpu_graph_launch(sipgpu_graph_t graph, sipgpu_stream_t s);
/* Number of kernel / copy nodes in the graph. */
int sipgpu_graph_nodes(sipgpu_graph_t graph, int *nodes);
int sipgpu_graph_destroy(sipgpu_graph_t graph);

/* Seeded synthetic data written in HBM, enqueued on `s`: n doubles of the
 * counter-based SplitMix64 sequence (seed, stream).  dist 0: uniform in
 * [a, b) from the top 53 bits of each output; dist 1: normal with standard
 * deviation a (Box-Muller over streams 2*stream+1 and 2*stream+2, cosines
 * into out[0, ceil(n/2)), sines after).  No reference counterpart: it
 * generates benchmark and test inputs in place. */
int sipgpu_fill_random(sipgpu_stream_t s, double *out, long long n, unsigned long long seed,
                       unsigned long long stream, int dist, double a, double b);

/* Measured device ceilings: FP64 MFMA throughput (TFLOP/s) of an MFMA-only
 * loop over all CUs, and HBM bandwidth (GB/s, read + write bytes) of a
 * streaming copy of `copy_bytes` bytes.  Allocates its own buffers. */
int sipgpu_measure_peaks(double *mfma_f64_tflops, double *copy_gbs, size_t copy_bytes);
/* Effective shader clock (MHz) during the last MFMA calibration. */
double sipgpu_measured_clock_mhz(void);
/* Measured LDS read bandwidth: bytes per shader clock per CU of 16-byte
 * reads by two workgroups of 8 or 16 waves per CU (the better of the two):
 * all bytes over the kernel's duration, at the shader clock its loops ran
 * at. */
int sipgpu_measure_lds(double *bytes_per_clock_per_cu);

/* Host-only introspection (no device work).  For one contraction shape:
 * mnk[0..2] = M, N, K of the product; o2n receives lo2n[1..lrank],
 * ro2n[1..rrank], do2n[1..drank] (the reference's determine_index_permutations
 * maps, 1-based, concatenated); the six offset tables are written when their
 * pointers are non-NULL (sizes K, M, K, N, M, N).  Returns 0 or
 * SIPGPU_EPATTERN / SIPGPU_ERANK / SIPGPU_ESIZE. */
int sipgpu_plan_tables(const int *ptrn, int lrank, const int *ldims, int rrank,
                       const int *rdims, int drank, const int *ddims,
                       long long *mnk, int *o2n, int *koffL, int *moffL,
                       int *koffR, int *noffR, int *moffD, int *noffD);
/* Host-only introspection of a permute plan: info[0] canonical rank,
 * info[1] leading tile dims, info[2] row mode, info[3] plain copy. */
int sipgpu_plan_permute(int rank, const int *dims, const int *dim_transp, int *info);
/* Host-only introspection of a slice plan: the canonical dims (extent-1
 * slice dims dropped, contiguous runs merged; at most 8), their slice
 * extents and tensor strides, and info[0..3] = tensor offset of the first
 * slice element, slice elements, tensor span touched, tensor elements. */
int sipgpu_plan_slice(int rank, const int *tens_ext, const int *slice_ext, const int *ext_beg,
                      int *canon_rank, int *ext, long long *tstride, long long *info);

#ifdef __cplusplus
}
#endif
#endif
