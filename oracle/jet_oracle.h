/*
 * jet_oracle.h -- CPU restatement of Jet's pairwise-contraction path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product (jet_amd, libjet_hip)
 * never links, imports or calls anything here.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the Jet 0.2.3 source tree).
 */
#ifndef JET_ORACLE_H
#define JET_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    JO_OK = 0,
    JO_ERR_NOT_POW2 = 1,   /* QFlex.hpp:333-336 */
    JO_ERR_BLOCKSIZE = 2,  /* QFlex.hpp:517-519 */
    JO_ERR_TOO_SMALL = 3,  /* QFlex.hpp:407-408, 475-478 */
    JO_ERR_MIN_DIMS = 4,   /* QFlex.hpp:432-434 */
    JO_ERR_ARG = 5,
    JO_ERR_ALLOC = 6,
};

/* Odometer permutation of a row-major tensor (Default.hpp:20-133).
 * Output axis i is input axis new_order[i].  elem = bytes per element. */
int jo_permute_default(size_t elem, int rank, const int64_t *shape,
                       const int32_t *new_order, const void *in, void *out);

/* QFlex binary-index multi-pass transpose (QFlex.hpp:20-539), restated with
 * integer bit labels.  Same plan (Right / Left / Left-Right-Left), same error
 * conditions.  nthreads <= 1 runs serially; more runs the passes with OpenMP
 * (QFlex.hpp:173-215). */
int jo_permute_qflex(size_t elem, int rank, const int64_t *shape,
                     const int32_t *new_order, const void *in, void *out,
                     size_t blocksize, size_t min_dims, int nthreads);

/* Number of passes (0..3) and their kinds ('R'/'L') that QFlex plans for this
 * permutation; returns an error code like jo_permute_qflex. */
int jo_qflex_plan_kinds(int rank, const int64_t *shape,
                        const int32_t *new_order, size_t blocksize,
                        size_t min_dims, char *kinds_out, int *npasses);

/* Row-major complex GEMM C[m,n] = sum_k A[m,k] B[k,n] with FP64 accumulation
 * (TensorHelpers.hpp:47-61 semantics: NoTrans/NoTrans, alpha=1, beta=0).
 * in_elem = 8 (complex64 operands) or 16 (complex128); output complex128. */
void jo_gemm(size_t in_elem, int64_t m, int64_t n, int64_t k, const void *A,
             const void *B, double *C, int nthreads);

/* Selected output entries of the same GEMM: rows[i], cols[i] -> out[2i..2i+1],
 * each an FP64 dot product in increasing k (used at sizes too big for jo_gemm).
 * lda/ldb are the row strides of A (m x k) and B (k x n) in elements. */
void jo_gemm_entries(size_t in_elem, int64_t k, const void *A, int64_t lda,
                     const void *B, int64_t ldb, int64_t count,
                     const int64_t *rows, const int64_t *cols, double *out,
                     int nthreads);

/* Tensor::ContractTensors (Tensor.hpp:708-752) over integer labels.
 * Output labels are left ++ right (rc, lc, sc filled; lc/sc need room for
 * ra+rb entries).  Inputs are complex64 (elem=8) or complex128 (elem=16);
 * C is complex128 (FP64 accumulation: the FP64 oracle). */
int jo_contract(size_t elem, int ra, const int32_t *la, const int64_t *sa,
                const void *A, int rb, const int32_t *lb, const int64_t *sb,
                const void *B, int *rc, int32_t *lc, int64_t *sc, double *C,
                int nthreads);

/* Neumaier-compensated sum of n complex128 values (re, im interleaved). */
void jo_sum_compensated(int64_t n, const double *x, double *out);

#ifdef __cplusplus
}
#endif

#endif
