/*
 * jet_oracle.c -- CPU restatement of Jet's pairwise-contraction path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product, never part of
 * it.  Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg only.
 *
 * Restated from (Jet 0.2.3 tree):
 *   include/jet/permute/Default.hpp:20-133   odometer permutation
 *   include/jet/permute/QFlex.hpp:20-539     binary multi-pass transpose
 *   include/jet/TensorHelpers.hpp:128-168    gemm / gemv / dotu dispatch
 *   include/jet/Tensor.hpp:708-752           ContractTensors index algebra
 *   include/jet/Utilities.hpp:316-412        vector set algebra
 */
#include "jet_oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Small integer-vector helpers restating Utilities.hpp:303-412.            */
/* ------------------------------------------------------------------------ */

static int in_vec(int32_t v, const int32_t *a, int n)
{
    for (int i = 0; i < n; i++)
        if (a[i] == v) return 1;
    return 0;
}

/* VectorIntersection (Utilities.hpp:316-327): elements of a also in b, in a's
 * order. */
static int vec_intersection(const int32_t *a, int na, const int32_t *b, int nb,
                            int32_t *out)
{
    int n = 0;
    for (int i = 0; i < na; i++)
        if (in_vec(a[i], b, nb)) out[n++] = a[i];
    return n;
}

/* VectorSubtraction (Utilities.hpp:358-369): elements of a not in b. */
static int vec_subtraction(const int32_t *a, int na, const int32_t *b, int nb,
                           int32_t *out)
{
    int n = 0;
    for (int i = 0; i < na; i++)
        if (!in_vec(a[i], b, nb)) out[n++] = a[i];
    return n;
}

static int vec_equal(const int32_t *a, const int32_t *b, int n)
{
    for (int i = 0; i < n; i++)
        if (a[i] != b[i]) return 0;
    return 1;
}

static int is_pow2(uint64_t v) { return v && !(v & (v - 1)); }

static int ilog2(uint64_t v)
{
    int r = 0;
    while (v > 1) {
        v >>= 1;
        r++;
    }
    return r;
}

/* ------------------------------------------------------------------------ */
/* Default.hpp:20-133 -- odometer permutation, any shape.                   */
/* ------------------------------------------------------------------------ */

int jo_permute_default(size_t elem, int rank, const int64_t *shape,
                       const int32_t *new_order, const void *in, void *out)
{
    if (rank <= 0) return JO_ERR_ARG; /* Default.hpp:38-39 */
    int64_t total = 1;
    for (int i = 0; i < rank; i++) total *= shape[i];

    /* map_old_to_new_idxpos (Default.hpp:43-53) */
    int32_t old_to_new[64];
    int64_t new_dims[64];
    for (int j = 0; j < rank; j++) {
        old_to_new[new_order[j]] = j;
        new_dims[j] = shape[new_order[j]];
    }
    /* super dimensions = row-major strides (Default.hpp:55-63) */
    int64_t old_stride[64], new_stride[64];
    old_stride[rank - 1] = 1;
    new_stride[rank - 1] = 1;
    for (int i = rank - 1; i > 0; i--) {
        old_stride[i - 1] = old_stride[i] * shape[i];
        new_stride[i - 1] = new_stride[i] * new_dims[i];
    }
    int64_t stride_in_new[64];
    for (int i = 0; i < rank; i++) stride_in_new[i] = new_stride[old_to_new[i]];

    /* Odometer over the old multi-index; element at old position po goes to
     * new position pn (Default.hpp:95-128). */
    int64_t counter[64];
    memset(counter, 0, sizeof(counter));
    const char *src = (const char *)in;
    char *dst = (char *)out;
    int64_t pn = 0;
    for (int64_t po = 0; po < total; po++) {
        memcpy(dst + (size_t)pn * elem, src + (size_t)po * elem, elem);
        for (int j = rank - 1; j >= 0; j--) {
            pn += stride_in_new[j];
            if (++counter[j] < shape[j]) break;
            pn -= stride_in_new[j] * counter[j];
            counter[j] = 0;
        }
    }
    return JO_OK;
}

/* ------------------------------------------------------------------------ */
/* QFlex.hpp -- binary-index multi-pass transpose.                          */
/* ------------------------------------------------------------------------ */

enum { PASS_RIGHT = 0, PASS_LEFT = 1 };

typedef struct {
    int kind;
    int nbits;          /* bits covered by the map */
    int64_t dim_left;
    int64_t dim_right;
    int64_t *map;       /* 2^nbits entries */
} qpass_t;

typedef struct {
    int npasses;
    qpass_t pass[3];
} qplan_t;

static void qplan_free(qplan_t *p)
{
    for (int i = 0; i < p->npasses; i++) free(p->pass[i].map);
    p->npasses = 0;
}

/* GenerateBinaryReorderingMap (QFlex.hpp:104-156) with dim = 2: the bit of the
 * old position for label i (i = 0 is the most significant) moves to the bit
 * position of its new slot. */
static int64_t *binary_reordering_map(const int32_t *old_ord,
                                      const int32_t *new_ord, int n)
{
    int32_t old_to_new[64];
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) {
            if (old_ord[i] == new_ord[j]) {
                old_to_new[i] = j;
                break;
            }
        }
    }
    int64_t size = (int64_t)1 << n;
    int64_t *map = (int64_t *)malloc(sizeof(int64_t) * (size_t)size);
    if (!map) return NULL;
    for (int64_t po = 0; po < size; po++) {
        int64_t pn = 0;
        for (int i = 0; i < n; i++) {
            int64_t bit = (po >> (n - 1 - i)) & 1;
            pn |= bit << (n - 1 - old_to_new[i]);
        }
        map[po] = pn;
    }
    return map;
}

/* PrecomputeRightTransposeData (QFlex.hpp:221-271) and
 * PrecomputeLeftTransposeData (QFlex.hpp:273-321). */
static int qplan_push(qplan_t *p, int kind, const int32_t *old_ord,
                      const int32_t *new_ord, int n, int64_t tensor_size)
{
    if (vec_equal(old_ord, new_ord, n)) return JO_OK; /* nothing to reorder */
    qpass_t *q = &p->pass[p->npasses];
    q->kind = kind;
    q->nbits = n;
    q->map = binary_reordering_map(old_ord, new_ord, n);
    if (!q->map) return JO_ERR_ALLOC;
    if (kind == PASS_RIGHT) {
        q->dim_right = (int64_t)1 << n;
        q->dim_left = tensor_size >> n;
    }
    else {
        q->dim_left = (int64_t)1 << n;
        q->dim_right = tensor_size >> n;
    }
    p->npasses++;
    return JO_OK;
}

/* PrecomputeFastTransposeData (QFlex.hpp:323-538). */
static int qflex_plan(int rank, const int64_t *shape, const int32_t *new_order,
                      size_t blocksize, size_t min_dims, qplan_t *plan)
{
    plan->npasses = 0;
    for (int i = 0; i < rank; i++)
        if (!is_pow2((uint64_t)shape[i])) return JO_ERR_NOT_POW2;

    int identity = 1;
    for (int i = 0; i < rank; i++)
        if (new_order[i] != i) identity = 0;
    if (identity) return JO_OK; /* no_transpose (QFlex.hpp:358-361) */

    /* Binary groups: old index i owns bit labels [first[i], first[i]+log2)
     * in most-significant-first order (QFlex.hpp:363-395). */
    int32_t first[64], logs[64];
    int nbits = 0;
    for (int i = 0; i < rank; i++) {
        first[i] = nbits;
        logs[i] = ilog2((uint64_t)shape[i]);
        nbits += logs[i];
    }
    int32_t old_bin[64], new_bin[64];
    int pos = 0;
    for (int i = 0; i < rank; i++)
        for (int j = 0; j < logs[i]; j++) old_bin[pos++] = first[i] + j;
    pos = 0;
    for (int i = 0; i < rank; i++) {
        int o = new_order[i];
        for (int j = 0; j < logs[o]; j++) new_bin[pos++] = first[o] + j;
    }
    int64_t tensor_size = (int64_t)1 << nbits;
    const int Lr = ilog2(blocksize);
    const int Rr = ilog2(min_dims);

    if (nbits <= Lr) /* QFlex.hpp:399-404 */
        return qplan_push(plan, PASS_RIGHT, old_bin, new_bin, nbits,
                          tensor_size);

    const int Ll = nbits - Lr;
    if (vec_equal(old_bin, new_bin, Ll)) /* QFlex.hpp:419-430 */
        return qplan_push(plan, PASS_RIGHT, old_bin + Ll, new_bin + Ll, Lr,
                          tensor_size);

    if (Rr == 0) return JO_ERR_MIN_DIMS; /* QFlex.hpp:432-434 */

    /* Matched suffix (QFlex.hpp:440-458). */
    int num_elems = 0;
    while (num_elems < nbits &&
           old_bin[nbits - 1 - num_elems] == new_bin[nbits - 1 - num_elems])
        num_elems++;
    if (num_elems > 0 && num_elems >= Rr)
        return qplan_push(plan, PASS_LEFT, old_bin, new_bin, nbits - num_elems,
                          tensor_size);

    /* Worst case: Left, Right, Left (QFlex.hpp:461-537). */
    if (nbits < Lr || nbits < Rr) return JO_ERR_TOO_SMALL;
    const int Rl = nbits - Rr;
    const int32_t *Rr_indices = new_bin + Rl; /* Rr entries */
    const int32_t *Rl_indices = old_bin;      /* Rl entries */

    int32_t rr_in_rl[64], rl_not_rr[64], rl_first[64];
    int n_in = vec_intersection(Rl_indices, Rl, Rr_indices, Rr, rr_in_rl);
    int n_not = vec_subtraction(Rl_indices, Rl, rr_in_rl, n_in, rl_not_rr);
    memcpy(rl_first, rl_not_rr, sizeof(int32_t) * (size_t)n_not);
    memcpy(rl_first + n_not, rr_in_rl, sizeof(int32_t) * (size_t)n_in);
    int rc = qplan_push(plan, PASS_LEFT, Rl_indices, rl_first, Rl, tensor_size);
    if (rc) return rc;

    int32_t lr_first[64], lr_second[64], tmp[64];
    int n_lr1 = 0;
    for (int i = Ll; i < Rl; i++) lr_first[n_lr1++] = rl_first[i];
    for (int i = Rl; i < nbits; i++) lr_first[n_lr1++] = old_bin[i];
    int n_sub = vec_subtraction(lr_first, n_lr1, Rr_indices, Rr, tmp);
    memcpy(lr_second, tmp, sizeof(int32_t) * (size_t)n_sub);
    memcpy(lr_second + n_sub, Rr_indices, sizeof(int32_t) * (size_t)Rr);
    if (n_sub + Rr != n_lr1) {
        qplan_free(plan);
        return JO_ERR_BLOCKSIZE; /* QFlex.hpp:517-519 */
    }
    rc = qplan_push(plan, PASS_RIGHT, lr_first, lr_second, n_lr1, tensor_size);
    if (rc) return rc;

    int32_t rl_second[64];
    int n_rl2 = 0;
    for (int i = 0; i < Ll; i++) rl_second[n_rl2++] = rl_first[i];
    for (int i = 0; i < Lr - Rr; i++) rl_second[n_rl2++] = lr_second[i];
    return qplan_push(plan, PASS_LEFT, rl_second, new_bin, Rl, tensor_size);
}

int jo_qflex_plan_kinds(int rank, const int64_t *shape,
                        const int32_t *new_order, size_t blocksize,
                        size_t min_dims, char *kinds_out, int *npasses)
{
    qplan_t plan;
    int rc = qflex_plan(rank, shape, new_order, blocksize, min_dims, &plan);
    if (rc) return rc;
    *npasses = plan.npasses;
    for (int i = 0; i < plan.npasses; i++)
        kinds_out[i] = plan.pass[i].kind == PASS_RIGHT ? 'R' : 'L';
    qplan_free(&plan);
    return JO_OK;
}

/* FastTranspose (QFlex.hpp:158-219). */
static void qflex_execute(size_t elem, const qplan_t *plan, char *data,
                          char *scratch, int64_t tensor_size, int nthreads)
{
    (void)nthreads;
    for (int p = 0; p < plan->npasses; p++) {
        const qpass_t *q = &plan->pass[p];
        const int64_t dl = q->dim_left, dr = q->dim_right;
        const int64_t *map = q->map;
        if (q->kind == PASS_RIGHT) {
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
#endif
            {
                char *tmp = (char *)malloc(elem * (size_t)dr);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
                for (int64_t pl = 0; pl < dl; pl++) {
                    char *chunk = data + (size_t)(pl * dr) * elem;
                    memcpy(tmp, chunk, elem * (size_t)dr);
                    for (int64_t pr = 0; pr < dr; pr++)
                        memcpy(chunk + (size_t)map[pr] * elem,
                               tmp + (size_t)pr * elem, elem);
                }
                free(tmp);
            }
        }
        else {
            memcpy(scratch, data, elem * (size_t)tensor_size);
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
#endif
            for (int64_t pl = 0; pl < dl; pl++)
                memcpy(data + (size_t)(map[pl] * dr) * elem,
                       scratch + (size_t)(pl * dr) * elem,
                       elem * (size_t)dr);
        }
    }
}

int jo_permute_qflex(size_t elem, int rank, const int64_t *shape,
                     const int32_t *new_order, const void *in, void *out,
                     size_t blocksize, size_t min_dims, int nthreads)
{
    qplan_t plan;
    int rc = qflex_plan(rank, shape, new_order, blocksize, min_dims, &plan);
    if (rc) return rc;
    int64_t total = 1;
    for (int i = 0; i < rank; i++) total *= shape[i];
    memcpy(out, in, elem * (size_t)total); /* data_out = data_ (QFlex.hpp:45) */
    char *scratch = NULL;
    for (int p = 0; p < plan.npasses; p++)
        if (plan.pass[p].kind == PASS_LEFT && !scratch) {
            scratch = (char *)malloc(elem * (size_t)total);
            if (!scratch) {
                qplan_free(&plan);
                return JO_ERR_ALLOC;
            }
        }
    qflex_execute(elem, &plan, (char *)out, scratch, total, nthreads);
    free(scratch);
    qplan_free(&plan);
    return JO_OK;
}

/* ------------------------------------------------------------------------ */
/* GEMM with FP64 accumulation (TensorHelpers.hpp:47-61, 128-168).          */
/* ------------------------------------------------------------------------ */

static inline void load_c(size_t elem, const void *p, int64_t i, double *re,
                          double *im)
{
    if (elem == 8) {
        const float *f = (const float *)p;
        *re = f[2 * i];
        *im = f[2 * i + 1];
    }
    else {
        const double *d = (const double *)p;
        *re = d[2 * i];
        *im = d[2 * i + 1];
    }
}

void jo_gemm(size_t in_elem, int64_t m, int64_t n, int64_t k, const void *A,
             const void *B, double *C, int nthreads)
{
    (void)nthreads;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (int64_t i = 0; i < m; i++) {
        double *crow = C + 2 * i * n;
        for (int64_t j = 0; j < n; j++) {
            crow[2 * j] = 0.0;
            crow[2 * j + 1] = 0.0;
        }
        for (int64_t kk = 0; kk < k; kk++) {
            double ar, ai;
            load_c(in_elem, A, i * k + kk, &ar, &ai);
            for (int64_t j = 0; j < n; j++) {
                double br, bi;
                load_c(in_elem, B, kk * n + j, &br, &bi);
                crow[2 * j] += ar * br - ai * bi;
                crow[2 * j + 1] += ar * bi + ai * br;
            }
        }
    }
}

void jo_gemm_entries(size_t in_elem, int64_t k, const void *A, int64_t lda,
                     const void *B, int64_t ldb, int64_t count,
                     const int64_t *rows, const int64_t *cols, double *out,
                     int nthreads)
{
    (void)nthreads;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (int64_t e = 0; e < count; e++) {
        double sr = 0.0, si = 0.0;
        for (int64_t kk = 0; kk < k; kk++) {
            double ar, ai, br, bi;
            load_c(in_elem, A, rows[e] * lda + kk, &ar, &ai);
            load_c(in_elem, B, kk * ldb + cols[e], &br, &bi);
            sr += ar * br - ai * bi;
            si += ar * bi + ai * br;
        }
        out[2 * e] = sr;
        out[2 * e + 1] = si;
    }
}

/* ------------------------------------------------------------------------ */
/* Tensor::ContractTensors (Tensor.hpp:708-752).                            */
/* ------------------------------------------------------------------------ */

static int64_t dim_of(int32_t label, const int32_t *l, const int64_t *s, int r)
{
    for (int i = 0; i < r; i++)
        if (l[i] == label) return s[i];
    return -1;
}

/* Transpose to a new label order (Tensor.hpp:578-612): a copy when the order
 * already matches, else a permutation.  The permutation result does not depend
 * on the backend (QFlex or Default), so the odometer is used here. */
static int transpose_to(size_t elem, int r, const int32_t *l,
                        const int64_t *s, const void *in,
                        const int32_t *new_l, void *out)
{
    int64_t total = 1;
    for (int i = 0; i < r; i++) total *= s[i];
    if (r == 0 || vec_equal(l, new_l, r)) {
        memcpy(out, in, elem * (size_t)total);
        return JO_OK;
    }
    int32_t order[64];
    for (int i = 0; i < r; i++)
        for (int j = 0; j < r; j++)
            if (l[j] == new_l[i]) order[i] = j;
    return jo_permute_default(elem, r, s, order, in, out);
}

int jo_contract(size_t elem, int ra, const int32_t *la, const int64_t *sa,
                const void *A, int rb, const int32_t *lb, const int64_t *sb,
                const void *B, int *rc, int32_t *lc, int64_t *sc, double *C,
                int nthreads)
{
    if (ra > 64 || rb > 64) return JO_ERR_ARG;
    int32_t left[64], right[64], common[64];
    int nl = vec_subtraction(la, ra, lb, rb, left);
    int nr = vec_subtraction(lb, rb, la, ra, right);
    int nc = vec_intersection(la, ra, lb, rb, common);

    int64_t m = 1, n = 1, k = 1;
    for (int i = 0; i < nl; i++) m *= dim_of(left[i], la, sa, ra);
    for (int i = 0; i < nr; i++) n *= dim_of(right[i], lb, sb, rb);
    for (int i = 0; i < nc; i++) k *= dim_of(common[i], la, sa, ra);

    /* a_new_ordering = left ++ common, b_new_ordering = common ++ right. */
    int32_t a_new[64], b_new[64];
    memcpy(a_new, left, sizeof(int32_t) * (size_t)nl);
    memcpy(a_new + nl, common, sizeof(int32_t) * (size_t)nc);
    memcpy(b_new, common, sizeof(int32_t) * (size_t)nc);
    memcpy(b_new + nc, right, sizeof(int32_t) * (size_t)nr);

    *rc = nl + nr;
    for (int i = 0; i < nl; i++) {
        lc[i] = left[i];
        sc[i] = dim_of(left[i], la, sa, ra);
    }
    for (int i = 0; i < nr; i++) {
        lc[nl + i] = right[i];
        sc[nl + i] = dim_of(right[i], lb, sb, rb);
    }

    void *At = malloc(elem * (size_t)(m * k));
    void *Bt = malloc(elem * (size_t)(k * n));
    if (!At || !Bt) {
        free(At);
        free(Bt);
        return JO_ERR_ALLOC;
    }
    int err = transpose_to(elem, ra, la, sa, A, a_new, At);
    if (!err) err = transpose_to(elem, rb, lb, sb, B, b_new, Bt);
    /* All four MultiplyTensorData cases (gemm, gemv NoTrans, gemv Trans,
     * dotu) are this product with m or n equal to 1. */
    if (!err) jo_gemm(elem, m, n, k, At, Bt, C, nthreads);
    free(At);
    free(Bt);
    return err;
}

/* ------------------------------------------------------------------------ */
/* Compensated (Neumaier) summation.                                        */
/* ------------------------------------------------------------------------ */

void jo_sum_compensated(int64_t n, const double *x, double *out)
{
    for (int part = 0; part < 2; part++) {
        double s = 0.0, c = 0.0;
        for (int64_t i = 0; i < n; i++) {
            double v = x[2 * i + part];
            double t = s + v;
            if ((s >= 0 ? s : -s) >= (v >= 0 ? v : -v))
                c += (s - t) + v;
            else
                c += (v - t) + s;
            s = t;
        }
        out[part] = s + c;
    }
}
