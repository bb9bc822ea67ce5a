/*
 * cpu_baseline.c -- the reference CPU contraction path, restated, for timing.
 *
 * TEST INFRASTRUCTURE ONLY: bench.py's cpu_baseline leg times this on the
 * host cores next to the GPU.  It follows the reference's CPU path step by
 * step:
 *   Tensor::ContractTensors index algebra      Tensor.hpp:708-752
 *   Tensor::Transpose: QFlex for power-of-2     Tensor.hpp:578-612,
 *     data, Default otherwise                   permute/QFlex.hpp, Default.hpp
 *   MultiplyTensorData -> cblas_?gemm / gemv /  TensorHelpers.hpp:128-168
 *     dotu_sub, row-major NoTrans/NoTrans
 * with CBLAS taken from an OpenBLAS shared library loaded at run time (the
 * image ships no cblas.h; scipy's bundled OpenBLAS exports the CBLAS
 * functions under a "scipy_" prefix).
 */
#include "jet_oracle.h"

#include <dlfcn.h>
#include <stdlib.h>
#include <string.h>

enum { RowMajor = 101, NoTrans = 111, Trans = 112 };

typedef void (*gemm_fn)(int, int, int, int, int, int, const void *, const void *,
                        int, const void *, int, const void *, void *, int);
typedef void (*gemv_fn)(int, int, int, int, const void *, const void *, int,
                        const void *, int, const void *, void *, int);
typedef void (*dotu_fn)(int, const void *, int, const void *, int, void *);
typedef void (*threads_fn)(int);

static gemm_fn p_gemm[2];
static gemv_fn p_gemv[2];
static dotu_fn p_dotu[2];
static threads_fn p_threads;

static void *sym(void *h, const char *prefix, const char *name)
{
    char buf[128];
    strncpy(buf, prefix, sizeof(buf) - 1);
    buf[sizeof(buf) - 1] = 0;
    strncat(buf, name, sizeof(buf) - strlen(buf) - 1);
    return dlsym(h, buf);
}

/* Returns 0 when every function was found. */
int jb_load_blas(const char *path, const char *prefix)
{
    void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return 1;
    p_gemm[0] = (gemm_fn)sym(h, prefix, "cblas_cgemm");
    p_gemm[1] = (gemm_fn)sym(h, prefix, "cblas_zgemm");
    p_gemv[0] = (gemv_fn)sym(h, prefix, "cblas_cgemv");
    p_gemv[1] = (gemv_fn)sym(h, prefix, "cblas_zgemv");
    p_dotu[0] = (dotu_fn)sym(h, prefix, "cblas_cdotu_sub");
    p_dotu[1] = (dotu_fn)sym(h, prefix, "cblas_zdotu_sub");
    p_threads = (threads_fn)sym(h, prefix, "openblas_set_num_threads");
    for (int i = 0; i < 2; i++)
        if (!p_gemm[i] || !p_gemv[i] || !p_dotu[i]) return 2;
    return 0;
}

void jb_set_blas_threads(int n)
{
    if (p_threads) p_threads(n);
}

static int in_vec(int32_t v, const int32_t *a, int n)
{
    for (int i = 0; i < n; i++)
        if (a[i] == v) return 1;
    return 0;
}

static int64_t dim_of(int32_t l, const int32_t *ls, const int64_t *s, int r)
{
    for (int i = 0; i < r; i++)
        if (ls[i] == l) return s[i];
    return 1;
}

/* Tensor::Transpose: a copy when unchanged, QFlex when every extent is a
 * power of two (falling back to Default if QFlex refuses), Default else. */
static int transpose(size_t elem, int r, const int32_t *l, const int64_t *s,
                     const void *in, const int32_t *nl, void *out, int nthreads)
{
    int64_t total = 1;
    int same = 1;
    for (int i = 0; i < r; i++) {
        total *= s[i];
        if (l[i] != nl[i]) same = 0;
    }
    if (r == 0 || same) {
        memcpy(out, in, elem * (size_t)total);
        return 0;
    }
    int32_t order[64];
    for (int i = 0; i < r; i++)
        for (int j = 0; j < r; j++)
            if (l[j] == nl[i]) order[i] = j;
    if (jo_permute_qflex(elem, r, s, order, in, out, 1024, 32, nthreads) == JO_OK)
        return 0;
    return jo_permute_default(elem, r, s, order, in, out);
}

int jb_contract(size_t elem, int ra, const int32_t *la, const int64_t *sa,
                const void *A, int rb, const int32_t *lb, const int64_t *sb,
                const void *B, void *C, int nthreads)
{
    if (ra > 64 || rb > 64 || !p_gemm[0]) return JO_ERR_ARG;
    const int z = elem == 16;
    int32_t left[64], right[64], common[64];
    int nl = 0, nr = 0, nc = 0;
    for (int i = 0; i < ra; i++) {
        if (in_vec(la[i], lb, rb))
            common[nc++] = la[i];
        else
            left[nl++] = la[i];
    }
    for (int i = 0; i < rb; i++)
        if (!in_vec(lb[i], la, ra)) right[nr++] = lb[i];
    int64_t m = 1, n = 1, k = 1;
    for (int i = 0; i < nl; i++) m *= dim_of(left[i], la, sa, ra);
    for (int i = 0; i < nr; i++) n *= dim_of(right[i], lb, sb, rb);
    for (int i = 0; i < nc; i++) k *= dim_of(common[i], la, sa, ra);
    int32_t a_new[64], b_new[64];
    memcpy(a_new, left, sizeof(int32_t) * (size_t)nl);
    memcpy(a_new + nl, common, sizeof(int32_t) * (size_t)nc);
    memcpy(b_new, common, sizeof(int32_t) * (size_t)nc);
    memcpy(b_new + nc, right, sizeof(int32_t) * (size_t)nr);
    void *At = malloc(elem * (size_t)(m * k));
    void *Bt = malloc(elem * (size_t)(k * n));
    if (!At || !Bt) {
        free(At);
        free(Bt);
        return JO_ERR_ALLOC;
    }
    int err = transpose(elem, ra, la, sa, A, a_new, At, nthreads);
    if (!err) err = transpose(elem, rb, lb, sb, B, b_new, Bt, nthreads);
    if (!err) {
        const double one[2] = {1.0, 0.0}, zero[2] = {0.0, 0.0};
        const float onef[2] = {1.0f, 0.0f}, zerof[2] = {0.0f, 0.0f};
        const void *alpha = z ? (const void *)one : (const void *)onef;
        const void *beta = z ? (const void *)zero : (const void *)zerof;
        const int lk = (int)(k > 1 ? k : 1), ln = (int)(n > 1 ? n : 1);
        if (nl > 0 && nr > 0)
            p_gemm[z](RowMajor, NoTrans, NoTrans, (int)m, (int)n, (int)k, alpha, At, lk, Bt, ln,
                      beta, C, ln);
        else if (nl > 0)
            p_gemv[z](RowMajor, NoTrans, (int)m, (int)k, alpha, At, lk, Bt, 1, beta, C, 1);
        else if (nr > 0)
            p_gemv[z](RowMajor, Trans, (int)k, (int)n, alpha, Bt, ln, At, 1, beta, C, 1);
        else
            p_dotu[z]((int)k, At, 1, Bt, 1, C);
    }
    free(At);
    free(Bt);
    return err;
}
