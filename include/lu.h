/* Dense real LU with partial pivoting, drop-in for the reference's
 * include/lu.h:12-253. The factorization runs on the GPU (blocked,
 * right-looking, FP64 MFMA trailing update); the factors are also copied
 * into lu->data so host code that reads them keeps working. */
#ifndef PEEC_LU_H
#define PEEC_LU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct
{
    int parallel_threads;      /* must be >= 1 (lu.c:164-168) */
    double pivot_tolerance;    /* relative to max|A|, default 1e-14 */
    size_t parallel_threshold; /* kept for layout; unused on the GPU */
} LUOptions;

typedef struct
{
    size_t n;
    double *data;       /* n*n row-major: unit-L below, U on/above diagonal */
    size_t *pivots;     /* row swapped with k at step k */
    size_t swap_count;
    double matrix_scale;
    int factorized;
} LUFactorization;

/* lu.h:122 */
void lu_options_default(LUOptions *options, int parallel_threads);
/* lu.h:131,137 */
void lu_init(LUFactorization *lu);
void lu_free(LUFactorization *lu);
/* lu.h:160 (lu.c:194-479): 0 ok, -1 bad args/memory, -2 singular */
int lu_factorize(const double *input, size_t n, const LUOptions *options,
                 LUFactorization *lu);
/* lu.h:191 (lu.c:603); x == b allowed */
int lu_solve(const LUFactorization *lu, const double *b, double *x);
/* lu.h:226 (lu.c:672); B and X hold rhs_count vectors of length n */
int lu_solve_multiple(const LUFactorization *lu, const double *B, double *X,
                      size_t rhs_count, int parallel_threads);
/* lu.h:247 (lu.c:732) */
double lu_determinant(const LUFactorization *lu);
/* lu.h:253 */
void lu_print_info(const LUFactorization *lu);

#ifdef __cplusplus
}
#endif

#endif
