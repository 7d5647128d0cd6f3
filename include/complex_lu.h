/* Dense complex LU with row/column max equilibration and partial pivoting,
 * drop-in for the reference's include/complex_lu.h:14-161. The
 * factorization runs on the GPU; factors, pivots and scales are returned in
 * the reference's host layout and a device copy is kept for the solves. */
#ifndef PEEC_COMPLEX_LU_H
#define PEEC_COMPLEX_LU_H

#include "complex_matrix.h"

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct
{
    int parallel_threads;
    double pivot_tolerance;    /* relative to max|A_eq|, default 1e-14 */
    size_t parallel_threshold;
} ComplexLUOptions;

typedef struct
{
    size_t n;
    Complex *data;        /* equilibrated L\U, row-major */
    size_t *pivots;
    size_t swap_count;
    double matrix_scale;  /* max|A_eq| */
    double *row_scale;    /* r_i = max_j |A_ij| */
    double *column_scale; /* c_j = max_i |A_ij / r_i| */
    int factorized;
} ComplexLUFactorization;

/* complex_lu.h:77 (defaults complex_lu.c:20-31) */
void complex_lu_options_default(ComplexLUOptions *options, int parallel_threads);
/* complex_lu.h:86,94 */
void complex_lu_init(ComplexLUFactorization *lu);
void complex_lu_free(ComplexLUFactorization *lu);
/* complex_lu.h:116 (complex_lu.c:208-550): 0 ok, -1 args/memory, -2 singular */
int complex_lu_factorize(const Complex *input, size_t n,
                         const ComplexLUOptions *options,
                         ComplexLUFactorization *lu);
/* complex_lu.h:131 (complex_lu.c:724-778); x == b allowed */
int complex_lu_solve(const ComplexLUFactorization *lu, const Complex *b,
                     Complex *x);
/* complex_lu.h:147 (complex_lu.c:786-840) */
int complex_lu_solve_multiple(const ComplexLUFactorization *lu,
                              const Complex *B, Complex *X, size_t rhs_count,
                              int parallel_threads);
/* complex_lu.h:159 */
void complex_lu_print_info(const ComplexLUFactorization *lu);

#ifdef __cplusplus
}
#endif

#endif
