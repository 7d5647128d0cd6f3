/* Complex scalar and dense row-major complex matrix, drop-in for the
 * reference's include/complex_matrix.h:21-40 (identical layouts). Only the
 * storage helpers the hot path needs are exported here. */
#ifndef PEEC_COMPLEX_MATRIX_H
#define PEEC_COMPLEX_MATRIX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct
{
    double re;
    double im;
} Complex;

typedef struct
{
    size_t rows;
    size_t cols;
    Complex *data; /* row-major */
} ComplexMatrix;

/* complex_matrix.h:115-127 storage helpers */
void complex_matrix_init(ComplexMatrix *M);
int complex_matrix_alloc(ComplexMatrix *M, size_t rows, size_t cols); /* zero-filled */
void complex_matrix_free(ComplexMatrix *M);
/* y = A x on the GPU; complex_matrix.h:201, complex_matrix.c:542 */
int complex_matrix_vector_multiply(const ComplexMatrix *A, const Complex *x,
                                   Complex *y, int parallel_threads);

#ifdef __cplusplus
}
#endif

#endif
