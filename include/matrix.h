/* Dense row-major real matrix, drop-in for the reference's
 * include/matrix.h:24-31 (identical layout). */
#ifndef PEEC_MATRIX_H
#define PEEC_MATRIX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct
{
    size_t rows;
    size_t cols;
    double *data; /* row-major */
} Matrix;

/* matrix.h:45-76 storage helpers */
void matrix_init(Matrix *M);
int matrix_alloc(Matrix *M, size_t rows, size_t cols); /* zero-filled */
void matrix_free(Matrix *M);

#ifdef __cplusplus
}
#endif

#endif
