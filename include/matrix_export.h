/*
 * Matrix Market interchange of the PEEC matrices (drop-in for the
 * reference's include/matrix_export.h, src/io/matrix_export.c:105-433; the
 * files are byte-identical to the reference's).
 *
 * L, P, PAT: dense "array real general" (column-major, %.17e); A and R:
 * "coordinate real general" (1-based). matrix_import_dense_mtx reads a dense
 * file back (no counterpart in the reference: cache and fixture exchange).
 */
#ifndef PEEC_MATRIX_EXPORT_H
#define PEEC_MATRIX_EXPORT_H

#include <stddef.h>

#include "inductance.h"
#include "matrix.h"
#include "mesh.h"
#include "potential.h"
#include "resistance.h"

#ifdef __cplusplus
extern "C" {
#endif

/* data: rows x cols row-major */
int matrix_export_dense_mtx(const char *filename, const double *data, size_t rows, size_t cols);

/* A[e, a] = -1, A[e, b] = +1 for edge e = a -> b */
int matrix_export_incidence_mtx(const char *filename, const Mesh *mesh);

/* every diagonal entry, zeros included */
int matrix_export_resistance_mtx(const char *filename, const ResistanceMatrix *R);

/* directory/{L,A,P,PAT,R}.mtx (the directory is created) */
int matrix_export_peec_set(const char *directory, const Mesh *mesh, const InductanceMatrix *L,
                           const PotentialMatrix *P, const Matrix *PAT, const ResistanceMatrix *R);

/* *data: malloc'd rows x cols row-major (free with free()); 0 or -1 */
int matrix_import_dense_mtx(const char *filename, double **data, size_t *rows, size_t *cols);

#ifdef __cplusplus
}
#endif

#endif
