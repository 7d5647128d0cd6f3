/* Edge-node incidence and P*A^T, drop-in for the reference's
 * include/incidence.h:48-245 (the parts on the harmonic path). */
#ifndef PEEC_INCIDENCE_H
#define PEEC_INCIDENCE_H

#include "matrix.h"
#include "mesh.h"
#include "potential.h"

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct
{
    size_t rows; /* Ne */
    size_t cols; /* Nv */
    double *data; /* dense row-major, A[e,a] = -1, A[e,b] = +1 */
} IncidenceMatrix;

/* incidence.h:63-95 (incidence.c:34-115) */
void incidence_init(IncidenceMatrix *A);
void incidence_free(IncidenceMatrix *A);
int incidence_build(const Mesh *mesh, IncidenceMatrix *A);
/* incidence.h:240-245 (incidence.c:417-590): PAT[j,e] = P[j,b_e] - P[j,a_e],
 * Nv x Ne, computed on the GPU */
int incidence_build_PAT(const Mesh *mesh, const PotentialMatrix *P,
                        Matrix *PAT, int parallel_threads);

#ifdef __cplusplus
}
#endif

#endif
